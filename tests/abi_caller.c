/* abi_caller.c — a plain C host of the library, the way a Rust or C program
 * would drive the solver through include/coffee_sim.h (no Python, no torch).
 *
 *   abi_caller            host-only entry points (presets, seeded workload,
 *                         bed lattice, SDF generation); no device needed
 *   abi_caller run        also creates a context and a simulation of the
 *                         small V60 workload at half size, steps 3 frames,
 *                         checkpoints, restores and steps again
 *
 * Prints one "abi caller ok ..." line and returns 0, or names the failing
 * call and returns 1.  Built and run by tests/test_abi.py. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "coffee_sim.h"

#define CHECK(call)                                                          \
  do {                                                                       \
    int rc_ = (call);                                                        \
    if (rc_ != CS_OK) {                                                      \
      fprintf(stderr, "%s failed: %d (%s)\n", #call, rc_, cs_last_error()); \
      return 1;                                                              \
    }                                                                        \
  } while (0)

static int host_part(void) {
  cs_settings s;
  for (int p = 0; p < 3; ++p) CHECK(cs_settings_preset(p, &s));
  if (cs_settings_preset(9, &s) != CS_ERR_INVALID) return fprintf(stderr, "bad preset accepted\n"), 1;
  cs_workload w;
  CHECK(cs_workload_build(CS_WORKLOAD_V60_SMALL, 0, 2, &w));
  if (w.scene.num_particles == 0 || w.settings.grid_dims[0] != 32) return fprintf(stderr, "workload\n"), 1;
  CHECK(cs_workload_free(&w));
  cs_filter_config f;
  cs_bed_config b;
  cs_filter_config_default(&f);
  cs_bed_config_seated_in_filter(&f, &b);
  CHECK(cs_settings_preset(CS_PRESET_DEFAULT_V60, &s));
  uint32_t n = 0;
  CHECK(cs_bed_init(&b, s.grid_dims, s.bounds_size, 0, NULL, NULL, NULL, NULL, &n));
  float* bp = (float*)malloc((size_t)n * 8 * sizeof(float));
  float* ba = (float*)malloc((size_t)n * 12 * sizeof(float));
  float* be = (float*)malloc((size_t)n * 8 * sizeof(float));
  CHECK(cs_bed_init(&b, s.grid_dims, s.bounds_size, n, bp, ba, be, NULL, &n));
  int phase_ok = n > 0 && ba[3] == 1.0f;
  free(bp);
  free(ba);
  free(be);
  if (!phase_ok) return fprintf(stderr, "bed phase\n"), 1;
  size_t nsdf = (size_t)s.sdf_res * s.sdf_res * s.sdf_res;
  float* sdf = (float*)malloc(nsdf * sizeof(float));
  CHECK(cs_generate_sdf(&s, sdf));
  int finite = 1;
  for (size_t i = 0; i < nsdf; ++i) finite &= isfinite(sdf[i]) != 0;
  free(sdf);
  if (!finite) return fprintf(stderr, "sdf not finite\n"), 1;
  printf("abi caller ok: host part, %u bed particles\n", n);
  return 0;
}

static int device_part(void) {
  cs_workload w;
  CHECK(cs_workload_build(CS_WORKLOAD_V60_SMALL, 0, 2, &w));
  cs_ctx* ctx = NULL;
  cs_sim* sim = NULL;
  CHECK(cs_ctx_create(0, &ctx));
  CHECK(cs_sim_create(ctx, &w.settings, &w.scene, &sim));
  for (int k = 0; k < 3; ++k) CHECK(cs_sim_step_frame(sim, w.frame_dt));
  size_t bytes = 0;
  CHECK(cs_sim_save(sim, NULL, 0, &bytes));
  void* blob = malloc(bytes);
  CHECK(cs_sim_save(sim, blob, bytes, &bytes));
  CHECK(cs_sim_step_frame(sim, w.frame_dt));
  uint32_t nw = 0, nb = 0, mp = 0;
  CHECK(cs_sim_get_counts(sim, &nw, &nb, &mp));
  float* a = (float*)malloc((size_t)nw * 32);
  float* c = (float*)malloc((size_t)nw * 32);
  CHECK(cs_sim_read_particles(sim, 0, nw, a));
  CHECK(cs_sim_load(sim, blob, bytes));
  CHECK(cs_sim_step_frame(sim, w.frame_dt));
  CHECK(cs_sim_read_particles(sim, 0, nw, c));
  int same = memcmp(a, c, (size_t)nw * 32) == 0;
  uint32_t m[CS_METRICS_SLOTS];
  CHECK(cs_sim_read_metrics(sim, m));
  cs_solve_stats st;
  CHECK(cs_sim_get_solve_stats(sim, &st));
  free(a);
  free(c);
  free(blob);
  const uint32_t pairs = w.settings.pressure_rbgs_pairs;
  CHECK(cs_sim_destroy(sim));
  CHECK(cs_ctx_destroy(ctx));
  CHECK(cs_workload_free(&w));  /* clears w */
  if (!same) return fprintf(stderr, "restored run differs\n"), 1;
  if (m[1] == 0 || st.iterations != pairs) return fprintf(stderr, "no solve: %u fluid cells, %u iterations\n", m[1], st.iterations), 1;
  printf("abi caller ok: %u particles, %u fluid cells, %u iterations, post-projection max|div| %.3f\n", nw, m[1],
         st.iterations, m[5] / 1024.0);
  return 0;
}

int main(int argc, char** argv) {
  if (host_part()) return 1;
  if (argc > 1 && strcmp(argv[1], "run") == 0) return device_part();
  return 0;
}
