// cs_batch.h — many independent scenes advanced by one launch per pass.
//
// Every scene keeps its own simulation buffers; a device-resident table of
// scene descriptors holds each scene's parameter block and buffer addresses.
// A batched pass launches the sum of the scenes' workgroup counts: workgroup
// g belongs to the last scene s with start[s] <= g and runs the single-scene
// kernel body on that scene's block g - start[s].  A workgroup never spans
// two scenes, so every scene computes exactly what its own launch would.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "cs_common.h"
#include "cs_solver.h"

struct cs_scene_dev {
  cs_params u;
  float* part;  // 8 floats per particle
  float* aff;   // 12 floats per particle
  int32_t* lanes;
  float* gv;
  float* gv2;
  const float* sdf;
  const uint8_t* cls;
  const int32_t* bed_id;
  const cs_bed_cell* bed;
  int8_t* kind;
  float* kf;
  int32_t* pres;
  int32_t* pres2;
  float* rhs;
  int32_t* pack;
  float* tgt;  // classify's target divergence per cell
  int32_t* prev_p;
  uint32_t* metrics;
  float* wc;
  uint32_t* brick_flag;
  uint32_t* brick_list;
  cs_solve_state* st;
  uint32_t* tile_count;
  uint32_t* tile_start;
  uint2* p2g_key;
  uint32_t* p2g_sorted;
  cs_emit_frame emit;      // this substep's emission (count, first slot, golden-angle disk frame)
  uint32_t warm_previous;  // seed the solve from prev_p
  uint32_t p2g_tiled;      // tile-binned P2G (else one thread per particle)
  double tol;
};

// Block-start tables, each nscenes + 1 entries (the last is the total).
enum cs_batch_space {
  CS_SPACE_CELLS = 0,      // 256 cells per workgroup (grid side^3 / 256)
  CS_SPACE_PARTICLES = 1,  // 256 particle slots per workgroup (up to max_particles)
  CS_SPACE_BRICKS = 2,     // one 8^3 brick / P2G tile per workgroup
  CS_SPACE_SCENES = 3,     // one workgroup per scene
  CS_SPACE_COUNT = 4
};

// Batched solve bookkeeping on the device.
struct cs_batch_solve {
  uint32_t total;       // entries in the compacted (scene, brick) list
  uint32_t live;        // scenes still solving
  uint32_t pad[2];
};

#define CS_BATCH_MAX_SCENES 1024u
#define CS_BATCH_ENTRY_SCENE_SHIFT 16u  // entry = scene << 16 | brick (bricks per scene < 65536)

// Launchers.  `starts` holds CS_SPACE_COUNT tables of (n + 1) words; `sc` is the
// device descriptor table; every launch covers all n scenes.
struct cs_batch_launch {
  hipStream_t s;
  const cs_scene_dev* sc;
  const uint32_t* starts;  // device
  const uint32_t* h_starts;  // host copy (grid sizes)
  uint32_t n;
};
inline uint32_t cs_batch_blocks(const cs_batch_launch& b, int space) { return b.h_starts[space * (b.n + 1) + b.n]; }

// Emitted particles appended (affine zeroed), lanes, metrics, P2G tile counts
// and brick flags cleared.
hipError_t cs_batch_launch_begin(const cs_batch_launch& b);
// P2G of every scene, tile-binned or one thread per particle as its descriptor says.
hipError_t cs_batch_launch_p2g(const cs_batch_launch& b);
hipError_t cs_batch_launch_grid_classify(const cs_batch_launch& b);
// State init + active-brick scan, then face weights and the previous-step start.
hipError_t cs_batch_launch_solve_setup(const cs_batch_launch& b);
// Compacts the (scene, brick) entries of the scenes still solving (all scenes when
// !skip_done) into `entries`, with per-scene offsets first[]/count[].
hipError_t cs_batch_launch_compact(const cs_batch_launch& b, bool skip_done, uint32_t* entries, uint32_t* first,
                                   uint32_t* count, cs_batch_solve* bs);
// Four half sweeps (colours color0, color0^1, ...) over the compacted entries from buffer
// `cur` (0: pres, 1: pres2) into the other, with `groups` persistent workgroups.
hipError_t cs_batch_launch_sweeps(const cs_batch_launch& b, const uint32_t* entries, const cs_batch_solve* bs,
                                  uint32_t groups, uint32_t cur, uint32_t color0, const uint32_t* sweep_table);
// Residual of buffer `cur` over the compacted entries (per-entry partials, `groups`
// persistent workgroups), then one workgroup per scene reduces its partials in
// the single-scene order and advances its solve state (mode 0/1 as
// cs_launch_residual_bricks).
hipError_t cs_batch_launch_residual(const cs_batch_launch& b, const uint32_t* entries, const uint32_t* first,
                                    const uint32_t* count, const cs_batch_solve* bs, uint32_t groups, uint32_t cur,
                                    double* partials, uint32_t mode, uint32_t iters, uint32_t launches);
// Solved pressure into pres (second buffer when the launch count is odd: by_state, or
// always when `odd`), and into prev_p for previous-step starts.
hipError_t cs_batch_launch_finalize(const cs_batch_launch& b, bool by_state, bool odd);
hipError_t cs_batch_launch_post(const cs_batch_launch& b);
// Solve states and fluid-cell counts of all scenes into `out` [n] and `fluid` [n].
hipError_t cs_batch_launch_gather_states(const cs_batch_launch& b, cs_solve_state* out, uint32_t* fluid);

#if defined(__HIPCC__)
// The scene of workgroup `g` in a space whose block-start table is `start`
// (n + 1 ascending words): the number of starts <= g, minus one, counted a
// wavefront of scenes at a time.  Every thread of the wavefront must call it;
// the result is uniform.
__device__ __forceinline__ uint32_t cs_batch_scene_of(const uint32_t* __restrict__ start, uint32_t n, uint32_t g) {
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t cnt = 0;
  for (uint32_t base = 0; base < n; base += 64u) {
    const uint32_t i = base + lane;
    const bool le = i < n && start[i] <= g;
    cnt += (uint32_t)__popcll(__ballot(le));
  }
  return cnt - 1u;
}
#endif
