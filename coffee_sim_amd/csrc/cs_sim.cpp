// cs_sim.cpp — C ABI of the MI355X MLS-MPM substep (include/coffee_sim.h).
//
// Mirrors MpmSim3D (crates/sim-wasm/src/mpm_3d/mod.rs:179-567): a handle owns
// its device buffers and stream, step_frame splits the frame into substeps and
// runs emit -> params -> clear -> P2G -> grid -> classify -> pressure solve ->
// project -> packing -> viscosity -> G2P on one HIP stream.
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "cs_ctx.h"
#include "cs_host.h"
#include "cs_kernels.h"
#include "cs_sim_impl.h"
#include "cs_solver.h"
#include "cs_unet_host.h"

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}
int hip_fail(hipError_t e, const char* what) {
  g_error = std::string(what) + ": " + hipGetErrorString(e);
  return CS_ERR_DEVICE;
}
#define CS_HIP(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hip_fail(e_, #call);  \
  } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
}

}  // namespace

void cs_set_last_error(const char* msg) { g_error = msg ? msg : ""; }
int cs_fail(int code, const char* msg) { return fail(code, msg); }
int cs_hip_fail(hipError_t e, const char* what) { return hip_fail(e, what); }



namespace {

void free_sim(cs_sim* m) {
  if (!m) return;
  void* dev[] = {m->part,       m->aff,        m->lanes,      m->gv,         m->gv2,    m->sdf,
                 m->cls,        m->bed_id,     m->bed,        m->kind,       m->kf,     m->pres,
                 m->rhs,        m->pack,       m->prev_p,     m->warm,       m->metrics, m->res_partials,
                 m->res_out,    m->pres2,      m->wc,         m->group_partials, m->sweep_table,
                 m->brick_flag, m->brick_list, m->brick_partials, m->dstate,       m->tile_count,
                 m->tile_start, m->tile_chunk, m->p2g_key,    m->p2g_sorted, m->render,     m->bed_part,
                 m->bed_aff,    m->bed_ext,    m->bed_lookup, m->bed_counter, m->bed_dw, m->tgt,
                 m->bed_snap,   m->alias4,
                 m->persist_fb, m->persist_pflags, m->persist_partials, m->persist_slot, m->persist_xcc,
                 m->mg_buf,     m->mg_flags,   m->mg_ops,     m->mg_row_buf, m->mg_row_counts, m->mg_pmap, m->mg_pnb,
                 m->mg_achild,  m->dparams};
  for (void* p : dev)
    if (p) (void)hipFree(p);
  if (m->h_res) (void)hipHostFree(m->h_res);
  if (m->h_state) (void)hipHostFree(m->h_state);
  if (m->h_done) (void)hipHostFree(m->h_done);
  if (m->h_pring) (void)hipHostFree(m->h_pring);
  if (m->h_records) (void)hipHostFree(m->h_records);
  if (m->h_halt) (void)hipHostFree(m->h_halt);
  if (m->sg_exec) (void)hipGraphExecDestroy(m->sg_exec);
  if (m->sg_graph) (void)hipGraphDestroy(m->sg_graph);
  for (auto& e : m->batch_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : m->chunk_ev)
    if (e) (void)hipEventDestroy(e);
  if (m->batch_exec) (void)hipGraphExecDestroy(m->batch_exec);
  if (m->batch_graph) (void)hipGraphDestroy(m->batch_graph);
  for (auto& e : m->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : m->kev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : m->mg_ev)
    if (e) (void)hipEventDestroy(e);
  delete m;
}

int upload_initial(cs_sim* m) {
  hipStream_t st = m->ctx->stream;
  CS_HIP(hipMemsetAsync(m->part, 0, (size_t)m->s.max_particles * 32, st));
  CS_HIP(hipMemsetAsync(m->aff, 0, (size_t)m->s.max_particles * 48, st));
  if (m->init_count) {
    CS_HIP(hipMemcpyAsync(m->part, m->init_particles.data(), (size_t)m->init_count * 32, hipMemcpyHostToDevice, st));
    CS_HIP(hipMemcpyAsync(m->aff, m->init_affine.data(), (size_t)m->init_count * 48, hipMemcpyHostToDevice, st));
  }
  CS_HIP(hipMemsetAsync(m->prev_p, 0, m->nc * sizeof(int32_t), st));
  if (m->num_bed) {  // the bed's grains and extraction state start over
    CS_HIP(hipMemcpyAsync(m->bed_part, m->init_bed_part.data(), m->init_bed_part.size() * sizeof(float),
                          hipMemcpyHostToDevice, st));
    CS_HIP(hipMemcpyAsync(m->bed_aff, m->init_bed_aff.data(), m->init_bed_aff.size() * sizeof(float),
                          hipMemcpyHostToDevice, st));
    CS_HIP(hipMemcpyAsync(m->bed_ext, m->init_bed_ext.data(), m->init_bed_ext.size() * sizeof(float),
                          hipMemcpyHostToDevice, st));
    CS_HIP(hipMemsetAsync(m->bed_dw, 0, (size_t)m->num_bed * 4 * sizeof(int32_t), st));
  }
  CS_HIP(hipStreamSynchronize(st));
  return CS_OK;
}

void mark(cs_sim* m, int pass) {
  if (m->timing) (void)hipEventRecord(m->ev[pass], m->ctx->stream);
}
void kmark(cs_sim* m, int which) {
  if (m->timing) (void)hipEventRecord(m->kev[which], m->ctx->stream);
}
double kelapsed_us(cs_sim* m, int a, int b) {
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, m->kev[a], m->kev[b]) != hipSuccess) return 0.0;
  return (double)ms * 1000.0;
}

// Residual of the current pressure; blocks on the stream.
int residual_now(cs_sim* m, const cs_params& u, double* out) {
  hipStream_t st = m->ctx->stream;
  CS_HIP(cs_launch_residual(st, u, m->kf, m->rhs, m->pres, m->res_partials, CS_RESIDUAL_BLOCKS, m->res_out));
  CS_HIP(hipMemcpyAsync(m->h_res, m->res_out, sizeof(double), hipMemcpyDeviceToHost, st));
  CS_HIP(hipStreamSynchronize(st));
  *out = m->h_res[0];
  return CS_OK;
}

// The attached network's pressure prediction from this substep's solve inputs,
// into m->warm (then injected into the fluid cells like an external start).
int predict_warm_start(cs_sim* m, const cs_params& u) {
  if (!m->warm_net) return fail(CS_ERR_STATE, "warm_start CS_WARM_LEARNED needs cs_sim_set_warm_net");
  cs_unet* net = m->warm_net;
  CS_HIP(cs_launch_unet_features(m->ctx->stream, u, m->kf, m->rhs, nullptr, m->bed_id, m->bed, cs_unet_rhs_scale(net),
                                 1.0f, cs_unet_feature_buffer(net), nullptr, nullptr, nullptr));
  return cs_unet_predict_device(net, m->warm);
}

// One-thread-per-cell solve over the whole grid (the reference's dispatch
// shape), kept as a cross-check of the brick solve.
int solve_dense(cs_sim* m, const cs_pblock& P) {
  const cs_params& u = P.u;
  if (m->s.rbgs_rescale_after) return fail(CS_ERR_INVALID, "rbgs_rescale_after needs the brick solver");
  if (m->s.rbgs_mg_cycles) return fail(CS_ERR_INVALID, "rbgs_mg_cycles needs the brick solver");
  hipStream_t st = m->ctx->stream;
  cs_solve_stats& S = m->stats;
  memset(&S, 0, sizeof(S));
  S.final_residual = NAN;
  m->final_residual_pending = false;
  bool external = false;
  if (m->s.warm_start == CS_WARM_PREVIOUS) {
    CS_HIP(cs_launch_warm_previous(st, P, m->kf, m->prev_p, m->pres));
  } else if (m->s.warm_start == CS_WARM_EXTERNAL && m->has_warm) {
    CS_HIP(cs_launch_warm_external(st, u, m->kf, m->warm, m->pres, m->metrics));
    external = true;
  } else if (m->s.warm_start == CS_WARM_LEARNED) {
    int rc = predict_warm_start(m, u);
    if (rc) return rc;
    CS_HIP(cs_launch_warm_external(st, u, m->kf, m->warm, m->pres, m->metrics));
    external = true;
  }
  if (m->s.rbgs_mode == CS_RBGS_FIXED_PAIRS) {
    kmark(m, 4);
    for (uint32_t k = 0; k < m->s.pressure_rbgs_pairs; ++k)
      CS_HIP(cs_launch_rbgs_pair(st, u, m->kf, m->rhs, m->pres, m->metrics));
    kmark(m, 5);
    if (m->timing) {
      CS_HIP(hipEventSynchronize(m->kev[5]));
      m->kt.rbgs_us += kelapsed_us(m, 4, 5);
      m->kt.rbgs_launches += 2ull * m->s.pressure_rbgs_pairs;
      m->kt.rbgs_cells_swept += 2ull * m->s.pressure_rbgs_pairs * m->nc;
    }
    S.iterations = m->s.pressure_rbgs_pairs;
    // residual of the final pressure, read lazily by cs_sim_get_solve_stats
    CS_HIP(cs_launch_residual(st, u, m->kf, m->rhs, m->pres, m->res_partials, CS_RESIDUAL_BLOCKS, m->res_out));
    CS_HIP(hipMemcpyAsync(m->h_res, m->res_out, sizeof(double), hipMemcpyDeviceToHost, st));
    m->final_residual_pending = true;
  } else {
    double r = 0.0;
    int rc = residual_now(m, u, &r);
    if (rc) return rc;
    if (external && r > (double)m->s.warm_fallback_residual) {
      // the supplied start is worse than the zero start: fall back
      CS_HIP(cs_launch_fill_i32(st, m->pres, m->nc, 0));
      S.fell_back = 1;
      if ((rc = residual_now(m, u, &r))) return rc;
    }
    S.initial_residual = r;
    const uint32_t every = m->s.rbgs_check_every ? m->s.rbgs_check_every : 8u;
    const uint32_t cap = m->s.rbgs_max_iters;
    const double tol = (double)m->s.rbgs_tolerance;
    uint32_t it = 0;
    while (!(r <= tol) && it < cap) {
      uint32_t chunk = every - (it % every);
      chunk = std::min(chunk, cap - it);
      kmark(m, 4);
      for (uint32_t k = 0; k < chunk; ++k)
        CS_HIP(cs_launch_rbgs_pair(st, u, m->kf, m->rhs, m->pres, m->metrics));
      kmark(m, 5);
      it += chunk;
      if ((rc = residual_now(m, u, &r))) return rc;
      if (m->timing) {
        m->kt.rbgs_us += kelapsed_us(m, 4, 5);
        m->kt.rbgs_launches += 2ull * chunk;
        m->kt.rbgs_cells_swept += 2ull * chunk * m->nc;
      }
    }
    S.iterations = it;
    S.final_residual = r;
    S.converged = r <= tol ? 1u : 0u;
  }
  if (m->s.warm_start == CS_WARM_PREVIOUS) CS_HIP(cs_launch_copy_i32(st, m->pres, m->prev_p, m->nc));
  return CS_OK;
}

int32_t* buf_of(cs_sim* m, int cur) { return cur ? m->pres2 : m->pres; }

// The residual test of the previous chunk, folded into the first two launches
// of the next one (multi-brick launches only): the first sums the residual of
// its input, the second decides.
struct FusedCheck {
  uint32_t iters;     // iterations of the chunk being judged
  uint32_t launches;  // its sweep launches
  double tol;
};

// Whether a chunk of `iters` iterations starts with two multi-brick launches.
bool can_fuse(const cs_sim* m, uint32_t iters) { return m->halo == 4 && m->multi && iters >= 4; }

// Enqueues `iters` red+black iterations as brick launches of `halo` half
// sweeps (single-iteration launches for any odd remainder); returns the number
// of launches.  Buffers alternate per launch starting from `*cur`.  With
// `check` (only when can_fuse holds) the first two launches carry the
// previous chunk's residual test.
int enqueue_sweeps(cs_sim* m, const cs_params& u, uint32_t nactive, uint32_t iters, uint32_t* launches, int* cur,
                   const FusedCheck* check = nullptr) {
  hipStream_t st = m->ctx->stream;
  int32_t* buf[2] = {m->pres, m->pres2};
  uint32_t n = 0;
  uint32_t left = iters;
  const uint32_t per = (uint32_t)m->halo / 2u;  // iterations per launch (halo 1 -> half iterations)
  while (left > 0) {
    if (m->halo >= 2 && left >= per) {
      if (m->halo == 4 && m->multi) {
        uint32_t mode = 0u, ci = 0u, cl = 0u;
        double tol = 0.0;
        if (check && n == 0) mode = CS_RES_COMPUTE;
        if (check && n == 1) {
          mode = CS_RES_DECIDE;
          ci = check->iters;
          cl = check->launches;
          tol = check->tol;
        }
        CS_HIP(cs_launch_rbgs_multi(st, u, nactive, m->multi_groups, 0u, m->brick_list, m->wc, m->rhs, buf[*cur],
                                    buf[*cur ^ 1], m->metrics, m->dstate, m->group_partials, mode, ci, cl, tol,
                                    m->sweep_table));
      } else {
        CS_HIP(cs_launch_rbgs_bricks(st, u, nactive, m->halo, m->threads, 0u, m->brick_list, m->wc, m->rhs,
                                     buf[*cur], buf[*cur ^ 1], m->metrics, m->dstate));
      }
      *cur ^= 1;
      ++n;
      left -= per;
    } else {
      int h = m->halo >= 2 ? 2 : 1;
      for (uint32_t c = 0; c < (uint32_t)(2 / h); ++c) {
        CS_HIP(cs_launch_rbgs_bricks(st, u, nactive, h, h == 1 ? 256 : m->threads, h == 1 ? c : 0u, m->brick_list,
                                     m->wc, m->rhs, buf[*cur], buf[*cur ^ 1], m->metrics, m->dstate));
        *cur ^= 1;
        ++n;
      }
      left -= 1;
    }
  }
  *launches = n;
  return CS_OK;
}

// Captures one batch of `batch_chunks` chunks of `every` iterations whose
// residual tests are folded into the sweeps (the steady state of a solve)
// into m->batch_exec.  Kernel arguments are the current substep's, so every
// solve that uses it records the batch again and updates the instantiated
// graph in place (a new instantiation only when the update is refused).  The
// batch has an even number of launches, so the pressure buffers are back in
// place after each replay.
int capture_batch(cs_sim* m, const cs_params& u, uint32_t nactive, uint32_t every, FusedCheck first) {
  hipStream_t st = m->ctx->stream;
  CS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int cur = 0;
  FusedCheck chk = first;
  int rc = CS_OK;
  for (uint32_t c = 0; c < m->batch_chunks && rc == CS_OK; ++c) {
    uint32_t launches = 0;
    rc = enqueue_sweeps(m, u, nactive, every, &launches, &cur, &chk);
    chk.iters = every;
    chk.launches = launches;
    m->batch_launches_per_chunk = launches;
  }
  hipGraph_t g = nullptr;
  const hipError_t ee = hipStreamEndCapture(st, &g);
  if (rc != CS_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ee != hipSuccess) return hip_fail(ee, "hipStreamEndCapture");
  if (cur != 0) {
    (void)hipGraphDestroy(g);
    return fail(CS_ERR_STATE, "captured batch has an odd number of sweep launches");
  }
  if (m->batch_exec) {
    hipGraphNode_t bad = nullptr;
    hipGraphExecUpdateResult why = hipGraphExecUpdateSuccess;
    if (hipGraphExecUpdate(m->batch_exec, g, &bad, &why) == hipSuccess && why == hipGraphExecUpdateSuccess) {
      if (m->batch_graph) (void)hipGraphDestroy(m->batch_graph);
      m->batch_graph = g;
      return CS_OK;
    }
    (void)hipGetLastError();
    (void)hipGraphExecDestroy(m->batch_exec);
    m->batch_exec = nullptr;
  }
  if (m->batch_graph) (void)hipGraphDestroy(m->batch_graph);
  m->batch_graph = g;
  CS_HIP(hipGraphInstantiate(&m->batch_exec, g, nullptr, nullptr, 0));
  return CS_OK;
}

// ---- multigrid start (cs_mg.hip) ---------------------------------------------------------
int ensure_mg(cs_sim* m) {
  if (m->mg_buf) return CS_OK;
  std::vector<cs_mg_level> lv = cs_mg_shape(m->s.grid_dims[0], m->s.grid_dims[1], m->s.grid_dims[2]);
  // the fine level brick-major: a workgroup's brick is one contiguous run per array
  auto cells = [](const cs_mg_level& l) {
    return l.bnx ? (size_t)l.bnx * l.bny * ((l.Z + 7u) / 8u) * 512u : (size_t)l.X * l.Y * l.Z;
  };
  if (m->mg_brick_major) {
    lv[0].bnx = (lv[0].X + 7u) / 8u;
    lv[0].bny = (lv[0].Y + 7u) / 8u;
  }
  size_t total = 0;
  for (const auto& l : lv) total += 6 * cells(l);
  CS_HIP(dalloc(&m->mg_buf, total));
  CS_HIP(dalloc(&m->mg_flags, 3));
  size_t off = 0;
  for (auto& l : lv) {
    const size_t n = cells(l);
    float* b = m->mg_buf + off;
    l.D = b;
    l.Wx = b + n;
    l.Wy = b + 2 * n;
    l.Wz = b + 3 * n;
    l.R = b + 4 * n;
    l.E = b + 5 * n;
    off += 6 * n;
  }
  m->mg_floats = total;
  m->mg_n0 = cells(lv[0]);
  const int nl = (int)lv.size();
  // the first coarse level small enough for one workgroup; its visits run whole
  m->mg_wg_first = nl;
  for (int l = 1; l < nl; ++l)
    if ((uint64_t)lv[l].X * lv[l].Y * lv[l].Z <= CS_MG_WG_CELLS) {
      m->mg_wg_first = l;
      break;
    }
  if (m->mg_wg_first < nl) {
    CS_HIP(cs_mg_wg_prepare(cs_mg_wg_lds_bytes(lv.data(), m->mg_wg_first, nl)));
    std::vector<cs_mg_op> ops;
    cs_mg_visit_ops(m->mg_wg_first, nl, &ops);
    if (ops.size() > CS_MG_MAX_OPS) return fail(CS_ERR_STATE, "multigrid: the coarse visit has too many operations");
    CS_HIP(dalloc(&m->mg_ops, ops.size()));
    CS_HIP(hipMemcpy(m->mg_ops, ops.data(), ops.size() * sizeof(cs_mg_op), hipMemcpyHostToDevice));
    m->mg_nops = (uint32_t)ops.size();
    m->mg_rows = m->mg_rows_allowed >= 1 && cs_mg_rows_fit(lv.data(), m->mg_wg_first, nl);
    m->mg_rows3_shape = m->mg_rows && m->mg_rows_allowed >= 2 && cs_mg_rows3_fit(lv.data(), m->mg_wg_first, nl);
    if (m->mg_rows) {
      auto cells = [&](int l) { return (size_t)lv[l].X * lv[l].Y * lv[l].Z; };
      const size_t np = m->mg_rows3_shape ? cells(m->mg_wg_first - 1) : 0;
      CS_HIP(dalloc(&m->mg_row_buf, 2 * (cells(m->mg_wg_first) + cells(nl - 1) + np)));
      CS_HIP(dalloc(&m->mg_row_counts, 6));
      if (m->mg_rows3_shape) {
        CS_HIP(dalloc(&m->mg_pmap, np));
        CS_HIP(dalloc(&m->mg_pnb, 2 * np));
        CS_HIP(dalloc(&m->mg_achild, 2 * cells(m->mg_wg_first)));
        CS_HIP(cs_mg_rows3_prepare());
      }
    }
  }
  for (auto& e : m->mg_ev) CS_HIP(hipEventCreate(&e));
  m->mg_levels = lv;
  return CS_OK;
}

// One W-cycle visit of level l; the levels from mg_wg_first down run in one
// workgroup, the finer ones as launches (level 0 over the active bricks).
int mg_visit(cs_sim* m, const cs_pblock& P, uint32_t nactive, int l) {
  hipStream_t st = m->ctx->stream;
  const int nl = (int)m->mg_levels.size();
  const uint32_t* skip = m->mg_flags;
  if (l == m->mg_wg_first - 1 && m->mg_rows3) {
    CS_HIP(cs_mg_rows3_visit(st, m->mg_levels[l], m->mg_levels[l + 1], m->mg_levels[l + 2], m->mg_row_buf,
                             m->mg_row_counts, m->mg_pnb, m->mg_achild, cs_mg_coarse_passes(nl), skip));
    return CS_OK;
  }
  if (l >= m->mg_wg_first && m->mg_rows) {
    CS_HIP(cs_mg_rows_visit(st, m->mg_levels[l], m->mg_levels[l + 1], m->mg_row_buf, m->mg_row_counts,
                            cs_mg_coarse_passes(nl), skip));
    return CS_OK;
  }
  if (l >= m->mg_wg_first) {
    cs_mg_level lv[CS_MG_MAX_LEVELS] = {};
    for (int k = 0; k < nl; ++k) lv[k] = m->mg_levels[k];
    CS_HIP(cs_mg_wg_visit(st, lv, m->mg_wg_first, nl, m->mg_ops, m->mg_nops, skip));
    return CS_OK;
  }
  const cs_mg_level& L = m->mg_levels[l];
  auto passes = [&](uint32_t n) -> int {
    for (uint32_t k = 0; k < n; ++k)
      for (uint32_t c = 0; c < 2u; ++c) {
        if (l == 0) CS_HIP(cs_mg_fine_smooth(st, P, nactive, m->brick_list, L, c, skip));
        else CS_HIP(cs_mg_smooth(st, L, c, skip));
      }
    return CS_OK;
  };
  if (l == nl - 1) return passes(cs_mg_coarse_passes(nl));
  int rc = passes(CS_MG_NU);
  if (rc) return rc;
  const cs_mg_level& C = m->mg_levels[l + 1];
  for (uint32_t g = 0; g < CS_MG_GAMMA; ++g) {
    // restriction (zeroes the coarse correction), the coarse visit, then the
    // prolongation folded into the first red pass
    if (l == 0) CS_HIP(cs_mg_fine_restrict(st, P, nactive, m->brick_list, L, C, skip));
    else CS_HIP(cs_mg_restrict(st, L, C, skip));
    rc = mg_visit(m, P, nactive, l + 1);
    if (rc) return rc;
    if (l == 0) {
      CS_HIP(cs_mg_fine_smooth_prolonged(st, P, nactive, m->brick_list, L, C, skip));
      CS_HIP(cs_mg_fine_smooth(st, P, nactive, m->brick_list, L, 1u, skip));
    } else {
      CS_HIP(cs_mg_smooth_prolonged(st, L, C, skip));
      CS_HIP(cs_mg_smooth(st, L, 1u, skip));
    }
    rc = passes(CS_MG_NU - 1u);
    if (rc) return rc;
  }
  return CS_OK;
}

// rbgs_mg_cycles W-cycles on the current start in m->pres, skipped on the
// device when the start is already within the tolerance.
int mg_start(cs_sim* m, const cs_pblock& P, uint32_t nactive) {
  int rc = ensure_mg(m);
  if (rc) return rc;
  hipStream_t st = m->ctx->stream;
  const double tol = (double)m->s.rbgs_tolerance;
  if (m->timing) CS_HIP(hipEventRecord(m->mg_ev[0], st));
  CS_HIP(cs_launch_residual_gate(st, P, nactive, m->brick_list, m->wc, m->rhs, m->pres, m->brick_partials,
                                 m->dstate, tol, m->mg_flags, true, m->s.rbgs_mg_cycles == 0));
  // kernels touch only the active bricks' fine cells: the fine correction and
  // every coarse level start from zero everywhere else
  const cs_mg_level& L0 = m->mg_levels[0];
  // (the fine correction is the last of the fine arrays, the coarse levels follow it: one clear)
  CS_HIP(cs_launch_fill_i32(st, reinterpret_cast<int32_t*>(L0.E), m->mg_floats - 5 * m->mg_n0, 0));
  const int nl = (int)m->mg_levels.size();
  const uint32_t* skip = m->mg_flags;
  // the level above the small ones joins their visit when its operator fits
  // CS_MG_P3_MAX cells for sure: its cells lie in the active bricks' parents
  m->mg_rows3 = false;
  if (m->mg_rows3_shape) {
    const int lp = m->mg_wg_first - 1;
    const uint64_t side = std::max<uint64_t>(1u, (uint64_t)CS_BRICK >> std::min(lp, 30));
    const cs_mg_level& P = m->mg_levels[lp];
    const uint64_t bound = std::min<uint64_t>((uint64_t)P.X * P.Y * P.Z, side * side * side * nactive);
    m->mg_rows3 = bound <= CS_MG_P3_MAX;
  }
  for (uint32_t c = 0; c < m->s.rbgs_mg_cycles; ++c) {
    CS_HIP(cs_mg_fine_setup(st, P, nactive, m->brick_list, m->wc, m->rhs, m->pres, L0, skip,
                            m->mg_rows ? m->mg_row_counts : nullptr));
    if (nl > 1) CS_HIP(cs_mg_fine_build(st, P, nactive, m->brick_list, L0, m->mg_levels[1], skip));
    for (int l = 1; l + 1 < nl; ++l) CS_HIP(cs_mg_build(st, m->mg_levels[l], m->mg_levels[l + 1], skip));
    if (m->mg_rows)
      CS_HIP(cs_mg_rows(st, m->mg_levels[m->mg_wg_first], m->mg_levels[nl - 1],
                        m->mg_rows3 ? &m->mg_levels[m->mg_wg_first - 1] : nullptr, m->mg_row_buf, m->mg_row_counts,
                        m->mg_pmap, m->mg_pnb, m->mg_achild, skip));
    rc = mg_visit(m, P, nactive, 0);
    if (rc) return rc;
    CS_HIP(cs_mg_fine_apply(st, P, nactive, m->brick_list, L0, m->pres, skip));
    // the cycle's residual test: later cycles return at once within the tolerance
    CS_HIP(cs_launch_residual_gate(st, P, nactive, m->brick_list, m->wc, m->rhs, m->pres, m->brick_partials,
                                   m->dstate, tol, m->mg_flags, false, c + 1 == m->s.rbgs_mg_cycles));
  }
  if (m->timing) {
    CS_HIP(hipEventRecord(m->mg_ev[1], st));
    m->mg_timed = true;
  }
  return CS_OK;
}
void mg_harvest(cs_sim* m) {
  if (!m->mg_timed) return;
  float ms = 0.0f;
  uint32_t ran = 0;
  if (hipEventSynchronize(m->mg_ev[1]) == hipSuccess && hipEventElapsedTime(&ms, m->mg_ev[0], m->mg_ev[1]) == hipSuccess &&
      hipMemcpy(&ran, m->mg_flags + 2, sizeof(uint32_t), hipMemcpyDeviceToHost) == hipSuccess) {
    m->kt.mg_us += (double)ms * 1000.0;
    m->kt.mg_starts += 1;
    m->kt.mg_cycles += ran;
  }
  m->mg_timed = false;
}

// The persistent solve's launch, after the multigrid start when one is asked
// for, sized for at most `cap` active bricks (the count itself is read on the
// device).  `cooperative` has the runtime refuse a launch whose workgroups
// cannot all be resident; a recorded substep uses a plain launch, whose width
// the host has checked against the same limit.
int persist_enqueue(cs_sim* m, const cs_pblock& P, uint32_t cap, bool cooperative) {
  const cs_params& u = P.u;
  hipStream_t st = m->ctx->stream;
  const bool fixed = m->s.rbgs_mode == CS_RBGS_FIXED_PAIRS;
  if (!fixed && m->s.rbgs_mg_cycles) {
    int rc = mg_start(m, P, cap);
    if (rc) return rc;
  }
  CS_HIP(cs_persist_prepare(st, u, m->brick_list, m->dstate, m->persist_slot, m->persist_pflags, m->persist_xcc));
  cs_persist_args a{};
  a.dp = P.d;
  a.list = m->brick_list;
  a.slot = m->persist_slot;
  a.xcc = m->persist_xcc;
  a.wc = reinterpret_cast<const float4*>(m->wc);
  a.rhs = m->rhs;
  a.p = m->pres;
  a.metrics = m->metrics;
  a.st = m->dstate;
  a.fb = m->persist_fb;
  a.pflags = m->persist_pflags;
  a.partials = m->persist_partials;
  a.max_runs = u.nbricks;
  a.fixed_pairs = fixed ? m->s.pressure_rbgs_pairs : 0u;
  a.every = m->s.rbgs_check_every ? m->s.rbgs_check_every : 8u;
  a.cap = m->s.rbgs_max_iters;
  a.tol = (double)m->s.rbgs_tolerance;
  a.spin_limit = 200000000ull;  // 2 s of the 100 MHz clock: far beyond any legitimate wait
  a.poll_delay = m->persist_poll_delay;
  // after a multigrid start the first residual test comes after one check interval
  a.min_gate = (!fixed && m->s.rbgs_mg_cycles) ? m->mg_flags + 1 : nullptr;
  a.min_iters = a.every;
  if (m->timing) CS_HIP(hipEventRecord(m->chunk_ev[0], st));
  CS_HIP(cs_persist_launch(st, a, cs_persist_groups(cap), cooperative));
  if (m->timing) CS_HIP(hipEventRecord(m->chunk_ev[1], st));
  return CS_OK;
}

// Epochs one persistent solve takes (the next solve's base moves by this).
uint32_t persist_span(const cs_sim* m) {
  const bool fixed = m->s.rbgs_mode == CS_RBGS_FIXED_PAIRS;
  return cs_persist_epochs(fixed ? m->s.pressure_rbgs_pairs : m->s.rbgs_max_iters);
}

// The whole solve as one launch of co-resident workgroups (cs_rbgs_persist.hip):
// same updates, same residual tests, same stopping iteration as the
// launch-per-chunk solve; the pressure is solved in place in m->pres.
int solve_persistent(cs_sim* m, const cs_pblock& P, uint32_t nactive, bool external) {
  hipStream_t st = m->ctx->stream;
  cs_solve_stats& S = m->stats;
  const bool fixed = m->s.rbgs_mode == CS_RBGS_FIXED_PAIRS;
  const double tol = (double)m->s.rbgs_tolerance;
  if (!fixed && external) {
    // a supplied start is judged by its residual before the solve (the fallback)
    CS_HIP(cs_launch_residual_bricks(st, P, nactive, m->brick_list, m->wc, m->rhs, m->pres, m->brick_partials,
                                     m->dstate, 0u, 0u, 0u, tol));
    CS_HIP(hipMemcpyAsync(m->h_state, m->dstate, sizeof(cs_solve_state), hipMemcpyDeviceToHost, st));
    CS_HIP(hipStreamSynchronize(st));
    if (m->h_state->initial_residual > (double)m->s.warm_fallback_residual) {
      CS_HIP(cs_launch_fill_i32(st, m->pres, m->nc, 0));
      S.fell_back = 1;
    }
  }
  int rc = persist_enqueue(m, P, nactive, true);
  if (rc) return rc;
  CS_HIP(hipMemcpyAsync(m->h_state, m->dstate, sizeof(cs_solve_state), hipMemcpyDeviceToHost, st));
  CS_HIP(hipStreamSynchronize(st));
  const cs_solve_state& hs = *m->h_state;
  if (hs.faults) return fail(CS_ERR_DEVICE, "persistent pressure solve: a workgroup gave up waiting for its neighbours");
  ++m->persist_solves;
  m->persist_epoch += persist_span(m);
  S.iterations = hs.iterations;
  S.initial_residual = fixed ? 0.0 : hs.initial_residual;
  S.final_residual = hs.residual;
  S.converged = hs.residual <= tol ? 1u : 0u;
  if (m->timing) {
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, m->chunk_ev[0], m->chunk_ev[1]);
    m->kt.rbgs_us += (double)ms * 1000.0;
    m->kt.rbgs_launches += 1;
    m->kt.rbgs_cells_swept += (uint64_t)hs.iterations * nactive * CS_BRICK * CS_BRICK * CS_BRICK;
  }
  return CS_OK;
}

// The brick solve of a recorded substep: the launches of solve_bricks' persistent
// path, each sized for `cap` active bricks, with no host decision between them;
// the solve's final state goes to the substep's record (m->h_records).
int solve_recorded(cs_sim* m, const cs_pblock& P, uint32_t cap) {
  hipStream_t st = m->ctx->stream;
  CS_HIP(cs_launch_brick_scan(st, P, m->brick_flag, m->brick_list, m->dstate));
  CS_HIP(cs_launch_face_weights(st, P, m->kf, m->wc));
  if (m->s.warm_start == CS_WARM_PREVIOUS) CS_HIP(cs_launch_warm_previous(st, P, m->kf, m->prev_p, m->pres));
  int rc = persist_enqueue(m, P, cap, false);
  if (rc) return rc;
  CS_HIP(cs_launch_solve_record(st, P.d, m->dstate, m->h_records, m->h_halt, CS_SOLVE_RECORDS));
  if (m->s.warm_start == CS_WARM_PREVIOUS) CS_HIP(cs_launch_copy_i32_gated(st, P.d, m->pres, m->prev_p, m->nc));
  return CS_OK;
}

int solve_bricks(cs_sim* m, const cs_pblock& P) {
  const cs_params& u = P.u;
  hipStream_t st = m->ctx->stream;
  cs_solve_stats& S = m->stats;
  memset(&S, 0, sizeof(S));
  S.final_residual = NAN;
  m->final_residual_pending = false;
  CS_HIP(cs_launch_brick_scan(st, P, m->brick_flag, m->brick_list, m->dstate));
  CS_HIP(cs_launch_face_weights(st, P, m->kf, m->wc));
  bool external = false;
  if (m->s.warm_start == CS_WARM_PREVIOUS) {
    CS_HIP(cs_launch_warm_previous(st, P, m->kf, m->prev_p, m->pres));
  } else if (m->s.warm_start == CS_WARM_EXTERNAL && m->has_warm) {
    CS_HIP(cs_launch_warm_external(st, u, m->kf, m->warm, m->pres, m->metrics));
    external = true;
  } else if (m->s.warm_start == CS_WARM_LEARNED) {
    int rc = predict_warm_start(m, u);
    if (rc) return rc;
    CS_HIP(cs_launch_warm_external(st, u, m->kf, m->warm, m->pres, m->metrics));
    external = true;
  }
  const double tol = (double)m->s.rbgs_tolerance;
  const bool fixed = m->s.rbgs_mode == CS_RBGS_FIXED_PAIRS;
  // the active brick count sizes the launches
  CS_HIP(hipMemcpyAsync(m->h_state, m->dstate, sizeof(cs_solve_state), hipMemcpyDeviceToHost, st));
  CS_HIP(hipStreamSynchronize(st));
  const uint32_t nactive = m->h_state->active_bricks;
  m->nactive = nactive;
  if (m->persist && !m->s.rbgs_rescale_after && nactive > 0 && nactive <= CS_PERSIST_MAX_BRICKS &&
      cs_persist_groups(nactive) <= m->persist_max_groups &&
      (!fixed || m->s.pressure_rbgs_pairs > 0)) {
    int rc = solve_persistent(m, P, nactive, external);
    if (rc) return rc;
    mg_harvest(m);
    if (m->s.warm_start == CS_WARM_PREVIOUS) CS_HIP(cs_launch_copy_i32(st, m->pres, m->prev_p, m->nc));
    return CS_OK;
  }
  int cur = 0;
  if (fixed) {
    uint32_t launches = 0;
    size_t ev0 = 0;
    if (m->timing) CS_HIP(hipEventRecord(m->chunk_ev[ev0], st));
    int rc = enqueue_sweeps(m, u, nactive, m->s.pressure_rbgs_pairs, &launches, &cur);
    if (rc) return rc;
    if (m->timing) CS_HIP(hipEventRecord(m->chunk_ev[ev0 + 1], st));
    // the final residual for the statistics (mode 0 also records launches = 0)
    CS_HIP(cs_launch_residual_bricks(st, P, nactive, m->brick_list, m->wc, m->rhs, buf_of(m, cur), m->brick_partials,
                                     m->dstate, 0u, 0u, 0u, tol));
    S.iterations = m->s.pressure_rbgs_pairs;
    if (cur != 0) CS_HIP(cs_launch_select_final(st, u, nactive, m->brick_list, m->pres, m->pres2, m->dstate, false));
    CS_HIP(hipMemcpyAsync(m->h_state, m->dstate, sizeof(cs_solve_state), hipMemcpyDeviceToHost, st));
    m->final_residual_pending = true;
    m->pending_from_state = true;
    if (m->timing) {
      CS_HIP(hipEventSynchronize(m->chunk_ev[ev0 + 1]));
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, m->chunk_ev[ev0], m->chunk_ev[ev0 + 1]);
      m->kt.rbgs_us += (double)ms * 1000.0;
      m->kt.rbgs_launches += launches;
      m->kt.rbgs_cells_swept += (uint64_t)launches * nactive * CS_BRICK * CS_BRICK * CS_BRICK;
    }
  } else {
    CS_HIP(cs_launch_residual_bricks(st, P, nactive, m->brick_list, m->wc, m->rhs, m->pres, m->brick_partials,
                                     m->dstate, 0u, 0u, 0u, tol));
    if (external) {
      CS_HIP(hipMemcpyAsync(m->h_state, m->dstate, sizeof(cs_solve_state), hipMemcpyDeviceToHost, st));
      CS_HIP(hipStreamSynchronize(st));
      if (m->h_state->initial_residual > (double)m->s.warm_fallback_residual) {
        // the supplied start is worse than the zero start: fall back
        CS_HIP(cs_launch_fill_i32(st, m->pres, m->nc, 0));
        S.fell_back = 1;
        CS_HIP(cs_launch_residual_bricks(st, P, nactive, m->brick_list, m->wc, m->rhs, m->pres, m->brick_partials,
                                         m->dstate, 0u, 0u, 0u, tol));
      }
    }
    const uint32_t every = m->s.rbgs_check_every ? m->s.rbgs_check_every : 8u;
    const uint32_t cap = m->s.rbgs_max_iters;
    const size_t pool = m->chunk_ev.size() / 2;  // event pairs, one per batch, >= 3 batches in flight
    // rbgs_rescale_after: that many untested iterations (an even number of
    // launches, so the pressure is back in m->pres), then the rescale; a start
    // already within the tolerance skips both (the kernels read the state).
    const uint32_t pre = m->s.rbgs_rescale_after;
    uint32_t pre_launches = 0;
    uint32_t enq = pre, batch = 0;
    uint32_t harvested = 0;  // batches whose sweep time has been added
    std::vector<uint32_t> batch_iter_end;  // iterations enqueued through each batch
    auto harvest = [&](uint32_t upto) {
      for (; harvested < upto; ++harvested) {
        size_t e = (size_t)(harvested % pool) * 2;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, m->chunk_ev[e], m->chunk_ev[e + 1]);
        m->kt.rbgs_us += (double)ms * 1000.0;
      }
    };
    // A batch of full chunks whose residual tests are all folded into the
    // sweeps is the same launch sequence every time: capture it once per
    // solve and replay it.
    const bool graph_ok = m->use_graphs && m->fuse_residual && can_fuse(m, every);
    bool graph_built = false;
    bool stop = false;
    bool pending = false;  // the last enqueued chunk still awaits its residual test
    FusedCheck prev{0u, 0u, tol};
    if (m->s.rbgs_mg_cycles) {
      // the multigrid start leaves the state's initial residual and `done` as
      // the oracle's loop sees them; its first chunk runs before any test (the
      // sweeps return at once when the start was within the tolerance)
      int rc = mg_start(m, P, nactive);
      if (rc) return rc;
      const uint32_t chunk = std::min(every, cap);
      uint32_t launches = 0;
      rc = enqueue_sweeps(m, u, nactive, chunk, &launches, &cur);
      if (rc) return rc;
      pending = true;
      prev.iters = chunk;
      prev.launches = launches;
      enq = chunk;
    }
    while (enq < cap && !stop) {
      const size_t e = (size_t)(batch % pool) * 2;
      if (m->timing) CS_HIP(hipEventRecord(m->chunk_ev[e], st));
      const bool steady = graph_ok && pending && cur == 0 && prev.iters == every &&
                          cap - enq >= m->batch_chunks * every;
      if (steady && !graph_built) {
        int rc = capture_batch(m, u, nactive, every, prev);
        if (rc) return rc;
        graph_built = true;
      }
      if (batch == 0 && pre > 0) {
        int rc = enqueue_sweeps(m, u, nactive, pre, &pre_launches, &cur);
        if (rc) return rc;
        if (cur != 0) return fail(CS_ERR_STATE, "rescale sweeps left the pressure in the second buffer");
        CS_HIP(cs_launch_rescale(st, u, nactive, m->brick_list, m->wc, m->rhs, m->kf, m->pres, m->brick_partials,
                                 m->dstate, m->metrics));
      }
      if (steady) {
        CS_HIP(hipGraphLaunch(m->batch_exec, st));
        enq += m->batch_chunks * every;
        prev.iters = every;
        prev.launches = m->batch_launches_per_chunk;
      } else {
        for (uint32_t c = 0; c < m->batch_chunks && enq < cap; ++c) {
          uint32_t chunk = std::min(every, cap - enq);
          const bool fuse = pending && m->fuse_residual && can_fuse(m, chunk);
          if (pending && !fuse) {
            CS_HIP(cs_launch_residual_bricks(st, P, nactive, m->brick_list, m->wc, m->rhs, buf_of(m, cur),
                                             m->brick_partials, m->dstate, 1u, prev.iters, prev.launches, tol));
          }
          uint32_t launches = 0;
          int rc = enqueue_sweeps(m, u, nactive, chunk, &launches, &cur, fuse ? &prev : nullptr);
          if (rc) return rc;
          pending = true;
          prev.iters = chunk;
          prev.launches = launches;
          enq += chunk;
        }
      }
      if (m->timing) CS_HIP(hipEventRecord(m->chunk_ev[e + 1], st));
      batch_iter_end.push_back(enq);
      CS_HIP(hipMemcpyAsync(&m->h_done[batch & 1u], &m->dstate->done, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
      CS_HIP(hipEventRecord(m->batch_ev[batch & 1u], st));
      if (batch > 0) {
        CS_HIP(hipEventSynchronize(m->batch_ev[(batch - 1) & 1u]));
        if (m->timing) harvest(batch);
        if (m->h_done[(batch - 1) & 1u]) stop = true;
      }
      ++batch;
    }
    // the last chunk's test (a no-op once the state says done)
    if (pending)
      CS_HIP(cs_launch_residual_bricks(st, P, nactive, m->brick_list, m->wc, m->rhs, buf_of(m, cur),
                                       m->brick_partials, m->dstate, 1u, prev.iters, prev.launches, tol));
    CS_HIP(hipMemcpyAsync(m->h_state, m->dstate, sizeof(cs_solve_state), hipMemcpyDeviceToHost, st));
    CS_HIP(hipStreamSynchronize(st));
    const cs_solve_state& hs = *m->h_state;
    // the untested iterations before the rescale ran unless the start was within the tolerance
    const bool pre_ran = pre > 0 && !(hs.initial_residual <= tol);
    const uint32_t total_iters = hs.iterations + (pre_ran ? pre : 0u);
    if (m->timing) {
      // batches after the converged one returned at once; count the executed ones
      uint32_t executed = 0;
      while (executed < batch && (executed == 0 || batch_iter_end[executed - 1] < total_iters)) ++executed;
      if (executed > harvested) harvest(executed);
    }
    S.iterations = total_iters;
    S.initial_residual = hs.initial_residual;
    S.final_residual = hs.residual;
    S.converged = hs.done;
    // the solved pressure is in pres2 when an odd number of launches ran
    CS_HIP(cs_launch_select_final(st, u, nactive, m->brick_list, m->pres, m->pres2, m->dstate, true));
    if (m->timing) {
      const uint32_t ran = hs.launches + (pre_ran ? pre_launches : 0u);
      m->kt.rbgs_launches += ran;
      m->kt.rbgs_cells_swept += (uint64_t)ran * nactive * CS_BRICK * CS_BRICK * CS_BRICK;
    }
  }
  mg_harvest(m);
  if (m->s.warm_start == CS_WARM_PREVIOUS) CS_HIP(cs_launch_copy_i32(st, m->pres, m->prev_p, m->nc));
  return CS_OK;
}

int solve(cs_sim* m, const cs_pblock& P) { return m->dense_solver ? solve_dense(m, P) : solve_bricks(m, P); }

// ---- one substep (step_frame's loop body, mod.rs:305-416) ------------------------------------
// What the host decides before a substep's first launch.
struct SubstepPlan {
  cs_dparams blk;   // the device block's host part
  bool epoch_wrap;  // the persistent solve's face buffers are cleared before the substep
};

// Host side of a substep: the emission (the flow model and the disk frame,
// inflow.rs:122-241; the particles are made on the device), the parameter block
// (write_uniforms, mod.rs:569-670) and the persistent solve's epoch base.
void plan_substep(cs_sim* m, float sub_dt, SubstepPlan* pl) {
  memset(pl, 0, sizeof(*pl));
  cs_dparams& b = pl->blk;
  // capacity counts the bed particles, which the reference's buffer holds first
  const uint32_t emitted = cs::emit_plan(m->s, m->inflow, sub_dt, m->num_water + m->num_bed, &b.ef, &m->total_dropped);
  b.ef.first = m->num_water;  // the water's own array
  b.ef.count = emitted;
  if (emitted) {
    m->num_water += emitted;
    m->total_emitted_mass += (float)emitted * (80.0f / 320.0f);
  }
  cs::make_params(m->s, m->num_water, m->num_bed_cells, sub_dt, &b.u);
  // epochs would leave the 31-bit range: forget every published word first
  pl->epoch_wrap = (uint64_t)m->persist_epoch + persist_span(m) >= 0x80000000ull;
  if (pl->epoch_wrap) m->persist_epoch = 1;
  b.epoch_base = m->persist_epoch;
  b.cap_bricks = 0xffffffffu;
  m->total_time += sub_dt;
}

// The block's host part into the next pinned ring slot, then one copy to the
// device block, in stream order.
int upload_block(cs_sim* m, const cs_dparams& b) {
  hipStream_t st = m->ctx->stream;
  if (m->pring_next == CS_PARAM_RING) {
    // every slot has been handed to a copy: let them all run before reuse
    CS_HIP(hipStreamSynchronize(st));
    m->pring_next = 0;
  }
  cs_dparams* h = m->h_pring + m->pring_next++;
  memcpy(h, &b, CS_DPARAMS_HOST_BYTES);
  CS_HIP(hipMemcpyAsync(m->dparams, h, CS_DPARAMS_HOST_BYTES, hipMemcpyHostToDevice, st));
  return CS_OK;
}

int clear_face_buffers(cs_sim* m) {
  CS_HIP(cs_launch_fill_i32(m->ctx->stream, reinterpret_cast<int32_t*>(m->persist_fb),
                            2 * (size_t)m->nbricks * CS_PERSIST_FACE_WORDS, 0));
  return CS_OK;
}

// Emission through classify (mod.rs:308-350); `emit_threads` covers the
// substep's emitted particles.
int substep_front(cs_sim* m, const cs_pblock& P, uint32_t emit_threads) {
  hipStream_t st = m->ctx->stream;
  if (emit_threads) CS_HIP(cs_launch_emit(st, P.d, emit_threads, m->part, m->aff));
  // clear (mod.rs:308-309) and metrics_clear (shader.rs:2282-2287)
  mark(m, P_P2G);
  CS_HIP(cs_launch_fill_i32_gated(st, P.d, m->lanes, m->nc * 6, 0));
  CS_HIP(cs_launch_fill_i32_gated(st, P.d, reinterpret_cast<int32_t*>(m->metrics), CS_METRICS_SLOTS, 0));
  // bed_lookup_clear / bed_lookup_scatter (mod.rs:322-331) and the bed cells' table
  if (m->num_bed)
    CS_HIP(cs_launch_bed_lookup(st, P, m->num_bed, m->bed_part, m->bed_aff, m->bed_ext, m->num_bed_cells,
                                m->bed_lookup, m->bed_counter, m->bed_id, m->bed));
  kmark(m, 0);
  if (m->p2g_tiled)
    CS_HIP(cs_launch_p2g_tiled(st, P, m->part, m->aff, m->lanes, m->metrics, m->tile_count, m->tile_start, m->tile_chunk,
                               m->p2g_key, m->p2g_sorted));
  else
    CS_HIP(cs_launch_p2g(st, P, m->part, m->aff, m->lanes, m->metrics));
  kmark(m, 1);
  mark(m, P_GRID);
  CS_HIP(cs_launch_grid(st, P, m->lanes, m->gv, m->sdf));
  mark(m, P_CLASSIFY);
  CS_HIP(cs_launch_classify(st, P, m->gv, m->cls, m->lanes, m->bed_id, m->bed, m->kind, m->kf, m->pres, m->rhs,
                            m->pack, m->tgt, m->metrics, m->sdf, m->brick_flag));
  mark(m, P_SOLVE);
  return CS_OK;
}

// project_pressure through bed_dynamics (mod.rs:359-409).
int substep_back(cs_sim* m, const cs_pblock& P, uint32_t bricks) {
  hipStream_t st = m->ctx->stream;
  mark(m, P_POST);
  // the bed particles' grains take the reaction of the water's velocity change
  // (bed_impulse_add, shader.rs:1484-1491 and 1870-1878)
  const cs_grains gr = m->num_bed ? cs_grains{m->bed_lookup, m->bed_aff, m->bed_dw, m->num_bed} : cs_grains{};
  CS_HIP(cs_launch_project(st, P, m->gv, m->kf, m->pres, m->bed_id, m->bed, m->sdf, gr));
  if (m->dense_solver)  // no brick list: the whole grid
    CS_HIP(cs_launch_post_div(st, P.u, m->gv, m->kf, m->cls, m->sdf, m->tgt, m->metrics));
  else
    CS_HIP(cs_launch_post_div_bricks(st, P, bricks, m->brick_list, m->gv, m->kf, m->cls, m->sdf, m->tgt,
                                     m->metrics));
  CS_HIP(cs_launch_pack(st, P, m->gv, m->kf, m->pack, m->sdf));
  CS_HIP(cs_launch_visc(st, P, m->gv, m->gv2, m->cls, m->sdf, m->alias4));
  mark(m, P_G2P);
  kmark(m, 2);
  if (m->p2g_tiled && m->g2p_tiled)  // the bins P2G made this substep
    CS_HIP(cs_launch_g2p_tiled(st, P, m->part, m->aff, m->gv2, m->bed_id, m->bed, m->sdf, gr, m->tile_count,
                               m->tile_start, m->tile_chunk, m->p2g_key, m->p2g_sorted));
  else
    CS_HIP(cs_launch_g2p(st, P, m->part, m->aff, m->gv2, m->bed_id, m->bed, m->sdf, gr));
  kmark(m, 3);
  // bed_coupling and extraction_advect (mod.rs:391-402), outside the G2P timing
  if (m->num_bed)
    CS_HIP(cs_launch_bed_coupling(st, P, m->part, m->aff, m->bed_lookup, m->num_bed, m->bed_ext, m->bed_dw));
  // bed_dynamics (mod.rs:404-409): the grains move under the water's drag,
  // the pressure gradient, packing contacts and the reaction impulses
  if (m->num_bed)
    CS_HIP(cs_launch_bed_dynamics(st, P, m->num_bed, m->bed_part, m->bed_aff, m->bed_ext, m->bed_dw, m->bed_snap,
                                  m->bed_lookup, m->gv2, m->pres, m->alias4, m->sdf));
  mark(m, P_END);
  return CS_OK;
}

int harvest_timing(cs_sim* m, const cs_params& u) {
  if (!m->timing) return CS_OK;
  CS_HIP(hipEventSynchronize(m->ev[P_END]));
  for (int p = 0; p < 6; ++p) {
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, m->ev[p], m->ev[p + 1]);
    m->pass_us[p] = ms * 1000.0f;
  }
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, m->ev[P_P2G], m->ev[P_END]);
  m->pass_us[6] = ms * 1000.0f;
  m->kt.p2g_us += kelapsed_us(m, 0, 1);
  m->kt.p2g_launches += 1;
  m->kt.p2g_particles += u.num_particles;
  m->kt.g2p_us += kelapsed_us(m, 2, 3);
  m->kt.g2p_launches += 1;
  m->kt.g2p_particles += u.num_particles;
  uint32_t slots[CS_METRICS_SLOTS];
  CS_HIP(hipMemcpy(slots, m->metrics, sizeof(slots), hipMemcpyDeviceToHost));
  m->kt.rbgs_iterations += m->stats.iterations;
  m->kt.rbgs_cell_updates += (uint64_t)m->stats.iterations * slots[CS_M_FLUID_CELLS];
  return CS_OK;
}

// A planned substep, launch by launch.  With `from_solve` the launches up to
// classify have already run (a recorded substep that halted there).
int run_substep(cs_sim* m, const SubstepPlan& pl, bool from_solve) {
  int rc = upload_block(m, pl.blk);
  if (rc) return rc;
  if (pl.epoch_wrap && !from_solve && (rc = clear_face_buffers(m))) return rc;
  const cs_pblock P{pl.blk.u, m->dparams};
  if (!from_solve && (rc = substep_front(m, P, pl.blk.ef.count))) return rc;
  if ((rc = solve(m, P))) return rc;
  if ((rc = substep_back(m, P, m->nactive))) return rc;
  return harvest_timing(m, P.u);
}

int substep(cs_sim* m, float sub_dt) {
  SubstepPlan pl;
  plan_substep(m, sub_dt, &pl);
  return run_substep(m, pl, false);
}

// ---- recorded substeps: one graph per simulation, replayed every substep -----------------------
// Active bricks a recorded substep's launches are sized for, or 0 when this
// simulation's substeps cannot be recorded: the persistent solve's limit and the
// multigrid visit's (its three-level visit is chosen when it fits this many).
uint32_t recorded_capacity(cs_sim* m) {
  // kernel timing brackets every pass with events: launch by launch
  if (!m->use_graphs || !m->substep_graphs || m->timing || m->dense_solver || !m->persist ||
      m->s.rbgs_rescale_after)
    return 0;
  if (m->s.warm_start != CS_WARM_ZERO && m->s.warm_start != CS_WARM_PREVIOUS) return 0;
  const bool fixed = m->s.rbgs_mode == CS_RBGS_FIXED_PAIRS;
  if (fixed && m->s.pressure_rbgs_pairs == 0) return 0;
  uint32_t cap = std::min<uint32_t>(CS_PERSIST_MAX_BRICKS, m->nbricks);
  if (!fixed && m->s.rbgs_mg_cycles && ensure_mg(m) == CS_OK && m->mg_rows3_shape) {
    // the three-level visit needs its top level's operator within CS_MG_P3_MAX
    // cells, which mg_start bounds by side^3 cells per active brick: give up the
    // last few bricks of capacity for it, not more
    const int lp = m->mg_wg_first - 1;
    const uint64_t side = std::max<uint64_t>(1u, (uint64_t)CS_BRICK >> std::min(lp, 30));
    const uint32_t fit = (uint32_t)(CS_MG_P3_MAX / (side * side * side));
    if (fit < cap && fit >= cap - cap / 8u) cap = fit;
  }
  if (m->substep_cap) cap = std::min(cap, m->substep_cap);
  if (cap == 0 || cs_persist_groups(cap) > m->persist_max_groups) return 0;
  return cap;
}

// Records the substep's launch sequence (emission through bed_dynamics, the
// solve as solve_recorded) unless the graph for this capacity and emission
// width exists.
int record_substep(cs_sim* m, const SubstepPlan& pl, uint32_t cap, uint32_t emit_threads) {
  if (m->sg_exec && m->sg_cap == cap && m->sg_emit == emit_threads) return CS_OK;
  const bool fixed = m->s.rbgs_mode == CS_RBGS_FIXED_PAIRS;
  if (!fixed && m->s.rbgs_mg_cycles) {
    int rc = ensure_mg(m);  // allocations stay outside the capture
    if (rc) return rc;
  }
  cs_params uc = pl.blk.u;
  uc.num_particles = m->s.max_particles;  // launches cover every slot; the kernels read the count
  const cs_pblock P{uc, m->dparams};
  hipStream_t st = m->ctx->stream;
  CS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int rc = substep_front(m, P, emit_threads);
  if (rc == CS_OK) rc = solve_recorded(m, P, cap);
  if (rc == CS_OK) rc = substep_back(m, P, cap);
  hipGraph_t g = nullptr;
  const hipError_t ee = hipStreamEndCapture(st, &g);
  if (rc != CS_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ee != hipSuccess) return hip_fail(ee, "hipStreamEndCapture (substep)");
  hipGraphExec_t ex = nullptr;
  const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (ie != hipSuccess) {
    (void)hipGraphDestroy(g);
    return hip_fail(ie, "hipGraphInstantiate (substep)");
  }
  if (m->sg_exec) (void)hipGraphExecDestroy(m->sg_exec);
  if (m->sg_graph) (void)hipGraphDestroy(m->sg_graph);
  m->sg_graph = g;
  m->sg_exec = ex;
  m->sg_cap = cap;
  m->sg_emit = emit_threads;
  return CS_OK;
}

// `n` substeps of sub_dt as replays of the recorded substep, at most
// CS_SOLVE_RECORDS between two waits.  A substep whose active bricks exceed the
// recorded capacity halts at its brick scan, and so does every later replay;
// the host then runs that substep from its solve on, and the rest of the
// batch, launch by launch.
int step_recorded(cs_sim* m, float sub_dt, uint32_t n, uint32_t cap) {
  hipStream_t st = m->ctx->stream;
  const bool fixed = m->s.rbgs_mode == CS_RBGS_FIXED_PAIRS;
  std::vector<SubstepPlan> plans;
  for (uint32_t i0 = 0; i0 < n; i0 += CS_SOLVE_RECORDS) {
    const uint32_t cnt = std::min<uint32_t>(n - i0, CS_SOLVE_RECORDS);
    plans.resize(cnt);
    uint32_t emit_max = 0;
    for (uint32_t k = 0; k < cnt; ++k) {
      plan_substep(m, sub_dt, &plans[k]);
      plans[k].blk.slot = k;
      plans[k].blk.cap_bricks = cap;
      m->persist_epoch += persist_span(m);  // the replay's solve takes its epochs now
      emit_max = std::max(emit_max, plans[k].blk.ef.count);
    }
    // emission widths in steps of 1024 threads, so that small changes of the
    // flow keep the recorded graph
    const uint32_t emit_threads = emit_max ? (emit_max + 1023u) / 1024u * 1024u : 0u;
    const uint32_t width = std::max(emit_threads, m->sg_exec && m->sg_cap == cap ? m->sg_emit : 0u);
    int rc = record_substep(m, plans[0], cap, width);
    if (rc) {
      // the runtime refused the recording: this simulation runs launch by launch
      m->substep_graphs = false;
      for (uint32_t k = 0; k < cnt; ++k) {
        plans[k].blk.cap_bricks = 0xffffffffu;
        if ((rc = run_substep(m, plans[k], false))) return rc;
      }
      continue;
    }
    for (uint32_t k = 0; k < cnt; ++k) {
      if ((rc = upload_block(m, plans[k].blk))) return rc;
      if (plans[k].epoch_wrap && (rc = clear_face_buffers(m))) return rc;
      CS_HIP(hipGraphLaunch(m->sg_exec, st));
    }
    // the replays wrote their records and the halt word straight into pinned host memory
    CS_HIP(hipStreamSynchronize(st));
    const uint32_t halted = *m->h_halt ? *m->h_halt - 1u : cnt;  // the halting substep's slot + 1
    for (uint32_t k = 0; k < halted; ++k)
      if (m->h_records[k].faults)
        return fail(CS_ERR_DEVICE, "persistent pressure solve: a workgroup gave up waiting for its neighbours");
    m->persist_solves += halted;
    ++m->recorded_batches;
    m->recorded_substeps += halted;
    if (halted > 0) {
      const cs_solve_state& hs = m->h_records[halted - 1];
      cs_solve_stats& S = m->stats;
      memset(&S, 0, sizeof(S));
      S.iterations = fixed ? m->s.pressure_rbgs_pairs : hs.iterations;
      S.initial_residual = fixed ? 0.0 : hs.initial_residual;
      S.final_residual = hs.residual;
      S.converged = hs.residual <= (double)m->s.rbgs_tolerance ? 1u : 0u;
      m->final_residual_pending = false;
      m->nactive = hs.active_bricks;
    }
    if (halted < cnt) {
      ++m->recorded_halts;
      CS_HIP(hipMemsetAsync(&m->dparams->halt, 0, sizeof(uint32_t), st));
      plans[halted].blk.cap_bricks = 0xffffffffu;
      if ((rc = run_substep(m, plans[halted], true))) return rc;
      for (uint32_t k = halted + 1; k < cnt; ++k) {
        plans[k].blk.cap_bricks = 0xffffffffu;
        if ((rc = run_substep(m, plans[k], false))) return rc;
      }
    }
  }
  return CS_OK;
}

}  // namespace

// ---------------------------------------------------------------------------------------
extern "C" {

const char* cs_last_error(void) { return g_error.c_str(); }

int cs_settings_preset(int preset, cs_settings* out) {
  if (!out) return fail(CS_ERR_INVALID, "null settings");
  int rc = cs::settings_preset(preset, out);
  if (rc) return fail(rc, "unknown preset");
  return CS_OK;
}

int cs_generate_sdf(const cs_settings* s, float* out) {
  if (!s || !out) return fail(CS_ERR_INVALID, "null argument");
  std::string why;
  if (cs::validate_settings(*s, &why)) return fail(CS_ERR_INVALID, why);
  cs::generate_sdf(*s, out);
  return CS_OK;
}

int cs_generate_sdf_class(const cs_settings* s, const float* sdf, uint8_t* out) {
  if (!s || !sdf || !out) return fail(CS_ERR_INVALID, "null argument");
  std::string why;
  if (cs::validate_settings(*s, &why)) return fail(CS_ERR_INVALID, why);
  cs::generate_sdf_class(*s, sdf, out);
  return CS_OK;
}

int cs_device_count(int* out) {
  if (!out) return fail(CS_ERR_INVALID, "null argument");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *out = n;
  return CS_OK;
}

int cs_ctx_create(int device, cs_ctx** out) {
  if (!out) return fail(CS_ERR_INVALID, "null output");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(CS_ERR_DEVICE, "no HIP device");
  if (device < 0 || device >= n) return fail(CS_ERR_INVALID, "device index out of range");
  CS_HIP(hipSetDevice(device));
  cs_ctx* c = new cs_ctx();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "hipStreamCreateWithFlags");
  }
  *out = c;
  return CS_OK;
}

int cs_ctx_destroy(cs_ctx* ctx) {
  if (!ctx) return CS_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return CS_OK;
}

int cs_sim_create(cs_ctx* ctx, const cs_settings* s, const cs_scene* scene, cs_sim** out) {
  if (!ctx || !s || !out) return fail(CS_ERR_INVALID, "null argument");
  *out = nullptr;
  std::string why;
  if (cs::validate_settings(*s, &why)) return fail(CS_ERR_INVALID, why);
  CS_HIP(hipSetDevice(ctx->device));
  cs_sim* m = new cs_sim();
  m->ctx = ctx;
  m->s = *s;
  m->gx = s->grid_dims[0];
  m->gy = s->grid_dims[1];
  m->gz = s->grid_dims[2];
  m->nc = (size_t)m->gx * m->gy * m->gz;
  const size_t nsdf = (size_t)s->sdf_res * s->sdf_res * s->sdf_res;
  // scene arrays (state.rs:178-182): uploaded when the scene supplies them,
  // otherwise generated on the device from the settings
  const float* sdf_src = scene ? scene->sdf : nullptr;
  const uint8_t* cls_src = scene ? scene->sdf_class : nullptr;
  std::vector<int32_t> bed_index;
  std::vector<cs_bed_cell> bed_table;
  if (s->has_bed_field && scene && scene->bed_lookup && scene->bed_records) {
    if (!cs::build_bed_table(*s, scene->bed_lookup, scene->bed_records, scene->num_bed_records, bed_index, bed_table)) {
      free_sim(m);
      return fail(CS_ERR_INVALID, "bed lookup index out of range");
    }
  } else {
    bed_index.assign(m->nc, -1);
  }
  m->num_bed_cells = (uint32_t)bed_table.size();
  if (scene && scene->num_bed_particles) {
    if (!scene->bed_particles || !scene->bed_affine || !scene->bed_extracts) {
      free_sim(m);
      return fail(CS_ERR_INVALID, "bed particles need bed_particles, bed_affine and bed_extracts");
    }
    if (s->has_bed_field) {
      free_sim(m);
      return fail(CS_ERR_INVALID, "a bed field and bed particles are exclusive");
    }
    if (scene->num_bed_particles > s->max_particles ||
        (scene->particles ? scene->num_particles : 0u) > s->max_particles - scene->num_bed_particles) {
      free_sim(m);
      return fail(CS_ERR_RANGE, "bed and water particles exceed max_particles");
    }
    m->num_bed = scene->num_bed_particles;
    // each bed particle stamps at most 27 cells
    m->num_bed_cells = (uint32_t)std::min<size_t>(m->nc, (size_t)27 * m->num_bed);
  }
  if (scene && scene->particles && scene->num_particles) {
    if (scene->num_particles > s->max_particles) {
      free_sim(m);
      return fail(CS_ERR_RANGE, "initial particles exceed max_particles");
    }
    m->init_count = scene->num_particles;
    m->init_particles.assign(scene->particles, scene->particles + (size_t)m->init_count * 8);
    if (scene->affine) m->init_affine.assign(scene->affine, scene->affine + (size_t)m->init_count * 12);
    else m->init_affine.assign((size_t)m->init_count * 12, 0.0f);
  }
  hipError_t e = hipSuccess;
  const size_t mp = s->max_particles;
#define ALLOC(ptr, n)                                         \
  if (e == hipSuccess) e = dalloc(&(ptr), (n));
  ALLOC(m->part, mp * 8);
  ALLOC(m->aff, mp * 12);
  ALLOC(m->lanes, m->nc * 6);
  ALLOC(m->gv, m->nc * 4);
  ALLOC(m->gv2, m->nc * 4);
  ALLOC(m->sdf, nsdf);
  ALLOC(m->cls, m->nc);
  ALLOC(m->bed_id, m->nc);
  ALLOC(m->bed, m->num_bed ? (size_t)m->num_bed_cells : bed_table.size());
  if (m->num_bed) {
    ALLOC(m->bed_part, (size_t)m->num_bed * 8);
    ALLOC(m->bed_aff, (size_t)m->num_bed * 12);
    ALLOC(m->bed_ext, (size_t)m->num_bed * 8);
    ALLOC(m->bed_lookup, m->nc);
    ALLOC(m->bed_counter, 1);
    ALLOC(m->bed_dw, (size_t)m->num_bed * 4);  // bed_delta: water lane, then impulse x/y/z
    ALLOC(m->bed_snap, (size_t)m->num_bed * 4);
    ALLOC(m->alias4, m->nc);
  }
  ALLOC(m->kind, m->nc);
  ALLOC(m->kf, m->nc);
  ALLOC(m->pres, m->nc);
  ALLOC(m->rhs, m->nc);
  ALLOC(m->pack, m->nc);
  ALLOC(m->tgt, m->nc);
  ALLOC(m->prev_p, m->nc);
  ALLOC(m->warm, m->nc);
  ALLOC(m->metrics, CS_METRICS_SLOTS);
  ALLOC(m->res_partials, 2 * CS_RESIDUAL_BLOCKS);
  ALLOC(m->res_out, 1);
  cs_params geom;
  cs::make_params(*s, 0, 0, 0.0f, &geom);
  ALLOC(m->pres2, m->nc);
  ALLOC(m->wc, m->nc * 4);
  ALLOC(m->group_partials, 2 * 1024);
  ALLOC(m->sweep_table, CS_SWEEP_TABLE_WORDS);
  ALLOC(m->tile_count, geom.nbricks);
  ALLOC(m->tile_start, geom.nbricks);
  ALLOC(m->tile_chunk, geom.nbricks + 1);
  ALLOC(m->p2g_key, mp);
  ALLOC(m->p2g_sorted, mp);
  ALLOC(m->brick_flag, geom.nbricks);
  ALLOC(m->brick_list, geom.nbricks);
  ALLOC(m->brick_partials, 2 * (size_t)geom.nbricks);
  ALLOC(m->dstate, 1);
  ALLOC(m->persist_fb, (size_t)geom.nbricks * CS_PERSIST_FACE_WORDS);
  ALLOC(m->persist_slot, geom.nbricks);
  ALLOC(m->persist_xcc, geom.nbricks);
  ALLOC(m->persist_pflags, (size_t)geom.nbricks * CS_PERSIST_FLAG_STRIDE);
  ALLOC(m->persist_partials, 4 * (size_t)geom.nbricks);
  ALLOC(m->dparams, 1);
#undef ALLOC
  m->nbricks = geom.nbricks;
  // solver variant (benchmarking and cross-checks; both give identical results)
  // a named layout of the launch-per-chunk solve turns the persistent solve off
  for (const char* k : {"CS_SOLVER", "CS_RBGS_HALO", "CS_RBGS_MULTI", "CS_RBGS_FUSE", "CS_RBGS_GROUPS",
                        "CS_RBGS_THREADS", "CS_RBGS_BATCH"})
    if (getenv(k)) m->persist = false;
  if (const char* v = getenv("CS_RBGS_PERSIST")) m->persist = atoi(v) != 0;
  if (const char* v = getenv("CS_PERSIST_POLL_DELAY")) m->persist_poll_delay = (uint32_t)std::min(std::max(atoi(v), 0), 256);
  if (m->persist && cs_persist_max_groups(m->ctx->device, &m->persist_max_groups) != hipSuccess)
    m->persist = false;
  if (const char* v = getenv("CS_SOLVER")) m->dense_solver = strcmp(v, "dense") == 0;
  if (const char* v = getenv("CS_RBGS_HALO")) {
    int h = atoi(v);
    if (h == 1 || h == 2 || h == 4) m->halo = h;
  }
  if (const char* v = getenv("CS_RBGS_MULTI")) m->multi = atoi(v) != 0;
  if (const char* v = getenv("CS_RBGS_FUSE")) m->fuse_residual = atoi(v) != 0;
  if (const char* v = getenv("CS_GRAPHS")) m->use_graphs = atoi(v) != 0;
  if (const char* v = getenv("CS_SUBSTEP_GRAPHS")) m->substep_graphs = atoi(v) != 0;
  if (const char* v = getenv("CS_SUBSTEP_CAP")) m->substep_cap = (uint32_t)std::max(atoi(v), 0);
  if (const char* v = getenv("CS_P2G")) m->p2g_tiled = strcmp(v, "atomic") != 0;
  if (const char* v = getenv("CS_G2P")) m->g2p_tiled = strcmp(v, "global") != 0;
  if (const char* v = getenv("CS_MG_ROWS")) m->mg_rows_allowed = atoi(v);
  if (const char* v = getenv("CS_MG_BRICKS")) m->mg_brick_major = atoi(v) != 0;
  {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, m->ctx->device) == hipSuccess && cus > 0)
      m->multi_groups = (uint32_t)cus;
  }
  if (const char* v = getenv("CS_RBGS_GROUPS")) {
    int g = atoi(v);
    if (g >= 1 && g <= 1024) m->multi_groups = (uint32_t)g;
  }
  if (const char* v = getenv("CS_RBGS_THREADS")) {
    int t = atoi(v);
    if (t == 256 || t == 512) m->threads = t;
  }
  if (m->halo == 1) m->threads = 256;
  if (const char* v = getenv("CS_RBGS_BATCH")) {
    int b = atoi(v);
    if (b >= 1 && b <= 64) m->batch_chunks = (uint32_t)b;
  }
  m->chunk_ev.assign((size_t)m->batch_chunks * 3 * 2, nullptr);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&m->h_res), sizeof(double));
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&m->h_state), sizeof(cs_solve_state));
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&m->h_done), 2 * sizeof(uint32_t));
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&m->h_pring), CS_PARAM_RING * sizeof(cs_dparams));
  if (e == hipSuccess)
    e = hipHostMalloc(reinterpret_cast<void**>(&m->h_records), CS_SOLVE_RECORDS * sizeof(cs_solve_state));
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&m->h_halt), sizeof(uint32_t));
  for (auto& ev : m->ev)
    if (e == hipSuccess) e = hipEventCreate(&ev);
  for (auto& ev : m->kev)
    if (e == hipSuccess) e = hipEventCreate(&ev);
  for (auto& ev : m->batch_ev)
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
  for (auto& ev : m->chunk_ev)
    if (e == hipSuccess) e = hipEventCreate(&ev);
  if (e != hipSuccess) {
    free_sim(m);
    return fail(e == hipErrorOutOfMemory ? CS_ERR_NOMEM : CS_ERR_DEVICE, std::string("allocation: ") + hipGetErrorString(e));
  }
  hipStream_t st = ctx->stream;
  // no surface word is current before the first persistent solve
  e = hipMemsetAsync(m->persist_fb, 0, (size_t)geom.nbricks * CS_PERSIST_FACE_WORDS * sizeof(uint64_t), st);
  // the device-written words of the parameter block start clear (no halt)
  if (e == hipSuccess) e = hipMemsetAsync(m->dparams, 0, sizeof(cs_dparams), st);
  std::vector<uint32_t> sweep_table(CS_SWEEP_TABLE_WORDS);
  cs_build_sweep_table(sweep_table.data());
  if (e == hipSuccess)
    e = hipMemcpyAsync(m->sweep_table, sweep_table.data(), sweep_table.size() * sizeof(uint32_t),
                       hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = sdf_src ? hipMemcpyAsync(m->sdf, sdf_src, nsdf * sizeof(float), hipMemcpyHostToDevice, st)
                : cs_launch_gen_sdf(st, *s, m->sdf);
  if (e == hipSuccess)
    e = cls_src ? hipMemcpyAsync(m->cls, cls_src, m->nc, hipMemcpyHostToDevice, st)
                : cs_launch_gen_sdf_class(st, *s, m->sdf, m->cls);
  if (e == hipSuccess) e = hipMemcpyAsync(m->bed_id, bed_index.data(), m->nc * sizeof(int32_t), hipMemcpyHostToDevice, st);
  if (e == hipSuccess && !bed_table.empty())
    e = hipMemcpyAsync(m->bed, bed_table.data(), bed_table.size() * sizeof(cs_bed_cell), hipMemcpyHostToDevice, st);
  if (e == hipSuccess && m->num_bed) {
    e = hipMemcpyAsync(m->bed_part, scene->bed_particles, (size_t)m->num_bed * 32, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(m->bed_aff, scene->bed_affine, (size_t)m->num_bed * 48, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(m->bed_ext, scene->bed_extracts, (size_t)m->num_bed * 32, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(m->bed_dw, 0, (size_t)m->num_bed * 4 * sizeof(int32_t), st);
    m->init_bed_ext.assign(scene->bed_extracts, scene->bed_extracts + (size_t)m->num_bed * 8);
    m->init_bed_part.assign(scene->bed_particles, scene->bed_particles + (size_t)m->num_bed * 8);
    m->init_bed_aff.assign(scene->bed_affine, scene->bed_affine + (size_t)m->num_bed * 12);
  }
  if (e == hipSuccess) e = hipMemsetAsync(m->gv, 0, m->nc * 16, st);
  if (e == hipSuccess) e = hipMemsetAsync(m->gv2, 0, m->nc * 16, st);
  if (e == hipSuccess) e = hipMemsetAsync(m->pres, 0, m->nc * 4, st);
  if (e == hipSuccess) e = hipMemsetAsync(m->pres2, 0, m->nc * 4, st);
  if (e == hipSuccess) e = hipMemsetAsync(m->kind, 0, m->nc, st);
  if (e == hipSuccess) e = hipMemsetAsync(m->lanes, 0, m->nc * 24, st);
  if (e == hipSuccess) e = hipMemsetAsync(m->metrics, 0, CS_METRICS_SLOTS * 4, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    free_sim(m);
    return hip_fail(e, "scene upload");
  }
  int rc = cs_sim_reset(m);
  if (rc) {
    free_sim(m);
    return rc;
  }
  *out = m;
  return CS_OK;
}

int cs_sim_destroy(cs_sim* m) {
  if (!m) return CS_OK;
  (void)hipSetDevice(m->ctx->device);
  (void)hipStreamSynchronize(m->ctx->stream);
  free_sim(m);
  return CS_OK;
}

int cs_sim_reset(cs_sim* m) {
  if (!m) return fail(CS_ERR_INVALID, "null sim");
  m->num_water = m->init_count;
  m->total_time = 0.0f;
  m->total_emitted_mass = 0.0f;
  m->total_dropped = 0;
  m->inflow = cs::Inflow();
  m->inflow.kettle_angle_deg = m->s.initial_kettle_angle_deg;
  m->has_warm = false;
  memset(&m->stats, 0, sizeof(m->stats));
  m->final_residual_pending = false;
  return upload_initial(m);
}

int cs_sim_step_frame(cs_sim* m, float dt) {
  if (!m) return fail(CS_ERR_INVALID, "null sim");
  if (!(dt >= 0.0f)) return fail(CS_ERR_INVALID, "dt must be a non-negative number");
  (void)hipSetDevice(m->ctx->device);
  dt = cs_min(dt, 1.0f / 30.0f);  // mod.rs:265
  const uint32_t n = std::max(m->s.substeps, 1u);
  const float sub_dt = dt / (float)n;
  // replays while the last substep's active bricks fit the recorded launches
  // (a scene beyond them would halt every replay at its brick scan)
  if (const uint32_t cap = recorded_capacity(m))
    if (m->nactive <= cap) return step_recorded(m, sub_dt, n, cap);
  for (uint32_t i = 0; i < n; ++i) {
    int rc = substep(m, sub_dt);
    if (rc) return rc;
  }
  return CS_OK;
}

int cs_sim_get_recorded_counts(cs_sim* m, uint64_t* replayed, uint64_t* halted) {
  if (!m || !replayed || !halted) return fail(CS_ERR_INVALID, "null argument");
  *replayed = m->recorded_substeps;
  *halted = m->recorded_halts;
  return CS_OK;
}

int cs_sim_synchronize(cs_sim* m) {
  if (!m) return fail(CS_ERR_INVALID, "null sim");
  CS_HIP(hipStreamSynchronize(m->ctx->stream));
  return CS_OK;
}

int cs_sim_set_kettle_angle(cs_sim* m, float a) {
  if (!m) return fail(CS_ERR_INVALID, "null sim");
  m->inflow.kettle_angle_deg = a;
  return CS_OK;
}
int cs_sim_set_spout_position(cs_sim* m, float x, float y, float z) {
  if (!m) return fail(CS_ERR_INVALID, "null sim");
  cs::spout_translate_origin_to(&m->s.spout, x, y, z);
  return CS_OK;
}
int cs_sim_set_spout_target(cs_sim* m, float x, float y, float z) {
  if (!m) return fail(CS_ERR_INVALID, "null sim");
  cs::spout_aim_at(&m->s.spout, x, y, z);
  return CS_OK;
}
int cs_sim_get_counts(cs_sim* m, uint32_t* w, uint32_t* b, uint32_t* mp) {
  if (!m) return fail(CS_ERR_INVALID, "null sim");
  if (w) *w = m->num_water;
  if (b) *b = m->num_bed;
  if (mp) *mp = m->s.max_particles;
  return CS_OK;
}
int cs_sim_get_time(cs_sim* m, float* t, float* mass, uint32_t* dropped) {
  if (!m) return fail(CS_ERR_INVALID, "null sim");
  if (t) *t = m->total_time;
  if (mass) *mass = m->total_emitted_mass;
  if (dropped) *dropped = m->total_dropped;
  return CS_OK;
}
int cs_sim_get_inflow(cs_sim* m, float* a, float* f, float* e) {
  if (!m) return fail(CS_ERR_INVALID, "null sim");
  if (a) *a = m->inflow.kettle_angle_deg;
  if (f) *f = m->inflow.flow_rate;
  if (e) *e = m->inflow.exit_speed;
  return CS_OK;
}

int cs_sim_write_particles(cs_sim* m, uint32_t off, uint32_t count, const float* p8, const float* a12) {
  if (!m) return fail(CS_ERR_INVALID, "null sim");
  if ((uint64_t)off + count > m->s.max_particles) return fail(CS_ERR_RANGE, "write past max_particles");
  if (count == 0) return CS_OK;
  hipStream_t st = m->ctx->stream;
  if (p8) CS_HIP(hipMemcpyAsync(m->part + (size_t)off * 8, p8, (size_t)count * 32, hipMemcpyHostToDevice, st));
  if (a12) CS_HIP(hipMemcpyAsync(m->aff + (size_t)off * 12, a12, (size_t)count * 48, hipMemcpyHostToDevice, st));
  CS_HIP(hipStreamSynchronize(st));
  return CS_OK;
}
int cs_sim_set_particle_count(cs_sim* m, uint32_t n) {
  if (!m) return fail(CS_ERR_INVALID, "null sim");
  // the reference's buffer holds the bed first, then the water (mod.rs:239-249),
  // so the water may use what the bed leaves of max_particles
  if (n > m->s.max_particles - m->num_bed) return fail(CS_ERR_RANGE, "count past max_particles less the bed");
  m->num_water = n;
  return CS_OK;
}

static int read_dev(cs_sim* m, void* dst, const void* src, size_t bytes) {
  if (!dst) return fail(CS_ERR_INVALID, "null destination");
  if (bytes == 0) return CS_OK;
  hipStream_t st = m->ctx->stream;
  CS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
  CS_HIP(hipStreamSynchronize(st));
  return CS_OK;
}

int cs_sim_read_particles(cs_sim* m, uint32_t off, uint32_t count, float* p8) {
  if (!m) return fail(CS_ERR_INVALID, "null sim");
  if ((uint64_t)off + count > m->s.max_particles) return fail(CS_ERR_RANGE, "read past max_particles");
  return read_dev(m, p8, m->part + (size_t)off * 8, (size_t)count * 32);
}
int cs_sim_read_affine(cs_sim* m, uint32_t off, uint32_t count, float* a12) {
  if (!m) return fail(CS_ERR_INVALID, "null sim");
  if ((uint64_t)off + count > m->s.max_particles) return fail(CS_ERR_RANGE, "read past max_particles");
  return read_dev(m, a12, m->aff + (size_t)off * 12, (size_t)count * 48);
}
int cs_sim_read_p2g_lanes(cs_sim* m, int32_t* l6) {
  if (!m) return fail(CS_ERR_INVALID, "null sim");
  return read_dev(m, l6, m->lanes, m->nc * 24);
}
int cs_sim_read_grid_vel(cs_sim* m, float* v4) {
  if (!m) return fail(CS_ERR_INVALID, "null sim");
  return read_dev(m, v4, m->gv2, m->nc * 16);
}
int cs_sim_read_pressure(cs_sim* m, int32_t* p, int32_t* kind) {
  if (!m) return fail(CS_ERR_INVALID, "null sim");
  int rc = read_dev(m, p, m->pres, m->nc * 4);
  if (rc || !kind) return rc;
  std::vector<int8_t> k(m->nc);
  if ((rc = read_dev(m, k.data(), m->kind, m->nc))) return rc;
  for (size_t i = 0; i < m->nc; ++i) kind[i] = k[i];
  return CS_OK;
}
int cs_sim_read_sdf(cs_sim* m, float* sdf, uint8_t* sdf_class) {
  if (!m) return fail(CS_ERR_INVALID, "null sim");
  int rc = CS_OK;
  if (sdf) rc = read_dev(m, sdf, m->sdf, (size_t)m->s.sdf_res * m->s.sdf_res * m->s.sdf_res * sizeof(float));
  if (rc == CS_OK && sdf_class) rc = read_dev(m, sdf_class, m->cls, m->nc);
  return rc;
}
int cs_sim_read_bed_lookup(cs_sim* m, int32_t* lookup) {
  if (!m || !lookup) return fail(CS_ERR_INVALID, "null argument");
  if (!m->num_bed) {
    for (size_t i = 0; i < m->nc; ++i) lookup[i] = -1;
    return CS_OK;
  }
  return read_dev(m, lookup, m->bed_lookup, m->nc * sizeof(int32_t));
}
int cs_sim_read_bed_extracts(cs_sim* m, float* extracts) {
  if (!m || !extracts) return fail(CS_ERR_INVALID, "null argument");
  if (!m->num_bed) return CS_OK;
  return read_dev(m, extracts, m->bed_ext, (size_t)m->num_bed * 32);
}
int cs_sim_read_bed_particles(cs_sim* m, float* particles, float* affine) {
  if (!m) return fail(CS_ERR_INVALID, "null sim");
  if (!m->num_bed) return CS_OK;
  if (particles) {
    const int rc = read_dev(m, particles, m->bed_part, (size_t)m->num_bed * 32);
    if (rc) return rc;
  }
  if (affine) return read_dev(m, affine, m->bed_aff, (size_t)m->num_bed * 48);
  return CS_OK;
}
int cs_sim_read_metrics(cs_sim* m, uint32_t* slots) {
  if (!m) return fail(CS_ERR_INVALID, "null sim");
  return read_dev(m, slots, m->metrics, CS_METRICS_SLOTS * 4);
}
int cs_sim_pack_render(cs_sim* m, float* out) {
  if (!m || !out) return fail(CS_ERR_INVALID, "null argument");
  (void)hipSetDevice(m->ctx->device);
  if (!m->render) CS_HIP(dalloc(&m->render, (size_t)m->s.max_particles * 8));
  cs_params u;
  // the reference's order: bed particles first, then the water (mod.rs:239-249)
  if (m->num_bed) {
    cs::make_params(m->s, m->num_bed, m->num_bed_cells, 1.0e-4f, &u);  // dt does not enter the packing
    CS_HIP(cs_launch_pack_render(m->ctx->stream, u, m->bed_part, m->bed_aff, m->num_bed, m->bed_ext, m->render));
  }
  cs::make_params(m->s, m->num_water, m->num_bed_cells, 1.0e-4f, &u);
  CS_HIP(cs_launch_pack_render(m->ctx->stream, u, m->part, m->aff, 0u, nullptr,
                               m->render + (size_t)m->num_bed * 8));
  return read_dev(m, out, m->render, ((size_t)m->num_bed + m->num_water) * 32);
}

int cs_sim_get_solve_stats(cs_sim* m, cs_solve_stats* out) {
  if (!m || !out) return fail(CS_ERR_INVALID, "null argument");
  if (m->final_residual_pending) {
    CS_HIP(hipStreamSynchronize(m->ctx->stream));
    m->stats.final_residual = m->pending_from_state ? m->h_state->residual : m->h_res[0];
    m->stats.converged = m->stats.final_residual <= (double)m->s.rbgs_tolerance ? 1u : 0u;
    m->final_residual_pending = false;
    m->pending_from_state = false;
  }
  uint32_t slots[CS_METRICS_SLOTS];
  int rc = cs_sim_read_metrics(m, slots);
  if (rc) return rc;
  m->stats.fluid_cells = slots[CS_M_FLUID_CELLS];
  *out = m->stats;
  return CS_OK;
}

int cs_sim_set_warm_start(cs_sim* m, const float* p) {
  if (!m || !p) return fail(CS_ERR_INVALID, "null argument");
  hipStream_t st = m->ctx->stream;
  CS_HIP(hipMemcpyAsync(m->warm, p, m->nc * sizeof(float), hipMemcpyHostToDevice, st));
  CS_HIP(hipStreamSynchronize(st));
  m->has_warm = true;
  return CS_OK;
}

// Checkpoint layout: header, then water particles, water affine, previous
// pressure, then (for a bed of particles) the bed particles, their affine
// state and BedExtracts, which bed_coupling / extraction_advect / bed_dynamics
// change every substep.
struct SaveHeader {
  uint32_t magic, version, num_water, max_particles;
  uint64_t nc;
  cs::Inflow inflow;
  float total_time, total_emitted_mass;
  uint32_t total_dropped, num_bed;
};
static const uint32_t kMagic = 0x43534d31u;  // "CSM1"
static const uint32_t kSaveVersion = 2u;

static size_t save_size(const cs_sim* m, uint32_t num_water) {
  return sizeof(SaveHeader) + (size_t)num_water * 80 + m->nc * 4 + (size_t)m->num_bed * (32 + 48 + 32);
}

int cs_sim_save(cs_sim* m, void* buf, size_t cap, size_t* written) {
  if (!m || !written) return fail(CS_ERR_INVALID, "null argument");
  size_t need = save_size(m, m->num_water);
  *written = need;
  if (!buf) return CS_OK;  // size query
  if (cap < need) return fail(CS_ERR_RANGE, "checkpoint buffer too small");
  SaveHeader h{};
  h.magic = kMagic;
  h.version = kSaveVersion;
  h.num_water = m->num_water;
  h.max_particles = m->s.max_particles;
  h.nc = m->nc;
  h.inflow = m->inflow;
  h.total_time = m->total_time;
  h.total_emitted_mass = m->total_emitted_mass;
  h.total_dropped = m->total_dropped;
  h.num_bed = m->num_bed;
  char* p = static_cast<char*>(buf);
  memcpy(p, &h, sizeof(h));
  p += sizeof(h);
  int rc = read_dev(m, p, m->part, (size_t)m->num_water * 32);
  if (rc) return rc;
  p += (size_t)m->num_water * 32;
  if ((rc = read_dev(m, p, m->aff, (size_t)m->num_water * 48))) return rc;
  p += (size_t)m->num_water * 48;
  if ((rc = read_dev(m, p, m->prev_p, m->nc * 4))) return rc;
  p += m->nc * 4;
  if (m->num_bed) {
    if ((rc = read_dev(m, p, m->bed_part, (size_t)m->num_bed * 32))) return rc;
    p += (size_t)m->num_bed * 32;
    if ((rc = read_dev(m, p, m->bed_aff, (size_t)m->num_bed * 48))) return rc;
    p += (size_t)m->num_bed * 48;
    if ((rc = read_dev(m, p, m->bed_ext, (size_t)m->num_bed * 32))) return rc;
  }
  return CS_OK;
}

int cs_sim_load(cs_sim* m, const void* buf, size_t size) {
  if (!m || !buf) return fail(CS_ERR_INVALID, "null argument");
  if (size < sizeof(SaveHeader)) return fail(CS_ERR_INVALID, "checkpoint truncated");
  SaveHeader h;
  memcpy(&h, buf, sizeof(h));
  if (h.magic != kMagic || h.version != kSaveVersion) return fail(CS_ERR_INVALID, "not a checkpoint of this version");
  if (h.nc != m->nc || h.max_particles != m->s.max_particles || h.num_bed != m->num_bed)
    return fail(CS_ERR_INVALID, "checkpoint does not match this simulation's settings or bed");
  if (h.num_water > m->s.max_particles - m->num_bed) return fail(CS_ERR_RANGE, "checkpoint water count past capacity");
  if (size < save_size(m, h.num_water)) return fail(CS_ERR_INVALID, "checkpoint truncated");
  const char* p = static_cast<const char*>(buf) + sizeof(h);
  hipStream_t st = m->ctx->stream;
  CS_HIP(hipMemsetAsync(m->part, 0, (size_t)m->s.max_particles * 32, st));
  CS_HIP(hipMemsetAsync(m->aff, 0, (size_t)m->s.max_particles * 48, st));
  CS_HIP(hipMemcpyAsync(m->part, p, (size_t)h.num_water * 32, hipMemcpyHostToDevice, st));
  p += (size_t)h.num_water * 32;
  CS_HIP(hipMemcpyAsync(m->aff, p, (size_t)h.num_water * 48, hipMemcpyHostToDevice, st));
  p += (size_t)h.num_water * 48;
  CS_HIP(hipMemcpyAsync(m->prev_p, p, m->nc * 4, hipMemcpyHostToDevice, st));
  p += m->nc * 4;
  if (m->num_bed) {
    CS_HIP(hipMemcpyAsync(m->bed_part, p, (size_t)m->num_bed * 32, hipMemcpyHostToDevice, st));
    p += (size_t)m->num_bed * 32;
    CS_HIP(hipMemcpyAsync(m->bed_aff, p, (size_t)m->num_bed * 48, hipMemcpyHostToDevice, st));
    p += (size_t)m->num_bed * 48;
    CS_HIP(hipMemcpyAsync(m->bed_ext, p, (size_t)m->num_bed * 32, hipMemcpyHostToDevice, st));
    // water credits and reaction impulses are consumed within a substep
    CS_HIP(cs_launch_fill_i32(st, m->bed_dw, (size_t)m->num_bed * 4, 0));
  }
  CS_HIP(hipStreamSynchronize(st));
  m->num_water = h.num_water;
  m->inflow = h.inflow;
  m->total_time = h.total_time;
  m->total_emitted_mass = h.total_emitted_mass;
  m->total_dropped = h.total_dropped;
  return CS_OK;
}

int cs_sim_enable_timing(cs_sim* m, int enable) {
  if (!m) return fail(CS_ERR_INVALID, "null sim");
  m->timing = enable != 0;
  return CS_OK;
}
int cs_sim_get_pass_times(cs_sim* m, float* us7) {
  if (!m || !us7) return fail(CS_ERR_INVALID, "null argument");
  memcpy(us7, m->pass_us, sizeof(m->pass_us));
  return CS_OK;
}

int cs_sim_get_kernel_timing(cs_sim* m, cs_kernel_timing* out) {
  if (!m || !out) return fail(CS_ERR_INVALID, "null argument");
  *out = m->kt;
  return CS_OK;
}
int cs_sim_reset_kernel_timing(cs_sim* m) {
  if (!m) return fail(CS_ERR_INVALID, "null sim");
  m->kt = cs_kernel_timing{};
  return CS_OK;
}

int cs_device_expf(cs_ctx* ctx, const float* in, float* out, uint32_t n) {
  if (!ctx || (!in && n) || (!out && n)) return fail(CS_ERR_INVALID, "null argument");
  if (n == 0) return CS_OK;
  CS_HIP(hipSetDevice(ctx->device));
  float *din = nullptr, *dout = nullptr;
  CS_HIP(dalloc(&din, n));
  hipError_t e = dalloc(&dout, n);
  if (e == hipSuccess) e = hipMemcpyAsync(din, in, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = cs_launch_expf(ctx->stream, din, dout, n);
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(din);
  if (dout) (void)hipFree(dout);
  if (e != hipSuccess) return hip_fail(e, "cs_device_expf");
  return CS_OK;
}
int cs_device_sincosf(cs_ctx* ctx, const float* in, float* sn, float* cs, uint32_t n) {
  if (!ctx || (n && (!in || !sn || !cs))) return fail(CS_ERR_INVALID, "null argument");
  if (n == 0) return CS_OK;
  CS_HIP(hipSetDevice(ctx->device));
  float* d = nullptr;
  CS_HIP(dalloc(&d, (size_t)n * 3));
  hipError_t e = hipMemcpyAsync(d, in, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = cs_launch_sincosf(ctx->stream, d, d + n, d + 2 * (size_t)n, n);
  if (e == hipSuccess) e = hipMemcpyAsync(sn, d + n, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(cs, d + 2 * (size_t)n, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return hip_fail(e, "cs_device_sincosf");
  return CS_OK;
}

}  // extern "C"

// ---- learned warm start: attachment and training captures ----------------------------------
int cs_sim_set_warm_net(cs_sim* m, cs_unet* net) {
  if (!m) return fail(CS_ERR_INVALID, "null simulation");
  if (net) {
    const int n = cs_unet_side(net);
    if ((uint32_t)n != m->gx || (uint32_t)n != m->gy || (uint32_t)n != m->gz)
      return fail(CS_ERR_INVALID, "the network's grid side differs from the simulation's grid");
    if (cs_unet_context(net) != m->ctx) return fail(CS_ERR_INVALID, "the network belongs to another context");
  }
  m->warm_net = net;
  return CS_OK;
}

int cs_unet_capture(cs_unet* net, cs_sim* m) {
  if (!net || !m) return fail(CS_ERR_INVALID, "null argument");
  const int n = cs_unet_side(net);
  if ((uint32_t)n != m->gx || (uint32_t)n != m->gy || (uint32_t)n != m->gz)
    return fail(CS_ERR_INVALID, "the network's grid side differs from the simulation's grid");
  if (cs_unet_context(net) != m->ctx) return fail(CS_ERR_INVALID, "the network belongs to another context");
  (void)hipSetDevice(m->ctx->device);
  cs_params u;
  cs::make_params(m->s, m->num_water, m->num_bed_cells, 0.0f, &u);
  float *f4 = nullptr, *target = nullptr;
  uint8_t* mask = nullptr;
  uint32_t* count = nullptr;
  int rc = cs_unet_new_sample(net, &f4, &target, &mask, &count);
  if (rc) return rc;
  CS_HIP(cs_launch_unet_features(m->ctx->stream, u, m->kf, m->rhs, m->pres, m->bed_id, m->bed, cs_unet_rhs_scale(net),
                                 1.0f / cs_unet_pressure_scale(net), f4, target, mask, count));
  return cs_unet_finish_sample(net);
}

int cs_sim_read_warm_start(cs_sim* m, float* out) {
  if (!m || !out) return fail(CS_ERR_INVALID, "null argument");
  (void)hipSetDevice(m->ctx->device);
  CS_HIP(hipMemcpyAsync(out, m->warm, m->nc * sizeof(float), hipMemcpyDeviceToHost, m->ctx->stream));
  CS_HIP(hipStreamSynchronize(m->ctx->stream));
  return CS_OK;
}
