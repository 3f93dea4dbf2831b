// cs_batch.cpp — many independent simulations advanced together, one launch
// per pass over every scene (the batched kernels of cs_kernels.hip,
// cs_p2g.hip and cs_solver.hip).
//
// A substep of the batch: the host runs each scene's flow model and writes
// its emission frame and parameter block into the descriptor table (one copy
// to the device; the begin pass generates the particles); then the fixed launch sequence up to the
// first residual test, the solve loop, and the fixed sequence after the
// solve.  The fixed sequences, and a steady-state group of solve chunks, keep
// the same arguments for the life of the batch (only the table's contents
// change), so they are captured as hipGraphs once and replayed every substep.
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "cs_batch.h"
#include "cs_sim_impl.h"

#define CS_BHIP(call)                                     \
  do {                                                    \
    hipError_t e_ = (call);                               \
    if (e_ != hipSuccess) return cs_hip_fail(e_, #call);  \
  } while (0)

struct cs_batch {
  cs_ctx* ctx = nullptr;
  std::vector<cs_sim*> sims;
  uint32_t n = 0;
  cs_batch_config cfg{};
  bool fixed = false;
  uint32_t every = 8, cap = 0, pairs = 0, substeps = 1;
  std::vector<uint32_t> p2g_tiled;
  // device
  cs_scene_dev* d_sc = nullptr;
  uint32_t* d_starts = nullptr;
  uint32_t* entries = nullptr;
  uint32_t* first = nullptr;
  uint32_t* count = nullptr;
  cs_batch_solve* bs = nullptr;
  double* partials = nullptr;
  cs_solve_state* d_states = nullptr;
  uint32_t* d_fluid = nullptr;
  // host
  cs_scene_dev* h_sc = nullptr;  // pinned
  std::vector<uint32_t> h_starts;
  cs_solve_state* h_states = nullptr;  // pinned [n]
  uint32_t* h_fluid = nullptr;         // pinned [n]
  cs_batch_solve* h_bs = nullptr;      // pinned [2]
  hipEvent_t poll_ev[2] = {};
  hipEvent_t upload_done = nullptr;
  hipEvent_t solve_ev[2] = {};  // around the solve loop
  hipGraph_t graph[3] = {};
  hipGraphExec_t exec[3] = {};  // before the solve loop, a group of chunks, after the solve
  bool captured = false;
  cs_batch_stats stats{};
};

namespace {

enum { G_PRE = 0, G_CHUNKS = 1, G_POST = 2 };

template <typename T>
hipError_t balloc(T** p, size_t count) {
  *p = nullptr;
  return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T));
}

void free_batch(cs_batch* b) {
  if (!b) return;
  for (int g = 0; g < 3; ++g) {
    if (b->exec[g]) (void)hipGraphExecDestroy(b->exec[g]);
    if (b->graph[g]) (void)hipGraphDestroy(b->graph[g]);
  }
  void* dev[] = {b->d_sc, b->d_starts, b->entries,  b->first,    b->count,
                 b->bs,   b->partials, b->d_states, b->d_fluid};
  for (void* p : dev)
    if (p) (void)hipFree(p);
  void* host[] = {b->h_sc, b->h_states, b->h_fluid, b->h_bs};
  for (void* p : host)
    if (p) (void)hipHostFree(p);
  for (auto& e : b->poll_ev)
    if (e) (void)hipEventDestroy(e);
  if (b->upload_done) (void)hipEventDestroy(b->upload_done);
  for (auto& e : b->solve_ev)
    if (e) (void)hipEventDestroy(e);
  delete b;
}

cs_batch_launch launch_of(const cs_batch* b) {
  return cs_batch_launch{b->ctx->stream, b->d_sc, b->d_starts, b->h_starts.data(), b->n};
}

// The static part of a scene's descriptor: buffer addresses and solve settings.
void fill_static(const cs_batch* b, uint32_t s, cs_scene_dev* d) {
  const cs_sim* m = b->sims[s];
  memset(d, 0, sizeof(*d));
  d->part = m->part;
  d->aff = m->aff;
  d->lanes = m->lanes;
  d->gv = m->gv;
  d->gv2 = m->gv2;
  d->sdf = m->sdf;
  d->cls = m->cls;
  d->bed_id = m->bed_id;
  d->bed = m->bed;
  d->kind = m->kind;
  d->kf = m->kf;
  d->pres = m->pres;
  d->pres2 = m->pres2;
  d->rhs = m->rhs;
  d->pack = m->pack;
  d->tgt = m->tgt;
  d->prev_p = m->prev_p;
  d->metrics = m->metrics;
  d->wc = m->wc;
  d->brick_flag = m->brick_flag;
  d->brick_list = m->brick_list;
  d->st = m->dstate;
  d->tile_count = m->tile_count;
  d->tile_start = m->tile_start;
  d->p2g_key = m->p2g_key;
  d->p2g_sorted = m->p2g_sorted;
  d->warm_previous = m->s.warm_start == CS_WARM_PREVIOUS ? 1u : 0u;
  d->p2g_tiled = b->p2g_tiled[s];
  d->tol = (double)m->s.rbgs_tolerance;
}

// Launches up to the solve loop: append + clears, P2G, grid, classify, solve
// set-up, and (tolerance mode) the initial residual of every scene followed by
// the first compaction of the scenes still solving.
int enqueue_pre(cs_batch* b) {
  const cs_batch_launch L = launch_of(b);
  CS_BHIP(cs_batch_launch_begin(L));
  CS_BHIP(cs_batch_launch_p2g(L));
  CS_BHIP(cs_batch_launch_grid_classify(L));
  CS_BHIP(cs_batch_launch_solve_setup(L));
  CS_BHIP(cs_batch_launch_compact(L, false, b->entries, b->first, b->count, b->bs));
  if (!b->fixed) {
    CS_BHIP(cs_batch_launch_residual(L, b->entries, b->first, b->count, b->bs, b->cfg.residual_groups, 0u,
                                     b->partials, 0u, 0u, 0u));
    CS_BHIP(cs_batch_launch_compact(L, true, b->entries, b->first, b->count, b->bs));
  }
  return CS_OK;
}

// One chunk of `iters` iterations (iters / 2 launches of four half sweeps)
// from buffer *cur, its residual test, and the compaction of the scenes that
// go on.
int enqueue_chunk(cs_batch* b, uint32_t iters, uint32_t* cur) {
  const cs_batch_launch L = launch_of(b);
  const uint32_t launches = iters / 2u;
  const uint32_t* table = b->sims[0]->sweep_table;
  for (uint32_t l = 0; l < launches; ++l) {
    CS_BHIP(cs_batch_launch_sweeps(L, b->entries, b->bs, b->cfg.sweep_groups, *cur, 0u, table));
    *cur ^= 1u;
  }
  CS_BHIP(cs_batch_launch_residual(L, b->entries, b->first, b->count, b->bs, b->cfg.residual_groups, *cur,
                                   b->partials, 1u, iters, launches));
  CS_BHIP(cs_batch_launch_compact(L, true, b->entries, b->first, b->count, b->bs));
  return CS_OK;
}

// Fixed mode: every scene sweeps pairs / 2 launches, then its residual.
int enqueue_fixed_solve(cs_batch* b, uint32_t* cur) {
  const cs_batch_launch L = launch_of(b);
  const uint32_t* table = b->sims[0]->sweep_table;
  for (uint32_t l = 0; l < b->pairs / 2u; ++l) {
    CS_BHIP(cs_batch_launch_sweeps(L, b->entries, b->bs, b->cfg.sweep_groups, *cur, 0u, table));
    *cur ^= 1u;
  }
  CS_BHIP(cs_batch_launch_residual(L, b->entries, b->first, b->count, b->bs, b->cfg.residual_groups, *cur,
                                   b->partials, 0u, 0u, 0u));
  return CS_OK;
}

int enqueue_post(cs_batch* b, uint32_t cur) {
  const cs_batch_launch L = launch_of(b);
  CS_BHIP(cs_batch_launch_finalize(L, !b->fixed, cur != 0u));
  CS_BHIP(cs_batch_launch_post(L));
  CS_BHIP(cs_batch_launch_gather_states(L, b->d_states, b->d_fluid));
  return CS_OK;
}

// Iterations enqueued by one replay of the chunk graph.
uint32_t group_iters(const cs_batch* b) { return b->cfg.check_batches * b->every; }

// Whether a group of chunks leaves the buffers where it found them (an even
// number of sweep launches), so one graph serves every replay.
bool chunk_graph_ok(const cs_batch* b) { return ((b->every / 2u) * b->cfg.check_batches) % 2u == 0u; }

template <typename F>
int capture(cs_batch* b, int which, F&& body) {
  hipStream_t st = b->ctx->stream;
  CS_BHIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const int rc = body();
  hipGraph_t g = nullptr;
  const hipError_t ee = hipStreamEndCapture(st, &g);
  if (rc != CS_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ee != hipSuccess) return cs_hip_fail(ee, "hipStreamEndCapture");
  b->graph[which] = g;
  CS_BHIP(hipGraphInstantiate(&b->exec[which], g, nullptr, nullptr, 0));
  return CS_OK;
}

int capture_all(cs_batch* b) {
  int rc = capture(b, G_PRE, [&] { return enqueue_pre(b); });
  if (rc) return rc;
  uint32_t cur = 0;
  if (b->fixed) {
    // the fixed solve is part of the sequence after the set-up
    rc = capture(b, G_POST, [&] {
      int r = enqueue_fixed_solve(b, &cur);
      return r ? r : enqueue_post(b, cur);
    });
    if (rc) return rc;
  } else {
    if (chunk_graph_ok(b)) {
      rc = capture(b, G_CHUNKS, [&] {
        uint32_t c = 0;
        for (uint32_t k = 0; k < b->cfg.check_batches; ++k) {
          int r = enqueue_chunk(b, b->every, &c);
          if (r) return r;
        }
        return CS_OK;
      });
      if (rc) return rc;
    }
    rc = capture(b, G_POST, [&] { return enqueue_post(b, 0u); });
    if (rc) return rc;
  }
  b->captured = true;
  return CS_OK;
}

// Emission frames and parameter blocks of every scene, uploaded with one copy
// (the particles themselves are generated by the begin pass).
int upload_substep(cs_batch* b, float sub_dt) {
  hipStream_t st = b->ctx->stream;
  CS_BHIP(hipEventSynchronize(b->upload_done));  // the pinned table is free again
  for (uint32_t s = 0; s < b->n; ++s) {
    cs_sim* m = b->sims[s];
    cs_scene_dev& d = b->h_sc[s];
    const uint32_t emitted = cs::emit_plan(m->s, m->inflow, sub_dt, m->num_water, &d.emit, &m->total_dropped);
    m->num_water += emitted;
    m->total_emitted_mass += (float)emitted * (80.0f / 320.0f);
    cs::make_params(m->s, m->num_water, m->num_bed_cells, sub_dt, &d.u);
    b->stats.particle_substeps += m->num_water;
  }
  CS_BHIP(hipMemcpyAsync(b->d_sc, b->h_sc, (size_t)b->n * sizeof(cs_scene_dev), hipMemcpyHostToDevice, st));
  CS_BHIP(hipEventRecord(b->upload_done, st));
  return CS_OK;
}

int batch_substep(cs_batch* b, float sub_dt) {
  hipStream_t st = b->ctx->stream;
  const auto t0 = std::chrono::steady_clock::now();
  int rc = upload_substep(b, sub_dt);
  if (rc) return rc;
  const bool graphs = b->cfg.use_graphs != 0u;
  if (graphs && !b->captured) {
    rc = capture_all(b);
    if (rc) return rc;
  }
  if (graphs) CS_BHIP(hipGraphLaunch(b->exec[G_PRE], st));
  else if ((rc = enqueue_pre(b))) return rc;
  CS_BHIP(hipEventRecord(b->solve_ev[0], st));
  uint32_t cur = 0;
  if (b->fixed) {
    if (graphs) CS_BHIP(hipGraphLaunch(b->exec[G_POST], st));
    else {
      if ((rc = enqueue_fixed_solve(b, &cur))) return rc;
      CS_BHIP(hipEventRecord(b->solve_ev[1], st));
      if ((rc = enqueue_post(b, cur))) return rc;
    }
    b->stats.sweep_launches += b->pairs / 2u;
    b->stats.residual_checks += 1u;
  } else {
    uint32_t enq = 0, batch = 0;
    bool stop = false;
    while (enq < b->cap && !stop) {
      if (graphs && b->exec[G_CHUNKS] && cur == 0u && b->cap - enq >= group_iters(b)) {
        CS_BHIP(hipGraphLaunch(b->exec[G_CHUNKS], st));
        enq += group_iters(b);
      } else {
        for (uint32_t c = 0; c < b->cfg.check_batches && enq < b->cap; ++c) {
          const uint32_t chunk = std::min(b->every, b->cap - enq);
          if ((rc = enqueue_chunk(b, chunk, &cur))) return rc;
          enq += chunk;
        }
      }
      CS_BHIP(hipMemcpyAsync(&b->h_bs[batch & 1u], b->bs, sizeof(cs_batch_solve), hipMemcpyDeviceToHost, st));
      CS_BHIP(hipEventRecord(b->poll_ev[batch & 1u], st));
      if (batch > 0) {
        CS_BHIP(hipEventSynchronize(b->poll_ev[(batch - 1) & 1u]));
        if (b->h_bs[(batch - 1) & 1u].live == 0u) stop = true;
      }
      ++batch;
    }
    CS_BHIP(hipEventRecord(b->solve_ev[1], st));
    // the post sequence was captured for buffers in place; finalize reads
    // each scene's launch parity from its state, so `cur` does not matter
    if (graphs) CS_BHIP(hipGraphLaunch(b->exec[G_POST], st));
    else if ((rc = enqueue_post(b, cur))) return rc;
  }
  CS_BHIP(hipMemcpyAsync(b->h_states, b->d_states, (size_t)b->n * sizeof(cs_solve_state), hipMemcpyDeviceToHost,
                         st));
  CS_BHIP(hipMemcpyAsync(b->h_fluid, b->d_fluid, (size_t)b->n * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  CS_BHIP(hipStreamSynchronize(st));
  if (!(b->fixed && graphs)) {  // the fixed solve sits inside the post graph, untimed
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, b->solve_ev[0], b->solve_ev[1]) == hipSuccess) b->stats.solve_us += ms * 1000.0;
  }
  uint32_t max_it = 0;
  for (uint32_t s = 0; s < b->n; ++s) {
    cs_sim* m = b->sims[s];
    const cs_solve_state& hs = b->h_states[s];
    cs_solve_stats& S = m->stats;
    memset(&S, 0, sizeof(S));
    if (b->fixed) {
      S.iterations = b->pairs;
      S.final_residual = hs.residual;
      S.converged = hs.residual <= (double)m->s.rbgs_tolerance ? 1u : 0u;
    } else {
      S.iterations = hs.iterations;
      S.initial_residual = hs.initial_residual;
      S.final_residual = hs.residual;
      S.converged = hs.done;
    }
    m->final_residual_pending = false;
    m->pending_from_state = false;
    m->total_time += sub_dt;
    max_it = std::max(max_it, S.iterations);
    b->stats.cell_updates += (uint64_t)S.iterations * b->h_fluid[s];
  }
  if (!b->fixed) {
    b->stats.sweep_launches += max_it / 2u;
    b->stats.residual_checks += 1u + (max_it + b->every - 1u) / b->every;
  }
  b->stats.max_iterations += max_it;
  b->stats.substeps += 1u;
  b->stats.substep_us += std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
  return CS_OK;
}

}  // namespace

extern "C" {

int cs_batch_default_config(cs_batch_config* out) {
  if (!out) return cs_fail(CS_ERR_INVALID, "null argument");
  memset(out, 0, sizeof(*out));
  out->use_graphs = 1u;
  return CS_OK;
}

int cs_batch_create(cs_ctx* ctx, cs_sim* const* sims, uint32_t count, const cs_batch_config* config,
                    cs_batch** out) {
  if (!ctx || !sims || !out) return cs_fail(CS_ERR_INVALID, "null argument");
  *out = nullptr;
  if (count == 0 || count > CS_BATCH_MAX_SCENES) return cs_fail(CS_ERR_INVALID, "a batch holds 1..1024 scenes");
  const cs_sim* m0 = sims[0];
  if (!m0) return cs_fail(CS_ERR_INVALID, "null simulation");
  const bool fixed = m0->s.rbgs_mode == CS_RBGS_FIXED_PAIRS;
  const uint32_t every = m0->s.rbgs_check_every ? m0->s.rbgs_check_every : 8u;
  for (uint32_t s = 0; s < count; ++s) {
    const cs_sim* m = sims[s];
    if (!m) return cs_fail(CS_ERR_INVALID, "null simulation");
    for (uint32_t k = 0; k < s; ++k)
      if (sims[k] == m) return cs_fail(CS_ERR_INVALID, "a simulation appears twice in the batch");
    if (m->ctx != ctx) return cs_fail(CS_ERR_INVALID, "every simulation of a batch must use the batch's context");
    if (m->dense_solver) return cs_fail(CS_ERR_INVALID, "batched scenes need the brick solver");
    if (m->s.warm_start != CS_WARM_ZERO && m->s.warm_start != CS_WARM_PREVIOUS)
      return cs_fail(CS_ERR_INVALID, "batched scenes start their solves from zero or the previous pressure");
    if (m->s.rbgs_rescale_after)
      return cs_fail(CS_ERR_INVALID, "batched scenes do not take rbgs_rescale_after");
    if (m->num_bed)
      return cs_fail(CS_ERR_INVALID, "batched scenes do not take bed particles");
    if ((m->s.rbgs_mode == CS_RBGS_FIXED_PAIRS) != fixed)
      return cs_fail(CS_ERR_INVALID, "every scene of a batch must use the same rbgs_mode");
    if (std::max(m->s.substeps, 1u) != std::max(m0->s.substeps, 1u))
      return cs_fail(CS_ERR_INVALID, "every scene of a batch must use the same substep count");
    if (m->nc == 0u) return cs_fail(CS_ERR_INVALID, "empty grid");
    const uint32_t nbx = (m->gx + CS_BRICK - 1) / CS_BRICK, nby = (m->gy + CS_BRICK - 1) / CS_BRICK,
                   nbz = (m->gz + CS_BRICK - 1) / CS_BRICK;
    if ((uint64_t)nbx * nby * nbz >= (1u << CS_BATCH_ENTRY_SCENE_SHIFT))
      return cs_fail(CS_ERR_INVALID, "a batched scene holds at most 65535 bricks");
    if (fixed) {
      if (m->s.pressure_rbgs_pairs != m0->s.pressure_rbgs_pairs || m->s.pressure_rbgs_pairs % 2u != 0u)
        return cs_fail(CS_ERR_INVALID, "batched fixed-mode scenes need one even pressure_rbgs_pairs");
    } else {
      const uint32_t e = m->s.rbgs_check_every ? m->s.rbgs_check_every : 8u;
      if (e != every || m->s.rbgs_max_iters != m0->s.rbgs_max_iters || every % 2u != 0u ||
          m->s.rbgs_max_iters % 2u != 0u)
        return cs_fail(CS_ERR_INVALID,
                       "batched scenes need one even rbgs_check_every and one even rbgs_max_iters");
    }
  }
  (void)hipSetDevice(ctx->device);
  cs_batch* b = new cs_batch();
  b->ctx = ctx;
  b->sims.assign(sims, sims + count);
  b->n = count;
  b->fixed = fixed;
  b->every = every;
  b->cap = m0->s.rbgs_max_iters;
  b->pairs = m0->s.pressure_rbgs_pairs;
  b->substeps = std::max(m0->s.substeps, 1u);
  b->p2g_tiled.resize(count);
  for (uint32_t s = 0; s < count; ++s) b->p2g_tiled[s] = sims[s]->p2g_tiled ? 1u : 0u;
  cs_batch_default_config(&b->cfg);
  if (config) b->cfg = *config;
  if (const char* v = getenv("CS_GRAPHS")) {  // as for a single scene: 0 submits every launch eagerly
    if (atoi(v) == 0) b->cfg.use_graphs = 0u;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || cus <= 0)
    cus = 256;
  if (b->cfg.sweep_groups == 0u) b->cfg.sweep_groups = (uint32_t)cus;
  if (b->cfg.residual_groups == 0u) b->cfg.residual_groups = 4u * (uint32_t)cus;
  if (b->cfg.check_batches == 0u) b->cfg.check_batches = 4u;
  if (b->cfg.sweep_groups > 4096u || b->cfg.residual_groups > 65536u || b->cfg.check_batches > 64u) {
    free_batch(b);
    return cs_fail(CS_ERR_INVALID, "batch configuration out of range");
  }
  // block-start tables of the batched spaces
  const uint32_t n1 = count + 1u;
  b->h_starts.assign((size_t)CS_SPACE_COUNT * n1, 0u);
  uint64_t tot[CS_SPACE_COUNT] = {};
  size_t total_bricks = 0;
  for (uint32_t s = 0; s < count; ++s) {
    const cs_sim* m = sims[s];
    const uint32_t nb = ((m->gx + CS_BRICK - 1) / CS_BRICK) * ((m->gy + CS_BRICK - 1) / CS_BRICK) *
                        ((m->gz + CS_BRICK - 1) / CS_BRICK);
    const uint64_t per[CS_SPACE_COUNT] = {(m->nc + 255u) / 256u, ((uint64_t)m->s.max_particles + 255u) / 256u, nb,
                                          1u};
    for (int k = 0; k < CS_SPACE_COUNT; ++k) {
      b->h_starts[(size_t)k * n1 + s] = (uint32_t)tot[k];
      tot[k] += per[k];
    }
    total_bricks += nb;
  }
  for (int k = 0; k < CS_SPACE_COUNT; ++k) {
    if (tot[k] > 0x7fffffffu) {
      free_batch(b);
      return cs_fail(CS_ERR_INVALID, "batch too large for one launch");
    }
    b->h_starts[(size_t)k * n1 + count] = (uint32_t)tot[k];
  }
  hipStream_t st = ctx->stream;
  hipError_t e = hipSuccess;
  auto ok = [&](hipError_t r) {
    if (e == hipSuccess) e = r;
  };
  ok(balloc(&b->d_sc, count));
  ok(balloc(&b->d_starts, b->h_starts.size()));
  ok(balloc(&b->entries, total_bricks));
  ok(balloc(&b->first, count));
  ok(balloc(&b->count, count));
  ok(balloc(&b->bs, 1));
  ok(balloc(&b->partials, 2 * total_bricks));
  ok(balloc(&b->d_states, count));
  ok(balloc(&b->d_fluid, count));
  ok(hipHostMalloc(reinterpret_cast<void**>(&b->h_fluid), (size_t)count * sizeof(uint32_t)));
  ok(hipEventCreate(&b->solve_ev[0]));
  ok(hipEventCreate(&b->solve_ev[1]));
  ok(hipHostMalloc(reinterpret_cast<void**>(&b->h_sc), (size_t)count * sizeof(cs_scene_dev)));
  ok(hipHostMalloc(reinterpret_cast<void**>(&b->h_states), (size_t)count * sizeof(cs_solve_state)));
  ok(hipHostMalloc(reinterpret_cast<void**>(&b->h_bs), 2 * sizeof(cs_batch_solve)));
  ok(hipEventCreateWithFlags(&b->poll_ev[0], hipEventDisableTiming));
  ok(hipEventCreateWithFlags(&b->poll_ev[1], hipEventDisableTiming));
  ok(hipEventCreateWithFlags(&b->upload_done, hipEventDisableTiming));
  if (e == hipSuccess) {
    for (uint32_t s = 0; s < count; ++s) fill_static(b, s, &b->h_sc[s]);
    ok(hipMemcpyAsync(b->d_starts, b->h_starts.data(), b->h_starts.size() * sizeof(uint32_t),
                      hipMemcpyHostToDevice, st));
    ok(hipEventRecord(b->upload_done, st));
    ok(hipStreamSynchronize(st));
  }
  if (e != hipSuccess) {
    free_batch(b);
    return cs_hip_fail(e, "cs_batch_create");
  }
  *out = b;
  return CS_OK;
}

int cs_batch_destroy(cs_batch* b) {
  if (!b) return CS_OK;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  free_batch(b);
  return CS_OK;
}

int cs_batch_set_scene_p2g(cs_batch* b, uint32_t scene, int tiled) {
  if (!b) return cs_fail(CS_ERR_INVALID, "null batch");
  if (scene >= b->n) return cs_fail(CS_ERR_INVALID, "scene index out of range");
  b->p2g_tiled[scene] = tiled ? 1u : 0u;
  // the descriptor table is written whole before every substep
  b->h_sc[scene].p2g_tiled = b->p2g_tiled[scene];
  return CS_OK;
}

int cs_batch_step_frame(cs_batch* b, float dt) {
  if (!b) return cs_fail(CS_ERR_INVALID, "null batch");
  if (!(dt >= 0.0f)) return cs_fail(CS_ERR_INVALID, "dt must be a non-negative number");
  (void)hipSetDevice(b->ctx->device);
  dt = cs_min(dt, 1.0f / 30.0f);  // mod.rs:265
  const float sub_dt = dt / (float)b->substeps;
  for (uint32_t i = 0; i < b->substeps; ++i) {
    int rc = batch_substep(b, sub_dt);
    if (rc) return rc;
  }
  return CS_OK;
}

int cs_batch_get_stats(cs_batch* b, cs_batch_stats* out) {
  if (!b || !out) return cs_fail(CS_ERR_INVALID, "null argument");
  *out = b->stats;
  return CS_OK;
}

int cs_batch_reset_stats(cs_batch* b) {
  if (!b) return cs_fail(CS_ERR_INVALID, "null batch");
  memset(&b->stats, 0, sizeof(b->stats));
  return CS_OK;
}

}  // extern "C"
