// cs_solver.hip — brick-sparse red-black Gauss-Seidel pressure solve.
//
// Restates pressure_update (shader.rs:1371-1439) over the fluid cells only.
// The grid is tiled into 8x8x8 bricks; classify flags the bricks holding a
// fluid cell and a single-block scan lists them in ascending order.  The face
// weights and diagonal of every cell are formed once per substep
// (k_face_weights).  One workgroup per active brick loads the brick plus an
// H-cell halo of pressure, weights and right-hand side into LDS and runs H
// half sweeps there:
// after half sweep t every cell at least t+1 cells from the loaded region's
// edge holds exactly the value the whole-grid sweep would give, so after H
// sweeps the brick itself is exact and is written out.  Pressure is
// double-buffered (every workgroup reads the state at launch start).  Within
// one colour each update reads only the other colour, so the order of cells
// and bricks never changes a result.
//
// Convergence (relative residual every `check_every` iterations) is decided on
// the device: each sweep launch and residual pass returns at once when the
// solve-state flag says done, so the host can enqueue ahead without reading
// the residual back every chunk.
#include <algorithm>
#include <cstdlib>

#include "cs_batch.h"
#include "cs_device.h"
#include "cs_solver.h"

namespace {

constexpr int BR = CS_BRICK;  // brick side

__device__ __forceinline__ void brick_origin(const cs_params& u, uint32_t b, int& x0, int& y0, int& z0) {
  uint32_t bx = b % u.nbx;
  uint32_t r = b / u.nbx;
  uint32_t by = r % u.nby;
  uint32_t bz = r / u.nby;
  x0 = (int)bx * BR;
  y0 = (int)by * BR;
  z0 = (int)bz * BR;
}

constexpr float KF_OFFGRID = -3.0f;

// Global index of local cell i of an R^3 region whose corner is (x0, y0, z0),
// clamped onto the grid, and whether the unclamped cell lies on the grid.  The
// staging loops load through the clamped index unconditionally and select
// afterwards, so every load of a thread is issued before the first wait.
template <int R>
__device__ __forceinline__ uint32_t region_cell(const cs_params& u, int i, int x0, int y0, int z0, bool& on) {
  const int lx = i % R, ly = (i / R) % R, lz = i / (R * R);
  const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
  on = (uint32_t)gx < u.gx && (uint32_t)gy < u.gy && (uint32_t)gz < u.gz;
  const uint32_t cx = (uint32_t)min(max(gx, 0), (int)u.gx - 1);
  const uint32_t cy = (uint32_t)min(max(gy, 0), (int)u.gy - 1);
  const uint32_t cz = (uint32_t)min(max(gz, 0), (int)u.gz - 1);
  return (cz * u.gy + cy) * u.gx + cx;
}

// Neighbour offsets in the order pressure_update visits them (shader.rs:1393-1420):
// -x, +x, -y, +y, -z, +z.
constexpr int NB_DIR[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};

// A fixed-point pressure as the update reads it (pressure_load, shader.rs:198).
__device__ __forceinline__ float p_float(const cs_params& u, int32_t p) { return (float)p * u.inv_fp_scale; }

// The LDS copy of a swept region holds each fixed-point pressure p as the
// float the update reads, p_float(p) = f32(p) * 2^-18.  Every pressure in the
// buffers is the truncation of a float times 2^18 (the store below), so f32(p)
// is exact and p = p_float(p) * 2^18 recovers it (fp_scale is the power of two
// 2^18, state.rs:13).
__device__ __forceinline__ float p_stage(const cs_params& u, int32_t p) { return p_float(u, p); }
__device__ __forceinline__ int32_t p_unstage(const cs_params& u, float v) { return (int32_t)(v * u.fp_scale); }

// One update of pressure_update (shader.rs:1386-1438) from precomputed face
// weights, on pressures held as p_float values.  `self` holds the +x, +y, +z
// face weights and the diagonal count of the cell; the -x, -y, -z weights are
// the +x, +y, +z weights of the neighbours on those sides (the face weight is
// symmetric).  A face the reference skips carries weight 0, whose product +-0
// leaves the sum unchanged, so the sum is the reference's bit for bit (up to
// the sign of an all-zero sum, which no stored pressure can tell apart).
// Returns p_float of the pressure the reference stores: the clamped value
// (|v| <= pressure_clamp = FP_VALUE_LIMIT - 1, mod.rs:592, so v * 2^18 fits
// an i32 and truncation equals the saturating conversion) truncated to the
// fixed-point grid.  The clamp is max-then-min: for the non-NaN bounds
// -c < c these return the same value as cs_clamp for every input, NaN included.
__device__ __forceinline__ float update_cell(const cs_params& u, float4 self, float srhs, float wxm, float wym,
                                             float wzm, const float* s_p, int li, int sx, int sy, bool& fired) {
  float sum = wxm * s_p[li - 1];
  sum += self.x * s_p[li + 1];
  sum += wym * s_p[li - sx];
  sum += self.y * s_p[li + sx];
  sum += wzm * s_p[li - sy];
  sum += self.z * s_p[li + sy];
  // count > 0 selects the quotient, else 0 (shader.rs:1422-1430); the division
  // runs either way so the update has no branch
  const float quot = (sum - srhs) / self.w;
  const float pn = self.w > 0.0f ? quot : 0.0f;
  const float cl = fminf(fmaxf(pn, -u.pressure_clamp), u.pressure_clamp);
  fired = cl != pn;
  return truncf(cl * u.fp_scale) * u.inv_fp_scale;
}

}  // namespace

// ---- face weights of the pressure system ------------------------------------------------
// For every cell: .x/.y/.z = the weight pressure_update gives the +x/+y/+z
// neighbour's pressure (pressure_face_weight, shader.rs:306-323, when that
// neighbour is fluid and the face is used; 0 otherwise), .w = the diagonal
// `count` accumulated over all six faces in the reference's order
// (shader.rs:1393-1420); .w = -1 marks a cell that is not fluid.  The kinds
// and fills do not change during a solve, so this runs once per substep.
__device__ __forceinline__ void face_weights_cell(const cs_params& u, const float* __restrict__ kf,
                                                  float4* __restrict__ wc, uint32_t idx) {
  if (idx >= u.nc) return;
  const float skf = kf[idx];
  if (!kf_fluid(skf)) {
    wc[idx] = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    return;
  }
  uint32_t ix, iy, iz;
  decompose(u, idx, ix, iy, iz);
  const float sf = kf_fill(skf);
  const bool sbed = kf_bed(skf);
  float count = 0.0f;
  float wplus[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int n = 0; n < 6; ++n) {
    const int nx = (int)ix + NB_DIR[n][0], ny = (int)iy + NB_DIR[n][1], nz = (int)iz + NB_DIR[n][2];
    float nkf = KF_OFFGRID;
    if (nx >= 0 && ny >= 0 && nz >= 0 && nx < (int)u.gx && ny < (int)u.gy && nz < (int)u.gz)
      nkf = kf[cell_index(u, (uint32_t)nx, (uint32_t)ny, (uint32_t)nz)];
    const bool nfluid = nkf >= 0.0f;
    const bool skip = (nkf == KF_OFFGRID) | (nkf == CS_KF_SOLID);
    float w = cs_min(sf, kf_fill(nkf));
    w = nfluid ? w : sf;
    w = (sbed | kf_bed(nkf)) ? 1.0f : w;
    const bool use = !skip && (w > 0.0f);
    count += use ? w : 0.0f;
    if (n & 1) wplus[n >> 1] = (use && nfluid) ? w : 0.0f;
  }
  wc[idx] = make_float4(wplus[0], wplus[1], wplus[2], count);
}
__global__ void __launch_bounds__(256) k_face_weights(const cs_dparams* __restrict__ dp, const float* __restrict__ kf,
                                                      float4* __restrict__ wc) {
  CS_PARAMS_FROM(dp);
  face_weights_cell(u, kf, wc, blockIdx.x * blockDim.x + threadIdx.x);
}

// The solve state's reset before a solve's first residual test.
__device__ __forceinline__ void state_init(cs_solve_state* st) {
  st->done = 0u;
  st->iterations = 0u;
  st->launches = 0u;
  st->pending_fires = 0u;
  st->residual = 0.0;
  st->initial_residual = 0.0;
}

// ---- active brick list ----------------------------------------------------------------
// Single block: ascending list of flagged bricks and their count.
__device__ __forceinline__ void brick_scan_block(const uint32_t* __restrict__ flags, uint32_t nbricks,
                                                 uint32_t* __restrict__ list, cs_solve_state* __restrict__ st,
                                                 uint32_t* s_cnt) {
  const uint32_t t = threadIdx.x;
  const uint32_t per = (nbricks + 1023u) / 1024u;
  const uint32_t b0 = t * per;
  const uint32_t b1 = min(b0 + per, nbricks);
  uint32_t c = 0;
  for (uint32_t b = b0; b < b1; ++b) c += flags[b] ? 1u : 0u;
  s_cnt[t] = c;
  __syncthreads();
  // inclusive Hillis-Steele scan
  for (uint32_t d = 1; d < 1024u; d <<= 1) {
    uint32_t v = t >= d ? s_cnt[t - d] : 0u;
    __syncthreads();
    s_cnt[t] += v;
    __syncthreads();
  }
  uint32_t pos = s_cnt[t] - c;
  for (uint32_t b = b0; b < b1; ++b)
    if (flags[b]) list[pos++] = b;
  if (t == 1023u) st->active_bricks = s_cnt[1023];
}
// The count also goes to the substep's parameter block, which halts when it
// exceeds the bricks the substep's launches were sized for.
__global__ void __launch_bounds__(1024) k_brick_scan(cs_dparams* __restrict__ dp, const uint32_t* __restrict__ flags,
                                                     uint32_t nbricks, uint32_t* __restrict__ list,
                                                     cs_solve_state* __restrict__ st) {
  __shared__ uint32_t s_cnt[1024];
  if (dp->halt) return;
  if (threadIdx.x == 0) state_init(st);
  brick_scan_block(flags, nbricks, list, st, s_cnt);
  if (threadIdx.x == 1023u) {
    const uint32_t n = s_cnt[1023];
    dp->nactive = n;
    if (n > dp->cap_bricks) dp->halt = dp->slot + 1u;  // which substep stopped
  }
}

// ---- red-black sweeps over LDS-resident regions --------------------------------------------
// A region is a brick plus an H-cell halo, R = 8 + 2H cells a side, staged as
// face weights, right-hand side and fixed-point pressure (24 B per cell).
template <int H, int NT>
struct Region {
  static constexpr int R = BR + 2 * H;  // even: BR is 8
  static constexpr int R3 = R * R * R;
  static constexpr int PER = (R3 + NT - 1) / NT;
  float4 wv[PER];
  float rv[PER];
  int32_t pv[PER];
  bool on[PER];

  // Issues every global load of the region at (x0, y0, z0) into registers.
  __device__ __forceinline__ void fetch(const cs_params& u, int x0, int y0, int z0, const float4* __restrict__ wc,
                                        const float* __restrict__ rhs, const int32_t* __restrict__ pin) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const uint32_t idx = region_cell<R>(u, min((int)threadIdx.x + k * NT, R3 - 1), x0, y0, z0, on[k]);
      wv[k] = wc[idx];
      rv[k] = rhs[idx];
      pv[k] = pin[idx];
    }
  }
  // Writes the fetched region to LDS; off-grid cells become non-fluid.
  __device__ __forceinline__ void stage(const cs_params& u, float4* s_wc, float* s_rhs, float* s_p) const {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int i = (int)threadIdx.x + k * NT;
      if (i < R3) {
        s_wc[i] = on[k] ? wv[k] : make_float4(0.0f, 0.0f, 0.0f, -1.0f);
        s_rhs[i] = on[k] ? rv[k] : 0.0f;
        s_p[i] = on[k] ? p_stage(u, pv[k]) : 0.0f;
      }
    }
  }

  // Whether local cell i lies within L1 distance H of the brick: the cells
  // the listed sweeps, the residual and the store ever read (2656 of 4096
  // for H = 4).  The others are neither fetched nor staged.
  __device__ __forceinline__ static bool in_ball(int i) {
    const int lx = i % R, ly = (i / R) % R, lz = i / (R * R);
    auto d = [](int l) { return max(0, max(H - l, l - (H + BR - 1))); };
    return d(lx) + d(ly) + d(lz) <= H;
  }
  // fetch/stage restricted to the ball.
  __device__ __forceinline__ void fetch_ball(const cs_params& u, int x0, int y0, int z0,
                                             const float4* __restrict__ wc, const float* __restrict__ rhs,
                                             const int32_t* __restrict__ pin) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int i = (int)threadIdx.x + k * NT;
      const uint32_t idx = region_cell<R>(u, min(i, R3 - 1), x0, y0, z0, on[k]);
      if (i < R3 && in_ball(i)) {
        wv[k] = wc[idx];
        rv[k] = rhs[idx];
        pv[k] = pin[idx];
      }
    }
  }
  __device__ __forceinline__ void stage_ball(const cs_params& u, float4* s_wc, float* s_rhs, float* s_p) const {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int i = (int)threadIdx.x + k * NT;
      if (i < R3 && in_ball(i)) {
        s_wc[i] = on[k] ? wv[k] : make_float4(0.0f, 0.0f, 0.0f, -1.0f);
        s_rhs[i] = on[k] ? rv[k] : 0.0f;
        s_p[i] = on[k] ? p_stage(u, pv[k]) : 0.0f;
      }
    }
  }
};

// The cells a thread updates in each half sweep, fixed for the whole launch.
// Half sweep s covers the cells at least s+1 cells inside the region; its
// j-th cell of a colour sits at (lo + 2m + parity, ly, lz) with
// (m, ly, lz) from j, so a thread keeps, per round of NT cells, the local
// index of the even-x candidate, the parity of ly + lz + lo and whether the
// cell is a brick cell for either x parity.  code = li0 | q << 16 |
// inner_even << 17 | inner_odd << 18 | valid << 19.
template <int H, int NT>
struct SweepMap {
  static constexpr int R = BR + 2 * H;
  static constexpr int rounds(int s) {
    const int span = R - 2 * (s + 1);
    return ((span / 2) * span * span + NT - 1) / NT;
  }
  static constexpr int total() {
    int t = 0;
    for (int s = 0; s < H; ++s) t += rounds(s);
    return t;
  }
  uint32_t code[total()];

  __device__ __forceinline__ void build() {
    int t = 0;
#pragma unroll
    for (int s = 0; s < H; ++s) {
      const int lo = s + 1;
      const int span = R - 2 * (s + 1);
      const int half = span / 2;
      const int n = half * span * span;
#pragma unroll
      for (int r = 0; r < rounds(s); ++r, ++t) {
        const int j = (int)threadIdx.x + r * NT;
        const int m = j % half;
        const int ly = lo + (j / half) % span;
        const int lz = lo + j / (half * span);
        const int lx0 = lo + 2 * m;
        const bool yz_in = ly >= H && lz >= H && ly < H + BR && lz < H + BR;
        const uint32_t inner_even = yz_in && lx0 >= H && lx0 < H + BR;
        const uint32_t inner_odd = yz_in && lx0 + 1 >= H && lx0 + 1 < H + BR;
        const uint32_t q = (uint32_t)(ly + lz + lo) & 1u;
        code[t] = j < n ? ((uint32_t)((lz * R + ly) * R + lx0) | q << 16 | inner_even << 17 | inner_odd << 18 | 1u << 19)
                        : 0u;
      }
    }
  }
};

// H half sweeps (colours color0, color0^1, ...) over a staged region whose
// corner cell has parity base_par.  Each update is branch-free: every LDS read
// is issued at once and a cell that is not fluid writes its own value back.
// Returns the clamp fires of the brick cells (each brick cell counted once per
// half sweep of its colour, as the whole-grid sweep counts it).
template <int H, int NT>
__device__ __forceinline__ uint32_t sweep_region(const cs_params& u, const SweepMap<H, NT>& map, uint32_t color0,
                                                 uint32_t base_par, const float4* s_wc, const float* s_rhs,
                                                 float* s_p) {
  constexpr int R = SweepMap<H, NT>::R;
  uint32_t fires = 0;
  int t = 0;
#pragma unroll
  for (int s = 0; s < H; ++s) {
    const uint32_t cpar = color0 ^ (uint32_t)(s & 1) ^ base_par;
#pragma unroll
    for (int r = 0; r < SweepMap<H, NT>::rounds(s); ++r, ++t) {
      const uint32_t code = map.code[t];
      if (code >> 19) {
        const uint32_t par = (cpar + (code >> 16)) & 1u;
        const int li = (int)(code & 0xffffu) + (int)par;
        const float4 self = s_wc[li];
        const float srhs = s_rhs[li];
        bool fired;
        const float np = update_cell(u, self, srhs, s_wc[li - 1].x, s_wc[li - R].y, s_wc[li - R * R].z, s_p, li, R,
                                     R * R, fired);
        const bool fluid = self.w >= 0.0f;
        if (fluid) s_p[li] = np;
        const uint32_t inner = (code >> (17u + par)) & 1u;
        fires += (fluid && fired) ? inner : 0u;
      }
    }
    __syncthreads();
  }
  return fires;
}

// Half sweeps over precomputed cell lists.  After half sweep s only the cells
// within L1 distance H-1-s of the brick still matter (one neighbour step per
// half sweep), so each sweep visits that diamond-shaped shell of the region
// instead of the whole cube: for H = 4 the four colour lists hold 980, 688,
// 448 and 256 cells, one per thread of a 1024-thread workgroup, where the
// cube needs 1372 + 864 + 500 + 256.  codes[s] = li | inner << 17 | valid << 19
// for this thread's cell of half sweep s (cs_build_sweep_table).
template <int H, int NT>
__device__ __forceinline__ uint32_t sweep_listed(const cs_params& u, const uint32_t (&codes)[H], const float4* s_wc,
                                                 const float* s_rhs, float* s_p) {
  constexpr int R = BR + 2 * H;
  uint32_t fires = 0;
#pragma unroll
  for (int s = 0; s < H; ++s) {
    const uint32_t code = codes[s];
    if (code >> 19) {
      const int li = (int)(code & 0xffffu);
      const float4 self = s_wc[li];
      const float srhs = s_rhs[li];
      bool fired;
      const float np = update_cell(u, self, srhs, s_wc[li - 1].x, s_wc[li - R].y, s_wc[li - R * R].z, s_p, li, R,
                                   R * R, fired);
      const bool fluid = self.w >= 0.0f;
      if (fluid) s_p[li] = np;
      fires += (fluid && fired) ? ((code >> 17) & 1u) : 0u;
    }
    __syncthreads();
  }
  return fires;
}

// Writes the brick cells of a swept region (corner x0, y0, z0) to `pout`.
template <int H, int NT>
__device__ __forceinline__ void store_brick(const cs_params& u, int x0, int y0, int z0, const float* s_p,
                                            int32_t* __restrict__ pout) {
  constexpr int R = Region<H, NT>::R;
  for (int j = threadIdx.x; j < BR * BR * BR; j += NT) {
    const int lx = H + j % BR, ly = H + (j / BR) % BR, lz = H + j / (BR * BR);
    const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
    if (gx < (int)u.gx && gy < (int)u.gy && gz < (int)u.gz)
      pout[((uint32_t)gz * u.gy + (uint32_t)gy) * u.gx + (uint32_t)gx] = p_unstage(u, s_p[(lz * R + ly) * R + lx]);
  }
}

// One workgroup per active brick.
template <int H, int NT>
__global__ void __launch_bounds__(NT) k_rbgs_brick(cs_params u, const uint32_t* __restrict__ list, uint32_t color0,
                                                   const float4* __restrict__ wc, const float* __restrict__ rhs,
                                                   const int32_t* __restrict__ pin, int32_t* __restrict__ pout,
                                                   uint32_t* __restrict__ metrics,
                                                   const cs_solve_state* __restrict__ st) {
  using Reg = Region<H, NT>;
  __shared__ float4 s_wc[Reg::R3];
  __shared__ float s_rhs[Reg::R3];
  __shared__ float s_p[Reg::R3];
  // both scalar loads in flight before the first wait
  const uint32_t done = st->done;
  const uint32_t brick = list[blockIdx.x];
  if (done) return;
  int x0, y0, z0;
  brick_origin(u, brick, x0, y0, z0);
  x0 -= H;
  y0 -= H;
  z0 -= H;
  Reg reg;
  reg.fetch(u, x0, y0, z0, wc, rhs, pin);
  SweepMap<H, NT> map;
  map.build();
  reg.stage(u, s_wc, s_rhs, s_p);
  __syncthreads();
  const uint32_t fires = sweep_region<H, NT>(u, map, color0, (uint32_t)(x0 + y0 + z0) & 1u, s_wc, s_rhs, s_p);
  store_brick<H, NT>(u, x0, y0, z0, s_p, pout);
  if (fires) atomicAdd(&metrics[CS_M_P_CLAMP], fires);
}

// ---- residual helpers ---------------------------------------------------------------------
// Relative residual from the sums of squares (A14: r = |b - Ap| / |b| over fluid
// cells, leaving out cells held at the pressure clamp, whose row the bound
// replaces).
__device__ __forceinline__ double relative_residual(double rr, double bb) {
  if (bb > 0.0) return sqrt(rr / bb);
  return rr > 0.0 ? __builtin_huge_val() : 0.0;
}

__device__ __forceinline__ void finalize_state(cs_solve_state* st, double rr, double bb, uint32_t mode,
                                               uint32_t iters, uint32_t launches, double tol) {
  const double r = relative_residual(rr, bb);
  st->residual = r;
  if (mode == 0u) {
    st->initial_residual = r;
    st->done = r <= tol ? 1u : 0u;
  } else {
    st->iterations += iters;
    st->launches += launches;
    if (r <= tol) st->done = 1u;
  }
}

// Sum over a wavefront in a fixed butterfly order (every lane gets the total).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// Squared residual and squared right-hand side of the brick cells of a staged
// region (the one-cell halo around the brick is staged too), summed over the
// workgroup in a fixed order; the totals are valid in thread 0.  `s_red`
// holds two doubles per wavefront.
template <int H, int NT>
__device__ __forceinline__ void brick_residual(const cs_params& u, const float4* s_wc, const float* s_rhs,
                                               const float* s_p, double* s_red, double& rr_out, double& bb_out) {
  constexpr int R = Region<H, NT>::R;
  constexpr int NW = NT / 64;
  double rr = 0.0, bb = 0.0;
  for (int j = (int)threadIdx.x; j < BR * BR * BR; j += NT) {
    const int li = ((H + j / (BR * BR)) * R + H + (j / BR) % BR) * R + H + j % BR;
    const float4 self = s_wc[li];
    const float count = self.w;
    // residual_rows order: -x, +x, -y, +y, -z, +z; unused faces carry weight 0
    const float wn[6] = {s_wc[li - 1].x, self.x, s_wc[li - R].y, self.y, s_wc[li - R * R].z, self.z};
    const int off[6] = {-1, 1, -R, R, -R * R, R * R};
    double ap_n = 0.0;
#pragma unroll
    for (int n = 0; n < 6; ++n) ap_n += (double)wn[n] * (double)s_p[li + off[n]];
    const double b = -(double)s_rhs[li];
    const double ap = (double)count * (double)s_p[li] - ap_n;
    if (count > 0.0f && fabsf(s_p[li]) != u.pressure_clamp) {
      rr += (b - ap) * (b - ap);
      bb += b * b;
    }
  }
  rr = wave_sum(rr);
  bb = wave_sum(bb);
  const int w = (int)threadIdx.x / 64;
  if ((threadIdx.x & 63u) == 0u) {
    s_red[2 * w] = rr;
    s_red[2 * w + 1] = bb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tr = 0.0, tb = 0.0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      tr += s_red[2 * k];
      tb += s_red[2 * k + 1];
    }
    rr_out = tr;
    bb_out = tb;
  }
}

// A fixed grid of workgroups (one per compute unit when the region fills the
// LDS), each walking a run of the active-brick list and fetching the next
// brick's region into registers while it sweeps the current one, so the
// memory latency of the staging loads is paid once per launch rather than
// once per brick.  The list is split into eight contiguous runs, one per
// group of workgroups with equal blockIdx.x % 8 (the XCD a workgroup is
// dispatched to), so spatially neighbouring bricks share an L2.
//
// The convergence test rides on the sweeps (`res_mode`):
//  CS_RES_COMPUTE  the residual of the launch's input pressure is summed from
//                  the staged regions before sweeping; each workgroup writes
//                  one (rr, bb) pair to partials[blockIdx.x] (zeros when it
//                  has no brick).
//  CS_RES_DECIDE   before anything else every workgroup sums the previous
//                  launch's gridDim.x partials in the same fixed order, so all
//                  reach the same verdict; workgroup 0 advances the solve
//                  state by (chunk_iters, chunk_launches).  When the residual
//                  meets `tol` the launch writes nothing: the converged
//                  pressure stays in its input's input buffer, which the
//                  launch count now points at.
template <int H, int NT>
__global__ void __launch_bounds__(NT) k_rbgs_multi(cs_params u, const uint32_t* __restrict__ list, uint32_t nactive,
                                                   uint32_t color0, const float4* __restrict__ wc,
                                                   const float* __restrict__ rhs, const int32_t* __restrict__ pin,
                                                   int32_t* __restrict__ pout, uint32_t* __restrict__ metrics,
                                                   cs_solve_state* __restrict__ st, double* __restrict__ partials,
                                                   uint32_t res_mode, uint32_t chunk_iters, uint32_t chunk_launches,
                                                   double tol, const uint32_t* __restrict__ sweep_table) {
  using Reg = Region<H, NT>;
  constexpr int NW = NT / 64;
  static_assert(H % 2 == 0, "the listed sweeps assume an even region corner");
  __shared__ float4 s_wc[Reg::R3];
  __shared__ float s_rhs[Reg::R3];
  __shared__ float s_pp[2][Reg::R3];  // pressure of alternate bricks: a store never races the next staging
  __shared__ double s_red[2 * NW];
  __shared__ uint32_t s_list[NT];
  __shared__ uint32_t s_stop;
  // this workgroup's bricks: run r = blockIdx.x % 8 of the list, every `per_run`-th from `slot`
  const uint32_t runs = gridDim.x >= 8u ? 8u : 1u;
  const uint32_t run = blockIdx.x % runs;
  const uint32_t per_run = gridDim.x / runs;  // the host launches a multiple of 8 (or fewer than 8) groups
  const uint32_t slot = blockIdx.x / runs;
  const uint32_t run_lo = (uint32_t)(((uint64_t)nactive * run) / runs);
  const uint32_t run_hi = (uint32_t)(((uint64_t)nactive * (run + 1u)) / runs);
  const uint32_t b0 = run_lo + slot;
  const uint32_t nmine = b0 < run_hi ? (run_hi - 1u - b0) / per_run + 1u : 0u;
  // Every load that does not wait on a verdict is issued first: the solve
  // state, the first brick and its region, and this workgroup's brick list
  // (one entry per thread, staged to LDS with the first region), so their
  // latencies overlap one another and the residual decision.
  const uint32_t done = st->done;
  uint32_t my_brick = 0;
  if (threadIdx.x < nmine) my_brick = list[b0 + threadIdx.x * per_run];
  int x0 = 0, y0 = 0, z0 = 0;
  Reg reg;
  if (nmine) {
    brick_origin(u, list[b0], x0, y0, z0);
    x0 -= H;
    y0 -= H;
    z0 -= H;
    reg.fetch_ball(u, x0, y0, z0, wc, rhs, pin);
  }
  // this thread's cell of each half sweep: colour color0 ^ (s & 1); the region
  // corner (brick origin - H, H even) has even parity, so local parity is colour
  uint32_t codes[H];
#pragma unroll
  for (int k = 0; k < H; ++k)
    codes[k] = sweep_table[((color0 ^ (uint32_t)(k & 1)) * H + (uint32_t)k) * NT + threadIdx.x];
  if (done) return;
  if (res_mode & CS_RES_DECIDE) {
    double rr = 0.0, bb = 0.0;
    for (uint32_t i = threadIdx.x; i < gridDim.x; i += NT) {
      rr += partials[2 * i];
      bb += partials[2 * i + 1];
    }
    rr = wave_sum(rr);
    bb = wave_sum(bb);
    const int w = (int)threadIdx.x / 64;
    if ((threadIdx.x & 63u) == 0u) {
      s_red[2 * w] = rr;
      s_red[2 * w + 1] = bb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double tr = 0.0, tb = 0.0;
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        tr += s_red[2 * k];
        tb += s_red[2 * k + 1];
      }
      const double r = relative_residual(tr, tb);
      const uint32_t stop = r <= tol ? 1u : 0u;
      if (blockIdx.x == 0) {
        st->residual = r;
        st->iterations += chunk_iters;
        st->launches += chunk_launches;
        // the computing launch's sweeps count only if the solve goes on past them
        const uint32_t pf = st->pending_fires;
        if (!stop && pf) atomicAdd(&metrics[CS_M_P_CLAMP], pf);
        st->pending_fires = 0u;
        if (stop) st->done = 1u;
      }
      s_stop = stop;
    }
    __syncthreads();
    if (s_stop) return;
  }
  const bool compute = (res_mode & CS_RES_COMPUTE) != 0u;
  if (nmine == 0) {
    if (compute && threadIdx.x == 0) {
      partials[2 * blockIdx.x] = 0.0;
      partials[2 * blockIdx.x + 1] = 0.0;
    }
    return;
  }
  s_list[threadIdx.x] = my_brick;  // read after the first staging barrier
  uint32_t fires = 0;
  double acc_rr = 0.0, acc_bb = 0.0;  // thread 0: this workgroup's residual sums
  for (uint32_t k = 0;; ++k) {
    float* s_p = s_pp[k & 1u];
    reg.stage_ball(u, s_wc, s_rhs, s_p);
    __syncthreads();
    // prefetch the next brick; its loads stay in flight through the sweeps
    const bool more = k + 1u < nmine;
    int nx0 = 0, ny0 = 0, nz0 = 0;
    if (more) {
      const uint32_t nb = k + 1u < (uint32_t)NT ? s_list[k + 1u] : list[b0 + (k + 1u) * per_run];
      brick_origin(u, nb, nx0, ny0, nz0);
      nx0 -= H;
      ny0 -= H;
      nz0 -= H;
      reg.fetch_ball(u, nx0, ny0, nz0, wc, rhs, pin);
    }
    if (compute) {
      double rr = 0.0, bb = 0.0;
      brick_residual<H, NT>(u, s_wc, s_rhs, s_p, s_red, rr, bb);
      if (threadIdx.x == 0) {
        acc_rr += rr;
        acc_bb += bb;
      }
      // s_red is reused by the next brick only after the sweeps' barriers
    }
    fires += sweep_listed<H, NT>(u, codes, s_wc, s_rhs, s_p);
    store_brick<H, NT>(u, x0, y0, z0, s_p, pout);
    if (!more) break;
    // no barrier: the next staging writes the other pressure buffer, and the
    // weights and right-hand side were last read before the final sweep barrier
    x0 = nx0;
    y0 = ny0;
    z0 = nz0;
  }
  if (compute && threadIdx.x == 0) {
    partials[2 * blockIdx.x] = acc_rr;
    partials[2 * blockIdx.x + 1] = acc_bb;
  }
  if (fires) atomicAdd(compute ? &st->pending_fires : &metrics[CS_M_P_CLAMP], fires);
}

// ---- residual over the active bricks ------------------------------------------------------
// One workgroup per brick stages the brick and a one-cell halo of (pressure,
// kind/fill) in LDS, forms b - A p for its fluid cells in double precision
// and writes a per-brick partial; a one-workgroup pass then sums the partials
// in brick order (a fixed order) and advances the solve state.  (Electing the
// last workgroup with an arrival counter instead serialises one device-scope
// atomic per brick on a single address, which costs more than the launch.)
// Sums of squared residual and squared right-hand side over one brick's fluid
// cells (one 256-thread workgroup); valid in s_rr[0], s_bb[0] on return.  With
// ENERGY the same rows give <p, A p> and <p, b> instead (the Galerkin rescale).
template <int NT, bool ENERGY = false>
__device__ __forceinline__ void brick_residual_sums(const cs_params& u, uint32_t brick, const float4* __restrict__ wc,
                                                    const float* __restrict__ rhs, const int32_t* __restrict__ p,
                                                    float4* s_wc, int32_t* s_p, double* s_rr, double* s_bb) {
  constexpr int R = BR + 2, R3 = R * R * R, PER = (R3 + NT - 1) / NT;
  constexpr int BPER = BR * BR * BR / NT;
  int x0, y0, z0;
  brick_origin(u, brick, x0, y0, z0);
  float rb[BPER];  // right-hand side of this thread's brick cells
  bool rb_on[BPER];
#pragma unroll
  for (int k = 0; k < BPER; ++k) {
    const uint32_t idx = region_cell<BR>(u, (int)threadIdx.x + k * NT, x0, y0, z0, rb_on[k]);
    rb[k] = rhs[idx];
  }
  x0 -= 1;
  y0 -= 1;
  z0 -= 1;
  {
    float4 wv[PER];
    int32_t pv[PER];
    bool on[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const uint32_t idx = region_cell<R>(u, min((int)threadIdx.x + k * NT, R3 - 1), x0, y0, z0, on[k]);
      wv[k] = wc[idx];
      pv[k] = p[idx];
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int i = (int)threadIdx.x + k * NT;
      if (i < R3) {
        s_wc[i] = on[k] ? wv[k] : make_float4(0.0f, 0.0f, 0.0f, -1.0f);
        s_p[i] = on[k] ? pv[k] : 0;
      }
    }
  }
  __syncthreads();
  double rr = 0.0, bb = 0.0;
#pragma unroll
  for (int k = 0; k < BPER; ++k) {
    const int j = (int)threadIdx.x + k * NT;
    const int lx = 1 + j % BR, ly = 1 + (j / BR) % BR, lz = 1 + j / (BR * BR);
    const int li = (lz * R + ly) * R + lx;
    const float4 self = s_wc[li];
    const float count = self.w;
    if (count <= 0.0f) continue;  // not fluid (-1), or a fluid cell with no used face
    if (fabsf(p_float(u, s_p[li])) == u.pressure_clamp) continue;  // held at the bound: not a row of the system
    // residual_rows order: -x, +x, -y, +y, -z, +z; unused faces carry weight 0
    const float wn[6] = {s_wc[li - 1].x, self.x, s_wc[li - R].y, self.y, s_wc[li - R * R].z, self.z};
    const int off[6] = {-1, 1, -R, R, -R * R, R * R};
    double ap_n = 0.0;
#pragma unroll
    for (int n = 0; n < 6; ++n) ap_n += (double)wn[n] * (double)p_float(u, s_p[li + off[n]]);
    const double b = -(double)rb[k];
    const double pi = (double)p_float(u, s_p[li]);
    const double ap = (double)count * pi - ap_n;
    if (ENERGY) {
      rr += pi * ap;
      bb += pi * b;
    } else {
      rr += (b - ap) * (b - ap);
      bb += b * b;
    }
  }
  s_rr[threadIdx.x] = rr;
  s_bb[threadIdx.x] = bb;
  __syncthreads();
  for (uint32_t st = NT / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      s_rr[threadIdx.x] += s_rr[threadIdx.x + st];
      s_bb[threadIdx.x] += s_bb[threadIdx.x + st];
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256) k_residual_bricks(const cs_dparams* __restrict__ dp,
                                                         const uint32_t* __restrict__ list,
                                                         const float4* __restrict__ wc, const float* __restrict__ rhs,
                                                         const int32_t* __restrict__ p, double* __restrict__ partials,
                                                         const cs_solve_state* __restrict__ st, uint32_t mode) {
  CS_PARAMS_FROM(dp);
  constexpr int R3 = (BR + 2) * (BR + 2) * (BR + 2), NT = 256;
  __shared__ float4 s_wc[R3];
  __shared__ int32_t s_p[R3];
  __shared__ double s_rr[NT], s_bb[NT];
  if (blockIdx.x >= dp->nactive) return;
  const uint32_t done = st->done;
  const uint32_t brick = list[blockIdx.x];
  if (mode != 0u && done) return;
  brick_residual_sums<NT>(u, brick, wc, rhs, p, s_wc, s_p, s_rr, s_bb);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = s_rr[0];
    partials[2 * blockIdx.x + 1] = s_bb[0];
  }
}

// Fixed-order sum of `n` per-brick partials, then the solve-state update.
// Sum of n (rr, bb) partials, thread-strided then a tree over 1024 threads;
// valid in s_rr[0], s_bb[0].
__device__ __forceinline__ void reduce_partials_block(const double* __restrict__ partials, uint32_t n, double* s_rr,
                                                      double* s_bb) {
  constexpr uint32_t NT = 1024;
  double rr = 0.0, bb = 0.0;
  for (uint32_t i = threadIdx.x; i < n; i += NT) {
    rr += partials[2 * i];
    bb += partials[2 * i + 1];
  }
  s_rr[threadIdx.x] = rr;
  s_bb[threadIdx.x] = bb;
  __syncthreads();
  for (uint32_t s = NT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      s_rr[threadIdx.x] += s_rr[threadIdx.x + s];
      s_bb[threadIdx.x] += s_bb[threadIdx.x + s];
    }
    __syncthreads();
  }
}
// The multigrid start's gate on the initial residual just finalised (mode 0).
// flags[0]: skip the next cycle; flags[1]: the start was within the tolerance
// (no cycle runs); flags[2]: cycles run so far.  The first test sets all
// three, later ones count the cycle just run and set flags[0].  After the
// start's last test (mode bit 1) the iteration runs (its first test comes
// after a chunk) unless the start was already within the tolerance.
__device__ __forceinline__ void mg_gate(cs_solve_state* st, double tol, uint32_t* flags, uint32_t mode) {
  const uint32_t within = st->initial_residual <= tol ? 1u : 0u;
  if (mode & 1u) {
    flags[1] = within;
    flags[2] = 0u;
  } else if (!flags[0]) {
    flags[2] += 1u;
  }
  flags[0] = flags[1] | within;
  if ((mode & 2u) && !flags[1]) st->done = 0u;
}
// One partial per active brick: the count is the brick scan's.  With `gate`,
// the multigrid start's gate follows in the same thread (gate_mode: bit 0 its
// first test, bit 1 its last).
__global__ void __launch_bounds__(1024) k_residual_reduce(const double* __restrict__ partials,
                                                          cs_solve_state* __restrict__ st, uint32_t mode,
                                                          uint32_t iters, uint32_t launches, double tol,
                                                          uint32_t* __restrict__ gate, uint32_t gate_mode) {
  __shared__ double s_rr[1024], s_bb[1024];
  if (mode != 0u && st->done) return;
  reduce_partials_block(partials, st->active_bricks, s_rr, s_bb);
  if (threadIdx.x == 0) {
    finalize_state(st, s_rr[0], s_bb[0], mode, iters, launches, tol);
    if (gate) mg_gate(st, tol, gate, gate_mode);
  }
}

// The residual of a system with no fluid cell (r = 0: nothing to solve).
__global__ void k_residual_empty(cs_solve_state* __restrict__ st, uint32_t mode, uint32_t iters, uint32_t launches,
                                 double tol, uint32_t* __restrict__ gate, uint32_t gate_mode) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (mode != 0u && st->done) return;
  finalize_state(st, 0.0, 0.0, mode, iters, launches, tol);
  if (gate) mg_gate(st, tol, gate, gate_mode);
}

// ---- one-time Galerkin rescale ----------------------------------------------------------------
// After rbgs_rescale_after iterations the error left is mostly smooth, and a
// start that carried a few-per-cent scale error still carries it.  The factor
// alpha = <b, p> / <p, A p> minimises the energy norm of the error of alpha p
// over alpha; it is formed over the residual's rows (fluid, non-empty
// stencil, not held at the clamp), rounded to a multiple of 2^-20 so that the
// summation order of the double sums cannot change it in practice, and kept in
// [1/16, 16] (1 when the sums are degenerate).
__device__ __forceinline__ float rescale_factor(double pap, double pb) {
  if (!(pap > 0.0)) return 1.0f;
  const double a = pb / pap;
  if (!(a >= 0.0625 && a <= 16.0)) return a > 16.0 ? 16.0f : (a < 0.0625 ? 0.0625f : 1.0f);
  return (float)(rint(a * 1048576.0) * (1.0 / 1048576.0));
}
__global__ void __launch_bounds__(256) k_rescale_bricks(cs_params u, const uint32_t* __restrict__ list,
                                                        const float4* __restrict__ wc, const float* __restrict__ rhs,
                                                        const int32_t* __restrict__ p, double* __restrict__ partials,
                                                        const cs_solve_state* __restrict__ st) {
  constexpr int R3 = (BR + 2) * (BR + 2) * (BR + 2), NT = 256;
  __shared__ float4 s_wc[R3];
  __shared__ int32_t s_p[R3];
  __shared__ double s_rr[NT], s_bb[NT];
  if (st->done) return;
  brick_residual_sums<NT, true>(u, list[blockIdx.x], wc, rhs, p, s_wc, s_p, s_rr, s_bb);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = s_rr[0];
    partials[2 * blockIdx.x + 1] = s_bb[0];
  }
}
__global__ void __launch_bounds__(1024) k_rescale_reduce(const double* __restrict__ partials, uint32_t n,
                                                         cs_solve_state* __restrict__ st) {
  __shared__ double s_rr[1024], s_bb[1024];
  if (st->done) return;
  reduce_partials_block(partials, n, s_rr, s_bb);
  if (threadIdx.x == 0) st->rescale = rescale_factor(s_rr[0], s_bb[0]);
}
// p <- alpha p in every fluid cell, clamped and quantised as pressure_update stores it.
__global__ void __launch_bounds__(256) k_rescale_apply(cs_params u, const float* __restrict__ kf,
                                                       int32_t* __restrict__ p, const cs_solve_state* __restrict__ st,
                                                       uint32_t* __restrict__ metrics) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= u.nc || st->done) return;
  if (!kf_fluid(kf[idx])) return;
  const float v = st->rescale * p_float(u, p[idx]);
  const float cl = cs_clamp(v, -u.pressure_clamp, u.pressure_clamp);
  if (cl != v) atomicAdd(&metrics[CS_M_P_CLAMP], 1u);
  p[idx] = cs_f2i(cl * u.fp_scale);
}

// Copies the brick cells of the second buffer back to the first: when an odd
// number of sweep launches ran (`by_state`: the launch count in the solve
// state decides), or unconditionally.  Cells outside the bricks are never
// written in the second buffer, so only brick cells are copied.
__global__ void __launch_bounds__(256) k_select_final(cs_params u, const uint32_t* __restrict__ list,
                                                      int32_t* __restrict__ a, const int32_t* __restrict__ b,
                                                      const cs_solve_state* __restrict__ st, uint32_t by_state) {
  if (by_state && (st->launches & 1u) == 0u) return;
  int x0, y0, z0;
  brick_origin(u, list[blockIdx.x], x0, y0, z0);
  for (int j = threadIdx.x; j < BR * BR * BR; j += blockDim.x) {
    int ix = x0 + j % BR, iy = y0 + (j / BR) % BR, iz = z0 + j / (BR * BR);
    if (ix >= (int)u.gx || iy >= (int)u.gy || iz >= (int)u.gz) continue;
    uint32_t idx = ((uint32_t)iz * u.gy + (uint32_t)iy) * u.gx + (uint32_t)ix;
    a[idx] = b[idx];
  }
}



// ---- batched solve: every scene of a batch in one launch per step ---------------------------
// The active bricks of the scenes still solving are compacted into one list of
// entries (scene << 16 | brick), scene by scene in ascending brick order.  The
// sweep launch walks that list like k_rbgs_multi walks one scene's list, and
// the residual is formed per brick with the single-scene helper and reduced
// per scene in the single-scene order, so each scene's solve equals its own
// unfused brick solve bit for bit.
#define CS_BATCH_SCENE(space)                                                                         \
  const uint32_t* st_ = starts + (space) * (n + 1u);                                                  \
  const uint32_t scene_ = cs_batch_scene_of(st_, n, blockIdx.x);                                      \
  const cs_scene_dev& d = sc[scene_];                                                                 \
  const uint32_t blk = blockIdx.x - st_[scene_];                                                      \
  const cs_params u = d.u

// Solve-state reset and active-brick scan (as k_brick_scan), one
// workgroup per scene.
__global__ void __launch_bounds__(1024) k_batch_scan(const cs_scene_dev* __restrict__ sc,
                                                     const uint32_t* __restrict__ starts, uint32_t n) {
  __shared__ uint32_t s_cnt[1024];
  CS_BATCH_SCENE(CS_SPACE_SCENES);
  (void)blk;
  if (threadIdx.x == 0) state_init(d.st);
  brick_scan_block(d.brick_flag, u.nbricks, d.brick_list, d.st, s_cnt);
}

// Face weights, and the previous step's pressure on fluid cells for scenes
// that start from it (k_face_weights + k_warm_previous).
__global__ void __launch_bounds__(256) k_batch_face_warm(const cs_scene_dev* __restrict__ sc,
                                                         const uint32_t* __restrict__ starts, uint32_t n) {
  CS_BATCH_SCENE(CS_SPACE_CELLS);
  const uint32_t idx = blk * 256u + threadIdx.x;
  face_weights_cell(u, d.kf, reinterpret_cast<float4*>(d.wc), idx);
  if (d.warm_previous && idx < u.nc && kf_fluid(d.kf[idx])) d.pres[idx] = d.prev_p[idx];
}

// One workgroup per scene: this scene's offset in the compacted list (the
// active-brick counts of the solving scenes before it), then its entries.
// The last workgroup also records the list length and the solving-scene count.
__global__ void __launch_bounds__(256) k_batch_compact(const cs_scene_dev* __restrict__ sc, uint32_t n,
                                                       uint32_t skip_done, uint32_t* __restrict__ entries,
                                                       uint32_t* __restrict__ first, uint32_t* __restrict__ count,
                                                       cs_batch_solve* __restrict__ bs) {
  __shared__ uint32_t s_sum[256], s_live[256];
  const uint32_t s = blockIdx.x;
  const bool last = s + 1u == n;
  uint32_t before = 0, live = 0;
  for (uint32_t k = threadIdx.x; k < (last ? n : s); k += 256u) {
    const cs_solve_state* st = sc[k].st;
    const bool on = !(skip_done && st->done);
    if (k < s) before += on ? st->active_bricks : 0u;
    live += on ? 1u : 0u;
  }
  s_sum[threadIdx.x] = before;
  s_live[threadIdx.x] = live;
  __syncthreads();
  for (uint32_t h = 128u; h > 0u; h >>= 1) {
    if (threadIdx.x < h) {
      s_sum[threadIdx.x] += s_sum[threadIdx.x + h];
      s_live[threadIdx.x] += s_live[threadIdx.x + h];
    }
    __syncthreads();
  }
  const cs_solve_state* mine = sc[s].st;
  const uint32_t cnt = (skip_done && mine->done) ? 0u : mine->active_bricks;
  const uint32_t f = s_sum[0];
  const uint32_t* list = sc[s].brick_list;
  for (uint32_t i = threadIdx.x; i < cnt; i += 256u) entries[f + i] = s << CS_BATCH_ENTRY_SCENE_SHIFT | list[i];
  if (threadIdx.x == 0) {
    first[s] = f;
    count[s] = cnt;
    if (last) {
      bs->total = f + cnt;
      bs->live = s_live[0];
    }
  }
}

// Where one entry's brick lives: the scene's parameters and buffers.
// The scene of an entry is the same in every lane, and its parameter block
// is read in place (scalar loads through a uniform pointer): a copy of the
// whole block per brick would live in scratch memory.
struct BatchBrick {
  const cs_params* u;
  const float4* wc;
  const float* rhs;
  const int32_t* pin;
  int32_t* pout;
  uint32_t* metrics;
  uint32_t brick;
};

__device__ __forceinline__ void batch_brick(const cs_scene_dev* __restrict__ sc, uint32_t entry, uint32_t cur,
                                            BatchBrick& b) {
  entry = __builtin_amdgcn_readfirstlane(entry);
  const cs_scene_dev& d = sc[entry >> CS_BATCH_ENTRY_SCENE_SHIFT];
  b.u = &d.u;
  b.wc = reinterpret_cast<const float4*>(d.wc);
  b.rhs = d.rhs;
  b.pin = cur ? d.pres2 : d.pres;
  b.pout = cur ? d.pres : d.pres2;
  b.metrics = d.metrics;
  b.brick = entry & ((1u << CS_BATCH_ENTRY_SCENE_SHIFT) - 1u);
}

// k_rbgs_multi over the compacted entries of all solving scenes (without the
// folded residual test): a fixed grid of workgroups, each walking a run of the
// list and prefetching the next entry's region while it sweeps the current one.
template <int H, int NT>
__global__ void __launch_bounds__(NT) k_batch_rbgs(const cs_scene_dev* __restrict__ sc,
                                                   const uint32_t* __restrict__ entries,
                                                   const cs_batch_solve* __restrict__ bs, uint32_t cur,
                                                   uint32_t color0, const uint32_t* __restrict__ sweep_table) {
  using Reg = Region<H, NT>;
  static_assert(H % 2 == 0, "the listed sweeps assume an even region corner");
  __shared__ float4 s_wc[Reg::R3];
  __shared__ float s_rhs[Reg::R3];
  __shared__ float s_pp[2][Reg::R3];  // pressure of alternate bricks: a store never races the next staging
  __shared__ uint32_t s_list[NT];
  const uint32_t total = bs->total;
  const uint32_t runs = gridDim.x >= 8u ? 8u : 1u;
  const uint32_t run = blockIdx.x % runs;
  const uint32_t per_run = gridDim.x / runs;
  const uint32_t slot = blockIdx.x / runs;
  const uint32_t run_lo = (uint32_t)(((uint64_t)total * run) / runs);
  const uint32_t run_hi = (uint32_t)(((uint64_t)total * (run + 1u)) / runs);
  const uint32_t e0 = run_lo + slot;
  const uint32_t nmine = e0 < run_hi ? (run_hi - 1u - e0) / per_run + 1u : 0u;
  if (nmine == 0) return;
  // this workgroup's entries, one per thread, staged to LDS with the first region
  uint32_t my_entry = 0;
  if (threadIdx.x < nmine) my_entry = entries[e0 + threadIdx.x * per_run];
  uint32_t entry = entries[e0];
  BatchBrick cb;
  batch_brick(sc, entry, cur, cb);
  int x0, y0, z0;
  brick_origin(*cb.u, cb.brick, x0, y0, z0);
  x0 -= H;
  y0 -= H;
  z0 -= H;
  Reg reg;
  reg.fetch_ball(*cb.u, x0, y0, z0, cb.wc, cb.rhs, cb.pin);
  uint32_t codes[H];
#pragma unroll
  for (int k = 0; k < H; ++k)
    codes[k] = sweep_table[((color0 ^ (uint32_t)(k & 1)) * H + (uint32_t)k) * NT + threadIdx.x];
  s_list[threadIdx.x] = my_entry;  // read after the first staging barrier
  for (uint32_t k = 0;; ++k) {
    float* s_p = s_pp[k & 1u];
    reg.stage_ball(*cb.u, s_wc, s_rhs, s_p);
    __syncthreads();
    const bool more = k + 1u < nmine;
    BatchBrick nbk;
    int nx0 = 0, ny0 = 0, nz0 = 0;
    if (more) {
      const uint32_t ne = k + 1u < (uint32_t)NT ? s_list[k + 1u] : entries[e0 + (k + 1u) * per_run];
      if ((ne >> CS_BATCH_ENTRY_SCENE_SHIFT) == (entry >> CS_BATCH_ENTRY_SCENE_SHIFT)) {
        nbk = cb;  // same scene: only the brick changes
        nbk.brick = ne & ((1u << CS_BATCH_ENTRY_SCENE_SHIFT) - 1u);
      } else {
        batch_brick(sc, ne, cur, nbk);
      }
      entry = ne;
      brick_origin(*nbk.u, nbk.brick, nx0, ny0, nz0);
      nx0 -= H;
      ny0 -= H;
      nz0 -= H;
      reg.fetch_ball(*nbk.u, nx0, ny0, nz0, nbk.wc, nbk.rhs, nbk.pin);
    }
    const uint32_t fires = sweep_listed<H, NT>(*cb.u, codes, s_wc, s_rhs, s_p);
    store_brick<H, NT>(*cb.u, x0, y0, z0, s_p, cb.pout);
    if (fires) atomicAdd(&cb.metrics[CS_M_P_CLAMP], fires);
    if (!more) break;
    // no barrier: the next staging writes the other pressure buffer
    cb = nbk;
    x0 = nx0;
    y0 = ny0;
    z0 = nz0;
  }
}

// Per-entry residual partials of buffer `cur` (a fixed grid striding over the list).
__global__ void __launch_bounds__(256) k_batch_residual(const cs_scene_dev* __restrict__ sc,
                                                        const uint32_t* __restrict__ entries,
                                                        const cs_batch_solve* __restrict__ bs, uint32_t cur,
                                                        double* __restrict__ partials) {
  constexpr int R3 = (BR + 2) * (BR + 2) * (BR + 2), NT = 256;
  __shared__ float4 s_wc[R3];
  __shared__ int32_t s_p[R3];
  __shared__ double s_rr[NT], s_bb[NT];
  const uint32_t total = bs->total;
  for (uint32_t e = blockIdx.x; e < total; e += gridDim.x) {
    const uint32_t entry = entries[e];
    const cs_scene_dev& d = sc[entry >> CS_BATCH_ENTRY_SCENE_SHIFT];
    const cs_params u = d.u;
    brick_residual_sums<NT>(u, entry & ((1u << CS_BATCH_ENTRY_SCENE_SHIFT) - 1u),
                            reinterpret_cast<const float4*>(d.wc), d.rhs, cur ? d.pres2 : d.pres, s_wc, s_p, s_rr,
                            s_bb);
    if (threadIdx.x == 0) {
      partials[2 * e] = s_rr[0];
      partials[2 * e + 1] = s_bb[0];
    }
  }
}

// One workgroup per scene: the scene's partials reduced as k_residual_reduce
// reduces them, then its solve state advanced.
__global__ void __launch_bounds__(1024) k_batch_residual_reduce(const cs_scene_dev* __restrict__ sc,
                                                                const uint32_t* __restrict__ first,
                                                                const uint32_t* __restrict__ count,
                                                                const double* __restrict__ partials, uint32_t mode,
                                                                uint32_t iters, uint32_t launches) {
  __shared__ double s_rr[1024], s_bb[1024];
  const uint32_t s = blockIdx.x;
  cs_solve_state* st = sc[s].st;
  if (mode != 0u && st->done) return;
  reduce_partials_block(partials + 2u * (size_t)first[s], count[s], s_rr, s_bb);
  if (threadIdx.x == 0) finalize_state(st, s_rr[0], s_bb[0], mode, iters, launches, sc[s].tol);
}

// The solved pressure into pres (from the second buffer when the launch count
// is odd), and into prev_p for scenes that start the next solve from it
// (k_select_final + the copy after the single-scene solve).
__global__ void __launch_bounds__(256) k_batch_finalize(const cs_scene_dev* __restrict__ sc,
                                                        const uint32_t* __restrict__ starts, uint32_t n,
                                                        uint32_t by_state, uint32_t odd_all) {
  CS_BATCH_SCENE(CS_SPACE_CELLS);
  const uint32_t idx = blk * 256u + threadIdx.x;
  if (idx >= u.nc) return;
  const bool odd = by_state ? (d.st->launches & 1u) != 0u : odd_all != 0u;
  uint32_t ix, iy, iz;
  decompose(u, idx, ix, iy, iz);
  const uint32_t brick = ((iz / BR) * u.nby + iy / BR) * u.nbx + ix / BR;
  int32_t v = d.pres[idx];
  if (odd && d.brick_flag[brick]) {
    v = d.pres2[idx];
    d.pres[idx] = v;
  }
  if (d.warm_previous) d.prev_p[idx] = v;
}

__global__ void k_batch_gather_states(const cs_scene_dev* __restrict__ sc, uint32_t n,
                                      cs_solve_state* __restrict__ out, uint32_t* __restrict__ fluid) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    out[s] = *sc[s].st;
    fluid[s] = sc[s].metrics[CS_M_FLUID_CELLS];
  }
}

// The solve state into the substep's record (the statistics of a replayed
// substep) and the block's halt word into `halt`; both may be host memory.
__global__ void k_solve_record(const cs_dparams* __restrict__ dp, const cs_solve_state* __restrict__ st,
                               cs_solve_state* __restrict__ rec, uint32_t* __restrict__ halt, uint32_t n) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint32_t h = dp->halt;
  *halt = h;
  if (h || dp->slot >= n) return;
  rec[dp->slot] = *st;
}

// ---- launchers ---------------------------------------------------------------------------
hipError_t cs_launch_brick_scan(hipStream_t s, const cs_pblock& P, const uint32_t* flags, uint32_t* list,
                                cs_solve_state* st) {
  hipLaunchKernelGGL(k_brick_scan, dim3(1), dim3(1024), 0, s, P.d, flags, P.u.nbricks, list, st);
  return hipGetLastError();
}

hipError_t cs_launch_solve_record(hipStream_t s, const cs_dparams* d, const cs_solve_state* st, cs_solve_state* rec,
                                  uint32_t* halt, uint32_t n) {
  hipLaunchKernelGGL(k_solve_record, dim3(1), dim3(64), 0, s, d, st, rec, halt, n);
  return hipGetLastError();
}

hipError_t cs_launch_face_weights(hipStream_t s, const cs_pblock& P, const float* kf, float* wc) {
  const cs_params& u = P.u;
  hipLaunchKernelGGL(k_face_weights, dim3((u.nc + 255u) / 256u), dim3(256), 0, s, P.d, kf,
                     reinterpret_cast<float4*>(wc));
  return hipGetLastError();
}

hipError_t cs_launch_rbgs_bricks(hipStream_t s, const cs_params& u, uint32_t nactive, int halo, int threads,
                                 uint32_t color0, const uint32_t* list, const float* wc, const float* rhs,
                                 const int32_t* pin, int32_t* pout, uint32_t* metrics, const cs_solve_state* st) {
  if (nactive == 0) return hipSuccess;
  const float4* w4 = reinterpret_cast<const float4*>(wc);
  const dim3 g(nactive);
#define CS_BRICK_LAUNCH(H, NT) \
  hipLaunchKernelGGL((k_rbgs_brick<H, NT>), g, dim3(NT), 0, s, u, list, color0, w4, rhs, pin, pout, metrics, st)
  if (halo == 1 && threads == 256) CS_BRICK_LAUNCH(1, 256);
  else if (halo == 2 && threads == 256) CS_BRICK_LAUNCH(2, 256);
  else if (halo == 2 && threads == 512) CS_BRICK_LAUNCH(2, 512);
  else if (halo == 4 && threads == 256) CS_BRICK_LAUNCH(4, 256);
  else if (halo == 4 && threads == 512) CS_BRICK_LAUNCH(4, 512);
  else return hipErrorInvalidValue;
#undef CS_BRICK_LAUNCH
  return hipGetLastError();
}

hipError_t cs_launch_rbgs_multi(hipStream_t s, const cs_params& u, uint32_t nactive, uint32_t groups,
                                uint32_t color0, const uint32_t* list, const float* wc, const float* rhs,
                                const int32_t* pin, int32_t* pout, uint32_t* metrics, cs_solve_state* st,
                                double* partials, uint32_t res_mode, uint32_t chunk_iters, uint32_t chunk_launches,
                                double tol, const uint32_t* sweep_table) {
  if (nactive == 0) return hipSuccess;
  if (groups >= 8u) groups &= ~7u;  // whole runs
  if (groups == 0u || groups > 1024u) return hipErrorInvalidValue;
  if ((res_mode & CS_RES_COMPUTE) && (res_mode & CS_RES_DECIDE)) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_rbgs_multi<CS_MULTI_HALO, CS_MULTI_THREADS>), dim3(groups), dim3(CS_MULTI_THREADS), 0, s, u, list, nactive, color0,
                     reinterpret_cast<const float4*>(wc), rhs, pin, pout, metrics, st, partials, res_mode,
                     chunk_iters, chunk_launches, tol, sweep_table);
  return hipGetLastError();
}

void cs_build_sweep_table(uint32_t* out) {
  constexpr int H = CS_MULTI_HALO, NT = CS_MULTI_THREADS, R = BR + 2 * H;
  auto dist = [](int l) { return std::max(0, std::max(H - l, l - (H + BR - 1))); };
  for (int c = 0; c < 2; ++c) {
    for (int sw = 0; sw < H; ++sw) {
      uint32_t* row = out + (size_t)(c * H + sw) * NT;
      std::fill(row, row + NT, 0u);
      int n = 0;
      for (int lz = 0; lz < R; ++lz)
        for (int ly = 0; ly < R; ++ly)
          for (int lx = 0; lx < R; ++lx) {
            const int d = dist(lx) + dist(ly) + dist(lz);
            if (d > H - 1 - sw || ((lx + ly + lz) & 1) != c) continue;
            const uint32_t inner = d == 0 ? 1u : 0u;
            if (n < NT) row[n] = (uint32_t)((lz * R + ly) * R + lx) | inner << 17 | 1u << 19;
            ++n;
          }
      if (n > NT) std::abort();  // cannot happen for H = 4: at most 980 cells per list
    }
  }
}

static hipError_t residual_bricks(hipStream_t s, const cs_pblock& P, uint32_t nactive, const uint32_t* list,
                                  const float* wc, const float* rhs, const int32_t* p, double* partials,
                                  cs_solve_state* st, uint32_t mode, uint32_t iters, uint32_t launches, double tol,
                                  uint32_t* gate, uint32_t gate_mode) {
  if (nactive > 0) {
    hipLaunchKernelGGL(k_residual_bricks, dim3(nactive), dim3(256), 0, s, P.d, list,
                       reinterpret_cast<const float4*>(wc), rhs, p, partials, st, mode);
    hipLaunchKernelGGL(k_residual_reduce, dim3(1), dim3(1024), 0, s, partials, st, mode, iters, launches, tol, gate,
                       gate_mode);
  } else
    hipLaunchKernelGGL(k_residual_empty, dim3(1), dim3(64), 0, s, st, mode, iters, launches, tol, gate, gate_mode);
  return hipGetLastError();
}
hipError_t cs_launch_residual_bricks(hipStream_t s, const cs_pblock& P, uint32_t nactive, const uint32_t* list,
                                     const float* wc, const float* rhs, const int32_t* p, double* partials,
                                     cs_solve_state* st, uint32_t mode, uint32_t iters, uint32_t launches,
                                     double tol) {
  return residual_bricks(s, P, nactive, list, wc, rhs, p, partials, st, mode, iters, launches, tol, nullptr, 0u);
}
hipError_t cs_launch_residual_gate(hipStream_t s, const cs_pblock& P, uint32_t nactive, const uint32_t* list,
                                   const float* wc, const float* rhs, const int32_t* p, double* partials,
                                   cs_solve_state* st, double tol, uint32_t* flags, bool first, bool last) {
  return residual_bricks(s, P, nactive, list, wc, rhs, p, partials, st, 0u, 0u, 0u, tol, flags,
                         (first ? 1u : 0u) | (last ? 2u : 0u));
}

hipError_t cs_launch_rescale(hipStream_t s, const cs_params& u, uint32_t nactive, const uint32_t* list,
                             const float* wc, const float* rhs, const float* kf, int32_t* p, double* partials,
                             cs_solve_state* st, uint32_t* metrics) {
  if (nactive == 0) return hipSuccess;
  hipLaunchKernelGGL(k_rescale_bricks, dim3(nactive), dim3(256), 0, s, u, list, reinterpret_cast<const float4*>(wc),
                     rhs, p, partials, st);
  hipLaunchKernelGGL(k_rescale_reduce, dim3(1), dim3(1024), 0, s, partials, nactive, st);
  hipLaunchKernelGGL(k_rescale_apply, dim3((u.nc + 255u) / 256u), dim3(256), 0, s, u, kf, p, st, metrics);
  return hipGetLastError();
}

hipError_t cs_launch_select_final(hipStream_t s, const cs_params& u, uint32_t nactive, const uint32_t* list,
                                  int32_t* a, const int32_t* b, const cs_solve_state* st, bool by_state) {
  if (nactive == 0) return hipSuccess;
  hipLaunchKernelGGL(k_select_final, dim3(nactive), dim3(256), 0, s, u, list, a, b, st, by_state ? 1u : 0u);
  return hipGetLastError();
}

// ---- batched launchers --------------------------------------------------------------------
hipError_t cs_batch_launch_solve_setup(const cs_batch_launch& b) {
  hipLaunchKernelGGL(k_batch_scan, dim3(b.n), dim3(1024), 0, b.s, b.sc, b.starts, b.n);
  const uint32_t cb = cs_batch_blocks(b, CS_SPACE_CELLS);
  if (cb) hipLaunchKernelGGL(k_batch_face_warm, dim3(cb), dim3(256), 0, b.s, b.sc, b.starts, b.n);
  return hipGetLastError();
}

hipError_t cs_batch_launch_compact(const cs_batch_launch& b, bool skip_done, uint32_t* entries, uint32_t* first,
                                   uint32_t* count, cs_batch_solve* bs) {
  hipLaunchKernelGGL(k_batch_compact, dim3(b.n), dim3(256), 0, b.s, b.sc, b.n, skip_done ? 1u : 0u, entries, first,
                     count, bs);
  return hipGetLastError();
}

hipError_t cs_batch_launch_sweeps(const cs_batch_launch& b, const uint32_t* entries, const cs_batch_solve* bs,
                                  uint32_t groups, uint32_t cur, uint32_t color0, const uint32_t* sweep_table) {
  if (groups >= 8u) groups &= ~7u;  // whole runs
  if (groups == 0u || groups > 4096u) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_batch_rbgs<CS_MULTI_HALO, CS_MULTI_THREADS>), dim3(groups), dim3(CS_MULTI_THREADS), 0, b.s,
                     b.sc, entries, bs, cur, color0, sweep_table);
  return hipGetLastError();
}

hipError_t cs_batch_launch_residual(const cs_batch_launch& b, const uint32_t* entries, const uint32_t* first,
                                    const uint32_t* count, const cs_batch_solve* bs, uint32_t groups, uint32_t cur,
                                    double* partials, uint32_t mode, uint32_t iters, uint32_t launches) {
  if (groups == 0u) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_batch_residual, dim3(groups), dim3(256), 0, b.s, b.sc, entries, bs, cur, partials);
  hipLaunchKernelGGL(k_batch_residual_reduce, dim3(b.n), dim3(1024), 0, b.s, b.sc, first, count, partials, mode, iters,
                     launches);
  return hipGetLastError();
}

hipError_t cs_batch_launch_finalize(const cs_batch_launch& b, bool by_state, bool odd) {
  const uint32_t cb = cs_batch_blocks(b, CS_SPACE_CELLS);
  if (cb)
    hipLaunchKernelGGL(k_batch_finalize, dim3(cb), dim3(256), 0, b.s, b.sc, b.starts, b.n, by_state ? 1u : 0u,
                       odd ? 1u : 0u);
  return hipGetLastError();
}

hipError_t cs_batch_launch_gather_states(const cs_batch_launch& b, cs_solve_state* out, uint32_t* fluid) {
  hipLaunchKernelGGL(k_batch_gather_states, dim3((b.n + 255u) / 256u), dim3(256), 0, b.s, b.sc, b.n, out, fluid);
  return hipGetLastError();
}
