// cs_solver.hip — brick-sparse red-black Gauss-Seidel pressure solve.
//
// Restates pressure_update (shader.rs:1371-1439) over the fluid cells only.
// The grid is tiled into 8x8x8 bricks; classify flags the bricks holding a
// fluid cell and a single-block scan lists them in ascending order.  One
// workgroup per active brick loads the brick plus an H-cell halo of pressure,
// kind/fill word and right-hand side into LDS and runs H half sweeps there:
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

// One update of pressure_update (shader.rs:1386-1438) on LDS-resident data,
// branch-free: all six neighbour words are read up front and the skips of
// the reference become selects.  Adding +0 to a sum that starts at +0 and
// only ever accumulates positive weights or nonzero products leaves it
// unchanged, so the result equals the branching form bit for bit.
__device__ __forceinline__ int32_t update_cell(const cs_params& u, float skf, float srhs, const int2* s_pk, int li,
                                               int sx, int sy, bool& fired) {
  const float sf = kf_fill(skf);
  const bool sbed = kf_bed(skf);
  const int off[6] = {-1, 1, -sx, sx, -sy, sy};
  int2 nb[6];
#pragma unroll
  for (int n = 0; n < 6; ++n) nb[n] = s_pk[li + off[n]];
  float count = 0.0f, sum = 0.0f;
#pragma unroll
  for (int n = 0; n < 6; ++n) {
    const float nkf = __int_as_float(nb[n].y);
    const bool nfluid = nkf >= 0.0f;
    const bool skip = (nkf == KF_OFFGRID) | (nkf == CS_KF_SOLID);
    const float nfill = nkf == CS_KF_BED ? 1.0f : nkf;
    float w = cs_min(sf, nfill);
    w = nfluid ? w : sf;
    w = (sbed | (nkf == CS_KF_BED)) ? 1.0f : w;
    const bool use = !skip && (w > 0.0f);
    count += use ? w : 0.0f;
    const float term = w * ((float)nb[n].x * u.inv_fp_scale);
    sum += (use && nfluid) ? term : 0.0f;
  }
  float pn = 0.0f;
  if (count > 0.0f) pn = (sum - srhs) / count;
  const float cl = cs_clamp(pn, -u.pressure_clamp, u.pressure_clamp);
  fired = cl != pn;
  return cs_f2i(cl * u.fp_scale);
}

}  // namespace

// ---- active brick list ----------------------------------------------------------------
// Single block: ascending list of flagged bricks and their count.
__global__ void __launch_bounds__(1024) k_brick_scan(const uint32_t* __restrict__ flags, uint32_t nbricks,
                                                     uint32_t* __restrict__ list, cs_solve_state* __restrict__ st) {
  __shared__ uint32_t s_cnt[1024];
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

template <int H>
__global__ void __launch_bounds__(256) k_rbgs_brick(cs_params u, const uint32_t* __restrict__ list,
                                                    uint32_t color0, const float* __restrict__ kf,
                                                    const float* __restrict__ rhs, const int32_t* __restrict__ pin,
                                                    int32_t* __restrict__ pout, uint32_t* __restrict__ metrics,
                                                    const cs_solve_state* __restrict__ st) {
  constexpr int R = BR + 2 * H;  // even: BR is 8
  constexpr int R3 = R * R * R;
  constexpr int NT = 256;
  constexpr int PER = (R3 + NT - 1) / NT;
  __shared__ int2 s_pk[R3];  // (pressure fixed point, kind/fill word)
  __shared__ float s_rhs[R3];
  if (st->done) return;
  int x0, y0, z0;
  brick_origin(u, list[blockIdx.x], x0, y0, z0);
  x0 -= H;
  y0 -= H;
  z0 -= H;
  const uint32_t base_par = (uint32_t)(x0 + y0 + z0) & 1u;
  {
    float kfv[PER], rv[PER];
    int32_t pv[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int i = (int)threadIdx.x + k * NT;
      kfv[k] = KF_OFFGRID;
      rv[k] = 0.0f;
      pv[k] = 0;
      if (i < R3) {
        const int lx = i % R, ly = (i / R) % R, lz = i / (R * R);
        const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
        if (gx >= 0 && gy >= 0 && gz >= 0 && gx < (int)u.gx && gy < (int)u.gy && gz < (int)u.gz) {
          const uint32_t idx = ((uint32_t)gz * u.gy + (uint32_t)gy) * u.gx + (uint32_t)gx;
          kfv[k] = kf[idx];
          rv[k] = rhs[idx];
          pv[k] = pin[idx];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int i = (int)threadIdx.x + k * NT;
      if (i < R3) {
        s_pk[i] = make_int2(pv[k], __float_as_int(kfv[k]));
        s_rhs[i] = rv[k];
      }
    }
  }
  __syncthreads();
  uint32_t fires = 0;
#pragma unroll
  for (int s = 0; s < H; ++s) {
    // cells of this half sweep's colour at least s+1 cells inside the region
    const uint32_t color = color0 ^ (uint32_t)(s & 1);
    const int lo = s + 1;
    const int span = R - 2 * (s + 1);  // even
    const int half = span / 2;
    const int n = half * span * span;
    for (int j = (int)threadIdx.x; j < n; j += NT) {
      const int m = j % half;
      const int ly = lo + (j / half) % span;
      const int lz = lo + j / (half * span);
      const int lx = lo + 2 * m + (int)((color + base_par + (uint32_t)(ly + lz + lo)) & 1u);
      const int li = (lz * R + ly) * R + lx;
      const float skf = __int_as_float(s_pk[li].y);
      if (!(skf >= 0.0f)) continue;
      bool fired;
      const int32_t np = update_cell(u, skf, s_rhs[li], s_pk, li, R, R * R, fired);
      s_pk[li].x = np;
      const bool interior = lx >= H && ly >= H && lz >= H && lx < H + BR && ly < H + BR && lz < H + BR;
      if (fired && interior) ++fires;  // each brick cell once, as the whole-grid sweep counts it
    }
    __syncthreads();
  }
  for (int j = threadIdx.x; j < BR * BR * BR; j += blockDim.x) {
    int lx = H + j % BR, ly = H + (j / BR) % BR, lz = H + j / (BR * BR);
    int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
    if (gx < (int)u.gx && gy < (int)u.gy && gz < (int)u.gz)
      pout[((uint32_t)gz * u.gy + (uint32_t)gy) * u.gx + (uint32_t)gx] = s_pk[(lz * R + ly) * R + lx].x;
  }
  if (fires) atomicAdd(&metrics[CS_M_P_CLAMP], fires);
}

// ---- residual over the active bricks ------------------------------------------------------
// One workgroup per brick stages the brick and a one-cell halo of (pressure,
// kind/fill) in LDS, forms b - A p for its fluid cells in double precision
// and writes a per-brick partial; the last workgroup to finish sums the
// partials in brick order (a fixed order) and advances the solve state as
// k_residual_finalize does.
__device__ __forceinline__ void finalize_state(cs_solve_state* st, double rr, double bb, uint32_t mode,
                                               uint32_t iters, uint32_t launches, double tol) {
  double r;
  if (bb > 0.0) r = sqrt(rr / bb);
  else r = rr > 0.0 ? __builtin_huge_val() : 0.0;
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

__global__ void __launch_bounds__(256) k_residual_bricks(cs_params u, const uint32_t* __restrict__ list,
                                                         const float* __restrict__ kf, const float* __restrict__ rhs,
                                                         const int32_t* __restrict__ p, double* __restrict__ partials,
                                                         cs_solve_state* __restrict__ st, uint32_t mode,
                                                         uint32_t iters, uint32_t launches, double tol) {
  constexpr int R = BR + 2, R3 = R * R * R, NT = 256, PER = (R3 + NT - 1) / NT;
  __shared__ int2 s_pk[R3];
  __shared__ double s_rr[NT], s_bb[NT];
  __shared__ uint32_t s_last;
  if (mode != 0u && st->done) return;
  int x0, y0, z0;
  brick_origin(u, list[blockIdx.x], x0, y0, z0);
  x0 -= 1;
  y0 -= 1;
  z0 -= 1;
  {
    float kfv[PER];
    int32_t pv[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int i = (int)threadIdx.x + k * NT;
      kfv[k] = KF_OFFGRID;
      pv[k] = 0;
      if (i < R3) {
        const int lx = i % R, ly = (i / R) % R, lz = i / (R * R);
        const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
        if (gx >= 0 && gy >= 0 && gz >= 0 && gx < (int)u.gx && gy < (int)u.gy && gz < (int)u.gz) {
          const uint32_t idx = ((uint32_t)gz * u.gy + (uint32_t)gy) * u.gx + (uint32_t)gx;
          kfv[k] = kf[idx];
          pv[k] = p[idx];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int i = (int)threadIdx.x + k * NT;
      if (i < R3) s_pk[i] = make_int2(pv[k], __float_as_int(kfv[k]));
    }
  }
  __syncthreads();
  double rr = 0.0, bb = 0.0;
  for (int j = (int)threadIdx.x; j < BR * BR * BR; j += NT) {
    const int lx = 1 + j % BR, ly = 1 + (j / BR) % BR, lz = 1 + j / (BR * BR);
    const int li = (lz * R + ly) * R + lx;
    const int2 self = s_pk[li];
    const float skf = __int_as_float(self.y);
    if (!(skf >= 0.0f)) continue;
    const float sf = kf_fill(skf);
    const bool sbed = kf_bed(skf);
    const int off[6] = {-1, 1, -R, R, -R * R, R * R};
    int2 nb[6];
#pragma unroll
    for (int n = 0; n < 6; ++n) nb[n] = s_pk[li + off[n]];
    float count = 0.0f;
    double ap_n = 0.0;
#pragma unroll
    for (int n = 0; n < 6; ++n) {
      const float nkf = __int_as_float(nb[n].y);
      const bool nfluid = nkf >= 0.0f;
      const bool skip = (nkf == KF_OFFGRID) | (nkf == CS_KF_SOLID);
      float w = cs_min(sf, nkf == CS_KF_BED ? 1.0f : nkf);
      w = nfluid ? w : sf;
      w = (sbed | (nkf == CS_KF_BED)) ? 1.0f : w;
      const bool use = !skip && (w > 0.0f);
      count += use ? w : 0.0f;
      const double t = (double)w * (double)((float)nb[n].x * u.inv_fp_scale);
      ap_n += (use && nfluid) ? t : 0.0;
    }
    if (count <= 0.0f) continue;
    const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
    const double b = -(double)rhs[((uint32_t)gz * u.gy + (uint32_t)gy) * u.gx + (uint32_t)gx];
    const double ap = (double)count * (double)((float)self.x * u.inv_fp_scale) - ap_n;
    rr += (b - ap) * (b - ap);
    bb += b * b;
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
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = s_rr[0];
    partials[2 * blockIdx.x + 1] = s_bb[0];
    __threadfence();
    const uint32_t prev = __hip_atomic_fetch_add(&st->arrivals, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    s_last = prev + 1u == gridDim.x ? 1u : 0u;
  }
  __syncthreads();
  if (!s_last) return;
  // last workgroup: fixed-order sum of the per-brick partials
  rr = 0.0;
  bb = 0.0;
  for (uint32_t i = threadIdx.x; i < gridDim.x; i += NT) {
    rr += __hip_atomic_load(&partials[2 * i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    bb += __hip_atomic_load(&partials[2 * i + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
  if (threadIdx.x == 0) {
    finalize_state(st, s_rr[0], s_bb[0], mode, iters, launches, tol);
    st->arrivals = 0u;
  }
}

// The residual of a system with no fluid cell (r = 0: nothing to solve).
__global__ void k_residual_empty(cs_solve_state* __restrict__ st, uint32_t mode, uint32_t iters, uint32_t launches,
                                 double tol) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (mode != 0u && st->done) return;
  finalize_state(st, 0.0, 0.0, mode, iters, launches, tol);
}

// Copies the brick cells of the second buffer back to the first when an odd
// number of sweep launches ran, so the solved pressure always ends in `a`.
__global__ void __launch_bounds__(256) k_select_final(cs_params u, const uint32_t* __restrict__ list,
                                                      int32_t* __restrict__ a, const int32_t* __restrict__ b,
                                                      const cs_solve_state* __restrict__ st) {
  if ((st->launches & 1u) == 0u) return;
  int x0, y0, z0;
  brick_origin(u, list[blockIdx.x], x0, y0, z0);
  for (int j = threadIdx.x; j < BR * BR * BR; j += blockDim.x) {
    int ix = x0 + j % BR, iy = y0 + (j / BR) % BR, iz = z0 + j / (BR * BR);
    if (ix >= (int)u.gx || iy >= (int)u.gy || iz >= (int)u.gz) continue;
    uint32_t idx = ((uint32_t)iz * u.gy + (uint32_t)iy) * u.gx + (uint32_t)ix;
    a[idx] = b[idx];
  }
}

__global__ void k_state_init(cs_solve_state* st) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    st->done = 0u;
    st->iterations = 0u;
    st->launches = 0u;
    st->arrivals = 0u;
    st->residual = 0.0;
    st->initial_residual = 0.0;
  }
}

// ---- launchers ---------------------------------------------------------------------------
hipError_t cs_launch_brick_scan(hipStream_t s, const cs_params& u, const uint32_t* flags, uint32_t* list,
                                cs_solve_state* st) {
  hipLaunchKernelGGL(k_state_init, dim3(1), dim3(64), 0, s, st);
  hipLaunchKernelGGL(k_brick_scan, dim3(1), dim3(1024), 0, s, flags, u.nbricks, list, st);
  return hipGetLastError();
}

hipError_t cs_launch_rbgs_bricks(hipStream_t s, const cs_params& u, uint32_t nactive, int halo, uint32_t color0,
                                 const uint32_t* list, const float* kf, const float* rhs, const int32_t* pin,
                                 int32_t* pout, uint32_t* metrics, const cs_solve_state* st) {
  if (nactive == 0) return hipSuccess;
  dim3 g(nactive), b(256);
  switch (halo) {
    case 1: hipLaunchKernelGGL(k_rbgs_brick<1>, g, b, 0, s, u, list, color0, kf, rhs, pin, pout, metrics, st); break;
    case 2: hipLaunchKernelGGL(k_rbgs_brick<2>, g, b, 0, s, u, list, color0, kf, rhs, pin, pout, metrics, st); break;
    case 4: hipLaunchKernelGGL(k_rbgs_brick<4>, g, b, 0, s, u, list, color0, kf, rhs, pin, pout, metrics, st); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t cs_launch_residual_bricks(hipStream_t s, const cs_params& u, uint32_t nactive, const uint32_t* list,
                                     const float* kf, const float* rhs, const int32_t* p, double* partials,
                                     cs_solve_state* st, uint32_t mode, uint32_t iters, uint32_t launches,
                                     double tol) {
  if (nactive > 0)
    hipLaunchKernelGGL(k_residual_bricks, dim3(nactive), dim3(256), 0, s, u, list, kf, rhs, p, partials, st, mode,
                       iters, launches, tol);
  else
    hipLaunchKernelGGL(k_residual_empty, dim3(1), dim3(64), 0, s, st, mode, iters, launches, tol);
  return hipGetLastError();
}

hipError_t cs_launch_select_final(hipStream_t s, const cs_params& u, uint32_t nactive, const uint32_t* list,
                                  int32_t* a, const int32_t* b, const cs_solve_state* st) {
  if (nactive == 0) return hipSuccess;
  hipLaunchKernelGGL(k_select_final, dim3(nactive), dim3(256), 0, s, u, list, a, b, st);
  return hipGetLastError();
}
