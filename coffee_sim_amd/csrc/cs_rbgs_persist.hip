// cs_rbgs_persist.hip — the whole red-black Gauss-Seidel pressure solve of a
// substep (pressure_update, shader.rs:1371-1453, looped as mod.rs:352-357 or to
// a residual tolerance) in ONE launch of co-resident workgroups.
//
// Why one launch: an iteration of this solve moves a few MB (about 185k fluid
// cells at the working size), far less than a kernel boundary costs at 8 TB/s,
// so a solve made of launches is bound by launch and first-load latency, not
// by HBM.  Device-wide barriers with acquire/release fences are no cheaper on
// a chip of 8 L2s (each release writes the XCD's L2 back, each acquire
// invalidates it).  Instead, each workgroup keeps one active 8^3 brick
// resident for the whole solve and talks only to the (at most six) workgroups
// of its face-neighbour bricks:
//
//  - Registers hold the static part of its two cells' rows (face weights,
//    diagonal, right-hand side): loaded once.
//  - LDS holds the brick's pressure with a one-cell halo (10^3 floats).
//  - Per half sweep of colour c: every thread updates its colour-c cell
//    exactly as pressure_update does (same operation order, so the result is
//    the reference's bit for bit).  A cell on the brick's surface is also
//    published as one 64-bit word (fixed-point pressure, epoch) with a
//    device-coherent store into the brick's face buffer (per face and colour
//    32 contiguous words, so a face moves in a few 64-byte segments); a halo
//    cell is read by polling the neighbour's word until it carries the
//    current epoch.  The data is its own flag, so a half sweep
//    costs one store-to-load latency and one barrier.  The next half sweep
//    reads only colour-c neighbours (red-black), so one word per cell is
//    enough: a neighbour rewrites colour c only at epoch e+2, after it has
//    read this brick's epoch-e+1 surface, which this brick publishes only
//    after its reads of epoch e.  Epochs grow across solves (a per-simulation
//    base), so words left by an earlier solve are never taken as current.
//  - Device-coherent here means relaxed agent-scope atomics (global loads and
//    stores with the device-scope cache policy): no L2 write-back or
//    invalidate is needed.
//  - Every `every` iterations (and at the cap) each workgroup sums its cells'
//    residual rows in double, publishes (rr, bb) and reads everyone's sums in
//    the same order, so all workgroups reach the same verdict and stop at the
//    same iteration, exactly where the oracle's loop stops.
//
// Every wait is bounded by a wall-clock limit; a workgroup that times out
// raises an abort flag that every other wait also watches, so the launch
// always drains.  The host reports the failure (it cannot happen while the
// launch is co-resident, which the cooperative launch guarantees).
#include "cs_device.h"
#include "cs_persist.h"

namespace {

constexpr int BR = CS_BRICK;           // brick side (8)
constexpr int PR = BR + 2;             // staged side: brick plus one halo cell
constexpr int PR2 = PR * PR;
constexpr int PR3 = PR * PR * PR;
constexpr int NT = CS_PERSIST_THREADS;  // 256: one cell of each colour per thread
static_assert(NT * 2 == BR * BR * BR, "two cells per thread");

__device__ __forceinline__ uint32_t ld_dev(const uint32_t* p) {
  return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t ld_dev(const int32_t* p) {
  return __hip_atomic_load(const_cast<int32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_dev(const uint64_t* p) {
  return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(int32_t* p, int32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Coherent within one XCD: a non-temporal store, and a load with the sc0 and
// nt policies, which misses the CU's L1 and is served by the XCD's shared L2.
// (A plain sc0 load can hit a stale line in the reading CU's L1.)  Used only
// between workgroups that read the same XCC id; `r` covers the face buffers.
__device__ __forceinline__ uint64_t ld_xcd(__amdgpu_buffer_rsrc_t r, int32_t word) {
  return __builtin_bit_cast(uint64_t, __builtin_amdgcn_raw_buffer_load_b64(r, word * 8, 0, 3 /* sc0 | nt */));
}
__device__ __forceinline__ void st_xcd(__amdgpu_buffer_rsrc_t r, int32_t word, uint64_t v) {
  typedef __attribute__((ext_vector_type(2))) unsigned u2;
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, word * 8, 0, 2 /* nt */);
}
// s_getreg of HW_REG_XCC_ID (20), bits [3:0]
__device__ __forceinline__ uint32_t xcc_id() { return __builtin_amdgcn_s_getreg(20 | (3 << 11)); }

// Waits until every memory operation this wavefront issued has completed (the
// stores have reached the device-coherent level).  The empty asm statements
// keep the compiler from moving memory accesses across the wait.
__device__ __forceinline__ void stores_complete() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0);
  asm volatile("" ::: "memory");
}

// 100 MHz constant clock
__device__ __forceinline__ uint64_t now_ticks() { return __builtin_amdgcn_s_memrealtime(); }

// One update of pressure_update (shader.rs:1386-1438) from the precomputed
// face weights: `self` = (+x, +y, +z weights, diagonal count); the minus-side
// weights are the neighbours' plus-side ones.  Same order as the reference
// (-x, +x, -y, +y, -z, +z); an unused face has weight 0 and adds +-0.
__device__ __forceinline__ float update(const cs_params& u, float4 self, float srhs, float wxm, float wym, float wzm,
                                        const float* s_p, int li, bool& fired) {
  float sum = wxm * s_p[li - 1];
  sum += self.x * s_p[li + 1];
  sum += wym * s_p[li - PR];
  sum += self.y * s_p[li + PR];
  sum += wzm * s_p[li - PR2];
  sum += self.z * s_p[li + PR2];
  const float quot = (sum - srhs) / self.w;
  const float pn = self.w > 0.0f ? quot : 0.0f;
  const float cl = fminf(fmaxf(pn, -u.pressure_clamp), u.pressure_clamp);
  fired = cl != pn;
  return truncf(cl * u.fp_scale) * u.inv_fp_scale;
}

// Sum over a wavefront in a fixed butterfly order (every lane gets the total).
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

__device__ __forceinline__ double rel_residual(double rr, double bb) {
  if (bb > 0.0) return sqrt(rr / bb);
  return rr > 0.0 ? __builtin_huge_val() : 0.0;
}

// Word k of colour c on face f (-x, +x, -y, +y, -z, +z) of a brick's face buffer.
__device__ __forceinline__ int face_word(uint32_t brick, int f, int c, int k) {
  return (int)(((brick * 6u + (uint32_t)f) * 2u + (uint32_t)c) * 32u + (uint32_t)k);
}

struct Cell {
  int li;        // index in the staged 10^3 region
  int32_t g;     // global cell index (-1: past the grid's edge)
  int fslot[3];  // per axis: its word in the face buffer of the brick face it lies on, or -1
  int fface[3];  // that face (0..5), or -1
  uint32_t floc; // bit j: the neighbour across face fface[j] shares this XCD
  float4 self;   // +x, +y, +z weights and diagonal (w < 0: not fluid)
  float rhs;     // c = dx^2 * rhs of the update
  float wxm, wym, wzm;
};

}  // namespace

// Workgroup w serves brick `run` of the active list; runs are dealt so that a
// contiguous part of the (ascending, hence z-major) list lands on one XCD
// (workgroup w runs on XCD w % 8).
__global__ void __launch_bounds__(NT, 4) k_rbgs_persist(cs_persist_args a) {
  CS_PARAMS_FROM(a.dp);
  const uint32_t t = threadIdx.x;
  const uint32_t nactive = a.st->active_bricks;
  // XCD slices of ceil(nactive / 8) runs; a launch sized for more bricks than
  // are active leaves its extra workgroups without a run
  const uint32_t per = (nactive + 7u) / 8u;
  if (blockIdx.x / 8u >= per) return;
  const uint32_t run = (blockIdx.x % 8u) * per + blockIdx.x / 8u;
  if (run >= nactive) return;
  const uint32_t epoch_base = a.dp->epoch_base;
  const uint32_t brick = a.list[run];
  const uint32_t bx = brick % u.nbx, by = (brick / u.nbx) % u.nby, bz = brick / (u.nbx * u.nby);
  const int x0 = (int)bx * BR, y0 = (int)by * BR, z0 = (int)bz * BR;

  __shared__ float s_p[PR3];
  __shared__ double s_red[2 * (NT / 64)];
  __shared__ uint32_t s_fail;
  __shared__ uint32_t s_verdict;  // 1: converged or capped, stop

  if (t == 0) s_fail = 0u;
  // ---- the two cells of this thread (colour c at x = 2m + ((y + z + c) & 1)) -----------------
  Cell cell[2];
  int hl[2];
  int32_t hg[2];
  int hf[2];     // face of the halo cell
  const int m = (int)(t & 3u), ly = (int)((t >> 2) & 7u), lz = (int)(t >> 5);
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int lx = 2 * m + ((ly + lz + c) & 1);  // brick corners are even, so parity = colour
    const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
    Cell& e = cell[c];
    e.li = ((lz + 1) * PR + ly + 1) * PR + lx + 1;
    {
      // the faces this cell lies on; word k = (b2 << 2) | (a1 >> 1) of face f, colour c,
      // where (a1, b2) are its in-face coordinates (the reader's halo mapping below)
      const int lc[3] = {lx, ly, lz};
#pragma unroll
      for (int axis = 0; axis < 3; ++axis) {
        const int a1 = lc[axis == 0 ? 1 : 0], b2 = lc[axis == 2 ? 1 : 2];
        const int k = (b2 << 2) | (a1 >> 1);
        e.fface[axis] = lc[axis] == 0 ? 2 * axis : lc[axis] == BR - 1 ? 2 * axis + 1 : -1;
        e.fslot[axis] = e.fface[axis] >= 0 ? face_word(brick, e.fface[axis], c, k) : -1;
      }
    }
    // a brick on the grid's far edge may reach past it: such cells are not fluid
    const bool on = gx < (int)u.gx && gy < (int)u.gy && gz < (int)u.gz;
    e.g = on ? (int32_t)(((uint32_t)gz * u.gy + (uint32_t)gy) * u.gx + (uint32_t)gx) : -1;
    e.self = on ? a.wc[e.g] : make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    e.rhs = on ? a.rhs[e.g] : 0.0f;
    e.wxm = on && gx > 0 ? a.wc[e.g - 1].x : 0.0f;
    e.wym = on && gy > 0 ? a.wc[e.g - (int32_t)u.gx].y : 0.0f;
    e.wzm = on && gz > 0 ? a.wc[e.g - (int32_t)(u.gx * u.gy)].z : 0.0f;
    // this thread's halo cell of colour c (threads 0..191: 32 per face) when it is
    // fluid, hence in an active brick whose workgroup publishes it every epoch
    hl[c] = -1;
    hg[c] = -1;
    hf[c] = 0;
    e.floc = 0u;
    if (t < 192u) {
      const int f = (int)(t >> 5), k = (int)(t & 31u);
      const int axis = f >> 1, side = f & 1;
      const int fc = side ? BR : -1;   // halo coordinate along the face axis
      const int b2 = k >> 2;           // second in-face coordinate
      const int a1 = 2 * (k & 3) + ((c + (fc & 1) + b2) & 1);  // parity (fc + a1 + b2) = c
      int l[3];
      l[axis] = fc;
      l[axis == 0 ? 1 : 0] = a1;
      l[axis == 2 ? 1 : 2] = b2;
      hl[c] = ((l[2] + 1) * PR + l[1] + 1) * PR + l[0] + 1;
      const int hx = x0 + l[0], hy = y0 + l[1], hz = z0 + l[2];
      if (hx >= 0 && hy >= 0 && hz >= 0 && hx < (int)u.gx && hy < (int)u.gy && hz < (int)u.gz) {
        const int32_t hgi = (int32_t)(((uint32_t)hz * u.gy + (uint32_t)hy) * u.gx + (uint32_t)hx);
        if (a.wc[hgi].w >= 0.0f) {
          // the neighbour across face f publishes this cell on its opposite face, same word
          const uint32_t nb = ((uint32_t)(hz / BR) * u.nby + (uint32_t)(hy / BR)) * u.nbx + (uint32_t)(hx / BR);
          hg[c] = face_word(nb, f ^ 1, c, k);
          hf[c] = f;
        }
      }
    }
  }
  // ---- this workgroup's XCD, read by the neighbours after the first check ---------------------
  const uint32_t my_xcc = xcc_id();
  if (t == 0) st_dev(&a.xcc[run], my_xcc + 1u);  // 0: not published

  // ---- stage the pressure region (device-coherent loads: other workgroups write it) ----------
  for (int i = (int)t; i < PR3; i += NT) {
    const int lx = i % PR - 1, lyy = (i / PR) % PR - 1, lzz = i / PR2 - 1;
    const int gx = x0 + lx, gy = y0 + lyy, gz = z0 + lzz;
    float v = 0.0f;
    if (gx >= 0 && gy >= 0 && gz >= 0 && gx < (int)u.gx && gy < (int)u.gy && gz < (int)u.gz)
      v = (float)ld_dev(&a.p[((uint32_t)gz * u.gy + (uint32_t)gy) * u.gx + (uint32_t)gx]) * u.inv_fp_scale;
    s_p[i] = v;
  }
  __syncthreads();

  uint32_t fires = 0;
  uint32_t it = 0;
  uint32_t epoch = 0;
  uint32_t checks = 0;
  const uint32_t limit = a.fixed_pairs ? a.fixed_pairs : a.cap;
  const uint32_t minit = (a.min_gate && *a.min_gate == 0u) ? a.min_iters : 0u;
  bool failed = false;

  // Residual of this brick's rows (residual_rows: fluid, non-empty, not held
  // at the clamp), published and reduced over every brick in run order.
  auto check = [&](uint32_t iters) -> bool {
    double rr = 0.0, bb = 0.0;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const Cell& e = cell[c];
      const float p = s_p[e.li];
      const float wn[6] = {e.wxm, e.self.x, e.wym, e.self.y, e.wzm, e.self.z};
      const int off[6] = {-1, 1, -PR, PR, -PR2, PR2};
      double apn = 0.0;
#pragma unroll
      for (int n = 0; n < 6; ++n) apn += (double)wn[n] * (double)s_p[e.li + off[n]];
      const double b = -(double)e.rhs;
      const double ap = (double)e.self.w * (double)p - apn;
      if (e.self.w > 0.0f && fabsf(p) != u.pressure_clamp) {
        rr += (b - ap) * (b - ap);
        bb += b * b;
      }
    }
    rr = wave_sum_d(rr);
    bb = wave_sum_d(bb);
    if ((t & 63u) == 0u) {
      s_red[2 * (t >> 6)] = rr;
      s_red[2 * (t >> 6) + 1] = bb;
    }
    __syncthreads();
    const uint32_t slot = checks & 1u;
    if (t == 0) {
      double r0 = 0.0, b0 = 0.0;
      for (int w = 0; w < NT / 64; ++w) {
        r0 += s_red[2 * w];
        b0 += s_red[2 * w + 1];
      }
      uint64_t* dst = reinterpret_cast<uint64_t*>(a.partials) + ((size_t)slot * a.max_runs + run) * 2;
      st_dev(dst, __double_as_longlong(r0));
      st_dev(dst + 1, __double_as_longlong(b0));
      stores_complete();
      st_dev(&a.pflags[(size_t)run * CS_PERSIST_FLAG_STRIDE], checks + 1u);
    }
    // Run 0's workgroup gathers every brick's sums (all of its loads in flight
    // together), reduces them in one fixed order and publishes the verdict as
    // one word (check number << 32 | stop); the others poll that word.
    const uint32_t want = checks + 1u;
    if (run == 0) {
      constexpr int QMAX = CS_PERSIST_MAX_BRICKS / NT;
      uint32_t fl[QMAX];
#pragma unroll
      for (int j = 0; j < QMAX; ++j) {
        const uint32_t q = t + (uint32_t)j * NT;
        fl[j] = q < nactive ? ld_dev(&a.pflags[(size_t)q * CS_PERSIST_FLAG_STRIDE]) : want;
      }
      const uint64_t t0 = now_ticks();
      uint32_t n = 0;
      for (;;) {
        bool ready = true;
#pragma unroll
        for (int j = 0; j < QMAX; ++j) ready &= fl[j] >= want;
        if (ready) break;
        if ((++n & 31u) == 0u && (ld_dev(&a.st->abort) || now_ticks() - t0 > a.spin_limit)) {
          failed = true;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int j = 0; j < QMAX; ++j) {
          const uint32_t q = t + (uint32_t)j * NT;
          if (fl[j] < want) fl[j] = ld_dev(&a.pflags[(size_t)q * CS_PERSIST_FLAG_STRIDE]);
        }
      }
      double tr = 0.0, tb = 0.0;
      if (!failed) {
        uint64_t pr[QMAX], pb[QMAX];
#pragma unroll
        for (int j = 0; j < QMAX; ++j) {
          const uint32_t q = t + (uint32_t)j * NT;
          const uint64_t* src = reinterpret_cast<const uint64_t*>(a.partials) + ((size_t)slot * a.max_runs + q) * 2;
          pr[j] = q < nactive ? ld_dev(src) : 0ull;
          pb[j] = q < nactive ? ld_dev(src + 1) : 0ull;
        }
#pragma unroll
        for (int j = 0; j < QMAX; ++j) {
          tr += __longlong_as_double(pr[j]);
          tb += __longlong_as_double(pb[j]);
        }
      }
      if (failed) atomicOr(&s_fail, 1u);
      tr = wave_sum_d(tr);
      tb = wave_sum_d(tb);
      __syncthreads();
      if ((t & 63u) == 0u) {
        s_red[2 * (t >> 6)] = tr;
        s_red[2 * (t >> 6) + 1] = tb;
      }
      __syncthreads();
      if (t == 0) {
        double r0 = 0.0, b0 = 0.0;
        for (int w = 0; w < NT / 64; ++w) {
          r0 += s_red[2 * w];
          b0 += s_red[2 * w + 1];
        }
        const double r = rel_residual(r0, b0);
        const bool stop = a.fixed_pairs ? iters >= a.fixed_pairs : ((r <= a.tol && iters >= minit) || iters >= a.cap);
        s_verdict = stop ? 1u : 0u;
        if (checks == 0) a.st->initial_residual = r;
        a.st->residual = r;
        a.st->iterations = iters;
        a.st->done = r <= a.tol ? 1u : 0u;
        stores_complete();
        st_dev(&a.st->verdict, (uint64_t)want << 32 | (stop ? 1u : 0u));
      }
    } else if (t == 0) {
      uint64_t v = ld_dev(&a.st->verdict);
      const uint64_t t0 = now_ticks();
      uint32_t n = 0;
      while ((uint32_t)(v >> 32) < want) {
        if ((++n & 31u) == 0u && (ld_dev(&a.st->abort) || now_ticks() - t0 > a.spin_limit)) {
          atomicOr(&s_fail, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
        v = ld_dev(&a.st->verdict);
      }
      s_verdict = (uint32_t)v & 1u;
    }
    ++checks;
    __syncthreads();
    return s_verdict != 0u || s_fail != 0u;
  };

  // the first test (the oracle's initial residual in tolerance mode) is also the
  // point after which every workgroup's XCD is published
  bool stop = check(0u);
  {
    uint32_t local = 0u;  // bit f: the brick across face f is served on this XCD
    const int bc[3] = {(int)bx, (int)by, (int)bz};
    const int nbd[3] = {(int)u.nbx, (int)u.nby, (int)u.nbz};
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      int q[3] = {bc[0], bc[1], bc[2]};
      q[f >> 1] += (f & 1) ? 1 : -1;
      if (q[f >> 1] < 0 || q[f >> 1] >= nbd[f >> 1]) continue;
      const int32_t sl = a.slot[((uint32_t)q[2] * u.nby + (uint32_t)q[1]) * u.nbx + (uint32_t)q[0]];
      if (sl >= 0 && ld_dev(&a.xcc[sl]) == my_xcc + 1u) local |= 1u << f;
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
#pragma unroll
      for (int j = 0; j < 3; ++j)
        if (cell[c].fface[j] >= 0 && ((local >> cell[c].fface[j]) & 1u)) cell[c].floc |= 1u << j;
      hf[c] = (local >> hf[c]) & 1u;  // from here on: 1 when the halo word is XCD-local
    }
  }
  // the face buffers as a raw buffer (32-bit data format), for the XCD-local words
  const __amdgpu_buffer_rsrc_t fbr =
      __builtin_amdgcn_make_buffer_rsrc(a.fb, 0, (int)(u.nbricks * CS_PERSIST_FACE_WORDS * 8u), 0x00020000);

  // one half sweep of colour c; returns false when a wait gave up
  auto half_sweep = [&](const Cell& e, int h_l, int32_t h_g, int h_local) -> bool {
    ++epoch;
    const uint32_t want = epoch_base + epoch;
    bool fired;
    const float np = update(u, e.self, e.rhs, e.wxm, e.wym, e.wzm, s_p, e.li, fired);
    if (e.self.w >= 0.0f) {
      s_p[e.li] = np;
      fires += fired ? 1u : 0u;
      const uint64_t word = (uint64_t)want << 32 | (uint32_t)(int32_t)(np * u.fp_scale);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        if (e.fslot[j] < 0) continue;
        if ((e.floc >> j) & 1u) st_xcd(fbr, e.fslot[j], word);
        else st_dev(&a.fb[e.fslot[j]], word);
      }
    }
    // the colour-c halo of this epoch, polled until its word carries the epoch
    if (h_g >= 0) {
      // the neighbour stores its word about when this one is stored: a first
      // poll issued at once would mostly come back stale and cost a round trip
      for (uint32_t d = 0; d < a.poll_delay; ++d) __builtin_amdgcn_s_sleep(1);
      uint64_t v = h_local ? ld_xcd(fbr, h_g) : ld_dev(&a.fb[h_g]);
      if ((uint32_t)(v >> 32) < want) {
        // the abort flag and the clock are looked at every 32nd poll only, so a
        // retry costs one load latency
        const uint64_t t0 = now_ticks();
        uint32_t n = 0;
        do {
          __builtin_amdgcn_s_sleep(1);
          v = h_local ? ld_xcd(fbr, h_g) : ld_dev(&a.fb[h_g]);
          if ((++n & 31u) == 0u && (ld_dev(&a.st->abort) || now_ticks() - t0 > a.spin_limit)) {
            atomicOr(&s_fail, 1u);
            break;
          }
        } while ((uint32_t)(v >> 32) < want);
      }
      s_p[h_l] = (float)(int32_t)(uint32_t)v * u.inv_fp_scale;
    }
    __syncthreads();
    return s_fail == 0u;
  };

  while (!stop && it < limit) {
    if (!half_sweep(cell[0], hl[0], hg[0], hf[0]) || !half_sweep(cell[1], hl[1], hg[1], hf[1])) break;
    ++it;
    if (a.fixed_pairs ? it == a.fixed_pairs : (it % a.every == 0u || it >= a.cap)) stop = check(it);
  }
  if (s_fail) {
    if (t == 0) {
      st_dev(&a.st->abort, 1u);
      atomicAdd(&a.st->faults, 1u);
    }
    return;
  }
  // the solution of this brick's cells (later passes read it with plain loads)
#pragma unroll
  for (int c = 0; c < 2; ++c)
    if (cell[c].self.w >= 0.0f) a.p[cell[c].g] = (int32_t)(s_p[cell[c].li] * u.fp_scale);
#pragma unroll
  for (int m2 = 32; m2 > 0; m2 >>= 1) fires += (uint32_t)__shfl_xor((int)fires, m2, 64);
  if ((t & 63u) == 0u && fires) atomicAdd(&a.metrics[CS_M_P_CLAMP], fires);
}

// Clears the slot map, the residual-check flags, the XCD table and the
// failure state of a solve.
// One workgroup: every brick's slot, check flag and XCD word cleared, then
// each listed brick's slot set (the clears land before the barrier).
__global__ void __launch_bounds__(1024) k_persist_prepare(const uint32_t* __restrict__ list,
                                                          cs_solve_state* __restrict__ st, uint32_t nbricks,
                                                          int32_t* __restrict__ slot, uint32_t* __restrict__ pflags,
                                                          uint32_t* __restrict__ xcc) {
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < nbricks; i += 1024u) {
    slot[i] = -1;
    pflags[(size_t)i * CS_PERSIST_FLAG_STRIDE] = 0u;
    xcc[i] = 0u;
  }
  if (t == 0) {
    st->abort = 0u;
    st->faults = 0u;
    st->verdict = 0ull;
  }
  __syncthreads();
  const uint32_t n = st->active_bricks;
  for (uint32_t i = t; i < n; i += 1024u) slot[list[i]] = (int32_t)i;
}

hipError_t cs_persist_prepare(hipStream_t s, const cs_params& u, const uint32_t* list, cs_solve_state* st,
                              int32_t* slot, uint32_t* pflags, uint32_t* xcc) {
  hipLaunchKernelGGL(k_persist_prepare, dim3(1), dim3(1024), 0, s, list, st, u.nbricks, slot, pflags, xcc);
  return hipGetLastError();
}

uint32_t cs_persist_groups(uint32_t nactive) { return (nactive + 7u) / 8u * 8u; }

hipError_t cs_persist_max_groups(int device, uint32_t* out) {
  int per_cu = 0, cus = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_rbgs_persist, NT, 0);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  if (e != hipSuccess) return e;
  *out = (uint32_t)per_cu * (uint32_t)cus;
  return hipSuccess;
}

hipError_t cs_persist_launch(hipStream_t s, const cs_persist_args& a, uint32_t groups, bool cooperative) {
  if (cooperative) {
    cs_persist_args copy = a;
    void* args[] = {&copy};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_rbgs_persist), dim3(groups), dim3(NT), args, 0,
                                      s);
  }
  hipLaunchKernelGGL(k_rbgs_persist, dim3(groups), dim3(NT), 0, s, a);
  return hipGetLastError();
}
