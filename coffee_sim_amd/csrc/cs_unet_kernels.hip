// cs_unet_kernels.hip — kernels of the learned pressure warm start (a 3-D U-Net).
//
// The reference has no learned component: it restarts the pressure from zero
// every substep (shader.rs:1187).  This network predicts the converged
// pressure from the solve's inputs so RBGS can start from it; the plain solve
// still decides convergence (its residual) and is the fallback.
//
// Activations are bf16, channels-last ([voxel][channel], voxel = (z*s + y)*s + x
// on a cubic grid of side s), accumulation is fp32.  3x3x3 convolutions are
// implicit GEMMs on v_mfma_f32_16x16x32_bf16: rows are 16 voxels of an 8^3
// brick, columns 16 output channels, and each 32-wide K step covers two taps
// x 16 input channels staged in LDS (a 10^3 halo region per 16-channel chunk).
// Lane l of an MFMA holds A[l%16][8(l/16)..+7], B[8(l/16)..+7][l%16] and
// D[4(l/16)+i][l%16] (checked on the device against a host matmul).
#include <hip/hip_runtime.h>

#include "cs_unet.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ uint16_t f2bf(float f) {
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0u;  // NaN
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ bool bf_pos(uint16_t h) { return !(h & 0x8000u) && (h & 0x7fffu); }

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// Zero the bf16 entries of `v` whose counterpart in `m` is not positive (the
// ReLU derivative of a layer's output applied to its gradient).
__device__ __forceinline__ uint4 relu_mask(uint4 v, uint4 m) {
  uint32_t a[4] = {v.x, v.y, v.z, v.w}, b[4] = {m.x, m.y, m.z, m.w};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t lo = bf_pos((uint16_t)(b[i] & 0xffffu)) ? 0x0000ffffu : 0u;
    const uint32_t hi = bf_pos((uint16_t)(b[i] >> 16)) ? 0xffff0000u : 0u;
    a[i] &= lo | hi;
  }
  return make_uint4(a[0], a[1], a[2], a[3]);
}

constexpr int HALO = 10;  // 8^3 brick + one voxel each side
constexpr int HV = HALO * HALO * HALO;

}  // namespace

// ---- 3x3x3 convolution (forward, and input gradient with flipped weights) ----------
// out[v][co] = act(bias[co] + sum_tap sum_ci in[v + tap][ci] * w[tap][co][ci]), zero
// padding, act = ReLU when `relu`.  With `mask` the input is used only where the
// mask (a layer output) is positive.  One workgroup per 8^3 brick of the s^3 grid;
// each of the 4 wavefronts owns 8 row tiles of 16 voxels.
template <int COUT>
__global__ void __launch_bounds__(256) k_unet_conv(const uint16_t* __restrict__ in, int cin, int s,
                                                   const uint16_t* __restrict__ w, const float* __restrict__ bias,
                                                   const uint16_t* __restrict__ mask, uint16_t* __restrict__ out,
                                                   int relu) {
  constexpr int NTL = COUT / 16;
  __shared__ uint4 s_in[HV * 2];  // [halo voxel][16 channels] as two 16-byte halves
  const int nb = s / 8;
  const int bid = (int)blockIdx.x;
  const int bx = bid % nb, by = (bid / nb) % nb, bz = bid / (nb * nb);
  const int x0 = bx * 8 - 1, y0 = by * 8 - 1, z0 = bz * 8 - 1;
  const int lane = (int)(threadIdx.x & 63u), wv = (int)(threadIdx.x >> 6);
  const int g = lane >> 4, r16 = lane & 15;
  f32x4 acc[8][NTL];
#pragma unroll
  for (int mi = 0; mi < 8; ++mi)
#pragma unroll
    for (int nt = 0; nt < NTL; ++nt) acc[mi][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int c0 = 0; c0 < cin; c0 += 16) {
    __syncthreads();
    for (int e = (int)threadIdx.x; e < HV * 2; e += 256) {
      const int hv = e >> 1, h = e & 1;
      const int gx = x0 + hv % HALO, gy = y0 + (hv / HALO) % HALO, gz = z0 + hv / (HALO * HALO);
      uint4 val = make_uint4(0u, 0u, 0u, 0u);
      if ((unsigned)gx < (unsigned)s && (unsigned)gy < (unsigned)s && (unsigned)gz < (unsigned)s) {
        const size_t off = (((size_t)gz * s + gy) * s + gx) * (size_t)cin + (size_t)(c0 + 8 * h);
        val = *reinterpret_cast<const uint4*>(in + off);
        if (mask) val = relu_mask(val, *reinterpret_cast<const uint4*>(mask + off));
      }
      s_in[e] = val;
    }
    __syncthreads();
    for (int ks = 0; ks < 14; ++ks) {
      const int tap = 2 * ks + (g >> 1);  // K rows 0..15: tap 2ks, 16..31: tap 2ks+1
      const bool tv = tap < 27;
      const int tt = tv ? tap : 26;
      const int dz = tt / 9 - 1, dy = (tt / 3) % 3 - 1, dx = tt % 3 - 1;
      const int ci0 = 8 * (g & 1);
      bf16x8 b[NTL];
#pragma unroll
      for (int nt = 0; nt < NTL; ++nt) {
        const uint4 wvec = *reinterpret_cast<const uint4*>(w + ((size_t)(tt * COUT + nt * 16 + r16) * cin + c0 + ci0));
        b[nt] = as_bf16x8(tv ? wvec : make_uint4(0u, 0u, 0u, 0u));
      }
#pragma unroll
      for (int mi = 0; mi < 8; ++mi) {
        const int vb = (wv * 8 + mi) * 16 + r16;
        const int vx = vb & 7, vy = (vb >> 3) & 7, vz = vb >> 6;
        const int hv = ((vz + dz + 1) * HALO + (vy + dy + 1)) * HALO + (vx + dx + 1);
        const uint4 av = s_in[hv * 2 + (g & 1)];
        const bf16x8 a = as_bf16x8(tv ? av : make_uint4(0u, 0u, 0u, 0u));
#pragma unroll
        for (int nt = 0; nt < NTL; ++nt) acc[mi][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[nt], acc[mi][nt], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int mi = 0; mi < 8; ++mi)
#pragma unroll
    for (int nt = 0; nt < NTL; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int vb = (wv * 8 + mi) * 16 + 4 * g + i;
        const int co = nt * 16 + r16;
        float v = acc[mi][nt][i] + (bias ? bias[co] : 0.0f);
        if (relu) v = fmaxf(v, 0.0f);
        const int gx = bx * 8 + (vb & 7), gy = by * 8 + ((vb >> 3) & 7), gz = bz * 8 + (vb >> 6);
        out[(((size_t)gz * s + gy) * s + gx) * COUT + co] = f2bf(v);
      }
}

// ---- weight gradient ----------------------------------------------------------------
// dW[tap][ci][co] = sum_v in[v + tap][ci] * gz[v][co], gz = gout masked by the
// layer output's ReLU derivative.  Grid (slot, chunk pair): a workgroup of
// three wavefronts sums all 27 taps (wavefront w owns dz = w - 1) for one
// 16-channel input chunk and one 16-channel output chunk over its run of
// bricks, and writes fp32 partials[slot][tap][ci][co]; a fixed-order
// reduction over slots follows.  Rows of the MFMA are input channels, columns
// output channels, K runs over 32 voxels (four x-rows of a brick).
//
// LDS holds both operands channel-major with x fastest, so a lane's eight K
// values (eight consecutive x) are one 16-byte read: the input halo as
// [ci][z 10][y 10][x 16] (x 0..9 used), the masked gradient as [co][512
// voxels].  The three x-shifted input fragments of a (dz, dy) row come from
// one 16-byte and one 4-byte read by byte alignment.  Row strides are padded
// by 16 bytes so the 16 lanes of a read hit distinct banks.
namespace {
constexpr int WG_XS = 16;                              // x slots per halo row
constexpr int WG_CSTR = HALO * HALO * WG_XS + 8;       // u16 per input channel (3216 bytes)
constexpr int WG_GSTR = 512 + 8;                       // u16 per output channel (1040 bytes)
constexpr int WG_THREADS = 192;
}  // namespace

__global__ void __launch_bounds__(WG_THREADS) k_unet_wgrad(const uint16_t* __restrict__ in, int cin,
                                                           const uint16_t* __restrict__ gout,
                                                           const uint16_t* __restrict__ outv, int cout, int s,
                                                           int bricks_per_slot, float* __restrict__ partial) {
  __shared__ __attribute__((aligned(16))) uint16_t s_in[16 * WG_CSTR];
  __shared__ __attribute__((aligned(16))) uint16_t s_g[16 * WG_GSTR];
  const int slot = (int)blockIdx.x;
  const int nco = cout / 16;
  const int cic = (int)blockIdx.y / nco, coc = (int)blockIdx.y % nco;
  const int nb = s / 8, nbricks = nb * nb * nb;
  const int tid = (int)threadIdx.x;
  const int lane = tid & 63, dzi = tid >> 6;  // wavefront = dz + 1
  const int g = lane >> 4, r16 = lane & 15;
  f32x4 acc[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) acc[k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const int b0 = slot * bricks_per_slot, b1 = min(b0 + bricks_per_slot, nbricks);
  for (int bid = b0; bid < b1; ++bid) {
    const int bx = bid % nb, by = (bid / nb) % nb, bz = bid / (nb * nb);
    __syncthreads();
    // input halo: pairs of x-adjacent voxels, 8 channels per task, one dword per channel
    for (int e = tid; e < HALO * HALO * (HALO / 2) * 2; e += WG_THREADS) {
      const int h = e & 1, pr = e >> 1;
      const int hx = 2 * (pr % (HALO / 2)), hy = (pr / (HALO / 2)) % HALO, hz = pr / (HALO * HALO / 2);
      const int gx = bx * 8 - 1 + hx, gy = by * 8 - 1 + hy, gz = bz * 8 - 1 + hz;
      uint4 a = make_uint4(0u, 0u, 0u, 0u), c = a;
      if ((unsigned)gy < (unsigned)s && (unsigned)gz < (unsigned)s) {
        const size_t row = ((size_t)gz * s + gy) * s;
        if ((unsigned)gx < (unsigned)s)
          a = *reinterpret_cast<const uint4*>(in + (row + gx) * (size_t)cin + cic * 16 + 8 * h);
        if ((unsigned)(gx + 1) < (unsigned)s)
          c = *reinterpret_cast<const uint4*>(in + (row + gx + 1) * (size_t)cin + cic * 16 + 8 * h);
      }
      const uint32_t aw[4] = {a.x, a.y, a.z, a.w}, cw[4] = {c.x, c.y, c.z, c.w};
      uint16_t* dst = s_in + (8 * h) * WG_CSTR + (hz * HALO + hy) * WG_XS + hx;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const uint32_t lo = (aw[j >> 1] >> (16 * (j & 1))) & 0xffffu, hi = (cw[j >> 1] >> (16 * (j & 1))) & 0xffffu;
        *reinterpret_cast<uint32_t*>(dst + j * WG_CSTR) = lo | (hi << 16);
      }
    }
    // masked output gradient of the brick, the same way
    for (int e = tid; e < 256 * 2; e += WG_THREADS) {
      const int h = e & 1, v = 2 * (e >> 1);
      const int gx = bx * 8 + (v & 7), gy = by * 8 + ((v >> 3) & 7), gz = bz * 8 + (v >> 6);
      const size_t off = (((size_t)gz * s + gy) * s + gx) * (size_t)cout + coc * 16 + 8 * h;
      const uint4 a = relu_mask(*reinterpret_cast<const uint4*>(gout + off), *reinterpret_cast<const uint4*>(outv + off));
      const uint4 c = relu_mask(*reinterpret_cast<const uint4*>(gout + off + cout),
                                *reinterpret_cast<const uint4*>(outv + off + cout));
      const uint32_t aw[4] = {a.x, a.y, a.z, a.w}, cw[4] = {c.x, c.y, c.z, c.w};
      uint16_t* dst = s_g + (8 * h) * WG_GSTR + v;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const uint32_t lo = (aw[j >> 1] >> (16 * (j & 1))) & 0xffffu, hi = (cw[j >> 1] >> (16 * (j & 1))) & 0xffffu;
        *reinterpret_cast<uint32_t*>(dst + j * WG_GSTR) = lo | (hi << 16);
      }
    }
    __syncthreads();
#pragma unroll 2
    for (int ks = 0; ks < 16; ++ks) {
      const int row = ks * 4 + g;  // x-row of the brick: voxels 8 row .. 8 row + 7
      const int vy = row & 7, vz = row >> 3;
      const bf16x8 b = as_bf16x8(*reinterpret_cast<const uint4*>(s_g + r16 * WG_GSTR + row * 8));
#pragma unroll
      for (int dyi = 0; dyi < 3; ++dyi) {
        const uint16_t* src = s_in + r16 * WG_CSTR + ((vz + dzi) * HALO + vy + dyi) * WG_XS;
        const uint4 w03 = *reinterpret_cast<const uint4*>(src);
        const uint32_t w4 = *reinterpret_cast<const uint32_t*>(src + 8);
        // x shifts -1, 0, +1: halo x = j, j + 1, j + 2 for the eight voxels j
        const uint4 a1 = make_uint4(__builtin_amdgcn_alignbyte(w03.y, w03.x, 2u),
                                    __builtin_amdgcn_alignbyte(w03.z, w03.y, 2u),
                                    __builtin_amdgcn_alignbyte(w03.w, w03.z, 2u),
                                    __builtin_amdgcn_alignbyte(w4, w03.w, 2u));
        const uint4 a2 = make_uint4(w03.y, w03.z, w03.w, w4);
        acc[dyi * 3 + 0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(w03), b, acc[dyi * 3 + 0], 0, 0, 0);
        acc[dyi * 3 + 1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(a1), b, acc[dyi * 3 + 1], 0, 0, 0);
        acc[dyi * 3 + 2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(a2), b, acc[dyi * 3 + 2], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    const int tap = dzi * 9 + t;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int ci = cic * 16 + 4 * g + i, co = coc * 16 + r16;
      partial[(((size_t)slot * 27 + tap) * cin + ci) * cout + co] = acc[t][i];
    }
  }
}

// Fixed-order sum over slots, in two passes so that enough loads are in
// flight: groups[g][e] = sum of slots g*per .. g*per + per - 1 of partial[.][e],
// then out[e] = sum_g groups[g][e].
__global__ void __launch_bounds__(256) k_unet_reduce_slots(const float* __restrict__ partial, int slots, int per,
                                                           size_t n, float* __restrict__ groups) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int k0 = (int)blockIdx.y * per, k1 = min(k0 + per, slots);
  float acc = 0.0f;
  for (int k = k0; k < k1; ++k) acc += partial[(size_t)k * n + e];
  groups[(size_t)blockIdx.y * n + e] = acc;
}
__global__ void __launch_bounds__(256) k_unet_reduce_groups(const float* __restrict__ groups, int ngroups, size_t n,
                                                            float* __restrict__ out) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float acc = 0.0f;
  for (int g = 0; g < ngroups; ++g) acc += groups[(size_t)g * n + e];
  out[e] = acc;
}

// Bias gradient, stage 1: part[g][co] = sum over the g-th share of voxels of
// gz[v][co] (gout where the layer output is positive).  Each thread reads
// 8 channels of a voxel as one 16-byte vector of gout and one of the output
// (lanes = cout / 8 threads per voxel, rows = 256 / lanes voxels per pass,
// two voxels in flight per thread); the rows are then summed per channel in
// row order through LDS.
__global__ void __launch_bounds__(256) k_unet_bias_grad(const uint16_t* __restrict__ gout,
                                                        const uint16_t* __restrict__ outv, int cout, size_t nv,
                                                        float* __restrict__ part) {
  __shared__ float s_sum[256 * 8];
  const int lanes = cout / 8, rows = 256 / lanes;
  const int t = (int)threadIdx.x, row = t / lanes, lane = t % lanes;
  float acc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] = 0.0f;
  if (row < rows) {
    const size_t stride = (size_t)gridDim.x * rows;
    auto add = [&](const uint4 g, const uint4 o) {
      const uint32_t gw[4] = {g.x, g.y, g.z, g.w}, ow[4] = {o.x, o.y, o.z, o.w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (bf_pos((uint16_t)(ow[q] & 0xffffu))) acc[2 * q] += bf2f((uint16_t)(gw[q] & 0xffffu));
        if (bf_pos((uint16_t)(ow[q] >> 16))) acc[2 * q + 1] += bf2f((uint16_t)(gw[q] >> 16));
      }
    };
    size_t v = (size_t)blockIdx.x * rows + row;
    for (; v + stride < nv; v += 2 * stride) {
      const size_t e0 = v * cout + 8 * lane, e1 = (v + stride) * cout + 8 * lane;
      const uint4 g0 = *reinterpret_cast<const uint4*>(gout + e0), o0 = *reinterpret_cast<const uint4*>(outv + e0);
      const uint4 g1 = *reinterpret_cast<const uint4*>(gout + e1), o1 = *reinterpret_cast<const uint4*>(outv + e1);
      add(g0, o0);
      add(g1, o1);
    }
    if (v < nv) {
      const size_t e0 = v * cout + 8 * lane;
      add(*reinterpret_cast<const uint4*>(gout + e0), *reinterpret_cast<const uint4*>(outv + e0));
    }
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) s_sum[t * 8 + j] = acc[j];
  __syncthreads();
  if (t < cout) {
    float sum = 0.0f;
    for (int r = 0; r < rows; ++r) sum += s_sum[(r * lanes + t / 8) * 8 + t % 8];
    part[(size_t)blockIdx.x * cout + t] = sum;
  }
}

// ---- pooling, upsampling, concatenation -------------------------------------------------
// One thread per voxel and 8-channel group (channel counts, strides and
// offsets are multiples of 8): 16-byte loads and stores, the per-channel
// arithmetic in the same order as one channel at a time.
namespace {
__device__ __forceinline__ void unpack8(uint4 v, float* f) {
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    f[2 * j] = bf2f((uint16_t)(w[j] & 0xffffu));
    f[2 * j + 1] = bf2f((uint16_t)(w[j] >> 16));
  }
}
__device__ __forceinline__ uint4 pack8(const float* f) {
  uint32_t w[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) w[j] = (uint32_t)f2bf(f[2 * j]) | ((uint32_t)f2bf(f[2 * j + 1]) << 16);
  return make_uint4(w[0], w[1], w[2], w[3]);
}
__device__ __forceinline__ uint4 ld8(const uint16_t* p) { return *reinterpret_cast<const uint4*>(p); }
__device__ __forceinline__ void st8(uint16_t* p, uint4 v) { *reinterpret_cast<uint4*>(p) = v; }
}  // namespace

// 2x average pool, [s^3][c] -> [(s/2)^3][c].
__global__ void __launch_bounds__(256) k_unet_pool(const uint16_t* __restrict__ in, int c, int s,
                                                   uint16_t* __restrict__ out) {
  const uint32_t h = (uint32_t)s / 2, cq = (uint32_t)c / 8;
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= h * h * h * cq) return;
  const uint32_t q = e % cq, vo = e / cq;
  const uint32_t x = vo % h, y = (vo / h) % h, z = vo / (h * h);
  float acc[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, f[8];
  for (uint32_t dz = 0; dz < 2; ++dz)
    for (uint32_t dy = 0; dy < 2; ++dy)
      for (uint32_t dx = 0; dx < 2; ++dx) {
        unpack8(ld8(in + ((((size_t)(2 * z + dz)) * s + 2 * y + dy) * s + 2 * x + dx) * c + 8 * q), f);
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[j] += f[j];
      }
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] *= 0.125f;
  st8(out + (size_t)vo * c + 8 * q, pack8(acc));
}

// Gradient of the pool plus a second contribution read from a channel window of
// another tensor: g_in[v][ch] = add[v][add_off + ch] + g_out[parent(v)][ch] / 8.
__global__ void __launch_bounds__(256) k_unet_pool_back(const uint16_t* __restrict__ g_out, int c, int s,
                                                        const uint16_t* __restrict__ add, int add_stride, int add_off,
                                                        uint16_t* __restrict__ g_in) {
  const uint32_t cq = (uint32_t)c / 8, us = (uint32_t)s, h = us / 2;
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= us * us * us * cq) return;
  const uint32_t q = e % cq, v = e / cq;
  const uint32_t x = v % us, y = (v / us) % us, z = v / (us * us);
  float gp[8], a[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  unpack8(ld8(g_out + ((((size_t)(z / 2)) * h + y / 2) * h + x / 2) * c + 8 * q), gp);
  if (add) unpack8(ld8(add + (size_t)v * add_stride + add_off + 8 * q), a);
#pragma unroll
  for (int j = 0; j < 8; ++j) a[j] = a[j] + gp[j] * 0.125f;
  st8(g_in + (size_t)v * c + 8 * q, pack8(a));
}

// Nearest 2x upsample into channels [off, off + c) of a concatenation with `stride` channels.
__global__ void __launch_bounds__(256) k_unet_up(const uint16_t* __restrict__ in, int c, int s,
                                                 uint16_t* __restrict__ cat, int stride, int off) {
  const uint32_t cq = (uint32_t)c / 8, S = 2u * (uint32_t)s;
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= S * S * S * cq) return;
  const uint32_t q = e % cq, v = e / cq;
  const uint32_t x = v % S, y = (v / S) % S, z = v / (S * S);
  st8(cat + (size_t)v * stride + off + 8 * q, ld8(in + ((((size_t)(z / 2)) * s + y / 2) * s + x / 2) * c + 8 * q));
}

// Gradient of the upsample: g_in[v'][ch] = sum of the 8 children's cat[child][off + ch].
__global__ void __launch_bounds__(256) k_unet_up_back(const uint16_t* __restrict__ g_cat, int stride, int off, int c,
                                                      int s, uint16_t* __restrict__ g_in) {
  const uint32_t cq = (uint32_t)c / 8, us = (uint32_t)s, S = 2u * us;
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= us * us * us * cq) return;
  const uint32_t q = e % cq, vo = e / cq;
  const uint32_t x = vo % us, y = (vo / us) % us, z = vo / (us * us);
  float acc[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, f[8];
  for (uint32_t dz = 0; dz < 2; ++dz)
    for (uint32_t dy = 0; dy < 2; ++dy)
      for (uint32_t dx = 0; dx < 2; ++dx) {
        unpack8(ld8(g_cat + ((((size_t)(2 * z + dz)) * S + 2 * y + dy) * S + 2 * x + dx) * stride + off + 8 * q), f);
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[j] += f[j];
      }
  st8(g_in + (size_t)vo * c + 8 * q, pack8(acc));
}

// Copies a [v][c] tensor into channels [off, off + c) of a concatenation.
__global__ void __launch_bounds__(256) k_unet_put(const uint16_t* __restrict__ in, int c, size_t nv,
                                                  uint16_t* __restrict__ cat, int stride, int off) {
  const uint32_t cq = (uint32_t)c / 8;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nv * cq) return;
  const size_t v = e / cq;
  const uint32_t q = (uint32_t)(e % cq);
  st8(cat + v * stride + off + 8 * q, ld8(in + v * c + 8 * q));
}

// ---- head, loss --------------------------------------------------------------------------
// pred[v] = sum_c x[v][c] * wh[c] + bh, x with 16 channels.
__global__ void k_unet_head(const uint16_t* __restrict__ x, const float* __restrict__ wh, size_t nv,
                            float* __restrict__ pred) {
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  float acc = 0.0f;
#pragma unroll
  for (int c = 0; c < 16; ++c) acc += bf2f(x[v * 16 + c]) * wh[c];
  pred[v] = acc + wh[16];
}

// Masked mean squared error: per-block partial sums of mask and squared error,
// and the prediction gradient 2 * mask * (pred - target) / count once `count`
// (the mask sum, computed on the host's request beforehand) is known.
__global__ void __launch_bounds__(256) k_unet_loss(const float* __restrict__ pred, const float* __restrict__ target,
                                                   const uint8_t* __restrict__ mask, size_t nv, float inv_count,
                                                   float* __restrict__ gpred, float* __restrict__ block_sse) {
  __shared__ float s_sum[256];
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  float se = 0.0f;
  if (v < nv) {
    const float d = mask[v] ? pred[v] - target[v] : 0.0f;
    se = d * d;
    gpred[v] = 2.0f * d * inv_count;
  }
  s_sum[threadIdx.x] = se;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) s_sum[threadIdx.x] += s_sum[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) block_sse[blockIdx.x] = s_sum[0];
}

// Gradient into the head: g_x[v][c] = gpred[v] * wh[c]; per-block partials of
// d wh[c] = sum_v gpred[v] x[v][c] and d bh = sum_v gpred[v].
__global__ void __launch_bounds__(256) k_unet_head_back(const uint16_t* __restrict__ x, const float* __restrict__ wh,
                                                        const float* __restrict__ gpred, size_t nv,
                                                        uint16_t* __restrict__ g_x, float* __restrict__ block_part) {
  __shared__ float s_part[17][256];
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const float gp = v < nv ? gpred[v] : 0.0f;
#pragma unroll
  for (int c = 0; c < 16; ++c) {
    const float xc = v < nv ? bf2f(x[v * 16 + c]) : 0.0f;
    if (v < nv) g_x[v * 16 + c] = f2bf(gp * wh[c]);
    s_part[c][threadIdx.x] = gp * xc;
  }
  s_part[16][threadIdx.x] = gp;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int c = 0; c < 17; ++c) s_part[c][threadIdx.x] += s_part[c][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < 17) block_part[blockIdx.x * 17 + threadIdx.x] = s_part[threadIdx.x][0];
}

// Fixed-order sum of `parts` rows of `width` values: one workgroup per column,
// a strided partial per thread, then a tree over the workgroup.
__global__ void __launch_bounds__(256) k_unet_sum_rows(const float* __restrict__ part, int parts, int width,
                                                       float* __restrict__ out) {
  __shared__ float s_sum[256];
  const int c = (int)blockIdx.x;
  float acc = 0.0f;
  for (int k = (int)threadIdx.x; k < parts; k += 256) acc += part[(size_t)k * width + c];
  s_sum[threadIdx.x] = acc;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) s_sum[threadIdx.x] += s_sum[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = s_sum[0];
}

// ---- parameters ---------------------------------------------------------------------------
// bf16 copies of one convolution's fp32 weights W[tap][ci][co]: wf[tap][co][ci]
// for the forward pass, wd[tap][ci][co] = W[26 - tap] for the input gradient.
__global__ void k_unet_prep_weights(const float* __restrict__ w, int cin, int cout, uint16_t* __restrict__ wf,
                                    uint16_t* __restrict__ wd) {
  const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (e >= 27 * cin * cout) return;
  const int co = e % cout, ci = (e / cout) % cin, tap = e / (cin * cout);
  const uint16_t h = f2bf(w[e]);
  wf[((size_t)tap * cout + co) * cin + ci] = h;
  wd[((size_t)(26 - tap) * cin + ci) * cout + co] = h;
}

// Adam (Kingma & Ba 2015) with bias correction folded into the step size.
__global__ void k_unet_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, size_t n, float lr_t, float b1, float b2, float eps) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr_t * mi / (sqrtf(vi) + eps);
}

// ---- inputs --------------------------------------------------------------------------------
// The network's input channels from host-side float features [v][4] (the
// remaining 12 of 16 channels are zero).
__global__ void k_unet_pack_features(const float* __restrict__ f4, size_t nv, uint16_t* __restrict__ x) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nv * 16) return;
  const size_t v = e / 16;
  const int c = (int)(e % 16);
  x[e] = c < 4 ? f2bf(f4[v * 4 + c]) : (uint16_t)0;
}

// out[v] = pred[v] * scale.
__global__ void k_unet_scale(const float* __restrict__ pred, size_t nv, float scale, float* __restrict__ out) {
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < nv) out[v] = pred[v] * scale;
}

// ---- launchers ------------------------------------------------------------------------------
namespace {
inline dim3 grid1(size_t n, int block = 256) { return dim3((unsigned)((n + block - 1) / block)); }
}  // namespace

hipError_t cs_unet_launch_conv(hipStream_t st, const uint16_t* in, int cin, int s, const uint16_t* w,
                               const float* bias, const uint16_t* mask, uint16_t* out, int cout, int relu) {
  const dim3 g((unsigned)((s / 8) * (s / 8) * (s / 8)));
  switch (cout) {
    case 16: hipLaunchKernelGGL(k_unet_conv<16>, g, dim3(256), 0, st, in, cin, s, w, bias, mask, out, relu); break;
    case 32: hipLaunchKernelGGL(k_unet_conv<32>, g, dim3(256), 0, st, in, cin, s, w, bias, mask, out, relu); break;
    case 48: hipLaunchKernelGGL(k_unet_conv<48>, g, dim3(256), 0, st, in, cin, s, w, bias, mask, out, relu); break;
    case 64: hipLaunchKernelGGL(k_unet_conv<64>, g, dim3(256), 0, st, in, cin, s, w, bias, mask, out, relu); break;
    case 96: hipLaunchKernelGGL(k_unet_conv<96>, g, dim3(256), 0, st, in, cin, s, w, bias, mask, out, relu); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

int cs_unet_wgrad_slots(int s, int cin, int cout) {
  // about CS_UNET_WGRAD_TARGET workgroups per launch (two resident per
  // compute unit); the partials then stay near that many * 27 * 256 floats
  const int nbricks = (s / 8) * (s / 8) * (s / 8);
  const int pairs = (cin / 16) * (cout / 16);
  const int want = pairs >= CS_UNET_WGRAD_TARGET ? 1 : CS_UNET_WGRAD_TARGET / pairs;
  return nbricks < want ? nbricks : want;
}

hipError_t cs_unet_launch_wgrad(hipStream_t st, const uint16_t* in, int cin, const uint16_t* gout,
                                const uint16_t* outv, int cout, int s, float* partial, float* bias_part, float* dw,
                                float* db) {
  // 16-byte vectors of 8 channels throughout; whole 8^3 bricks
  if (cin % 16 != 0 || cout % 16 != 0 || cout > 256 || s % 8 != 0 ||
      ((uintptr_t)in | (uintptr_t)gout | (uintptr_t)outv) % 16u != 0u)
    return hipErrorInvalidValue;
  const int nbricks = (s / 8) * (s / 8) * (s / 8);
  const int slots = cs_unet_wgrad_slots(s, cin, cout);
  const int bricks_per_slot = (nbricks + slots - 1) / slots;
  hipLaunchKernelGGL(k_unet_wgrad, dim3((unsigned)slots, (unsigned)((cin / 16) * (cout / 16))), dim3(WG_THREADS), 0,
                     st, in, cin, gout, outv, cout, s, bricks_per_slot, partial);
  const size_t n = (size_t)27 * cin * cout;
  const int per = CS_UNET_WGRAD_PER_GROUP, ngroups = (slots + per - 1) / per;
  float* groups = partial + (size_t)slots * n;
  hipLaunchKernelGGL(k_unet_reduce_slots, dim3(grid1(n).x, (unsigned)ngroups), dim3(256), 0, st, partial, slots, per,
                     n, groups);
  hipLaunchKernelGGL(k_unet_reduce_groups, grid1(n), dim3(256), 0, st, groups, ngroups, n, dw);
  hipLaunchKernelGGL(k_unet_bias_grad, dim3((unsigned)CS_UNET_BIAS_ROWS), dim3(256), 0, st,
                     gout, outv, cout, (size_t)s * s * s, bias_part);
  hipLaunchKernelGGL(k_unet_sum_rows, dim3((unsigned)cout), dim3(256), 0, st, bias_part, CS_UNET_BIAS_ROWS, cout, db);
  return hipGetLastError();
}

// The pooling, upsampling and concatenation kernels move 8 channels per 16-byte access.
static bool vec8_ok(int c, int stride, int off, const void* a, const void* b) {
  return c % 8 == 0 && stride % 8 == 0 && off % 8 == 0 && (((uintptr_t)a | (uintptr_t)b) % 16u) == 0u;
}
hipError_t cs_unet_launch_pool(hipStream_t st, const uint16_t* in, int c, int s, uint16_t* out) {
  if (!vec8_ok(c, 8, 0, in, out)) return hipErrorInvalidValue;
  const size_t n = (size_t)(s / 2) * (s / 2) * (s / 2) * (c / 8);
  hipLaunchKernelGGL(k_unet_pool, grid1(n), dim3(256), 0, st, in, c, s, out);
  return hipGetLastError();
}
hipError_t cs_unet_launch_pool_back(hipStream_t st, const uint16_t* g_out, int c, int s, const uint16_t* add,
                                    int add_stride, int add_off, uint16_t* g_in) {
  if (!vec8_ok(c, add ? add_stride : 8, add ? add_off : 0, g_out, g_in) || ((uintptr_t)add % 16u) != 0u)
    return hipErrorInvalidValue;
  const size_t n = (size_t)s * s * s * (c / 8);
  hipLaunchKernelGGL(k_unet_pool_back, grid1(n), dim3(256), 0, st, g_out, c, s, add, add_stride, add_off, g_in);
  return hipGetLastError();
}
hipError_t cs_unet_launch_up(hipStream_t st, const uint16_t* in, int c, int s, uint16_t* cat, int stride, int off) {
  if (!vec8_ok(c, stride, off, in, cat)) return hipErrorInvalidValue;
  const size_t n = (size_t)(2 * s) * (2 * s) * (2 * s) * (c / 8);
  hipLaunchKernelGGL(k_unet_up, grid1(n), dim3(256), 0, st, in, c, s, cat, stride, off);
  return hipGetLastError();
}
hipError_t cs_unet_launch_up_back(hipStream_t st, const uint16_t* g_cat, int stride, int off, int c, int s,
                                  uint16_t* g_in) {
  if (!vec8_ok(c, stride, off, g_cat, g_in)) return hipErrorInvalidValue;
  const size_t n = (size_t)s * s * s * (c / 8);
  hipLaunchKernelGGL(k_unet_up_back, grid1(n), dim3(256), 0, st, g_cat, stride, off, c, s, g_in);
  return hipGetLastError();
}
hipError_t cs_unet_launch_put(hipStream_t st, const uint16_t* in, int c, size_t nv, uint16_t* cat, int stride,
                              int off) {
  if (!vec8_ok(c, stride, off, in, cat)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_unet_put, grid1(nv * (c / 8)), dim3(256), 0, st, in, c, nv, cat, stride, off);
  return hipGetLastError();
}
hipError_t cs_unet_launch_head(hipStream_t st, const uint16_t* x, const float* wh, size_t nv, float* pred) {
  hipLaunchKernelGGL(k_unet_head, grid1(nv), dim3(256), 0, st, x, wh, nv, pred);
  return hipGetLastError();
}
hipError_t cs_unet_launch_loss(hipStream_t st, const float* pred, const float* target, const uint8_t* mask, size_t nv,
                               float inv_count, float* gpred, float* block_sse, float* sse) {
  const dim3 g = grid1(nv);
  hipLaunchKernelGGL(k_unet_loss, g, dim3(256), 0, st, pred, target, mask, nv, inv_count, gpred, block_sse);
  hipLaunchKernelGGL(k_unet_sum_rows, dim3(1), dim3(256), 0, st, block_sse, (int)g.x, 1, sse);
  return hipGetLastError();
}
hipError_t cs_unet_launch_head_back(hipStream_t st, const uint16_t* x, const float* wh, const float* gpred, size_t nv,
                                    uint16_t* g_x, float* block_part, float* dwh) {
  const dim3 g = grid1(nv);
  hipLaunchKernelGGL(k_unet_head_back, g, dim3(256), 0, st, x, wh, gpred, nv, g_x, block_part);
  hipLaunchKernelGGL(k_unet_sum_rows, dim3(17), dim3(256), 0, st, block_part, (int)g.x, 17, dwh);
  return hipGetLastError();
}
hipError_t cs_unet_launch_prep_weights(hipStream_t st, const float* w, int cin, int cout, uint16_t* wf, uint16_t* wd) {
  hipLaunchKernelGGL(k_unet_prep_weights, grid1((size_t)27 * cin * cout), dim3(256), 0, st, w, cin, cout, wf, wd);
  return hipGetLastError();
}
hipError_t cs_unet_launch_adam(hipStream_t st, float* p, const float* g, float* m, float* v, size_t n, float lr_t,
                               float b1, float b2, float eps) {
  hipLaunchKernelGGL(k_unet_adam, grid1(n), dim3(256), 0, st, p, g, m, v, n, lr_t, b1, b2, eps);
  return hipGetLastError();
}
hipError_t cs_unet_launch_scale(hipStream_t st, const float* pred, size_t nv, float scale, float* out) {
  hipLaunchKernelGGL(k_unet_scale, grid1(nv), dim3(256), 0, st, pred, nv, scale, out);
  return hipGetLastError();
}
hipError_t cs_unet_launch_pack_features(hipStream_t st, const float* f4, size_t nv, uint16_t* x) {
  hipLaunchKernelGGL(k_unet_pack_features, grid1(nv * 16), dim3(256), 0, st, f4, nv, x);
  return hipGetLastError();
}
