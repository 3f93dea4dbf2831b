// cs_kernels.hip — gfx950 kernels of the MLS-MPM substep.
//
// Each kernel restates one or more WGSL passes of the reference solver
// (crates/sim-wasm/src/mpm_3d/shader.rs) with the same binary32 operation
// order, so a pass run here equals the same pass on the CPU bit for bit.
// Passes that only touch their own cell are fused into the pass before them
// where the reference's pass order allows it (mod.rs:339-383):
//   k_grid      = grid_update + boundary_project
//   k_project   = project_pressure + boundary_project
//   k_pack      = packing_apply + boundary_project (packing_prepare is folded
//                 into k_classify: it only reads the cell's kind and volumes)
//   k_visc      = viscosity_prepare + viscosity_apply + boundary_project,
//                 double-buffered instead of the reference's lane scratch.
#include "cs_common.h"
#include "cs_kernels.h"

#include "cs_batch.h"
#include "cs_device.h"
#include "cs_bed.h"
#include "cs_sdf.h"

// ---------------------------------------------------------------------------------
__global__ void k_fill_i32(int32_t* __restrict__ p, size_t n, int32_t v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}
// The same, skipped when the substep's parameter block is halted.
__global__ void k_fill_i32_gated(const cs_dparams* __restrict__ dp, int32_t* __restrict__ p, size_t n, int32_t v) {
  if (dp->halt) return;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// ---- P2G (shader.rs:955-1052), one thread per particle, global i32 atomics -------
// The layout of the reference dispatch; the tiled P2G of cs_p2g.hip is the
// default and gives the same lanes bit for bit (CS_P2G=atomic selects this).
__global__ void __launch_bounds__(256) k_p2g_atomic(const cs_dparams* __restrict__ dp, const float4* __restrict__ part,
                                                    const float4* __restrict__ aff, int32_t* __restrict__ lanes,
                                                    uint32_t* __restrict__ metrics) {
  CS_PARAMS_FROM(dp);
  p2g_atomic_particle(u, part, aff, lanes, metrics, blockIdx.x * blockDim.x + threadIdx.x);
}

// ---- grid_update + boundary_project (shader.rs:1056-1082, 1630-1679) -------------
__device__ __forceinline__ void grid_cell(const cs_params& u, const int32_t* __restrict__ lanes,
                                              float4* __restrict__ gvel, const float* __restrict__ sdf, uint32_t idx) {
  if (idx >= u.nc) return;
  float inv_fp = u.inv_fp_scale;
  float mass = (float)lanes[idx] * inv_fp;
  if (mass < 1e-6f) {
    gvel[idx] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  cs_f3 v = f3((float)lanes[u.nc + idx] * inv_fp / mass, (float)lanes[2u * u.nc + idx] * inv_fp / mass,
               (float)lanes[3u * u.nc + idx] * inv_fp / mass);
  v.y += u.gravity * u.dt;
  v = cap_speed(u, v);
  uint32_t ix, iy, iz;
  decompose(u, idx, ix, iy, iz);
  v = boundary_apply(u, sdf, v, cell_center(u, ix, iy, iz));
  gvel[idx] = make_float4(v.x, v.y, v.z, mass);
}
__global__ void __launch_bounds__(256) k_grid(const cs_dparams* __restrict__ dp, const int32_t* __restrict__ lanes,
                                              float4* __restrict__ gvel, const float* __restrict__ sdf) {
  CS_PARAMS_FROM(dp);
  grid_cell(u, lanes, gvel, sdf, blockIdx.x * blockDim.x + threadIdx.x);
}

// The central-difference divergence of grid_vel at a cell with the ghost
// mirror on solid and off-grid faces (shader.rs:1278-1335); solidity from the
// cached class mask or live SDF samples, as use_sdf_cache selects.
__device__ __forceinline__ float cell_divergence(const cs_params& u, const float4* __restrict__ gvel,
                                                 const uint8_t* __restrict__ cls, const float* __restrict__ sdf,
                                                 uint32_t idx, uint32_t ix, uint32_t iy, uint32_t iz, cs_f3 cc) {
  const bool cache = u.use_sdf_cache != 0u;
  float4 sv = gvel[idx];
  float vxm = -sv.x, vxp = -sv.x, vym = -sv.y, vyp = -sv.y, vzm = -sv.z, vzp = -sv.z;
  uint32_t plane = u.gx * u.gy;
  const bool bxm = ix > 0u, bxp = ix + 1u < u.gx, bym = iy > 0u, byp = iy + 1u < u.gy, bzm = iz > 0u,
             bzp = iz + 1u < u.gz;
  if (cache) {
    // every neighbour's class byte and velocity component loaded at once (indices
    // clamped onto the cell itself off the grid), then selected
    const uint32_t nxm = bxm ? idx - 1u : idx, nxp = bxp ? idx + 1u : idx, nym = bym ? idx - u.gx : idx;
    const uint32_t nyp = byp ? idx + u.gx : idx, nzm = bzm ? idx - plane : idx, nzp = bzp ? idx + plane : idx;
    const float* gf = reinterpret_cast<const float*>(gvel);
    const uint8_t cxm = cls[nxm], cxp = cls[nxp], cym = cls[nym], cyp = cls[nyp], czm = cls[nzm], czp = cls[nzp];
    const float gxm = gf[4 * (size_t)nxm], gxp = gf[4 * (size_t)nxp], gym = gf[4 * (size_t)nym + 1];
    const float gyp = gf[4 * (size_t)nyp + 1], gzm = gf[4 * (size_t)nzm + 2], gzp = gf[4 * (size_t)nzp + 2];
    if (bxm && !cxm) vxm = gxm;
    if (bxp && !cxp) vxp = gxp;
    if (bym && !cym) vym = gym;
    if (byp && !cyp) vyp = gyp;
    if (bzm && !czm) vzm = gzm;
    if (bzp && !czp) vzp = gzp;
  } else {
    float d = u.dx;
#define CS_OPEN(px, py, pz) (sample_sdf(u, sdf, f3(px, py, pz)) >= 0.0f)
    if (bxm && CS_OPEN(cc.x + -d, cc.y + 0.0f, cc.z + 0.0f)) vxm = gvel[idx - 1u].x;
    if (bxp && CS_OPEN(cc.x + d, cc.y + 0.0f, cc.z + 0.0f)) vxp = gvel[idx + 1u].x;
    if (bym && CS_OPEN(cc.x + 0.0f, cc.y + -d, cc.z + 0.0f)) vym = gvel[idx - u.gx].y;
    if (byp && CS_OPEN(cc.x + 0.0f, cc.y + d, cc.z + 0.0f)) vyp = gvel[idx + u.gx].y;
    if (bzm && CS_OPEN(cc.x + 0.0f, cc.y + 0.0f, cc.z + -d)) vzm = gvel[idx - plane].z;
    if (bzp && CS_OPEN(cc.x + 0.0f, cc.y + 0.0f, cc.z + d)) vzp = gvel[idx + plane].z;
#undef CS_OPEN
  }
  float div = 0.5f * u.inv_dx * ((vxp - vxm) + (vyp - vym) + (vzp - vzm));
  return div;
}

// ---- classify_cells (shader.rs:1181-1367) + packing_prepare (shader.rs:1566-1583) ----
struct cs_cls_metrics {
  uint32_t max_div, max_res, fluid, div_clamp, p_clamp;
};
// Writes kind, the kind/fill word, pressure 0, the right-hand side c = dx*dx*rhs
// of the pressure update (shader.rs:1436-1437, from the quantised divergence)
// and the packing pressure.
__device__ __forceinline__ void classify_cell(const cs_params& u, const float4* __restrict__ gvel,
                                                  const uint8_t* __restrict__ cls, const int32_t* __restrict__ lanes,
                                                  const int32_t* __restrict__ bed_id,
                                                  const cs_bed_cell* __restrict__ bed, int8_t* __restrict__ kind_out,
                                                  float* __restrict__ kf_out, int32_t* __restrict__ pres,
                                                  float* __restrict__ rhs, int32_t* __restrict__ pack,
                                                  float* __restrict__ tgt,
                                                  uint32_t* __restrict__ metrics, const float* __restrict__ sdf,
                                                  uint32_t* __restrict__ brick_flag, uint32_t idx, cs_cls_metrics& cm) {
  if (idx >= u.nc) return;
  float occ = u.nominal_mass * 0.1f;
  float mass = gvel[idx].w;
  pres[idx] = 0;
  uint32_t ix, iy, iz;
  decompose(u, idx, ix, iy, iz);
  cs_f3 cc = cell_center(u, ix, iy, iz);
  // solid test from the cached class mask (use_sdf_cache, mod.rs:128) or from
  // live SDF samples at the cell and neighbour centres (shader.rs:1207-1211, 1286-1333)
  const bool cache = u.use_sdf_cache != 0u;
  bool self_solid = cache ? cls[idx] != 0 : sample_sdf(u, sdf, cc) < 0.0f;
  if (self_solid || mass <= occ) {
    kind_out[idx] = self_solid ? CS_CELL_SOLID : CS_CELL_AIR;
    kf_out[idx] = self_solid ? CS_KF_SOLID : CS_KF_AIR;
    rhs[idx] = 0.0f;
    pack[idx] = 0;
    return;
  }
  int kind;
  const int32_t my_bed = bed_id[idx];
  {
    // a face off the grid or toward a low-mass neighbour makes a surface cell:
    // the six neighbour masses loaded at once (indices clamped onto the cell)
    const uint32_t plane = u.gx * u.gy;
    const bool bxm = ix > 0u, bxp = ix + 1u < u.gx, bym = iy > 0u, byp = iy + 1u < u.gy, bzm = iz > 0u,
               bzp = iz + 1u < u.gz;
    const float m0 = gvel[bxm ? idx - 1u : idx].w, m1 = gvel[bxp ? idx + 1u : idx].w;
    const float m2 = gvel[bym ? idx - u.gx : idx].w, m3 = gvel[byp ? idx + u.gx : idx].w;
    const float m4 = gvel[bzm ? idx - plane : idx].w, m5 = gvel[bzp ? idx + plane : idx].w;
    const bool air = !bxm || !bxp || !bym || !byp || !bzm || !bzp || m0 <= occ || m1 <= occ || m2 <= occ ||
                     m3 <= occ || m4 <= occ || m5 <= occ;
    kind = my_bed >= 0 ? CS_CELL_BED : (air ? CS_CELL_SURFACE : CS_CELL_INTERIOR);
  }
  float div = cell_divergence(u, gvel, cls, sdf, idx, ix, iy, iz, cc);
  float rest = (float)lanes[4u * u.nc + idx] * u.inv_fp_scale;
  float cur = (float)lanes[5u * u.nc + idx] * u.inv_fp_scale;
  float target = volume_target_div(u, rest, cur);
  float cv = u.dx * u.dx * u.dx;
  if (kind == CS_CELL_BED) {
    const cs_bed_cell& b = bed[my_bed];
    // bed_pore_projection_target_divergence (shader.rs:621-666) with the
    // particle-independent factors taken from the bed table
    float pore = target;
    float overfill = cs_max(cur - b.pore_capacity, 0.0f) / cs_max(cv, 1e-8f);
    pore = cs_max(pore, overfill * b.overfill_resp);
    target = (pore - b.solid_fraction * b.solid_div) / b.porosity;
  }
  float dval = div - target;
  tgt[idx] = target;
  float dcl = cs_clamp(dval, -u.div_clamp, u.div_clamp);
  uint32_t fires_d = dcl != dval ? 1u : 0u;
  int32_t dfp = cs_f2i(dcl * u.fp_scale);
  float fill;
  if (kind == CS_CELL_SURFACE) {
    float dep = cs_max(rest, cur) / cs_max(cv, 1e-8f);
    fill = cs_clamp(dep, 0.0f, 1.0f);
  } else {
    fill = 1.0f;
  }
  float r = (float)dfp * u.inv_fp_scale * fill / cs_max(u.dt, 1e-6f);
  rhs[idx] = u.dx * u.dx * r;
  kind_out[idx] = (int8_t)kind;
  kf_out[idx] = kind == CS_CELL_BED ? CS_KF_BED : (kind == CS_CELL_INTERIOR ? 1.0f : fill);
  // packing_prepare
  uint32_t fires_p = 0;
  if (kind == CS_CELL_BED) {
    pack[idx] = 0;
  } else {
    float packed = cs_max(rest, cur) / cs_max(cv, 1e-8f);
    float over = cs_max(packed - cs_min(u.max_rest_fraction, 1.0f), 0.0f);
    float pv = u.bulk_K * over;
    float pcl = cs_clamp(pv, 0.0f, u.pressure_clamp);
    fires_p = pcl != pv ? 1u : 0u;
    pack[idx] = cs_f2i(pcl * u.fp_scale);
  }
  cm.max_div = cs_f2u(cs_clamp(fabsf(div) * u.metrics_div_fp_scale, 0.0f, 2147483648.0f));
  cm.max_res = cs_f2u(cs_clamp(fabsf(dval) * u.metrics_div_fp_scale, 0.0f, 2147483648.0f));
  cm.fluid = 1u;
  cm.div_clamp = fires_d;
  cm.p_clamp = fires_p;
  brick_flag[((iz / CS_BRICK) * u.nby + iy / CS_BRICK) * u.nbx + ix / CS_BRICK] = 1u;
}
// classify_cell for NC cells per lane (idx0, idx0 + stride, ...), then their
// metrics (max |div|, max |div - target|, fluid cells, clamp fires) reduced
// over the wavefront and added with one atomic per slot and workgroup (of at
// most 256 threads): the slots are a max and counts, so the result does not
// depend on how the cells were grouped, and NC cells a lane cut the atomics on
// the five addresses NC-fold.  Every thread of the workgroup must call it.
template <int NC>
__device__ __forceinline__ void classify_wave(const cs_params& u, const float4* __restrict__ gvel,
                                              const uint8_t* __restrict__ cls, const int32_t* __restrict__ lanes,
                                              const int32_t* __restrict__ bed_id,
                                              const cs_bed_cell* __restrict__ bed, int8_t* __restrict__ kind_out,
                                              float* __restrict__ kf_out, int32_t* __restrict__ pres,
                                              float* __restrict__ rhs, int32_t* __restrict__ pack,
                                              float* __restrict__ tgt, uint32_t* __restrict__ metrics,
                                              const float* __restrict__ sdf, uint32_t* __restrict__ brick_flag,
                                              uint32_t idx0, uint32_t stride) {
  cs_cls_metrics cm{};
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    cs_cls_metrics c{};
    classify_cell(u, gvel, cls, lanes, bed_id, bed, kind_out, kf_out, pres, rhs, pack, tgt, metrics, sdf, brick_flag,
                  idx0 + (uint32_t)j * stride, c);
    cm.max_div = max(cm.max_div, c.max_div);
    cm.max_res = max(cm.max_res, c.max_res);
    cm.fluid += c.fluid;
    cm.div_clamp += c.div_clamp;
    cm.p_clamp += c.p_clamp;
  }
  for (int off = 32; off > 0; off >>= 1) {
    cm.max_div = max(cm.max_div, (uint32_t)__shfl_xor((int)cm.max_div, off));
    cm.max_res = max(cm.max_res, (uint32_t)__shfl_xor((int)cm.max_res, off));
    cm.fluid += (uint32_t)__shfl_xor((int)cm.fluid, off);
    cm.div_clamp += (uint32_t)__shfl_xor((int)cm.div_clamp, off);
    cm.p_clamp += (uint32_t)__shfl_xor((int)cm.p_clamp, off);
  }
  // then over the workgroup's wavefronts: atomics on one address are
  // serialised device-wide, so one per workgroup and slot
  __shared__ cs_cls_metrics s_cm[4];
  const uint32_t wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63u) == 0u) s_cm[wave] = cm;
  __syncthreads();
  if (threadIdx.x == 0u) {
    for (uint32_t w = 1; w < (blockDim.x >> 6); ++w) {
      cm.max_div = max(cm.max_div, s_cm[w].max_div);
      cm.max_res = max(cm.max_res, s_cm[w].max_res);
      cm.fluid += s_cm[w].fluid;
      cm.div_clamp += s_cm[w].div_clamp;
      cm.p_clamp += s_cm[w].p_clamp;
    }
    if (cm.fluid) {
      atomicMax(&metrics[CS_M_MAX_ABS_DIV], cm.max_div);
      atomicMax(&metrics[CS_M_PRE_MAX_ABS_RES], cm.max_res);
      atomicAdd(&metrics[CS_M_FLUID_CELLS], cm.fluid);
      if (cm.div_clamp) atomicAdd(&metrics[CS_M_DIV_CLAMP], cm.div_clamp);
      if (cm.p_clamp) atomicAdd(&metrics[CS_M_P_CLAMP], cm.p_clamp);
    }
  }
}
constexpr uint32_t kClassifyCells = 4;  // cells per thread of k_classify, a workgroup's stride apart
__global__ void __launch_bounds__(256) k_classify(const cs_dparams* __restrict__ dp, const float4* __restrict__ gvel,
                                                  const uint8_t* __restrict__ cls, const int32_t* __restrict__ lanes,
                                                  const int32_t* __restrict__ bed_id,
                                                  const cs_bed_cell* __restrict__ bed, int8_t* __restrict__ kind_out,
                                                  float* __restrict__ kf_out, int32_t* __restrict__ pres,
                                                  float* __restrict__ rhs, int32_t* __restrict__ pack,
                                                  float* __restrict__ tgt,
                                                  uint32_t* __restrict__ metrics, const float* __restrict__ sdf,
                                                  uint32_t* __restrict__ brick_flag) {
  CS_PARAMS_FROM(dp);
  classify_wave<kClassifyCells>(u, gvel, cls, lanes, bed_id, bed, kind_out, kf_out, pres, rhs, pack, tgt, metrics,
                                sdf, brick_flag, blockIdx.x * blockDim.x * kClassifyCells + threadIdx.x, blockDim.x);
}

// ---- warm starts (injected after classify zeroed the pressure) -------------------
__device__ __forceinline__ void warm_previous_cell(const cs_params& u, const float* __restrict__ kf, const int32_t* __restrict__ prev,
                                int32_t* __restrict__ pres, uint32_t idx) {
  if (idx >= u.nc) return;
  if (kf_fluid(kf[idx])) pres[idx] = prev[idx];
}
__global__ void k_warm_previous(const cs_dparams* __restrict__ dp,
                                const float* __restrict__ kf, const int32_t* __restrict__ prev,
                                int32_t* __restrict__ pres) {
  CS_PARAMS_FROM(dp);
  warm_previous_cell(u, kf, prev, pres, blockIdx.x * blockDim.x + threadIdx.x);
}
__global__ void k_warm_external(cs_params u, const float* __restrict__ kf, const float* __restrict__ warm,
                                int32_t* __restrict__ pres, uint32_t* __restrict__ metrics) {
  uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= u.nc) return;
  if (!kf_fluid(kf[idx])) return;
  float v = warm[idx];
  float cl = cs_clamp(v, -u.pressure_clamp, u.pressure_clamp);
  if (cl != v) atomicAdd(&metrics[CS_M_P_CLAMP], 1u);
  pres[idx] = cs_f2i(cl * u.fp_scale);
}
__global__ void k_copy_i32(const int32_t* __restrict__ src, int32_t* __restrict__ dst, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}
__global__ void k_copy_i32_gated(const cs_dparams* __restrict__ dp,
                                 const int32_t* __restrict__ src, int32_t* __restrict__ dst, size_t n) {
  if (dp->halt) return;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// ---- red-black Gauss-Seidel half sweep (pressure_update, shader.rs:1371-1453) ------
__global__ void __launch_bounds__(256) k_rbgs_half(cs_params u, uint32_t parity, const float* __restrict__ kf,
                                                   const float* __restrict__ rhs, int32_t* __restrict__ pres,
                                                   uint32_t* __restrict__ metrics) {
  uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= u.nc) return;
  uint32_t ix, iy, iz;
  decompose(u, idx, ix, iy, iz);
  if (((ix + iy + iz) & 1u) != parity) return;
  float skf = kf[idx];
  if (!kf_fluid(skf)) return;  // non-fluid pressure is already 0
  float sf = kf_fill(skf);
  float count = 0.0f, sum = 0.0f;
  for (int n = 0; n < 6; ++n) {
    int x = (int)ix + c_off6[n][0], y = (int)iy + c_off6[n][1], z = (int)iz + c_off6[n][2];
    if (!on_grid(u, x, y, z)) continue;
    uint32_t ni = cell_index(u, x, y, z);
    float nkf = kf[ni];
    if (kf_solid(nkf)) continue;
    float w = face_w(skf, sf, nkf);
    if (w <= 0.0f) continue;
    count += w;
    if (kf_fluid(nkf)) sum += w * ((float)pres[ni] * u.inv_fp_scale);
  }
  float pn = 0.0f;
  if (count > 0.0f) pn = (sum - rhs[idx]) / count;
  float cl = cs_clamp(pn, -u.pressure_clamp, u.pressure_clamp);
  if (cl != pn) atomicAdd(&metrics[CS_M_P_CLAMP], 1u);
  pres[idx] = cs_f2i(cl * u.fp_scale);
}

// ---- residual ||b - A p|| / ||b|| over fluid cells (per-block partials) ------------
__global__ void __launch_bounds__(256) k_residual(cs_params u, const float* __restrict__ kf,
                                                  const float* __restrict__ rhs, const int32_t* __restrict__ pres,
                                                  double* __restrict__ partials) {
  __shared__ double s_rr[256], s_bb[256];
  uint32_t tid = threadIdx.x;
  double rr = 0.0, bb = 0.0;
  for (uint32_t idx = blockIdx.x * blockDim.x + tid; idx < u.nc; idx += gridDim.x * blockDim.x) {
    float skf = kf[idx];
    if (!kf_fluid(skf)) continue;
    uint32_t ix, iy, iz;
    decompose(u, idx, ix, iy, iz);
    float sf = kf_fill(skf);
    float count = 0.0f;
    double nb = 0.0;
    for (int n = 0; n < 6; ++n) {
      int x = (int)ix + c_off6[n][0], y = (int)iy + c_off6[n][1], z = (int)iz + c_off6[n][2];
      if (!on_grid(u, x, y, z)) continue;
      uint32_t ni = cell_index(u, x, y, z);
      float nkf = kf[ni];
      if (kf_solid(nkf)) continue;
      float w = face_w(skf, sf, nkf);
      if (w <= 0.0f) continue;
      count += w;
      if (kf_fluid(nkf)) nb += (double)w * (double)((float)pres[ni] * u.inv_fp_scale);
    }
    if (count <= 0.0f) continue;
    if (fabsf((float)pres[idx] * u.inv_fp_scale) == u.pressure_clamp) continue;  // the bound replaces the row
    double ap = (double)count * (double)((float)pres[idx] * u.inv_fp_scale) - nb;
    double b = -(double)rhs[idx];
    rr += (b - ap) * (b - ap);
    bb += b * b;
  }
  s_rr[tid] = rr;
  s_bb[tid] = bb;
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (tid < s) {
      s_rr[tid] += s_rr[tid + s];
      s_bb[tid] += s_bb[tid + s];
    }
    __syncthreads();
  }
  if (tid == 0) {
    partials[2 * blockIdx.x] = s_rr[0];
    partials[2 * blockIdx.x + 1] = s_bb[0];
  }
}
__global__ void k_residual_final(const double* __restrict__ partials, uint32_t nblocks, double* __restrict__ out) {
  __shared__ double s_rr[256], s_bb[256];
  uint32_t tid = threadIdx.x;
  double rr = 0.0, bb = 0.0;
  for (uint32_t i = tid; i < nblocks; i += blockDim.x) {
    rr += partials[2 * i];
    bb += partials[2 * i + 1];
  }
  s_rr[tid] = rr;
  s_bb[tid] = bb;
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (tid < s) {
      s_rr[tid] += s_rr[tid + s];
      s_bb[tid] += s_bb[tid + s];
    }
    __syncthreads();
  }
  if (tid == 0) {
    double r;
    if (s_bb[0] > 0.0) r = sqrt(s_rr[0] / s_bb[0]);
    else r = s_rr[0] > 0.0 ? __builtin_huge_val() : 0.0;
    out[0] = r;
  }
}

// ---- bed reaction impulses (shader.rs:458-515) ------------------------------------------
__device__ __forceinline__ bool grain_solid(const cs_grains& gr, int32_t id) {  // is_valid_bed_solid_idx
  return id >= 0 && (uint32_t)id < gr.n && !(gr.aff[12 * (size_t)id + 3] < 0.5f);
}
// bed_impulse_delta_add_neighborhood: one share per stamped cell of the 3x3x3
// block, each share clamped per component and added in fixed point
__device__ void grain_impulse(const cs_params& u, const cs_grains& gr, int cx, int cy, int cz, cs_f3 imp) {
  float count = 0.0f;
  for (int di = -1; di <= 1; ++di)
    for (int dj = -1; dj <= 1; ++dj)
      for (int dk = -1; dk <= 1; ++dk) {
        const int x = cx + di, y = cy + dj, z = cz + dk;
        if (!on_grid(u, x, y, z)) continue;
        if (grain_solid(gr, gr.lookup[cell_index(u, x, y, z)])) count += 1.0f;
      }
  if (count <= 0.0f) return;
  const cs_f3 share = f3(imp.x / count, imp.y / count, imp.z / count);
  const float cap = u.nominal_mass * u.vel_cap * CS_BED_REACTION_IMPULSE_CAP;
  const int32_t ix = cs_f2i(cs_clamp(share.x, -cap, cap) * u.fp_scale);
  const int32_t iy = cs_f2i(cs_clamp(share.y, -cap, cap) * u.fp_scale);
  const int32_t iz = cs_f2i(cs_clamp(share.z, -cap, cap) * u.fp_scale);
  for (int di = -1; di <= 1; ++di)
    for (int dj = -1; dj <= 1; ++dj)
      for (int dk = -1; dk <= 1; ++dk) {
        const int x = cx + di, y = cy + dj, z = cz + dk;
        if (!on_grid(u, x, y, z)) continue;
        const int32_t id = gr.lookup[cell_index(u, x, y, z)];
        if (!grain_solid(gr, id)) continue;
        atomicAdd(&gr.delta[1u * gr.n + (uint32_t)id], ix);
        atomicAdd(&gr.delta[2u * gr.n + (uint32_t)id], iy);
        atomicAdd(&gr.delta[3u * gr.n + (uint32_t)id], iz);
      }
}

// ---- project_pressure + boundary_project (shader.rs:1457-1543, 1630-1679) ---------
__device__ __forceinline__ void project_cell(const cs_params& u, float4* __restrict__ gvel, const float* __restrict__ kf,
                                                 const int32_t* __restrict__ pres, const int32_t* __restrict__ bed_id,
                                                 const cs_bed_cell* __restrict__ bed, const float* __restrict__ sdf,
                                                 const cs_grains& gr, uint32_t idx) {
  if (idx >= u.nc) return;
  float4 gv = gvel[idx];
  if (gv.w < 1e-6f) return;
  uint32_t ix, iy, iz;
  decompose(u, idx, ix, iy, iz);
  cs_f3 v = f3(gv.x, gv.y, gv.z);
  float skf = kf[idx];
  if (kf_fluid(skf)) {
    float ph = (float)pres[idx] * u.inv_fp_scale;
    float sf = kf_fill(skf);
    float pn[6];
    for (int n = 0; n < 6; ++n) {
      int x = (int)ix + c_off6[n][0], y = (int)iy + c_off6[n][1], z = (int)iz + c_off6[n][2];
      float val = ph;
      if (on_grid(u, x, y, z)) {
        uint32_t ni = cell_index(u, x, y, z);
        float nkf = kf[ni];
        if (!kf_solid(nkf)) {
          float np = kf_fluid(nkf) ? (float)pres[ni] * u.inv_fp_scale : 0.0f;
          val = cs_mix(ph, np, face_w(skf, sf, nkf));
        }
      }
      pn[n] = val;
    }
    float h = 0.5f * u.inv_dx;
    cs_f3 grad = f3(h * (pn[1] - pn[0]), h * (pn[3] - pn[2]), h * (pn[5] - pn[4]));
    v = v - grad * u.dt;
    if (kf_bed(skf)) {
      int32_t id = bed_id[idx];
      if (id >= 0 && (uint32_t)id < u.num_bed_cells) {
        const cs_bed_cell& b = bed[id];
        float rate = u.water_nu / b.k_record;
        float damping = 1.0f / (1.0f + cs_max(rate * u.dt, 0.0f));
        cs_f3 bv = f3(b.bmv[0], b.bmv[1], b.bmv[2]);
        const cs_f3 before = v;
        v = bv + (v - bv) * damping;
        if (gr.n) {  // the reaction on the grains (shader.rs:1533-1535)
          const cs_f3 wi = (v - before) * gv.w;
          grain_impulse(u, gr, (int)ix, (int)iy, (int)iz, f3(-wi.x, -wi.y, -wi.z) * CS_BED_REACTION_ALPHA);
        }
      }
    }
    v = cap_speed(u, v);
  }
  v = boundary_apply(u, sdf, v, cell_center(u, ix, iy, iz));
  gvel[idx] = make_float4(v.x, v.y, v.z, gv.w);
}
__global__ void __launch_bounds__(256) k_project(const cs_dparams* __restrict__ dp,
                                                 float4* __restrict__ gvel, const float* __restrict__ kf,
                                                 const int32_t* __restrict__ pres, const int32_t* __restrict__ bed_id,
                                                 const cs_bed_cell* __restrict__ bed, const float* __restrict__ sdf,
                                                 cs_grains gr) {
  CS_PARAMS_FROM(dp);
  project_cell(u, gvel, kf, pres, bed_id, bed, sdf, gr, blockIdx.x * blockDim.x + threadIdx.x);
}

// ---- post-projection divergence (no reference pass) ---------------------------------
// The classify stencil on the projected velocity of the fluid cells: max|div|
// and max|div - target| into metric slots 5 and 6.
__device__ __forceinline__ void post_div_cell(const cs_params& u, const float4* __restrict__ gvel,
                                              const float* __restrict__ kf, const uint8_t* __restrict__ cls,
                                              const float* __restrict__ sdf, const float* __restrict__ tgt,
                                              uint32_t idx, uint32_t& fd, uint32_t& fr) {
  if (idx < u.nc && kf_fluid(kf[idx])) {
    uint32_t ix, iy, iz;
    decompose(u, idx, ix, iy, iz);
    const float div = cell_divergence(u, gvel, cls, sdf, idx, ix, iy, iz, cell_center(u, ix, iy, iz));
    fd = max(fd, cs_f2u(cs_clamp(fabsf(div) * u.metrics_div_fp_scale, 0.0f, 2147483648.0f)));
    fr = max(fr, cs_f2u(cs_clamp(fabsf(div - tgt[idx]) * u.metrics_div_fp_scale, 0.0f, 2147483648.0f)));
  }
}
// The maxima of the workgroup (of at most 256 threads) into metric slots 5 and
// 6, one atomic each.  Every thread of the workgroup must call it.
__device__ __forceinline__ void post_div_flush(uint32_t* __restrict__ metrics, uint32_t fd, uint32_t fr) {
  for (int off = 32; off > 0; off >>= 1) {
    fd = max(fd, (uint32_t)__shfl_xor((int)fd, off));
    fr = max(fr, (uint32_t)__shfl_xor((int)fr, off));
  }
  __shared__ uint32_t s_f[2][4];
  const uint32_t wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63u) == 0u) {
    s_f[0][wave] = fd;
    s_f[1][wave] = fr;
  }
  __syncthreads();
  if (threadIdx.x == 0u) {
    for (uint32_t w = 1; w < (blockDim.x >> 6); ++w) {
      fd = max(fd, s_f[0][w]);
      fr = max(fr, s_f[1][w]);
    }
    if (fd) atomicMax(&metrics[CS_M_POST_MAX_ABS_DIV], fd);
    if (fr) atomicMax(&metrics[CS_M_POST_MAX_ABS_RES], fr);
  }
}
__global__ void __launch_bounds__(256) k_post_div(cs_params u, const float4* __restrict__ gvel,
                                                  const float* __restrict__ kf, const uint8_t* __restrict__ cls,
                                                  const float* __restrict__ sdf, const float* __restrict__ tgt,
                                                  uint32_t* __restrict__ metrics) {
  uint32_t fd = 0u, fr = 0u;
  post_div_cell(u, gvel, kf, cls, sdf, tgt, blockIdx.x * blockDim.x + threadIdx.x, fd, fr);
  post_div_flush(metrics, fd, fr);
}
// The same over the active bricks only (every fluid cell lies in one): one
// workgroup of 256 per brick, two cells a thread.
__global__ void __launch_bounds__(256) k_post_div_bricks(const cs_dparams* __restrict__ dp,
                                                         const uint32_t* __restrict__ list,
                                                         const float4* __restrict__ gvel, const float* __restrict__ kf,
                                                         const uint8_t* __restrict__ cls, const float* __restrict__ sdf,
                                                         const float* __restrict__ tgt, uint32_t* __restrict__ metrics) {
  CS_PARAMS_FROM(dp);
  if (blockIdx.x >= dp->nactive) return;
  const uint32_t b = list[blockIdx.x];
  const uint32_t x0 = (b % u.nbx) * CS_BRICK, y0 = ((b / u.nbx) % u.nby) * CS_BRICK, z0 = (b / (u.nbx * u.nby)) * CS_BRICK;
  uint32_t fd = 0u, fr = 0u;
  for (uint32_t k = threadIdx.x; k < (uint32_t)(CS_BRICK * CS_BRICK * CS_BRICK); k += 256u) {
    const uint32_t x = x0 + (k % CS_BRICK), y = y0 + ((k / CS_BRICK) % CS_BRICK), z = z0 + k / (CS_BRICK * CS_BRICK);
    const uint32_t idx = x < u.gx && y < u.gy && z < u.gz ? cell_index(u, x, y, z) : u.nc;  // u.nc: no cell
    post_div_cell(u, gvel, kf, cls, sdf, tgt, idx, fd, fr);
  }
  post_div_flush(metrics, fd, fr);
}

// ---- packing_apply + boundary_project (shader.rs:1585-1626, 1630-1679) ------------
__device__ __forceinline__ void pack_cell(const cs_params& u, float4* __restrict__ gvel, const float* __restrict__ kf,
                                              const int32_t* __restrict__ pack, const float* __restrict__ sdf, uint32_t idx) {
  if (idx >= u.nc) return;
  float4 gv = gvel[idx];
  if (gv.w < 1e-6f) return;
  uint32_t ix, iy, iz;
  decompose(u, idx, ix, iy, iz);
  cs_f3 v = f3(gv.x, gv.y, gv.z);
  float skf = kf[idx];
  if (gv.w > u.nominal_mass * 0.1f && kf_fluid(skf)) {
    float ph = (float)pack[idx] * u.inv_fp_scale;
    float pn[6];
    for (int n = 0; n < 6; ++n) {
      int x = (int)ix + c_off6[n][0], y = (int)iy + c_off6[n][1], z = (int)iz + c_off6[n][2];
      float val = ph;
      if (on_grid(u, x, y, z)) {
        uint32_t ni = cell_index(u, x, y, z);
        float nkf = kf[ni];
        if (!kf_solid(nkf)) val = kf_fluid(nkf) ? (float)pack[ni] * u.inv_fp_scale : 0.0f;
      }
      pn[n] = val;
    }
    float h = 0.5f * u.inv_dx;
    cs_f3 grad = f3(h * (pn[1] - pn[0]), h * (pn[3] - pn[2]), h * (pn[5] - pn[4]));
    v = v - grad * u.dt;
    v = cap_speed(u, v);
  }
  v = boundary_apply(u, sdf, v, cell_center(u, ix, iy, iz));
  gvel[idx] = make_float4(v.x, v.y, v.z, gv.w);
}
__global__ void __launch_bounds__(256) k_pack(const cs_dparams* __restrict__ dp,
                                              float4* __restrict__ gvel, const float* __restrict__ kf,
                                              const int32_t* __restrict__ pack, const float* __restrict__ sdf) {
  CS_PARAMS_FROM(dp);
  pack_cell(u, gvel, kf, pack, sdf, blockIdx.x * blockDim.x + threadIdx.x);
}

// ---- viscosity_prepare + viscosity_apply + boundary_project (shader.rs:1086-1177) --
__device__ __forceinline__ void visc_cell(const cs_params& u, const float4* __restrict__ gin, float4* __restrict__ gout,
                                              const uint8_t* __restrict__ cls, const float* __restrict__ sdf,
                                              uint8_t* __restrict__ alias4, uint32_t idx) {
  if (idx >= u.nc) return;
  float occ = u.nominal_mass * 0.1f;
  float4 gv = gin[idx];
  cs_f3 vh = f3(gv.x, gv.y, gv.z);
  cs_f3 vs = vh;
  if (gv.w > occ && u.viscosity > 0.0f && gv.w > u.nominal_mass * 2.0f) {
    uint32_t ix, iy, iz;
    decompose(u, idx, ix, iy, iz);
    float alpha = cs_clamp(u.viscosity * u.dt / cs_max(u.dx * u.dx, 1e-6f), 0.0f, 0.14f);
    cs_f3 nsum = f3(0.0f, 0.0f, 0.0f);
    float wsum = 0.0f;
    uint32_t nfl = 0;
    for (int n = 0; n < 6; ++n) {
      int x = (int)ix + c_off6[n][0], y = (int)iy + c_off6[n][1], z = (int)iz + c_off6[n][2];
      if (!on_grid(u, x, y, z)) { wsum += 1.0f; continue; }
      uint32_t ni = cell_index(u, x, y, z);
      if (cls[ni]) { wsum += 1.0f; continue; }
      float4 ng = gin[ni];
      if (ng.w <= occ) continue;
      float nw = cs_clamp(ng.w / cs_max(cs_max(gv.w, u.nominal_mass), 1e-6f), 0.0f, 1.0f);
      nsum = nsum + f3(ng.x, ng.y, ng.z) * nw;
      wsum += nw;
      nfl += 1u;
    }
    if (!(nfl < 3u || wsum <= 1e-6f)) {
      cs_f3 avg = nsum / wsum;
      float blend = cs_clamp(alpha * wsum, 0.0f, 0.65f);
      vs = mix3(vh, avg, blend);
    }
  }
  // velocity_scratch_store / load round trip (shader.rs:266-283)
  vs = cap_speed(u, vs);
  const int32_t lz = cs_f2i(vs.z * u.fp_scale);
  // The scratch's z lane is the lane the reference reads as the cell kind
  // until the next clear (shader.rs:158-164): bed_dynamics' pressure gradient
  // takes a cell as solid when it holds CELL_SOLID.
  if (alias4) alias4[idx] = lz == CS_CELL_SOLID ? 1u : 0u;
  if (gv.w <= occ) {
    gout[idx] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  uint32_t ix, iy, iz;
  decompose(u, idx, ix, iy, iz);
  cs_f3 v = f3((float)cs_f2i(vs.x * u.fp_scale) * u.inv_fp_scale, (float)cs_f2i(vs.y * u.fp_scale) * u.inv_fp_scale,
               (float)lz * u.inv_fp_scale);
  v = boundary_apply(u, sdf, v, cell_center(u, ix, iy, iz));
  gout[idx] = make_float4(v.x, v.y, v.z, gv.w);
}
__global__ void __launch_bounds__(256) k_visc(const cs_dparams* __restrict__ dp,
                                              const float4* __restrict__ gin, float4* __restrict__ gout,
                                              const uint8_t* __restrict__ cls, const float* __restrict__ sdf,
                                              uint8_t* __restrict__ alias4) {
  CS_PARAMS_FROM(dp);
  visc_cell(u, gin, gout, cls, sdf, alias4, blockIdx.x * blockDim.x + threadIdx.x);
}

// ---- G2P (shader.rs:1683-1881) ------------------------------------------------------
// Where G2P reads an on-grid node's grid velocity and bed-record index: the
// global arrays, or a tile's footprint staged in LDS.  Every on-grid node of a
// particle binned to a tile lies in that tile's footprint: the bin is the tile
// of the base node clamped to the grid, so a base inside the grid is at most 7
// cells past the tile's first cell (its nodes at most 9), a base below the grid
// reaches nodes 0 and 1 only, and a base beyond it reaches no on-grid node.
struct G2PGlobal {
  const float4* __restrict__ gvel;
  const int32_t* __restrict__ bed_id;
  __device__ __forceinline__ float4 gv(uint32_t ci, int, int, int) const { return gvel[ci]; }
  __device__ __forceinline__ int32_t bid(uint32_t ci, int, int, int) const { return bed_id[ci]; }
};
struct G2PTile {
  const float4* s_gv;
  const int32_t* s_bid;
  int ox, oy, oz;  // the footprint's first node
  __device__ __forceinline__ int local(int cx, int cy, int cz) const {
    return ((cz - oz) * FP + (cy - oy)) * FP + (cx - ox);
  }
  __device__ __forceinline__ float4 gv(uint32_t, int cx, int cy, int cz) const { return s_gv[local(cx, cy, cz)]; }
  __device__ __forceinline__ int32_t bid(uint32_t, int cx, int cy, int cz) const { return s_bid[local(cx, cy, cz)]; }
};
template <class Grid>
__device__ __forceinline__ void g2p_particle(const cs_params& u, float4* __restrict__ part, float4* __restrict__ aff,
                                             const Grid& grid, const cs_bed_cell* __restrict__ bed,
                                             const float* __restrict__ sdf, const cs_grains& gr, uint32_t pid) {
  if (pid >= u.num_particles) return;
  float4 P = part[2 * pid], Vv = part[2 * pid + 1];
  float phase = aff[3 * pid].w;
  if (!(phase < 0.5f)) return;
  float mass_p = Vv.w;
  if (mass_p <= u.nominal_mass * 0.10f) {
    part[2 * pid + 1] = make_float4(Vv.x, Vv.y, Vv.z, 0.0f);
    return;
  }
  cs_f3 xp = f3(P.x, P.y, P.z);
  float J_old = P.w;
  cs_f3 gp = (xp - f3(u.origin[0], u.origin[1], u.origin[2])) * u.inv_dx;
  int bx = cs_f2i(floorf(gp.x - 0.5f)), by = cs_f2i(floorf(gp.y - 0.5f)), bz = cs_f2i(floorf(gp.z - 0.5f));
  cs_f3 fx = f3(gp.x - (float)bx, gp.y - (float)by, gp.z - (float)bz);
  float wx[3], wy[3], wz[3];
  wx[0] = 0.5f * (1.5f - fx.x) * (1.5f - fx.x);
  wx[1] = 0.75f - (fx.x - 1.0f) * (fx.x - 1.0f);
  wx[2] = 0.5f * (fx.x - 0.5f) * (fx.x - 0.5f);
  wy[0] = 0.5f * (1.5f - fx.y) * (1.5f - fx.y);
  wy[1] = 0.75f - (fx.y - 1.0f) * (fx.y - 1.0f);
  wy[2] = 0.5f * (fx.y - 0.5f) * (fx.y - 0.5f);
  wz[0] = 0.5f * (1.5f - fx.z) * (1.5f - fx.z);
  wz[1] = 0.75f - (fx.z - 1.0f) * (fx.z - 1.0f);
  wz[2] = 0.5f * (fx.z - 0.5f) * (fx.z - 0.5f);
  cs_f3 nv = f3(0, 0, 0), C0 = f3(0, 0, 0), C1 = f3(0, 0, 0), C2 = f3(0, 0, 0), bvs = f3(0, 0, 0);
  float supported = 0.0f, local_mass = 0.0f, bed_overlap = 0.0f, bed_k = 0.0f;
  float B = 4.0f * u.inv_dx * u.inv_dx;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) {
        int cx = bx + i, cy = by + j, cz = bz + k;
        if (!on_grid(u, cx, cy, cz)) continue;
        float w = wx[i] * wy[j] * wz[k];
        uint32_t ci = cell_index(u, (uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
        float4 g = grid.gv(ci, cx, cy, cz);
        cs_f3 gv = f3(g.x, g.y, g.z);
        cs_f3 dpos = (f3((float)i, (float)j, (float)k) - fx) * u.dx;
        if (g.w > 1e-6f) {
          nv = nv + gv * w;
          float wB = w * B;
          C0 = C0 + (gv * wB) * dpos.x;
          C1 = C1 + (gv * wB) * dpos.y;
          C2 = C2 + (gv * wB) * dpos.z;
          supported += w;
          local_mass += w * g.w;
        }
        if (u.num_bed_cells > 0u) {
          int32_t id = grid.bid(ci, cx, cy, cz);
          if (id >= 0 && (uint32_t)id < u.num_bed_cells) {
            const cs_bed_cell& b = bed[id];
            bed_overlap += w;
            bvs = bvs + f3(b.v_record[0], b.v_record[1], b.v_record[2]) * w;
            bed_k += w * b.k_record;
          }
        }
      }
  float support_ratio = cs_clamp(supported, 0.0f, 1.0f);
  float j_support = support_ratio;
  if (supported > 1e-6f) {
    float inv = 1.0f / supported;
    nv = nv * inv;
    C0 = C0 * inv;
    C1 = C1 * inv;
    C2 = C2 * inv;
  }
  bool in_cup = false;
  if (u.geo.cup_enabled) {
    float lim = u.geo.cup_radius + u.contact_offset;
    in_cup = xp.y < u.geo.cup_top_y && (xp.x * xp.x + xp.z * xp.z) < lim * lim;
  }
  float dense_ratio = cs_clamp(local_mass / cs_max(u.nominal_mass * 4.0f, 1e-6f), 0.0f, 1.0f);
  float porous = cs_clamp(bed_overlap, 0.0f, 1.0f);
  cs_f3 pbv = f3(0, 0, 0);
  float pk = u.min_k;
  if (bed_overlap > 1e-6f) {
    float ib = 1.0f / bed_overlap;
    pbv = bvs * ib;
    pk = cs_max(bed_k * ib, u.min_k);
  }
  if (support_ratio < 0.999f) {
    cs_f3 ball = f3(Vv.x, Vv.y + u.gravity * u.dt, Vv.z);
    float gain = cs_mix(in_cup ? 1.15f : 1.35f, 0.18f, porous);
    float cap = cs_mix(in_cup ? 0.95f : 0.995f, 0.16f, porous);
    float damp = cs_mix(1.0f, 0.12f, dense_ratio);
    float pres = cs_clamp((1.0f - support_ratio) * gain * damp, 0.0f, cap);
    nv = mix3(nv, ball, pres);
    float ad = 1.0f - pres * 0.75f;
    C0 = C0 * ad;
    C1 = C1 * ad;
    C2 = C2 * ad;
  }
  bool airborne = porous <= 0.05f && sample_sdf(u, sdf, xp) > u.contact_offset * 2.0f;
  if (airborne) {
    float dr = cs_clamp(local_mass / cs_max(u.nominal_mass * 4.0f, 1e-6f), 0.0f, 1.0f);
    j_support = cs_min(j_support, dr);
    cs_f3 ball = f3(Vv.x, Vv.y + u.gravity * u.dt, Vv.z);
    float gain = in_cup ? 0.72f : 0.98f;
    float cap = in_cup ? 0.88f : 0.98f;
    float pres = cs_clamp((1.0f - dr) * gain, 0.0f, cap);
    nv = mix3(nv, ball, pres);
    float ad = 1.0f - pres * 0.65f;
    C0 = C0 * ad;
    C1 = C1 * ad;
    C2 = C2 * ad;
  }
  if (porous > 1e-4f) {
    cs_f3 rel = nv - pbv;
    float rate = u.water_nu / pk;
    float dd = 1.0f / (1.0f + cs_max(rate * u.dt * porous, 0.0f));
    float rs = len3(rel);
    float idd = 1.0f / (1.0f + cs_max(rs * u.dt * u.inv_dx * porous * 0.35f, 0.0f));
    const cs_f3 before = nv;
    nv = pbv + (rel * dd) * idd;
    if (gr.n) {  // the reaction on the grains (shader.rs:1837-1843)
      const cs_f3 wi = (nv - before) * mass_p;
      grain_impulse(u, gr, cs_f2i(floorf((xp.x - u.origin[0]) * u.inv_dx)),
                    cs_f2i(floorf((xp.y - u.origin[1]) * u.inv_dx)), cs_f2i(floorf((xp.z - u.origin[2]) * u.inv_dx)),
                    f3(-wi.x, -wi.y, -wi.z) * CS_BED_REACTION_ALPHA);
    }
    float ad = cs_mix(1.0f, dd * idd, porous);
    C0 = C0 * ad;
    C1 = C1 * ad;
    C2 = C2 * ad;
    j_support = cs_max(j_support, porous);
  }
  float tr = C0.x + C1.y + C2.z;
  float Jc = cs_clamp(J_old > 0.0f ? J_old : 1.0f, 0.40f, 2.00f);
  float Ja = cs_clamp(Jc * cs_expf(cs_clamp(u.dt * tr, -0.35f, 0.35f)), 0.40f, 2.00f);
  float Jb = cs_smoothstep(0.35f, 0.85f, cs_clamp(j_support, 0.0f, 1.0f));
  float Jn = cs_clamp(cs_mix(Jc, Ja, Jb), 0.40f, 2.00f);
  cs_f3 npos = xp + nv * u.dt;
  cs_f3 mid = mix3(xp, npos, 0.5f);
  contact c = sdf_contact(u, sdf, mid, nv);
  nv = c.vel;
  c = sdf_contact(u, sdf, npos, nv);
  npos = c.pos;
  nv = c.vel;
  float margin = u.dx * 0.5f;
  npos.x = cs_clamp(npos.x, u.origin[0] + margin, u.bmax[0] - margin);
  npos.y = cs_clamp(npos.y, u.origin[1] + margin, u.bmax[1] - margin);
  npos.z = cs_clamp(npos.z, u.origin[2] + margin, u.bmax[2] - margin);
  part[2 * pid] = make_float4(npos.x, npos.y, npos.z, Jn);
  part[2 * pid + 1] = make_float4(nv.x, nv.y, nv.z, mass_p);
  aff[3 * pid] = make_float4(C0.x, C0.y, C0.z, phase);
  aff[3 * pid + 1] = make_float4(C1.x, C1.y, C1.z, 0.0f);
  aff[3 * pid + 2] = make_float4(C2.x, C2.y, C2.z, 0.0f);
}
__global__ void __launch_bounds__(256) k_g2p(const cs_dparams* __restrict__ dp,
                                             float4* __restrict__ part, float4* __restrict__ aff,
                                             const float4* __restrict__ gvel, const int32_t* __restrict__ bed_id,
                                             const cs_bed_cell* __restrict__ bed, const float* __restrict__ sdf,
                                             cs_grains gr) {
  CS_PARAMS_FROM(dp);
  g2p_particle(u, part, aff, G2PGlobal{gvel, bed_id}, bed, sdf, gr, blockIdx.x * blockDim.x + threadIdx.x);
}

// G2P over the tile bins P2G made this substep (the particles have not moved
// since): one workgroup per chunk of at most kChunk particles of a tile, the
// tile's footprint of grid velocities and bed-record indices staged in LDS.
// Every particle gives the same result as in k_g2p.
__global__ void __launch_bounds__(256) k_g2p_tile_chunks(const cs_dparams* __restrict__ dp,
                                                         float4* __restrict__ part, float4* __restrict__ aff,
                                                         const float4* __restrict__ gvel,
                                                         const int32_t* __restrict__ bed_id,
                                                         const cs_bed_cell* __restrict__ bed,
                                                         const float* __restrict__ sdf, cs_grains gr,
                                                         const uint32_t* __restrict__ tile_count,
                                                         const uint32_t* __restrict__ start,
                                                         const uint32_t* __restrict__ chunk,
                                                         const uint32_t* __restrict__ sorted) {
  CS_PARAMS_FROM(dp);
  __shared__ float4 s_gv[FP3];
  __shared__ int32_t s_bid[FP3];
  const uint32_t w = blockIdx.x;
  if (w >= chunk[u.nbricks]) return;
  uint32_t lo = 0, hi = u.nbricks;  // chunk[lo] <= w < chunk[hi]
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (chunk[mid] <= w) lo = mid;
    else hi = mid;
  }
  const uint32_t off = (w - chunk[lo]) * kChunk;
  const uint32_t n = min(kChunk, tile_count[lo] - off);
  const uint32_t first = start[lo] + off;
  const int ox = (int)((lo % u.nbx) * TB), oy = (int)(((lo / u.nbx) % u.nby) * TB),
            oz = (int)((lo / (u.nbx * u.nby)) * TB);
  const bool beds = u.num_bed_cells > 0u;
  for (uint32_t li = threadIdx.x; li < (uint32_t)FP3; li += 256u) {
    const int gx = ox + (int)(li % FP), gy = oy + (int)((li / FP) % FP), gz = oz + (int)(li / (FP * FP));
    const bool on = on_grid(u, gx, gy, gz);
    const uint32_t ci = on ? cell_index(u, (uint32_t)gx, (uint32_t)gy, (uint32_t)gz) : 0u;
    s_gv[li] = on ? gvel[ci] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    s_bid[li] = on && beds ? bed_id[ci] : -1;
  }
  __syncthreads();
  const G2PTile grid{s_gv, s_bid, ox, oy, oz};
  for (uint32_t i = threadIdx.x; i < n; i += 256u) g2p_particle(u, part, aff, grid, bed, sdf, gr, sorted[first + i]);
}
// The particles P2G left unbinned (not water, or at most a tenth of the
// nominal mass): G2P changes only their mass.
__global__ void __launch_bounds__(256) k_g2p_unbinned(const cs_dparams* __restrict__ dp, float4* __restrict__ part,
                                                      float4* __restrict__ aff, const float4* __restrict__ gvel,
                                                      const int32_t* __restrict__ bed_id,
                                                      const cs_bed_cell* __restrict__ bed,
                                                      const float* __restrict__ sdf, cs_grains gr,
                                                      const uint2* __restrict__ key) {
  CS_PARAMS_FROM(dp);
  const uint32_t pid = blockIdx.x * blockDim.x + threadIdx.x;
  if (pid >= u.num_particles || key[pid].x != kNoTile) return;
  g2p_particle(u, part, aff, G2PGlobal{gvel, bed_id}, bed, sdf, gr, pid);
}

// ---- prepare_render (shader.rs:2242-2276) --------------------------------------------
// RenderParticle (shader.rs:40-43) per particle: (pos.xyz, colour) and (radius,
// 0, 0, 0); an inactive particle is parked below the scene.  A bed particle
// (pid < num_bed) takes the bed radius and colours by -1 - saturation, the
// saturation read from its BedExtract (extract.w) when `bed_ext` is given.
__global__ void __launch_bounds__(256) k_pack_render(cs_params u, const float4* __restrict__ part,
                                                     const float4* __restrict__ aff, uint32_t num_bed,
                                                     const float* __restrict__ bed_ext,
                                                     float4* __restrict__ out) {
  const uint32_t pid = blockIdx.x * blockDim.x + threadIdx.x;
  if (pid >= u.num_particles) return;
  const float4 P = part[2 * pid], Vv = part[2 * pid + 1];
  const float phase = aff[3 * pid].w;
  if (Vv.w <= u.nominal_mass * 0.10f) {
    out[2 * pid] = make_float4(0.0f, -1e6f, 0.0f, -999.0f);
    out[2 * pid + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  float color_t, radius = u.dx * CS_WATER_SAMPLE_RADIUS_DX;
  if (phase < 0.5f) {
    const float speed = len3(f3(Vv.x, Vv.y, Vv.z));
    color_t = cs_clamp(speed / 10.0f, 0.0f, 2.0f);
  } else {
    float sat = 0.0f;
    if (pid < num_bed) {
      if (bed_ext) sat = bed_ext[8 * (size_t)pid + 7];
      radius = u.dx * CS_BED_SAMPLE_RADIUS_DX;
    }
    color_t = -1.0f - sat;
  }
  out[2 * pid] = make_float4(P.x, P.y, P.z, color_t);
  out[2 * pid + 1] = make_float4(radius, 0.0f, 0.0f, 0.0f);
}

// ---- bed of particles: lookup and table (shader.rs:2291-2335, 380-449, 621-666) -------
// bed_lookup_clear, plus the table slot counter of k_bed_table.
__global__ void __launch_bounds__(256) k_bed_lookup_clear(const cs_dparams* __restrict__ dp, uint32_t nc,
                                                          int32_t* __restrict__ lookup, uint32_t* __restrict__ counter) {
  if (dp->halt) return;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *counter = 0u;
  if (i < nc) lookup[i] = -1;
}

// bed_lookup_scatter: every non-water bed particle stamps its id into the 3x3x3
// cells around world_to_cell(pos) (shader.rs:668-671) with atomicMax, so the
// highest id wins whatever the order.
__global__ void __launch_bounds__(256) k_bed_lookup_scatter(const cs_dparams* __restrict__ dp, uint32_t nbed,
                                                            const float4* __restrict__ bed_part,
                                                            const float4* __restrict__ bed_aff,
                                                            int32_t* __restrict__ lookup) {
  CS_PARAMS_FROM(dp);
  const uint32_t bid = blockIdx.x * blockDim.x + threadIdx.x;
  if (bid >= nbed) return;
  if (bed_aff[3 * bid].w < 0.5f) return;  // is_water_phase (PHASE_WATER_MAX = 0.5)
  const float4 P = bed_part[2 * bid];
  // floor, then i32(): saturating (the clamp keeps the conversion defined)
  auto cell = [&](float p, float o) {
    const float g = floorf((p - o) * u.inv_dx);
    return (int)fminf(fmaxf(g, -2147483648.0f), 2147483520.0f);
  };
  const int cx = cell(P.x, u.origin[0]), cy = cell(P.y, u.origin[1]), cz = cell(P.z, u.origin[2]);
  for (int di = -1; di <= 1; ++di)
    for (int dj = -1; dj <= 1; ++dj)
      for (int dk = -1; dk <= 1; ++dk) {
        const int x = cx + di, y = cy + dj, z = cz + dk;
        if (x < 0 || y < 0 || z < 0 || x >= (int)u.gx || y >= (int)u.gy || z >= (int)u.gz) continue;
        atomicMax(&lookup[((uint32_t)z * u.gy + (uint32_t)y) * u.gx + (uint32_t)x], (int32_t)bid);
      }
}

// The table entry of every stamped cell, from the stamping particle's record
// (cs_bed_particle_record), in a slot taken from `counter`; bed_id maps each
// cell to its slot (-1 elsewhere).  The slots' order varies from run to run;
// every reader goes through bed_id, so no result does.
__global__ void __launch_bounds__(256) k_bed_table(const cs_dparams* __restrict__ dp,
                                                   const int32_t* __restrict__ lookup,
                                                   const float* __restrict__ bed_part,
                                                   const float* __restrict__ bed_ext, uint32_t capacity,
                                                   uint32_t* __restrict__ counter, int32_t* __restrict__ bed_id,
                                                   cs_bed_cell* __restrict__ table) {
  CS_PARAMS_FROM(dp);
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= u.nc) return;
  const int32_t id = lookup[c];
  if (id < 0) {
    bed_id[c] = -1;
    return;
  }
  uint32_t ix, iy, iz;
  decompose(u, c, ix, iy, iz);
  const cs_bed_record rec = cs_bed_particle_record(bed_ext + 8 * (size_t)id, bed_part + 8 * (size_t)id);
  auto at = [&](int x, int y, int z) -> int32_t {
    if (x < 0 || y < 0 || z < 0 || x >= (int)u.gx || y >= (int)u.gy || z >= (int)u.gz) return -2;
    return lookup[((uint32_t)z * u.gy + (uint32_t)y) * u.gx + (uint32_t)x];
  };
  auto vel = [&](int32_t k, float* v) {
    v[0] = bed_part[8 * (size_t)k + 4];
    v[1] = bed_part[8 * (size_t)k + 5];
    v[2] = bed_part[8 * (size_t)k + 6];
  };
  cs_bed_cell t;
  cs_bed_cell_build(u, u.gx, u.gy, u.gz, (int)ix, (int)iy, (int)iz, rec, at, vel, &t);
  const uint32_t slot = atomicAdd(counter, 1u);
  if (slot < capacity) {
    table[slot] = t;
    bed_id[c] = (int32_t)slot;
  } else {
    bed_id[c] = -1;  // cannot happen: at most 27 stamped cells per bed particle
  }
}

// ---- bed_coupling and extraction_advect (shader.rs:1885-1987) -----------------------------
// bed_water_delta_add (shader.rs:455-460): fixed-point water credit of one bed particle.
__device__ __forceinline__ void bed_water_add(const cs_params& u, int32_t* __restrict__ dw, int32_t idx, float amount) {
  if (amount <= 0.0f) return;
  atomicAdd(&dw[idx], cs_f2i(amount * u.fp_scale));
}

// bed_coupling over the water: a particle in a stamped cell gives part of its
// mass to the stamping bed particle's pore water, 35 % of it shared among the
// distinct bed particles of the six face neighbours (deposit_absorbed_bed_water,
// shader.rs:524-576).  The filter-paper branch (shader.rs:1904-1911) needs
// filter_absorption_rate > 0; the reference sets it to 0 (brew_config.rs:68),
// so it is not carried.  Every credit is an integer atomic and each particle
// reads only the bed state of the previous extraction pass, so the result does
// not depend on the order.
__global__ void __launch_bounds__(256) k_bed_coupling(const cs_dparams* __restrict__ dp, float4* __restrict__ part,
                                                      const float4* __restrict__ aff,
                                                      const int32_t* __restrict__ lookup, uint32_t nbed,
                                                      const float* __restrict__ bed_ext, float max_sat,
                                                      int32_t* __restrict__ dw) {
  CS_PARAMS_FROM(dp);
  const uint32_t pid = blockIdx.x * blockDim.x + threadIdx.x;
  if (pid >= u.num_particles) return;
  if (!(aff[3 * pid].w < 0.5f)) return;  // is_water_phase
  const float4 P = part[2 * pid];
  float4 V = part[2 * pid + 1];
  auto cell = [&](float p, float o) {
    const float g = floorf((p - o) * u.inv_dx);
    return (int)fminf(fmaxf(g, -2147483648.0f), 2147483520.0f);
  };
  const int cx = cell(P.x, u.origin[0]), cy = cell(P.y, u.origin[1]), cz = cell(P.z, u.origin[2]);
  if (cx < 0 || cy < 0 || cz < 0 || cx >= (int)u.gx || cy >= (int)u.gy || cz >= (int)u.gz) return;
  const float mass_p = V.w;
  const int32_t home = lookup[((uint32_t)cz * u.gy + (uint32_t)cy) * u.gx + (uint32_t)cx];
  if (home < 0 || (uint32_t)home >= nbed) return;
  const float pore = bed_ext[8 * (size_t)home + 0], saturation = bed_ext[8 * (size_t)home + 7];
  const float capacity = cs_max(max_sat - pore, 0.0f);
  if (capacity <= 1e-6f) return;
  const float abs_rate = CS_BED_ABSORPTION_RATE * (1.0f - saturation) * u.dt;
  const float speed = len3(f3(V.x, V.y, V.z));
  const float inactive = u.nominal_mass * 0.10f;
  float absorbed = cs_min(cs_min(mass_p * cs_clamp(abs_rate, 0.0f, 0.25f), mass_p * 0.5f), capacity);
  const float remaining_after_partial = mass_p - absorbed;
  if (remaining_after_partial > 0.0f && remaining_after_partial <= u.nominal_mass * 0.22f) {
    absorbed = cs_min(mass_p, capacity);
  } else if (saturation > 0.55f && speed < 1.35f) {
    const float almost = cs_min(mass_p, capacity);
    if (mass_p - almost <= u.nominal_mass * 0.35f) absorbed = almost;
  }
  if (mass_p - absorbed <= inactive && absorbed < mass_p) absorbed = cs_max(mass_p - inactive * 1.05f, 0.0f);
  if (absorbed <= 1e-6f) return;
  const float remaining = mass_p - absorbed;
  part[2 * pid + 1] = remaining <= inactive ? make_float4(0.0f, 0.0f, 0.0f, 0.0f)
                                            : make_float4(V.x, V.y, V.z, remaining);
  // deposit_absorbed_bed_water
  const float neighbor_share = absorbed * 0.35f;
  const float home_share = absorbed - neighbor_share;
  bed_water_add(u, dw, home, home_share);
  int32_t ids[6];
  uint32_t n = 0;
  const int off[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  for (int k = 0; k < 6; ++k) {
    const int x = cx + off[k][0], y = cy + off[k][1], z = cz + off[k][2];
    if (x < 0 || y < 0 || z < 0 || x >= (int)u.gx || y >= (int)u.gy || z >= (int)u.gz) continue;
    const int32_t id = lookup[((uint32_t)z * u.gy + (uint32_t)y) * u.gx + (uint32_t)x];
    if (id < 0 || id == home || (uint32_t)id >= nbed) continue;
    bool dup = false;
    for (uint32_t j = 0; j < n; ++j) dup = dup || ids[j] == id;
    if (!dup) ids[n++] = id;
  }
  if (n == 0) {
    bed_water_add(u, dw, home, neighbor_share);
    return;
  }
  const float each = neighbor_share / (float)n;
  for (uint32_t j = 0; j < n; ++j) bed_water_add(u, dw, ids[j], each);
}

// extraction_advect: each bed particle takes (and clears) its water credit into
// its pore water and saturation, then moves extractable to dissolved mass.
__global__ void __launch_bounds__(256) k_extraction_advect(const cs_dparams* __restrict__ dp,
                                                           uint32_t nbed, float max_sat,
                                                           int32_t* __restrict__ dw, float* __restrict__ bed_ext) {
  CS_PARAMS_FROM(dp);
  const uint32_t bid = blockIdx.x * blockDim.x + threadIdx.x;
  if (bid >= nbed) return;
  float* be = bed_ext + 8 * (size_t)bid;
  const float absorbed = (float)dw[bid] * u.inv_fp_scale;
  dw[bid] = 0;
  if (absorbed > 0.0f) {
    be[0] += absorbed;
    be[7] = cs_clamp(be[0] / cs_max(max_sat, 1e-6f), 0.0f, 1.0f);
  }
  const float sat = be[7];
  if (sat > 0.01f) {
    const float flux = CS_EXTRACTION_RATE * be[4] * sat * u.dt;
    be[4] = cs_max(be[4] - flux, 0.0f);
    be[5] += flux;
  }
}

hipError_t cs_launch_bed_coupling(hipStream_t st, const cs_pblock& P, float* part, const float* aff,
                                  const int32_t* lookup, uint32_t nbed, float* bed_ext, int32_t* dw) {
  const cs_params& u = P.u;
  if (nbed == 0) return hipSuccess;
  const float max_sat = CS_TARGET_BED_RETENTION_ML * 80.0f / (float)nbed;  // mod.rs:577-581
  if (u.num_particles)
    hipLaunchKernelGGL(k_bed_coupling, dim3((u.num_particles + 255) / 256), dim3(256), 0, st, P.d,
                       reinterpret_cast<float4*>(part), reinterpret_cast<const float4*>(aff), lookup, nbed, bed_ext,
                       max_sat, dw);
  hipLaunchKernelGGL(k_extraction_advect, dim3((nbed + 255) / 256), dim3(256), 0, st, P.d, nbed, max_sat, dw, bed_ext);
  return hipGetLastError();
}

// ---- bed_dynamics (shader.rs:1991-2238) ------------------------------------------------
// pressure_gradient_at_cell (shader.rs:238-259) as bed_dynamics sees it: the
// solved pressure, with a neighbour taken as solid (mirror) when the lane the
// reference reads as its kind holds CELL_SOLID; by now that lane holds the
// viscosity scratch (alias4, written by the viscosity pass).
__device__ __forceinline__ float grad_p_or_mirror(const cs_params& u, const int32_t* __restrict__ pres,
                                                  const uint8_t* __restrict__ alias4, int x, int y, int z, float mirror) {
  if (!on_grid(u, x, y, z)) return mirror;
  const uint32_t ci = cell_index(u, x, y, z);
  if (alias4[ci]) return mirror;
  return (float)pres[ci] * u.inv_fp_scale;
}
__device__ cs_f3 pressure_gradient_at_cell(const cs_params& u, const int32_t* __restrict__ pres,
                                           const uint8_t* __restrict__ alias4, int x, int y, int z) {
  if (!on_grid(u, x, y, z)) return f3(0.0f, 0.0f, 0.0f);
  const float ph = (float)pres[cell_index(u, x, y, z)] * u.inv_fp_scale;
  const float pxm = grad_p_or_mirror(u, pres, alias4, x - 1, y, z, ph);
  const float pxp = grad_p_or_mirror(u, pres, alias4, x + 1, y, z, ph);
  const float pym = grad_p_or_mirror(u, pres, alias4, x, y - 1, z, ph);
  const float pyp = grad_p_or_mirror(u, pres, alias4, x, y + 1, z, ph);
  const float pzm = grad_p_or_mirror(u, pres, alias4, x, y, z - 1, ph);
  const float pzp = grad_p_or_mirror(u, pres, alias4, x, y, z + 1, ph);
  const float h = 0.5f * u.inv_dx;
  return f3(h * (pxp - pxm), h * (pyp - pym), h * (pzp - pzm));
}

__device__ __forceinline__ int floor_cell(const cs_params& u, float p, int axis) {
  return cs_f2i(floorf((p - u.origin[axis]) * u.inv_dx));
}

// resolve_coffee_particle_packing (shader.rs:825-877) against the grains'
// positions from before the pass (`snap`): the reference reads them while
// they move, which depends on thread scheduling.
__device__ contact coffee_packing(const cs_params& u, const int32_t* __restrict__ lookup, uint32_t nbed,
                                  const float4* __restrict__ snap, cs_f3 pos, cs_f3 vel, uint32_t self) {
  cs_f3 cd = f3(0.0f, 0.0f, 0.0f);
  float closest = 1e9f;
  bool found = false;
  const int cx = floor_cell(u, pos.x, 0), cy = floor_cell(u, pos.y, 1), cz = floor_cell(u, pos.z, 2);
  for (int di = -1; di <= 1; ++di)
    for (int dj = -1; dj <= 1; ++dj)
      for (int dk = -1; dk <= 1; ++dk) {
        const int x = cx + di, y = cy + dj, z = cz + dk;
        if (!on_grid(u, x, y, z)) continue;
        const int32_t o = lookup[cell_index(u, x, y, z)];
        if (o < 0 || (uint32_t)o == self || (uint32_t)o >= nbed) continue;
        const float4 q = snap[o];
        const cs_f3 d = pos - f3(q.x, q.y, q.z);
        const float dist = len3(d);
        if (dist < closest) {
          closest = dist;
          cd = d;
          found = true;
        }
      }
  const float min_sep = bed_particle_radius(u) * 1.78f;
  if (found && closest < min_sep) {
    cs_f3 n = f3(0.0f, 1.0f, 0.0f);
    if (closest > 1e-5f) n = cd / closest;
    const float pen = min_sep - closest;
    pos = pos + n * cs_min(pen * 0.65f, u.dx * 0.22f);
    const float vn = dot3(vel, n);
    if (vn < 0.0f) vel = vel - n * vn;
  }
  return contact{pos, vel};
}

__global__ void __launch_bounds__(256) k_bed_snapshot(uint32_t nbed, const float4* __restrict__ bed_part,
                                                      float4* __restrict__ snap) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nbed) snap[b] = bed_part[2 * b];
}

// One grain: its reaction impulse, then either the suspended drift with the
// water or the anchored, plastically compacting matrix motion around its rest
// position; bed_damping 8.5, bed_compaction_rate 5.5, bed_impact 8.0
// (mod.rs:644-649, brew_config.rs:69-70).
__global__ void __launch_bounds__(256) k_bed_dynamics(const cs_dparams* __restrict__ dp,
                                                      uint32_t nbed, float4* __restrict__ bed_part,
                                                      float4* __restrict__ bed_aff, float* __restrict__ bed_ext,
                                                      int32_t* __restrict__ dw, const float4* __restrict__ snap,
                                                      const int32_t* __restrict__ lookup,
                                                      const float4* __restrict__ gvel, const int32_t* __restrict__ pres,
                                                      const uint8_t* __restrict__ alias4,
                                                      const float* __restrict__ sdf) {
  CS_PARAMS_FROM(dp);
  const uint32_t bid = blockIdx.x * blockDim.x + threadIdx.x;
  if (bid >= nbed) return;
  const float bed_damping = 8.5f, bed_compaction = 5.5f, bed_impact = 8.0f;
  const float dx = u.dx, dt = u.dt;
  const float floor_y = coffee_filter_floor_y(u);
  const float4 P = bed_part[2 * bid], Vv = bed_part[2 * bid + 1];
  const float4 A0 = bed_aff[3 * bid], A1 = bed_aff[3 * bid + 1];
  float* be = bed_ext + 8 * (size_t)bid;
  float phase = A0.w;
  cs_f3 rest = f3(A1.x, A1.y, A1.z);
  const cs_f3 pos = f3(P.x, P.y, P.z);
  const float mass_p = Vv.w;
  // bed_impulse_delta_exchange (shader.rs:517-522)
  cs_f3 brv = f3((float)dw[1u * nbed + bid] * u.inv_fp_scale, (float)dw[2u * nbed + bid] * u.inv_fp_scale,
                 (float)dw[3u * nbed + bid] * u.inv_fp_scale);
  dw[1u * nbed + bid] = 0;
  dw[2u * nbed + bid] = 0;
  dw[3u * nbed + bid] = 0;
  brv = brv / cs_max(mass_p, u.nominal_mass * 0.25f);
  const float max_rs = cs_min(u.vel_cap * 0.012f, dx / cs_max(dt, 1e-6f) * 0.02f);
  const float rs = len3(brv);
  if (rs > max_rs && rs > 1e-6f) brv = brv * (max_rs / rs);
  // the water around the grain
  const cs_f3 gp = (pos - f3(u.origin[0], u.origin[1], u.origin[2])) * u.inv_dx;
  const int bx = cs_f2i(floorf(gp.x - 0.5f)), by = cs_f2i(floorf(gp.y - 0.5f)), bz = cs_f2i(floorf(gp.z - 0.5f));
  const cs_f3 fx = f3(gp.x - (float)bx, gp.y - (float)by, gp.z - (float)bz);
  float wx[3], wy[3], wz[3];
  wx[0] = 0.5f * (1.5f - fx.x) * (1.5f - fx.x);
  wx[1] = 0.75f - (fx.x - 1.0f) * (fx.x - 1.0f);
  wx[2] = 0.5f * (fx.x - 0.5f) * (fx.x - 0.5f);
  wy[0] = 0.5f * (1.5f - fx.y) * (1.5f - fx.y);
  wy[1] = 0.75f - (fx.y - 1.0f) * (fx.y - 1.0f);
  wy[2] = 0.5f * (fx.y - 0.5f) * (fx.y - 0.5f);
  wz[0] = 0.5f * (1.5f - fx.z) * (1.5f - fx.z);
  wz[1] = 0.75f - (fx.z - 1.0f) * (fx.z - 1.0f);
  wz[2] = 0.5f * (fx.z - 0.5f) * (fx.z - 0.5f);
  cs_f3 water_v = f3(0.0f, 0.0f, 0.0f);
  float water_mass = 0.0f, asw = 0.0f, lasw = 0.0f;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int cx = bx + i, cy = by + j, cz = bz + k;
        if (!on_grid(u, cx, cy, cz)) continue;
        const float w = wx[i] * wy[j] * wz[k];
        const uint32_t ci = cell_index(u, cx, cy, cz);
        const float4 g = gvel[ci];
        water_v = water_v + f3(g.x, g.y, g.z) * w;
        water_mass += w * g.w;
        const int32_t sb = lookup[ci];
        if (sb >= 0 && (uint32_t)sb != bid) {
          asw += w;
          const float ccy = u.origin[1] + ((float)cy + 0.5f) * dx;
          if (ccy <= pos.y + dx * 0.25f) lasw += w;
        }
      }
  if (phase >= 1.5f) {  // suspended grain (shader.rs:2065-2105)
    const float ws = cs_clamp(water_mass / cs_max(u.nominal_mass * 2.0f, 1e-6f), 0.0f, 1.0f);
    const float drag = cs_clamp((8.0f + 18.0f * ws) * dt, 0.0f, 1.0f);
    const cs_f3 unsup = (f3(Vv.x, Vv.y, Vv.z) + brv) + f3(0.0f, u.gravity * dt * 0.85f, 0.0f);
    const float settle = cs_mix(1.10f, 0.32f, ws);
    const cs_f3 carried = water_v + f3(0.0f, -settle, 0.0f);
    cs_f3 sv = mix3(unsup, carried, drag);
    const float sp = len3(sv);
    if (sp > u.vel_cap) sv = sv * (u.vel_cap / sp);
    cs_f3 spos = pos + sv * dt;
    const contact c = sdf_contact(u, sdf, spos, sv, true);
    spos = c.pos;
    sv = c.vel;
    bed_part[2 * bid] = make_float4(spos.x, spos.y, spos.z, P.w);
    bed_part[2 * bid + 1] = make_float4(sv.x, sv.y, sv.z, mass_p);
    const bool anchored = lasw > 0.06f || asw > 0.20f;
    const bool dense = asw > 0.42f;
    const bool redeposit = (spos.y <= floor_y + dx * 0.35f && ws < 0.35f) ||
                           (anchored && sv.y <= 0.20f && (ws < 0.55f || dense));
    if (redeposit) phase = 1.0f;
    bed_aff[3 * bid] = make_float4(A0.x, A0.y, A0.z, phase);
    bed_aff[3 * bid + 1] = make_float4(spos.x, spos.y, spos.z, 0.0f);
    be[3] *= 0.985f;
    return;
  }
  // anchored grain (shader.rs:2107-2237)
  const float sat = be[7];
  const float wl = cs_clamp(water_mass / cs_max(u.nominal_mass * 1.5f, 1e-6f), 0.0f, 1.0f);
  const float wet = 1.0f + sat * 0.40f;
  const float mob = cs_clamp((0.30f + wl * 0.62f) * wet, 0.0f, 1.0f);
  const float sf = cs_clamp((rest.y + 3.0f) / 3.5f, 0.12f, 1.0f);
  const float damping = cs_clamp(1.0f - bed_damping * dt, 0.0f, 1.0f);
  cs_f3 vel = f3(Vv.x, Vv.y, Vv.z) + brv;
  const float sr = cs_clamp(1.0f - asw / 0.55f, 0.0f, 1.0f);
  const float ed = cs_clamp((2.0f + 5.0f * sat) * wl * sr * dt, 0.0f, 0.08f);
  vel = mix3(vel, water_v + f3(0.0f, -cs_mix(0.95f, 0.35f, wl), 0.0f), ed);
  const float ls = sqrtf(water_v.x * water_v.x + water_v.z * water_v.z);
  const float ldx = ls > 1e-6f ? water_v.x / ls : 0.0f, ldz = ls > 1e-6f ? water_v.z / ls : 0.0f;
  const float lss = cs_min(ls, fabsf(cs_min(water_v.y, 0.0f)) * 0.18f);
  const cs_f3 impact = f3(ldx * lss * 0.08f, cs_min(water_v.y, 0.0f) * 0.62f, ldz * lss * 0.08f);
  vel = vel + (((impact * bed_impact) * mob) * sf) * dt;
  const cs_f3 pg = pressure_gradient_at_cell(u, pres, alias4, floor_cell(u, pos.x, 0), floor_cell(u, pos.y, 1),
                                             floor_cell(u, pos.z, 2));
  const float sfx = -pg.x, sfz = -pg.z;
  const float sps = sqrtf(sfx * sfx + sfz * sfz);
  const float sdx = sps > 1e-6f ? sfx / sps : 0.0f, sdz = sps > 1e-6f ? sfz / sps : 0.0f;
  const float drive = cs_min(sps, u.vel_cap / cs_max(dt, 1e-6f));
  const float pl = cs_clamp(water_mass * (0.25f + sat), 0.0f, 1.0f);
  const float slx = sdx * drive * pl * 0.08f, slz = sdz * drive * pl * 0.08f;
  vel.x += slx * bed_impact * mob * sf * dt;
  vel.z += slz * bed_impact * mob * sf * dt;
  vel = vel * damping;
  cs_f3 np = pos + vel * dt;
  cs_f3 off = np - rest;
  const float ll = sqrtf(off.x * off.x + off.z * off.z);
  const float lpt = dx * 0.22f;
  if (ll > lpt && ll > 1e-6f) {
    const float ox = off.x / ll, oz = off.z / ll;
    const float excess = ll - lpt;
    const float lp = cs_clamp((0.04f + sat * 0.35f + mob * 0.45f + pl * 0.25f) * sf * bed_compaction * dt, 0.0f, 0.18f);
    rest.x += ox * excess * lp;
    rest.z += oz * excess * lp;
    off = np - rest;
  }
  const float cll = sqrtf(off.x * off.x + off.z * off.z);
  const float ml = dx * 0.9f * sf;
  if (cll > ml && cll > 1e-6f) {
    const float ox = off.x / cll, oz = off.z / cll;
    off.x = ox * ml;
    off.z = oz * ml;
    vel.x *= 0.4f;
    vel.z *= 0.4f;
  }
  off.y = cs_clamp(off.y, -dx * (1.75f * sf + 0.2f), dx * 0.04f);
  np = rest + off;
  const contact pc = coffee_packing(u, lookup, nbed, snap, np, vel, bid);
  np = pc.pos;
  vel = pc.vel;
  const float comp = cs_max(rest.y - np.y, 0.0f);
  const float pth = dx * 0.18f;
  if (comp > pth) {
    const float ex = comp - pth;
    const float plast = cs_clamp((0.05f + sat * 0.35f + mob * 0.55f) * sf * bed_compaction * dt, 0.0f, 0.18f);
    rest.y -= ex * plast;
  }
  const bool detach = sat > 0.82f && pl > 0.65f && wl > 0.65f && sr > 0.35f && (comp > dx * 0.65f || cll > dx * 0.65f);
  const bool entrained = sat > 0.58f && wl > 0.74f && sr > 0.55f && len3(water_v) > 0.95f && np.y > floor_y + dx * 0.70f;
  const bool isolated = sat > 0.55f && water_mass < u.nominal_mass * 0.75f && asw < 0.04f && lasw < 0.03f &&
                        np.y > floor_y + dx * 0.50f;
  if (detach || entrained || isolated) {
    phase = 2.0f;
    rest = np;
  }
  const contact c = sdf_contact(u, sdf, np, vel, true);
  np = c.pos;
  vel = c.vel;
  bed_part[2 * bid] = make_float4(np.x, np.y, np.z, P.w);
  bed_part[2 * bid + 1] = make_float4(vel.x, vel.y, vel.z, mass_p);
  bed_aff[3 * bid] = make_float4(A0.x, A0.y, A0.z, phase);
  bed_aff[3 * bid + 1] = make_float4(rest.x, rest.y, rest.z, 0.0f);
  const float gc = cs_max((rest.y - np.y) / cs_max(dx, 1e-6f), 0.0f);
  be[3] = cs_max(be[3] * 0.995f, gc);
}

hipError_t cs_launch_bed_dynamics(hipStream_t st, const cs_pblock& P, uint32_t nbed, float* bed_part, float* bed_aff,
                                  float* bed_ext, int32_t* dw, float* snap, const int32_t* lookup, const float* gvel,
                                  const int32_t* pres, const uint8_t* alias4, const float* sdf) {
  const cs_params& u = P.u;
  if (nbed == 0) return hipSuccess;
  const dim3 g((nbed + 255u) / 256u);
  hipLaunchKernelGGL(k_bed_snapshot, g, dim3(256), 0, st, nbed, reinterpret_cast<const float4*>(bed_part),
                     reinterpret_cast<float4*>(snap));
  hipLaunchKernelGGL(k_bed_dynamics, g, dim3(256), 0, st, P.d, nbed, reinterpret_cast<float4*>(bed_part),
                     reinterpret_cast<float4*>(bed_aff), bed_ext, dw, reinterpret_cast<const float4*>(snap), lookup,
                     reinterpret_cast<const float4*>(gvel), pres, alias4, sdf);
  return hipGetLastError();
}

hipError_t cs_launch_bed_lookup(hipStream_t st, const cs_pblock& P, uint32_t nbed, const float* bed_part,
                                const float* bed_aff, const float* bed_ext, uint32_t capacity, int32_t* lookup,
                                uint32_t* counter, int32_t* bed_id, cs_bed_cell* table) {
  const cs_params& u = P.u;
  const uint32_t nc = (uint32_t)u.nc;
  hipLaunchKernelGGL(k_bed_lookup_clear, dim3((nc + 255) / 256), dim3(256), 0, st, P.d, nc, lookup, counter);
  if (nbed)
    hipLaunchKernelGGL(k_bed_lookup_scatter, dim3((nbed + 255) / 256), dim3(256), 0, st, P.d, nbed,
                       reinterpret_cast<const float4*>(bed_part), reinterpret_cast<const float4*>(bed_aff), lookup);
  hipLaunchKernelGGL(k_bed_table, dim3((nc + 255) / 256), dim3(256), 0, st, P.d, lookup, bed_part, bed_ext, capacity,
                     counter, bed_id, table);
  return hipGetLastError();
}

// ---- exp check kernel (used by the verification entry point) -----------------------
__global__ void k_expf(const float* __restrict__ in, float* __restrict__ out, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = cs_expf(in[i]);
}
__global__ void k_sincosf(const float* __restrict__ in, float* __restrict__ sn, float* __restrict__ cs, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) cs_sincosf(in[i], &sn[i], &cs[i]);
}

// ---- scene generation on the device (state.rs:206-344) ----------------------------------------
// One thread per SDF voxel, then one per grid cell of the class mask; the
// rules are cs_sdf.h's, shared with the host generator.
__global__ void __launch_bounds__(256) k_gen_sdf(cs_settings s, float* __restrict__ out) {
  const uint32_t n = s.sdf_res;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)n * n * n) return;
  const uint32_t ix = (uint32_t)(idx % n), iy = (uint32_t)((idx / n) % n), iz = (uint32_t)(idx / ((size_t)n * n));
  out[idx] = cs_sdf_voxel(s, ix, iy, iz);
}
__global__ void __launch_bounds__(256) k_gen_sdf_class(cs_settings s, const float* __restrict__ sdf,
                                                       uint8_t* __restrict__ out) {
  const uint32_t gx = s.grid_dims[0], gy = s.grid_dims[1], gz = s.grid_dims[2];
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)gx * gy * gz) return;
  const uint32_t ix = (uint32_t)(idx % gx), iy = (uint32_t)((idx / gx) % gy), iz = (uint32_t)(idx / ((size_t)gx * gy));
  out[idx] = cs_sdf_class_cell(s, sdf, ix, iy, iz);
}

// ---- inflow emitter on the device (inflow.rs:195-214; the frame and count come from the host) ----
// One thread per emitted particle: its 32 B particle record from the golden-angle
// disk and a zero affine record (48 B), at slot f.first + i.
__global__ void __launch_bounds__(256) k_emit(const cs_dparams* __restrict__ dp, float* __restrict__ part,
                                              float* __restrict__ aff) {
  if (dp->halt) return;
  const cs_emit_frame f = dp->ef;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.count) return;
  float p8[8];
  cs_emit_particle(f, i, p8);
  float4* pd = reinterpret_cast<float4*>(part + (size_t)(f.first + i) * 8);
  pd[0] = make_float4(p8[0], p8[1], p8[2], p8[3]);
  pd[1] = make_float4(p8[4], p8[5], p8[6], p8[7]);
  float4* ad = reinterpret_cast<float4*>(aff + (size_t)(f.first + i) * 12);
  const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  ad[0] = z;
  ad[1] = z;
  ad[2] = z;
}

// ---- inputs and targets of the learned warm start (no reference counterpart) ---------
// Four features per cell as the network reads them: rhs * rhs_scale and liquid
// fill for fluid cells (0 elsewhere), the solid mask, and the bed porosity; with
// `target`, also the converged pressure / pressure_scale and the mask of the
// loss (fluid cells whose pressure is not held at the clamp), counting the
// masked cells into `count`.
__global__ void __launch_bounds__(256) k_unet_features(cs_params u, const float* __restrict__ kf,
                                                       const float* __restrict__ rhs, const int32_t* __restrict__ pres,
                                                       const int32_t* __restrict__ bed_id,
                                                       const cs_bed_cell* __restrict__ bed, float rhs_scale,
                                                       float inv_p_scale, float4* __restrict__ f4,
                                                       float* __restrict__ target, uint8_t* __restrict__ mask,
                                                       uint32_t* __restrict__ count) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  bool fluid = false, use = false;
  if (idx < u.nc) {
    const float skf = kf[idx];
    fluid = kf_fluid(skf);
    const int32_t b = bed_id[idx];
    const float poro = (b >= 0 && (uint32_t)b < u.num_bed_cells) ? bed[b].porosity : 0.0f;
    f4[idx] = make_float4(fluid ? rhs[idx] * rhs_scale : 0.0f, fluid ? kf_fill(skf) : 0.0f,
                          skf == CS_KF_SOLID ? 1.0f : 0.0f, poro);
    if (target) {
      // a cell held at the pressure bound is not a row of the solved system
      // (its value is the bound, not a solution), so it is no training target
      const float p = (float)pres[idx] * u.inv_fp_scale;
      use = fluid && fabsf(p) != u.pressure_clamp;
      target[idx] = fluid ? p * inv_p_scale : 0.0f;
      mask[idx] = use ? 1u : 0u;
    }
  }
  if (target) {
    const uint64_t bal = __ballot(use);
    if ((threadIdx.x & 63u) == 0u && bal) atomicAdd(count, (uint32_t)__popcll(bal));
  }
}

// ---- batched passes: one launch over every scene of a batch ------------------------------
// Workgroup g of a batched launch runs block g - start[s] of scene s's own
// launch (cs_batch_scene_of), with the scene's parameter block copied from the
// descriptor table.
#define CS_BATCH_SCENE(space)                                                                         \
  const uint32_t* st_ = starts + (space) * (n + 1u);                                                  \
  const uint32_t scene_ = cs_batch_scene_of(st_, n, blockIdx.x);                                      \
  const cs_scene_dev& d = sc[scene_];                                                                 \
  const uint32_t blk = blockIdx.x - st_[scene_];                                                      \
  const cs_params u = d.u;                                                                            \
  const uint32_t tid = blk * 256u + threadIdx.x

// Emitted particles from each scene's disk frame into the particle arrays
// (affine zeroed), grid lanes, metric slots, P2G tile counts and brick flags
// to zero: the emission and buffer clears at the start of a substep.
__global__ void __launch_bounds__(256) k_batch_begin(const cs_scene_dev* __restrict__ sc,
                                                     const uint32_t* __restrict__ starts, uint32_t n) {
  CS_BATCH_SCENE(CS_SPACE_CELLS);
  const uint32_t span = (st_[scene_ + 1u] - st_[scene_]) * 256u;
  for (uint32_t i = tid; i < d.emit.count; i += span) {
    float p8[8];
    cs_emit_particle(d.emit, i, p8);
    float4* dp = reinterpret_cast<float4*>(d.part) + 2u * ((size_t)d.emit.first + i);
    float4* da = reinterpret_cast<float4*>(d.aff) + 3u * ((size_t)d.emit.first + i);
    dp[0] = make_float4(p8[0], p8[1], p8[2], p8[3]);
    dp[1] = make_float4(p8[4], p8[5], p8[6], p8[7]);
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    da[0] = z;
    da[1] = z;
    da[2] = z;
  }
  if (tid < u.nc) {
#pragma unroll
    for (int l = 0; l < 6; ++l) d.lanes[(size_t)l * u.nc + tid] = 0;
  }
  if (tid < u.nbricks) {
    d.tile_count[tid] = 0u;
    d.brick_flag[tid] = 0u;
  }
  if (tid < CS_METRICS_SLOTS) d.metrics[tid] = 0u;
}

__global__ void __launch_bounds__(256) k_batch_grid(const cs_scene_dev* __restrict__ sc,
                                                    const uint32_t* __restrict__ starts, uint32_t n) {
  CS_BATCH_SCENE(CS_SPACE_CELLS);
  grid_cell(u, d.lanes, reinterpret_cast<float4*>(d.gv), d.sdf, tid);
}

__global__ void __launch_bounds__(256) k_batch_classify(const cs_scene_dev* __restrict__ sc,
                                                        const uint32_t* __restrict__ starts, uint32_t n) {
  CS_BATCH_SCENE(CS_SPACE_CELLS);
  classify_wave<1>(u, reinterpret_cast<const float4*>(d.gv), d.cls, d.lanes, d.bed_id, d.bed, d.kind, d.kf, d.pres,
                   d.rhs, d.pack, d.tgt, d.metrics, d.sdf, d.brick_flag, tid, 0u);
}

__global__ void __launch_bounds__(256) k_batch_project(const cs_scene_dev* __restrict__ sc,
                                                       const uint32_t* __restrict__ starts, uint32_t n) {
  CS_BATCH_SCENE(CS_SPACE_CELLS);
  project_cell(u, reinterpret_cast<float4*>(d.gv), d.kf, d.pres, d.bed_id, d.bed, d.sdf, cs_grains{}, tid);
}

// the post-projection metric reads the neighbours' projected velocities, so
// packing (which changes them) waits for it
__global__ void __launch_bounds__(256) k_batch_post_div(const cs_scene_dev* __restrict__ sc,
                                                        const uint32_t* __restrict__ starts, uint32_t n) {
  CS_BATCH_SCENE(CS_SPACE_CELLS);
  uint32_t fd = 0u, fr = 0u;
  post_div_cell(u, reinterpret_cast<const float4*>(d.gv), d.kf, d.cls, d.sdf, d.tgt, tid, fd, fr);
  post_div_flush(d.metrics, fd, fr);
}

__global__ void __launch_bounds__(256) k_batch_pack(const cs_scene_dev* __restrict__ sc,
                                                    const uint32_t* __restrict__ starts, uint32_t n) {
  CS_BATCH_SCENE(CS_SPACE_CELLS);
  pack_cell(u, reinterpret_cast<float4*>(d.gv), d.kf, d.pack, d.sdf, tid);
}

__global__ void __launch_bounds__(256) k_batch_visc(const cs_scene_dev* __restrict__ sc,
                                                    const uint32_t* __restrict__ starts, uint32_t n) {
  CS_BATCH_SCENE(CS_SPACE_CELLS);
  visc_cell(u, reinterpret_cast<const float4*>(d.gv), reinterpret_cast<float4*>(d.gv2), d.cls, d.sdf, nullptr, tid);
}

__global__ void __launch_bounds__(256) k_batch_g2p(const cs_scene_dev* __restrict__ sc,
                                                   const uint32_t* __restrict__ starts, uint32_t n) {
  CS_BATCH_SCENE(CS_SPACE_PARTICLES);
  g2p_particle(u, reinterpret_cast<float4*>(d.part), reinterpret_cast<float4*>(d.aff),
               G2PGlobal{reinterpret_cast<const float4*>(d.gv2), d.bed_id}, d.bed, d.sdf, cs_grains{}, tid);
}

// ---- launchers -------------------------------------------------------------------------
namespace {
constexpr uint32_t kBlock = 256;
inline dim3 blocks_for(size_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }
}  // namespace

hipError_t cs_launch_fill_i32(hipStream_t s, int32_t* p, size_t n, int32_t v) {
  if (n == 0) return hipSuccess;
  size_t nb = (n + kBlock - 1) / kBlock;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)nb), dim3(kBlock), 0, s, p, n, v);
  return hipGetLastError();
}
hipError_t cs_launch_copy_i32(hipStream_t s, const int32_t* src, int32_t* dst, size_t n) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_i32, blocks_for(n), dim3(kBlock), 0, s, src, dst, n);
  return hipGetLastError();
}
hipError_t cs_launch_fill_i32_gated(hipStream_t s, const cs_dparams* d, int32_t* p, size_t n, int32_t v) {
  if (n == 0) return hipSuccess;
  size_t nb = (n + kBlock - 1) / kBlock;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(k_fill_i32_gated, dim3((unsigned)nb), dim3(kBlock), 0, s, d, p, n, v);
  return hipGetLastError();
}
hipError_t cs_launch_copy_i32_gated(hipStream_t s, const cs_dparams* d, const int32_t* src, int32_t* dst, size_t n) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_i32_gated, blocks_for(n), dim3(kBlock), 0, s, d, src, dst, n);
  return hipGetLastError();
}
hipError_t cs_launch_p2g(hipStream_t s, const cs_pblock& P, const float* part, const float* aff, int32_t* lanes,
                         uint32_t* metrics) {
  const cs_params& u = P.u;
  if (u.num_particles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_p2g_atomic, blocks_for(u.num_particles), dim3(kBlock), 0, s, P.d,
                     reinterpret_cast<const float4*>(part), reinterpret_cast<const float4*>(aff), lanes, metrics);
  return hipGetLastError();
}
hipError_t cs_launch_grid(hipStream_t s, const cs_pblock& P, const int32_t* lanes, float* gvel, const float* sdf) {
  const cs_params& u = P.u;
  hipLaunchKernelGGL(k_grid, blocks_for(u.nc), dim3(kBlock), 0, s, P.d, lanes, reinterpret_cast<float4*>(gvel), sdf);
  return hipGetLastError();
}
hipError_t cs_launch_classify(hipStream_t s, const cs_pblock& P, const float* gvel, const uint8_t* cls,
                              const int32_t* lanes, const int32_t* bed_id, const cs_bed_cell* bed, int8_t* kind,
                              float* kf, int32_t* pres, float* rhs, int32_t* pack, float* tgt, uint32_t* metrics,
                              const float* sdf, uint32_t* brick_flag) {
  const cs_params& u = P.u;
  hipLaunchKernelGGL(k_fill_i32_gated, dim3((u.nbricks + kBlock - 1) / kBlock), dim3(kBlock), 0, s, P.d,
                     reinterpret_cast<int32_t*>(brick_flag), (size_t)u.nbricks, 0);
  hipLaunchKernelGGL(k_classify, dim3((unsigned)((u.nc + kBlock * kClassifyCells - 1) / (kBlock * kClassifyCells))),
                     dim3(kBlock), 0, s, P.d, reinterpret_cast<const float4*>(gvel), cls,
                     lanes, bed_id, bed, kind, kf, pres, rhs, pack, tgt, metrics, sdf, brick_flag);
  return hipGetLastError();
}
hipError_t cs_launch_warm_previous(hipStream_t s, const cs_pblock& P, const float* kf, const int32_t* prev,
                                   int32_t* pres) {
  const cs_params& u = P.u;
  hipLaunchKernelGGL(k_warm_previous, blocks_for(u.nc), dim3(kBlock), 0, s, P.d, kf, prev, pres);
  return hipGetLastError();
}
hipError_t cs_launch_warm_external(hipStream_t s, const cs_params& u, const float* kf, const float* warm,
                                   int32_t* pres, uint32_t* metrics) {
  hipLaunchKernelGGL(k_warm_external, blocks_for(u.nc), dim3(kBlock), 0, s, u, kf, warm, pres, metrics);
  return hipGetLastError();
}
hipError_t cs_launch_rbgs_pair(hipStream_t s, const cs_params& u, const float* kf, const float* rhs, int32_t* pres,
                               uint32_t* metrics) {
  hipLaunchKernelGGL(k_rbgs_half, blocks_for(u.nc), dim3(kBlock), 0, s, u, 0u, kf, rhs, pres, metrics);
  hipLaunchKernelGGL(k_rbgs_half, blocks_for(u.nc), dim3(kBlock), 0, s, u, 1u, kf, rhs, pres, metrics);
  return hipGetLastError();
}
hipError_t cs_launch_residual(hipStream_t s, const cs_params& u, const float* kf, const float* rhs,
                              const int32_t* pres, double* partials, uint32_t max_blocks, double* out) {
  size_t nb = (u.nc + kBlock - 1) / kBlock;
  if (nb > max_blocks) nb = max_blocks;
  hipLaunchKernelGGL(k_residual, dim3((unsigned)nb), dim3(kBlock), 0, s, u, kf, rhs, pres, partials);
  hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(kBlock), 0, s, partials, (uint32_t)nb, out);
  return hipGetLastError();
}
hipError_t cs_launch_project(hipStream_t s, const cs_pblock& P, float* gvel, const float* kf, const int32_t* pres,
                             const int32_t* bed_id, const cs_bed_cell* bed, const float* sdf, const cs_grains& gr) {
  const cs_params& u = P.u;
  hipLaunchKernelGGL(k_project, blocks_for(u.nc), dim3(kBlock), 0, s, P.d, reinterpret_cast<float4*>(gvel), kf, pres,
                     bed_id, bed, sdf, gr);
  return hipGetLastError();
}
hipError_t cs_launch_post_div(hipStream_t s, const cs_params& u, const float* gvel, const float* kf, const uint8_t* cls,
                              const float* sdf, const float* tgt, uint32_t* metrics) {
  hipLaunchKernelGGL(k_post_div, blocks_for(u.nc), dim3(kBlock), 0, s, u, reinterpret_cast<const float4*>(gvel), kf,
                     cls, sdf, tgt, metrics);
  return hipGetLastError();
}
hipError_t cs_launch_post_div_bricks(hipStream_t s, const cs_pblock& P, uint32_t nactive, const uint32_t* list,
                                     const float* gvel, const float* kf, const uint8_t* cls, const float* sdf,
                                     const float* tgt, uint32_t* metrics) {
  const cs_params& u = P.u;
  if (!nactive) return hipSuccess;
  hipLaunchKernelGGL(k_post_div_bricks, dim3(nactive), dim3(256), 0, s, P.d,
                     list, reinterpret_cast<const float4*>(gvel),
                     kf, cls, sdf, tgt, metrics);
  return hipGetLastError();
}
hipError_t cs_launch_pack(hipStream_t s, const cs_pblock& P, float* gvel, const float* kf, const int32_t* pack,
                          const float* sdf) {
  const cs_params& u = P.u;
  hipLaunchKernelGGL(k_pack, blocks_for(u.nc), dim3(kBlock), 0, s, P.d, reinterpret_cast<float4*>(gvel), kf, pack, sdf);
  return hipGetLastError();
}
hipError_t cs_launch_visc(hipStream_t s, const cs_pblock& P, const float* gin, float* gout, const uint8_t* cls,
                          const float* sdf, uint8_t* alias4) {
  const cs_params& u = P.u;
  hipLaunchKernelGGL(k_visc, blocks_for(u.nc), dim3(kBlock), 0, s, P.d, reinterpret_cast<const float4*>(gin),
                     reinterpret_cast<float4*>(gout), cls, sdf, alias4);
  return hipGetLastError();
}
hipError_t cs_launch_g2p(hipStream_t s, const cs_pblock& P, float* part, float* aff, const float* gvel,
                         const int32_t* bed_id, const cs_bed_cell* bed, const float* sdf, const cs_grains& gr) {
  const cs_params& u = P.u;
  if (u.num_particles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_g2p, blocks_for(u.num_particles), dim3(kBlock), 0, s, P.d, reinterpret_cast<float4*>(part),
                     reinterpret_cast<float4*>(aff), reinterpret_cast<const float4*>(gvel), bed_id, bed, sdf, gr);
  return hipGetLastError();
}
hipError_t cs_launch_g2p_tiled(hipStream_t s, const cs_pblock& P, float* part, float* aff, const float* gvel,
                               const int32_t* bed_id, const cs_bed_cell* bed, const float* sdf, const cs_grains& grains,
                               const uint32_t* tile_count, const uint32_t* tile_start, const uint32_t* tile_chunk,
                               const uint2* key, const uint32_t* sorted) {
  const cs_params& u = P.u;
  if (u.num_particles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_g2p_unbinned, blocks_for(u.num_particles), dim3(kBlock), 0, s, P.d,
                     reinterpret_cast<float4*>(part), reinterpret_cast<float4*>(aff),
                     reinterpret_cast<const float4*>(gvel), bed_id, bed, sdf, grains, key);
  // at most one partial chunk per tile beyond the full ones
  const uint32_t items = u.nbricks + (u.num_particles + kChunk - 1u) / kChunk;
  hipLaunchKernelGGL(k_g2p_tile_chunks, dim3(items), dim3(256), 0, s, P.d, reinterpret_cast<float4*>(part),
                     reinterpret_cast<float4*>(aff), reinterpret_cast<const float4*>(gvel), bed_id, bed, sdf, grains,
                     tile_count, tile_start, tile_chunk, sorted);
  return hipGetLastError();
}
hipError_t cs_launch_pack_render(hipStream_t s, const cs_params& u, const float* part, const float* aff,
                                 uint32_t num_bed, const float* bed_ext, float* out) {
  if (u.num_particles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pack_render, blocks_for(u.num_particles), dim3(kBlock), 0, s, u,
                     reinterpret_cast<const float4*>(part), reinterpret_cast<const float4*>(aff), num_bed, bed_ext,
                     reinterpret_cast<float4*>(out));
  return hipGetLastError();
}
hipError_t cs_launch_expf(hipStream_t s, const float* in, float* out, uint32_t n) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_expf, blocks_for(n), dim3(kBlock), 0, s, in, out, n);
  return hipGetLastError();
}
hipError_t cs_launch_sincosf(hipStream_t s, const float* in, float* sn, float* cs, uint32_t n) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sincosf, blocks_for(n), dim3(kBlock), 0, s, in, sn, cs, n);
  return hipGetLastError();
}
hipError_t cs_launch_gen_sdf(hipStream_t st, const cs_settings& s, float* out) {
  const size_t n = (size_t)s.sdf_res * s.sdf_res * s.sdf_res;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gen_sdf, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, s, out);
  return hipGetLastError();
}
hipError_t cs_launch_gen_sdf_class(hipStream_t st, const cs_settings& s, const float* sdf, uint8_t* out) {
  const size_t n = (size_t)s.grid_dims[0] * s.grid_dims[1] * s.grid_dims[2];
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gen_sdf_class, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, s, sdf, out);
  return hipGetLastError();
}
hipError_t cs_launch_emit(hipStream_t s, const cs_dparams* d, uint32_t threads, float* part, float* aff) {
  if (threads == 0) return hipSuccess;
  hipLaunchKernelGGL(k_emit, dim3((threads + 255u) / 256u), dim3(256), 0, s, d, part, aff);
  return hipGetLastError();
}
hipError_t cs_launch_unet_features(hipStream_t s, const cs_params& u, const float* kf, const float* rhs,
                                   const int32_t* pres, const int32_t* bed_id, const cs_bed_cell* bed, float rhs_scale,
                                   float inv_p_scale, float* f4, float* target, uint8_t* mask, uint32_t* count) {
  hipLaunchKernelGGL(k_unet_features, blocks_for(u.nc), dim3(kBlock), 0, s, u, kf, rhs, pres, bed_id, bed, rhs_scale,
                     inv_p_scale, reinterpret_cast<float4*>(f4), target, mask, count);
  return hipGetLastError();
}

// ---- batched launchers --------------------------------------------------------------------
#define CS_BATCH_LAUNCH(kernel, space)                                                              \
  do {                                                                                            \
    const uint32_t nb_ = cs_batch_blocks(b, space);                                               \
    if (nb_) hipLaunchKernelGGL(kernel, dim3(nb_), dim3(kBlock), 0, b.s, b.sc, b.starts, b.n);     \
  } while (0)

hipError_t cs_batch_launch_begin(const cs_batch_launch& b) {
  CS_BATCH_LAUNCH(k_batch_begin, CS_SPACE_CELLS);
  return hipGetLastError();
}

hipError_t cs_batch_launch_grid_classify(const cs_batch_launch& b) {
  CS_BATCH_LAUNCH(k_batch_grid, CS_SPACE_CELLS);
  CS_BATCH_LAUNCH(k_batch_classify, CS_SPACE_CELLS);
  return hipGetLastError();
}

hipError_t cs_batch_launch_post(const cs_batch_launch& b) {
  CS_BATCH_LAUNCH(k_batch_project, CS_SPACE_CELLS);
  CS_BATCH_LAUNCH(k_batch_post_div, CS_SPACE_CELLS);
  CS_BATCH_LAUNCH(k_batch_pack, CS_SPACE_CELLS);
  CS_BATCH_LAUNCH(k_batch_visc, CS_SPACE_CELLS);
  CS_BATCH_LAUNCH(k_batch_g2p, CS_SPACE_PARTICLES);
  return hipGetLastError();
}
