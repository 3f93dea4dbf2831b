// cs_mg.hip — the multigrid start of the pressure solve (rbgs_mg_cycles; no
// reference counterpart).
//
// The red-black iteration of pressure_update (shader.rs:1371-1453) removes the
// error of a start at a rate that falls with its wavelength: on the pour's
// fluid, which touches air in only a few percent of its rows, the smooth part
// of the error takes thousands of iterations.  Before that iteration, a few
// W-cycles of a truncated aggregation multigrid remove it:
//
//  - free cells: fluid cells with a non-empty row whose pressure is inside the
//    clamp, or at it with pressure_update's unclamped value back inside.  Any
//    other cell at the clamp keeps its pressure; its correction is held at
//    zero, as a Dirichlet cell of the correction's equation.
//  - level 0: the error equation A e = b - A p over the free cells; level l+1
//    has one cell per 2x2x2 block of level l, with the Galerkin operator of
//    piecewise-constant aggregation: the children's diagonals less twice
//    their internal couplings, and the sum of the couplings across each face.
//    Levels are added until no side exceeds 8.
//  - a visit of a level: CS_MG_NU red-black passes, then CS_MG_GAMMA times
//    {restrict the residual, zero and visit the coarser level, add
//    CS_MG_SCALE times its correction, CS_MG_NU passes}; the coarsest level
//    takes CS_MG_COARSE passes.  A cycle ends with p <- clamp(p + e), stored in
//    fixed point, and the residual test: the cycles stop once the start is
//    within the tolerance (decided on the device).  The iteration that follows
//    makes its first test after a chunk, so the clamp's own updates always run.
//
// Every per-cell sum has one fixed order (faces -x, +x, -y, +y, -z, +z;
// children dz, dy, dx ascending), and a red-black pass reads only the other
// colour, so the result does not depend on how cells are dealt to threads:
// the oracle's serial loops give the same bits.  The fine level runs over the
// active bricks; the levels down to 16^3 are whole-grid launches; the smaller
// levels of a visit run inside one workgroup, from an operation list.
#include "cs_device.h"
#include "cs_mg.h"

namespace {

constexpr int BR = CS_BRICK;

__device__ __forceinline__ size_t lidx(const cs_mg_level& L, uint32_t x, uint32_t y, uint32_t z) {
  return ((size_t)z * L.Y + y) * L.X + x;
}

// The six neighbour terms of a cell against values v, summed in face order.
__device__ __forceinline__ float offsum(const cs_mg_level& L, const float* v, uint32_t x, uint32_t y, uint32_t z,
                                        size_t i) {
  const size_t sy = L.X, sz = (size_t)L.X * L.Y;
  const float t0 = x > 0u ? L.Wx[i - 1] * v[i - 1] : 0.0f;
  const float t1 = x + 1u < L.X ? L.Wx[i] * v[i + 1] : 0.0f;
  const float t2 = y > 0u ? L.Wy[i - sy] * v[i - sy] : 0.0f;
  const float t3 = y + 1u < L.Y ? L.Wy[i] * v[i + sy] : 0.0f;
  const float t4 = z > 0u ? L.Wz[i - sz] * v[i - sz] : 0.0f;
  const float t5 = z + 1u < L.Z ? L.Wz[i] * v[i + sz] : 0.0f;
  float s = t0;
  s += t1;
  s += t2;
  s += t3;
  s += t4;
  s += t5;
  return s;
}

__device__ __forceinline__ void smooth_cell(const cs_mg_level& L, uint32_t x, uint32_t y, uint32_t z,
                                            uint32_t colour) {
  if (((x + y + z) & 1u) != colour) return;
  const size_t i = lidx(L, x, y, z);
  const float d = L.D[i];
  if (!(d > 0.0f)) return;
  L.E[i] = (L.R[i] + offsum(L, L.E, x, y, z, i)) / d;
}

// Residual of level f summed over the children of coarse cell (X, Y, Z).
__device__ __forceinline__ void restrict_cell(const cs_mg_level& f, const cs_mg_level& c, uint32_t X, uint32_t Y,
                                              uint32_t Z) {
  float acc = 0.0f;
  for (uint32_t dz = 0; dz < 2u; ++dz)
    for (uint32_t dy = 0; dy < 2u; ++dy)
      for (uint32_t dx = 0; dx < 2u; ++dx) {
        const uint32_t x = 2u * X + dx, y = 2u * Y + dy, z = 2u * Z + dz;
        if (x >= f.X || y >= f.Y || z >= f.Z) continue;
        const size_t i = lidx(f, x, y, z);
        if (f.D[i] > 0.0f) {
          const float s = offsum(f, f.E, x, y, z, i);
          acc += f.R[i] - (f.D[i] * f.E[i] - s);
        }
      }
  c.R[lidx(c, X, Y, Z)] = acc;
}

// Galerkin operator of coarse cell (X, Y, Z) from level f.
__device__ __forceinline__ void build_cell(const cs_mg_level& f, const cs_mg_level& c, uint32_t X, uint32_t Y,
                                           uint32_t Z) {
  float d = 0.0f, wx = 0.0f, wy = 0.0f, wz = 0.0f;
  for (uint32_t dz = 0; dz < 2u; ++dz)
    for (uint32_t dy = 0; dy < 2u; ++dy)
      for (uint32_t dx = 0; dx < 2u; ++dx) {
        const uint32_t x = 2u * X + dx, y = 2u * Y + dy, z = 2u * Z + dz;
        if (x >= f.X || y >= f.Y || z >= f.Z) continue;
        d += f.D[lidx(f, x, y, z)];
      }
  for (uint32_t dz = 0; dz < 2u; ++dz)
    for (uint32_t dy = 0; dy < 2u; ++dy)
      for (uint32_t dx = 0; dx < 2u; ++dx) {
        const uint32_t x = 2u * X + dx, y = 2u * Y + dy, z = 2u * Z + dz;
        if (x >= f.X || y >= f.Y || z >= f.Z) continue;
        const size_t i = lidx(f, x, y, z);
        if (dx == 0u && x + 1u < f.X) d -= 2.0f * f.Wx[i];
        if (dy == 0u && y + 1u < f.Y) d -= 2.0f * f.Wy[i];
        if (dz == 0u && z + 1u < f.Z) d -= 2.0f * f.Wz[i];
        if (dx == 1u) wx += f.Wx[i];
        if (dy == 1u) wy += f.Wy[i];
        if (dz == 1u) wz += f.Wz[i];
      }
  const size_t I = lidx(c, X, Y, Z);
  c.D[I] = d > 0.0f ? d : 0.0f;
  c.Wx[I] = wx;
  c.Wy[I] = wy;
  c.Wz[I] = wz;
}

__device__ __forceinline__ void prolong_cell(const cs_mg_level& f, const cs_mg_level& c, uint32_t x, uint32_t y,
                                             uint32_t z) {
  const size_t i = lidx(f, x, y, z);
  if (f.D[i] > 0.0f) f.E[i] = f.E[i] + CS_MG_SCALE * c.E[lidx(c, x >> 1, y >> 1, z >> 1)];
}

__device__ __forceinline__ void level_cell(const cs_mg_level& L, uint32_t n, uint32_t& x, uint32_t& y, uint32_t& z) {
  x = n % L.X;
  const uint32_t r = n / L.X;
  y = r % L.Y;
  z = r / L.Y;
}

__device__ __forceinline__ void brick_origin(const cs_params& u, uint32_t b, uint32_t& x0, uint32_t& y0, uint32_t& z0) {
  const uint32_t bx = b % u.nbx, r = b / u.nbx;
  x0 = bx * BR;
  y0 = (r % u.nby) * BR;
  z0 = (r / u.nby) * BR;
}

// pressure_update's neighbour sum of fine cell (x, y, z) (shader.rs:1393-1420):
// -x, +x, -y, +y, -z, +z, with the minus-side weights read from the neighbours
__device__ __forceinline__ float fine_sum(const cs_params& u, const float4* __restrict__ wc,
                                          const int32_t* __restrict__ p, uint32_t x, uint32_t y, uint32_t z, size_t i,
                                          float4 w) {
  const size_t sy = u.gx, sz = (size_t)u.gx * u.gy;
  const float t0 = x > 0u ? wc[i - 1].x * ((float)p[i - 1] * u.inv_fp_scale) : 0.0f;
  const float t1 = x + 1u < u.gx ? w.x * ((float)p[i + 1] * u.inv_fp_scale) : 0.0f;
  const float t2 = y > 0u ? wc[i - sy].y * ((float)p[i - sy] * u.inv_fp_scale) : 0.0f;
  const float t3 = y + 1u < u.gy ? w.y * ((float)p[i + sy] * u.inv_fp_scale) : 0.0f;
  const float t4 = z > 0u ? wc[i - sz].z * ((float)p[i - sz] * u.inv_fp_scale) : 0.0f;
  const float t5 = z + 1u < u.gz ? w.z * ((float)p[i + sz] * u.inv_fp_scale) : 0.0f;
  float sum = t0;
  sum += t1;
  sum += t2;
  sum += t3;
  sum += t4;
  sum += t5;
  return sum;
}

// A free cell: fluid with a non-empty row, and inside the clamp or at it with
// pressure_update's unclamped value back inside (so the update would move it).
__device__ __forceinline__ bool is_free(const cs_params& u, const float4* __restrict__ wc,
                                        const float* __restrict__ rhs, const int32_t* __restrict__ p, uint32_t x,
                                        uint32_t y, uint32_t z, size_t i) {
  const float4 w = wc[i];
  if (!(w.w > 0.0f)) return false;
  const float pf = (float)p[i] * u.inv_fp_scale;
  if (fabsf(pf) < u.pressure_clamp) return true;
  const float v = (fine_sum(u, wc, p, x, y, z, i, w) - rhs[i]) / w.w;
  return pf > 0.0f ? v < u.pressure_clamp : v > -u.pressure_clamp;
}

// ---- fine level, one workgroup of 256 per active brick (two cells a thread) ----------
__global__ void __launch_bounds__(256) k_mg_fine_setup(cs_params u, const uint32_t* __restrict__ list,
                                                       const float4* __restrict__ wc, const float* __restrict__ rhs,
                                                       const int32_t* __restrict__ p, cs_mg_level L,
                                                       const uint32_t* __restrict__ skip) {
  if (*skip) return;
  uint32_t x0, y0, z0;
  brick_origin(u, list[blockIdx.x], x0, y0, z0);
  for (uint32_t k = threadIdx.x; k < (uint32_t)(BR * BR * BR); k += 256u) {
    const uint32_t x = x0 + (k & 7u), y = y0 + ((k >> 3) & 7u), z = z0 + (k >> 6);
    if (x >= u.gx || y >= u.gy || z >= u.gz) continue;
    const size_t i = lidx(L, x, y, z);
    const size_t sy = u.gx, sz = (size_t)u.gx * u.gy;
    const float4 w = wc[i];
    float d = 0.0f, wx = 0.0f, wy = 0.0f, wz = 0.0f, r = 0.0f;
    if (is_free(u, wc, rhs, p, x, y, z, i)) {
      d = w.w;
      if (x + 1u < u.gx && is_free(u, wc, rhs, p, x + 1u, y, z, i + 1)) wx = w.x;
      if (y + 1u < u.gy && is_free(u, wc, rhs, p, x, y + 1u, z, i + sy)) wy = w.y;
      if (z + 1u < u.gz && is_free(u, wc, rhs, p, x, y, z + 1u, i + sz)) wz = w.z;
      r = (fine_sum(u, wc, p, x, y, z, i, w) - rhs[i]) - w.w * ((float)p[i] * u.inv_fp_scale);
    }
    L.D[i] = d;
    L.Wx[i] = wx;
    L.Wy[i] = wy;
    L.Wz[i] = wz;
    L.R[i] = r;
    L.E[i] = 0.0f;
  }
}

__global__ void __launch_bounds__(256) k_mg_fine_smooth(cs_params u, const uint32_t* __restrict__ list, cs_mg_level L,
                                                        uint32_t colour, const uint32_t* __restrict__ skip) {
  if (*skip) return;
  uint32_t x0, y0, z0;
  brick_origin(u, list[blockIdx.x], x0, y0, z0);
  // thread t takes the two cells of colour `colour` in row pair t (a row of 8 holds 4 of each colour)
  const uint32_t t = threadIdx.x;  // 256 = 64 rows x 4
  const uint32_t row = t >> 2, q = t & 3u;
  const uint32_t y = y0 + (row & 7u), z = z0 + (row >> 3);
  const uint32_t x = x0 + 2u * q + ((colour + y + z + x0) & 1u);
  if (x < u.gx && y < u.gy && z < u.gz) smooth_cell(L, x, y, z, colour);
}

__global__ void __launch_bounds__(64) k_mg_fine_restrict(cs_params u, const uint32_t* __restrict__ list, cs_mg_level f,
                                                         cs_mg_level c, const uint32_t* __restrict__ skip) {
  if (*skip) return;
  uint32_t x0, y0, z0;
  brick_origin(u, list[blockIdx.x], x0, y0, z0);
  const uint32_t k = threadIdx.x;
  const uint32_t X = (x0 >> 1) + (k & 3u), Y = (y0 >> 1) + ((k >> 2) & 3u), Z = (z0 >> 1) + (k >> 4);
  if (X < c.X && Y < c.Y && Z < c.Z) restrict_cell(f, c, X, Y, Z);
}

__global__ void __launch_bounds__(64) k_mg_fine_build(cs_params u, const uint32_t* __restrict__ list, cs_mg_level f,
                                                      cs_mg_level c, const uint32_t* __restrict__ skip) {
  if (*skip) return;
  uint32_t x0, y0, z0;
  brick_origin(u, list[blockIdx.x], x0, y0, z0);
  const uint32_t k = threadIdx.x;
  const uint32_t X = (x0 >> 1) + (k & 3u), Y = (y0 >> 1) + ((k >> 2) & 3u), Z = (z0 >> 1) + (k >> 4);
  if (X < c.X && Y < c.Y && Z < c.Z) build_cell(f, c, X, Y, Z);
}

__global__ void __launch_bounds__(256) k_mg_fine_prolong(cs_params u, const uint32_t* __restrict__ list, cs_mg_level f,
                                                         cs_mg_level c, const uint32_t* __restrict__ skip) {
  if (*skip) return;
  uint32_t x0, y0, z0;
  brick_origin(u, list[blockIdx.x], x0, y0, z0);
  for (uint32_t k = threadIdx.x; k < (uint32_t)(BR * BR * BR); k += 256u) {
    const uint32_t x = x0 + (k & 7u), y = y0 + ((k >> 3) & 7u), z = z0 + (k >> 6);
    if (x < u.gx && y < u.gy && z < u.gz) prolong_cell(f, c, x, y, z);
  }
}

__global__ void __launch_bounds__(256) k_mg_fine_apply(cs_params u, const uint32_t* __restrict__ list, cs_mg_level L,
                                                       int32_t* __restrict__ p, const uint32_t* __restrict__ skip) {
  if (*skip) return;
  uint32_t x0, y0, z0;
  brick_origin(u, list[blockIdx.x], x0, y0, z0);
  for (uint32_t k = threadIdx.x; k < (uint32_t)(BR * BR * BR); k += 256u) {
    const uint32_t x = x0 + (k & 7u), y = y0 + ((k >> 3) & 7u), z = z0 + (k >> 6);
    if (x >= u.gx || y >= u.gy || z >= u.gz) continue;
    const size_t i = lidx(L, x, y, z);
    if (!(L.D[i] > 0.0f)) continue;
    const float pn = (float)p[i] * u.inv_fp_scale + L.E[i];
    const float cl = fminf(fmaxf(pn, -u.pressure_clamp), u.pressure_clamp);
    p[i] = cs_f2i(cl * u.fp_scale);
  }
}

// ---- dense levels --------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_mg_smooth(cs_mg_level L, uint32_t colour, const uint32_t* __restrict__ skip) {
  if (*skip) return;
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= L.X * L.Y * L.Z) return;
  uint32_t x, y, z;
  level_cell(L, n, x, y, z);
  smooth_cell(L, x, y, z, colour);
}
__global__ void __launch_bounds__(256) k_mg_restrict(cs_mg_level f, cs_mg_level c, const uint32_t* __restrict__ skip) {
  if (*skip) return;
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= c.X * c.Y * c.Z) return;
  uint32_t X, Y, Z;
  level_cell(c, n, X, Y, Z);
  restrict_cell(f, c, X, Y, Z);
}
__global__ void __launch_bounds__(256) k_mg_build(cs_mg_level f, cs_mg_level c, const uint32_t* __restrict__ skip) {
  if (*skip) return;
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= c.X * c.Y * c.Z) return;
  uint32_t X, Y, Z;
  level_cell(c, n, X, Y, Z);
  build_cell(f, c, X, Y, Z);
}
__global__ void __launch_bounds__(256) k_mg_prolong(cs_mg_level f, cs_mg_level c, const uint32_t* __restrict__ skip) {
  if (*skip) return;
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= f.X * f.Y * f.Z) return;
  uint32_t x, y, z;
  level_cell(f, n, x, y, z);
  prolong_cell(f, c, x, y, z);
}
__global__ void __launch_bounds__(256) k_mg_zero(cs_mg_level L, const uint32_t* __restrict__ skip) {
  if (*skip) return;
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < L.X * L.Y * L.Z) L.E[n] = 0.0f;
}

// ---- one workgroup runs a whole visit of the small levels --------------------------------
struct mg_levels {
  cs_mg_level l[CS_MG_MAX_LEVELS];
};
__global__ void __launch_bounds__(CS_MG_WG_THREADS) k_mg_wg_visit(mg_levels lv, const cs_mg_op* __restrict__ ops,
                                                                  uint32_t nops, const uint32_t* __restrict__ skip) {
  if (*skip) return;
  const uint32_t t = threadIdx.x;
  for (uint32_t o = 0; o < nops; ++o) {
    const cs_mg_op op = ops[o];
    const cs_mg_level& L = lv.l[op.level];
    if (op.kind == CS_MG_OP_SMOOTH0 || op.kind == CS_MG_OP_SMOOTH1) {
      const uint32_t n = L.X * L.Y * L.Z;
      for (uint32_t k = t; k < n; k += CS_MG_WG_THREADS) {
        uint32_t x, y, z;
        level_cell(L, k, x, y, z);
        smooth_cell(L, x, y, z, op.kind == CS_MG_OP_SMOOTH0 ? 0u : 1u);
      }
    } else if (op.kind == CS_MG_OP_RESTRICT) {
      const cs_mg_level& C = lv.l[op.level + 1];
      const uint32_t n = C.X * C.Y * C.Z;
      for (uint32_t k = t; k < n; k += CS_MG_WG_THREADS) {
        uint32_t X, Y, Z;
        level_cell(C, k, X, Y, Z);
        restrict_cell(L, C, X, Y, Z);
      }
    } else if (op.kind == CS_MG_OP_ZERO) {
      const uint32_t n = L.X * L.Y * L.Z;
      for (uint32_t k = t; k < n; k += CS_MG_WG_THREADS) L.E[k] = 0.0f;
    } else {  // CS_MG_OP_PROLONG
      const cs_mg_level& C = lv.l[op.level + 1];
      const uint32_t n = L.X * L.Y * L.Z;
      for (uint32_t k = t; k < n; k += CS_MG_WG_THREADS) {
        uint32_t x, y, z;
        level_cell(L, k, x, y, z);
        prolong_cell(L, C, x, y, z);
      }
    }
    // the next operation reads what this one wrote (global memory, one workgroup)
    __threadfence_block();
    __syncthreads();
  }
}

// flags[0]: skip the next cycle; flags[1]: the start was within the tolerance
// (no cycle ran).  The first gate sets both, later ones only flags[0].
__global__ void k_mg_gate(const cs_solve_state* __restrict__ st, double tol, uint32_t* __restrict__ flags,
                          uint32_t first) {
  const uint32_t within = st->initial_residual <= tol ? 1u : 0u;
  if (first) flags[1] = within;
  flags[0] = flags[1] | within;
}
// After the cycles: the iteration runs (its first test comes after a chunk)
// unless the start was already within the tolerance.
__global__ void k_mg_finish(cs_solve_state* __restrict__ st, const uint32_t* __restrict__ flags) {
  if (!flags[1]) st->done = 0u;
}

inline dim3 cells_grid(uint32_t n) { return dim3((n + 255u) / 256u); }
inline uint32_t cells(const cs_mg_level& L) { return L.X * L.Y * L.Z; }

}  // namespace

std::vector<cs_mg_level> cs_mg_shape(uint32_t gx, uint32_t gy, uint32_t gz) {
  std::vector<cs_mg_level> out;
  uint32_t X = gx, Y = gy, Z = gz;
  for (;;) {
    cs_mg_level l{};
    l.X = X;
    l.Y = Y;
    l.Z = Z;
    out.push_back(l);
    if ((X <= CS_MG_COARSEST_SIDE && Y <= CS_MG_COARSEST_SIDE && Z <= CS_MG_COARSEST_SIDE) ||
        out.size() == (size_t)CS_MG_MAX_LEVELS)
      break;
    X = (X + 1u) / 2u;
    Y = (Y + 1u) / 2u;
    Z = (Z + 1u) / 2u;
  }
  return out;
}

void cs_mg_visit_ops(int level, int nlevels, std::vector<cs_mg_op>* out) {
  auto pass = [&](int l) {
    out->push_back(cs_mg_op{CS_MG_OP_SMOOTH0, (uint8_t)l});
    out->push_back(cs_mg_op{CS_MG_OP_SMOOTH1, (uint8_t)l});
  };
  if (level == nlevels - 1) {
    for (uint32_t k = 0; k < CS_MG_COARSE; ++k) pass(level);
    return;
  }
  for (uint32_t k = 0; k < CS_MG_NU; ++k) pass(level);
  for (uint32_t g = 0; g < CS_MG_GAMMA; ++g) {
    out->push_back(cs_mg_op{CS_MG_OP_RESTRICT, (uint8_t)level});
    out->push_back(cs_mg_op{CS_MG_OP_ZERO, (uint8_t)(level + 1)});
    cs_mg_visit_ops(level + 1, nlevels, out);
    out->push_back(cs_mg_op{CS_MG_OP_PROLONG, (uint8_t)level});
    for (uint32_t k = 0; k < CS_MG_NU; ++k) pass(level);
  }
}

hipError_t cs_mg_fine_setup(hipStream_t s, const cs_params& u, uint32_t nactive, const uint32_t* list,
                            const float* wc, const float* rhs, const int32_t* p, cs_mg_level L0, const uint32_t* skip) {
  if (!nactive) return hipSuccess;
  hipLaunchKernelGGL(k_mg_fine_setup, dim3(nactive), dim3(256), 0, s, u, list, reinterpret_cast<const float4*>(wc), rhs,
                     p, L0, skip);
  return hipGetLastError();
}
hipError_t cs_mg_fine_smooth(hipStream_t s, const cs_params& u, uint32_t nactive, const uint32_t* list,
                             cs_mg_level L0, uint32_t colour, const uint32_t* skip) {
  if (!nactive) return hipSuccess;
  hipLaunchKernelGGL(k_mg_fine_smooth, dim3(nactive), dim3(256), 0, s, u, list, L0, colour, skip);
  return hipGetLastError();
}
hipError_t cs_mg_fine_restrict(hipStream_t s, const cs_params& u, uint32_t nactive, const uint32_t* list,
                               cs_mg_level L0, cs_mg_level L1, const uint32_t* skip) {
  if (!nactive) return hipSuccess;
  hipLaunchKernelGGL(k_mg_fine_restrict, dim3(nactive), dim3(64), 0, s, u, list, L0, L1, skip);
  return hipGetLastError();
}
hipError_t cs_mg_fine_build(hipStream_t s, const cs_params& u, uint32_t nactive, const uint32_t* list,
                            cs_mg_level L0, cs_mg_level L1, const uint32_t* skip) {
  if (!nactive) return hipSuccess;
  hipLaunchKernelGGL(k_mg_fine_build, dim3(nactive), dim3(64), 0, s, u, list, L0, L1, skip);
  return hipGetLastError();
}
hipError_t cs_mg_fine_prolong(hipStream_t s, const cs_params& u, uint32_t nactive, const uint32_t* list,
                              cs_mg_level L0, cs_mg_level L1, const uint32_t* skip) {
  if (!nactive) return hipSuccess;
  hipLaunchKernelGGL(k_mg_fine_prolong, dim3(nactive), dim3(256), 0, s, u, list, L0, L1, skip);
  return hipGetLastError();
}
hipError_t cs_mg_fine_apply(hipStream_t s, const cs_params& u, uint32_t nactive, const uint32_t* list,
                            cs_mg_level L0, int32_t* p, const uint32_t* skip) {
  if (!nactive) return hipSuccess;
  hipLaunchKernelGGL(k_mg_fine_apply, dim3(nactive), dim3(256), 0, s, u, list, L0, p, skip);
  return hipGetLastError();
}
hipError_t cs_mg_smooth(hipStream_t s, cs_mg_level L, uint32_t colour, const uint32_t* skip) {
  hipLaunchKernelGGL(k_mg_smooth, cells_grid(cells(L)), dim3(256), 0, s, L, colour, skip);
  return hipGetLastError();
}
hipError_t cs_mg_restrict(hipStream_t s, cs_mg_level f, cs_mg_level c, const uint32_t* skip) {
  hipLaunchKernelGGL(k_mg_restrict, cells_grid(cells(c)), dim3(256), 0, s, f, c, skip);
  return hipGetLastError();
}
hipError_t cs_mg_build(hipStream_t s, cs_mg_level f, cs_mg_level c, const uint32_t* skip) {
  hipLaunchKernelGGL(k_mg_build, cells_grid(cells(c)), dim3(256), 0, s, f, c, skip);
  return hipGetLastError();
}
hipError_t cs_mg_prolong(hipStream_t s, cs_mg_level f, cs_mg_level c, const uint32_t* skip) {
  hipLaunchKernelGGL(k_mg_prolong, cells_grid(cells(f)), dim3(256), 0, s, f, c, skip);
  return hipGetLastError();
}
hipError_t cs_mg_zero(hipStream_t s, cs_mg_level L, const uint32_t* skip) {
  hipLaunchKernelGGL(k_mg_zero, cells_grid(cells(L)), dim3(256), 0, s, L, skip);
  return hipGetLastError();
}
hipError_t cs_mg_wg_visit(hipStream_t s, const cs_mg_level* levels, const cs_mg_op* ops, uint32_t nops,
                          const uint32_t* skip) {
  mg_levels lv{};
  for (int l = 0; l < CS_MG_MAX_LEVELS; ++l) lv.l[l] = levels[l];
  hipLaunchKernelGGL(k_mg_wg_visit, dim3(1), dim3(CS_MG_WG_THREADS), 0, s, lv, ops, nops, skip);
  return hipGetLastError();
}
hipError_t cs_mg_gate(hipStream_t s, const cs_solve_state* st, double tol, uint32_t* flags, bool first) {
  hipLaunchKernelGGL(k_mg_gate, dim3(1), dim3(1), 0, s, st, tol, flags, first ? 1u : 0u);
  return hipGetLastError();
}
hipError_t cs_mg_finish(hipStream_t s, cs_solve_state* st, const uint32_t* flags) {
  hipLaunchKernelGGL(k_mg_finish, dim3(1), dim3(1), 0, s, st, flags);
  return hipGetLastError();
}
