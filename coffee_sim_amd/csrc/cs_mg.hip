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
//    takes cs_mg_coarse_passes(levels) passes.  A cycle ends with p <- clamp(p + e), stored in
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
#include <type_traits>

#include "cs_device.h"
#include "cs_kernels.h"
#include "cs_mg.h"

namespace {

constexpr int BR = CS_BRICK;

// A level held in LDS: the same fields as cs_mg_level with LDS pointers, so
// every access compiles to a ds load or store rather than a flat one.
typedef __attribute__((address_space(3))) float sfloat;
struct slevel {
  uint32_t X, Y, Z, mX, mY;
  uint32_t HX, mHX;  // cells of one colour per row at most, ceil(X / 2), and its division magic
  sfloat* D;
  sfloat* Wx;
  sfloat* Wy;
  sfloat* Wz;
  sfloat* R;
  sfloat* E;
};

// The per-cell steps below take either kind of level (cs_mg_level or slevel).
template <class LV>
__device__ __forceinline__ size_t lidx(const LV& L, uint32_t x, uint32_t y, uint32_t z) {
  return ((size_t)z * L.Y + y) * L.X + x;
}
// A level in HBM may be brick-major (the fine level: a brick's cells, which
// its workgroup visits together, are one contiguous 2 KB run per array).
__device__ __forceinline__ size_t lidx(const cs_mg_level& L, uint32_t x, uint32_t y, uint32_t z) {
  if (L.bnx)
    return ((((size_t)(z >> 3) * L.bny + (y >> 3)) * L.bnx + (x >> 3)) << 9) + ((z & 7u) << 6) + ((y & 7u) << 3) +
           (x & 7u);
  return ((size_t)z * L.Y + y) * L.X + x;
}

// The six neighbour terms of a cell against the corrections, summed in face order.
// Neighbour indices clamped onto the grid (the cell itself past an edge): every
// load is issued unconditionally and only the product is selected, so a cell's
// loads are in flight together instead of one branch-and-wait at a time.
struct nbr6 {
  size_t xm, xp, ym, yp, zm, zp;
  bool bxm, bxp, bym, byp, bzm, bzp;
};
__device__ __forceinline__ nbr6 neighbours(uint32_t X, uint32_t Y, uint32_t Z, uint32_t x, uint32_t y, uint32_t z,
                                           size_t i) {
  const size_t sy = X, sz = (size_t)X * Y;
  nbr6 n;
  n.bxm = x > 0u;
  n.bxp = x + 1u < X;
  n.bym = y > 0u;
  n.byp = y + 1u < Y;
  n.bzm = z > 0u;
  n.bzp = z + 1u < Z;
  n.xm = n.bxm ? i - 1 : i;
  n.xp = n.bxp ? i + 1 : i;
  n.ym = n.bym ? i - sy : i;
  n.yp = n.byp ? i + sy : i;
  n.zm = n.bzm ? i - sz : i;
  n.zp = n.bzp ? i + sz : i;
  return n;
}

// A level's neighbour indices (lidx of the clamped neighbours).
template <class LV>
__device__ __forceinline__ nbr6 lneighbours(const LV& L, uint32_t x, uint32_t y, uint32_t z, size_t i) {
  return neighbours(L.X, L.Y, L.Z, x, y, z, i);
}
__device__ __forceinline__ nbr6 lneighbours(const cs_mg_level& L, uint32_t x, uint32_t y, uint32_t z, size_t i) {
  if (!L.bnx) return neighbours(L.X, L.Y, L.Z, x, y, z, i);
  nbr6 n;
  n.bxm = x > 0u;
  n.bxp = x + 1u < L.X;
  n.bym = y > 0u;
  n.byp = y + 1u < L.Y;
  n.bzm = z > 0u;
  n.bzp = z + 1u < L.Z;
  n.xm = n.bxm ? lidx(L, x - 1u, y, z) : i;
  n.xp = n.bxp ? lidx(L, x + 1u, y, z) : i;
  n.ym = n.bym ? lidx(L, x, y - 1u, z) : i;
  n.yp = n.byp ? lidx(L, x, y + 1u, z) : i;
  n.zm = n.bzm ? lidx(L, x, y, z - 1u) : i;
  n.zp = n.bzp ? lidx(L, x, y, z + 1u) : i;
  return n;
}

template <class LV>
__device__ __forceinline__ float offsum(const LV& L, uint32_t x, uint32_t y, uint32_t z, size_t i) {
  const auto v = L.E;
  const nbr6 n = lneighbours(L, x, y, z, i);
  const float w0 = L.Wx[n.xm], w1 = L.Wx[i], w2 = L.Wy[n.ym], w3 = L.Wy[i], w4 = L.Wz[n.zm], w5 = L.Wz[i];
  const float e0 = v[n.xm], e1 = v[n.xp], e2 = v[n.ym], e3 = v[n.yp], e4 = v[n.zm], e5 = v[n.zp];
  const float t0 = n.bxm ? w0 * e0 : 0.0f;
  const float t1 = n.bxp ? w1 * e1 : 0.0f;
  const float t2 = n.bym ? w2 * e2 : 0.0f;
  const float t3 = n.byp ? w3 * e3 : 0.0f;
  const float t4 = n.bzm ? w4 * e4 : 0.0f;
  const float t5 = n.bzp ? w5 * e5 : 0.0f;
  float s = t0;
  s += t1;
  s += t2;
  s += t3;
  s += t4;
  s += t5;
  return s;
}

template <class LV>
__device__ __forceinline__ void smooth_cell(const LV& L, uint32_t x, uint32_t y, uint32_t z, uint32_t colour) {
  if (((x + y + z) & 1u) != colour) return;
  const size_t i = lidx(L, x, y, z);
  const float d = L.D[i];
  if (!(d > 0.0f)) return;
  L.E[i] = (L.R[i] + offsum(L, x, y, z, i)) / d;
}

// Residual of level f summed over the children of coarse cell (X, Y, Z).
template <class F, class C>
__device__ __forceinline__ void restrict_cell(const F& f, const C& c, uint32_t X, uint32_t Y, uint32_t Z) {
  float acc = 0.0f;
  for (uint32_t dz = 0; dz < 2u; ++dz)
    for (uint32_t dy = 0; dy < 2u; ++dy)
      for (uint32_t dx = 0; dx < 2u; ++dx) {
        const uint32_t x = 2u * X + dx, y = 2u * Y + dy, z = 2u * Z + dz;
        if (x >= f.X || y >= f.Y || z >= f.Z) continue;
        const size_t i = lidx(f, x, y, z);
        if (f.D[i] > 0.0f) {
          const float s = offsum(f, x, y, z, i);
          acc += f.R[i] - (f.D[i] * f.E[i] - s);
        }
      }
  const size_t I = lidx(c, X, Y, Z);
  c.R[I] = acc;
  c.E[I] = 0.0f;  // the coarse visit starts from a zero correction
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

// The correction of fine cell (x, y, z) after adding CS_MG_SCALE times the
// coarse correction (cells outside the operator keep theirs).
template <class F, class C>
__device__ __forceinline__ float prolonged(const F& f, const C& c, uint32_t x, uint32_t y, uint32_t z, size_t i) {
  const float e = f.E[i];
  const float d = f.D[i];
  const float ec = c.E[lidx(c, x >> 1, y >> 1, z >> 1)];
  return d > 0.0f ? e + CS_MG_SCALE * ec : e;
}
// Prolongation and the red pass in one: a red cell's update reads its
// neighbours' prolonged corrections.  The black cells' own prolonged values are
// read by nothing else before the black pass overwrites them, so the result
// equals prolonging every cell and then running the red pass.
template <class F, class C>
__device__ __forceinline__ void smooth_prolonged_cell(const F& f, const C& c, uint32_t x, uint32_t y, uint32_t z) {
  if (((x + y + z) & 1u) != 0u) return;
  const size_t i = lidx(f, x, y, z);
  const float d = f.D[i];
  if (!(d > 0.0f)) return;
  const nbr6 n = lneighbours(f, x, y, z, i);
  const float w0 = f.Wx[n.xm], w1 = f.Wx[i], w2 = f.Wy[n.ym], w3 = f.Wy[i], w4 = f.Wz[n.zm], w5 = f.Wz[i];
  // neighbour coordinates clamped like the indices (a clamped term is discarded)
  const float e0 = prolonged(f, c, n.bxm ? x - 1u : x, y, z, n.xm);
  const float e1 = prolonged(f, c, n.bxp ? x + 1u : x, y, z, n.xp);
  const float e2 = prolonged(f, c, x, n.bym ? y - 1u : y, z, n.ym);
  const float e3 = prolonged(f, c, x, n.byp ? y + 1u : y, z, n.yp);
  const float e4 = prolonged(f, c, x, y, n.bzm ? z - 1u : z, n.zm);
  const float e5 = prolonged(f, c, x, y, n.bzp ? z + 1u : z, n.zp);
  const float t0 = n.bxm ? w0 * e0 : 0.0f;
  const float t1 = n.bxp ? w1 * e1 : 0.0f;
  const float t2 = n.bym ? w2 * e2 : 0.0f;
  const float t3 = n.byp ? w3 * e3 : 0.0f;
  const float t4 = n.bzm ? w4 * e4 : 0.0f;
  const float t5 = n.bzp ? w5 * e5 : 0.0f;
  float s = t0;
  s += t1;
  s += t2;
  s += t3;
  s += t4;
  s += t5;
  f.E[i] = (f.R[i] + s) / d;
}

// k / d with m = floor((2^32 - 1) / d): the multiply-high is floor(k / d) or one
// less for k < 2^31, and one correction step settles it; r = k mod d.
__device__ __forceinline__ uint32_t div_magic(uint32_t k, uint32_t d, uint32_t m, uint32_t& r) {
  uint32_t q = __umulhi(k, m);
  r = k - q * d;
  if (r >= d) {
    ++q;
    r -= d;
  }
  return q;
}
template <class LV>
__device__ __forceinline__ void level_cell(const LV& L, uint32_t n, uint32_t& x, uint32_t& y, uint32_t& z) {
  const uint32_t r = div_magic(n, L.X, L.mX, x);
  z = div_magic(r, L.Y, L.mY, y);
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
  const nbr6 n = neighbours(u.gx, u.gy, u.gz, x, y, z, i);
  const float w0 = wc[n.xm].x, w2 = wc[n.ym].y, w4 = wc[n.zm].z;
  const float p0 = (float)p[n.xm] * u.inv_fp_scale, p1 = (float)p[n.xp] * u.inv_fp_scale;
  const float p2 = (float)p[n.ym] * u.inv_fp_scale, p3 = (float)p[n.yp] * u.inv_fp_scale;
  const float p4 = (float)p[n.zm] * u.inv_fp_scale, p5 = (float)p[n.zp] * u.inv_fp_scale;
  const float t0 = n.bxm ? w0 * p0 : 0.0f;
  const float t1 = n.bxp ? w.x * p1 : 0.0f;
  const float t2 = n.bym ? w2 * p2 : 0.0f;
  const float t3 = n.byp ? w.y * p3 : 0.0f;
  const float t4 = n.bzm ? w4 * p4 : 0.0f;
  const float t5 = n.bzp ? w.z * p5 : 0.0f;
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
__global__ void __launch_bounds__(256) k_mg_fine_setup(const cs_dparams* __restrict__ dp,
                                                       const uint32_t* __restrict__ list,
                                                       const float4* __restrict__ wc, const float* __restrict__ rhs,
                                                       const int32_t* __restrict__ p, cs_mg_level L,
                                                       const uint32_t* __restrict__ skip,
                                                       uint32_t* __restrict__ row_counts) {
  CS_PARAMS_FROM(dp);
  // the cycle's row lists start empty (cs_mg_rows appends to them)
  if (row_counts && blockIdx.x == 0u && threadIdx.x < 6u) row_counts[threadIdx.x] = 0u;
  if (blockIdx.x >= dp->nactive) return;
  if (*skip) return;
  uint32_t x0, y0, z0;
  brick_origin(u, list[blockIdx.x], x0, y0, z0);
  for (uint32_t k = threadIdx.x; k < (uint32_t)(BR * BR * BR); k += 256u) {
    const uint32_t x = x0 + (k & 7u), y = y0 + ((k >> 3) & 7u), z = z0 + (k >> 6);
    if (x >= u.gx || y >= u.gy || z >= u.gz) continue;
    const size_t i = ((size_t)z * u.gy + y) * u.gx + x;  // the grid's arrays
    const size_t li = lidx(L, x, y, z);                   // the level's
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
    L.D[li] = d;
    L.Wx[li] = wx;
    L.Wy[li] = wy;
    L.Wz[li] = wz;
    L.R[li] = r;
    L.E[li] = 0.0f;
  }
}

__global__ void __launch_bounds__(256) k_mg_fine_smooth(const cs_dparams* __restrict__ dp,
                                                        const uint32_t* __restrict__ list, cs_mg_level L,
                                                        uint32_t colour, const uint32_t* __restrict__ skip) {
  CS_PARAMS_FROM(dp);
  if (blockIdx.x >= dp->nactive) return;
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

__global__ void __launch_bounds__(64) k_mg_fine_restrict(const cs_dparams* __restrict__ dp,
                                                         const uint32_t* __restrict__ list, cs_mg_level f,
                                                         cs_mg_level c, const uint32_t* __restrict__ skip) {
  CS_PARAMS_FROM(dp);
  if (blockIdx.x >= dp->nactive) return;
  if (*skip) return;
  uint32_t x0, y0, z0;
  brick_origin(u, list[blockIdx.x], x0, y0, z0);
  const uint32_t k = threadIdx.x;
  const uint32_t X = (x0 >> 1) + (k & 3u), Y = (y0 >> 1) + ((k >> 2) & 3u), Z = (z0 >> 1) + (k >> 4);
  if (X < c.X && Y < c.Y && Z < c.Z) restrict_cell(f, c, X, Y, Z);
}

__global__ void __launch_bounds__(64) k_mg_fine_build(const cs_dparams* __restrict__ dp,
                                                      const uint32_t* __restrict__ list, cs_mg_level f,
                                                      cs_mg_level c, const uint32_t* __restrict__ skip) {
  CS_PARAMS_FROM(dp);
  if (blockIdx.x >= dp->nactive) return;
  if (*skip) return;
  uint32_t x0, y0, z0;
  brick_origin(u, list[blockIdx.x], x0, y0, z0);
  const uint32_t k = threadIdx.x;
  const uint32_t X = (x0 >> 1) + (k & 3u), Y = (y0 >> 1) + ((k >> 2) & 3u), Z = (z0 >> 1) + (k >> 4);
  if (X < c.X && Y < c.Y && Z < c.Z) build_cell(f, c, X, Y, Z);
}

__global__ void __launch_bounds__(256) k_mg_fine_smooth_prolonged(const cs_dparams* __restrict__ dp,
                                                                  const uint32_t* __restrict__ list,
                                                                  cs_mg_level f, cs_mg_level c,
                                                                  const uint32_t* __restrict__ skip) {
  CS_PARAMS_FROM(dp);
  if (blockIdx.x >= dp->nactive) return;
  if (*skip) return;
  uint32_t x0, y0, z0;
  brick_origin(u, list[blockIdx.x], x0, y0, z0);
  const uint32_t t = threadIdx.x;
  const uint32_t row = t >> 2, q = t & 3u;
  const uint32_t y = y0 + (row & 7u), z = z0 + (row >> 3);
  const uint32_t x = x0 + 2u * q + ((y + z + x0) & 1u);
  if (x < u.gx && y < u.gy && z < u.gz) smooth_prolonged_cell(f, c, x, y, z);
}

__global__ void __launch_bounds__(256) k_mg_fine_apply(const cs_dparams* __restrict__ dp,
                                                       const uint32_t* __restrict__ list, cs_mg_level L,
                                                       int32_t* __restrict__ p, const uint32_t* __restrict__ skip) {
  CS_PARAMS_FROM(dp);
  if (blockIdx.x >= dp->nactive) return;
  if (*skip) return;
  uint32_t x0, y0, z0;
  brick_origin(u, list[blockIdx.x], x0, y0, z0);
  for (uint32_t k = threadIdx.x; k < (uint32_t)(BR * BR * BR); k += 256u) {
    const uint32_t x = x0 + (k & 7u), y = y0 + ((k >> 3) & 7u), z = z0 + (k >> 6);
    if (x >= u.gx || y >= u.gy || z >= u.gz) continue;
    const size_t i = ((size_t)z * u.gy + y) * u.gx + x, li = lidx(L, x, y, z);
    if (!(L.D[li] > 0.0f)) continue;
    const float pn = (float)p[i] * u.inv_fp_scale + L.E[li];
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
__global__ void __launch_bounds__(256) k_mg_smooth_prolonged(cs_mg_level f, cs_mg_level c,
                                                             const uint32_t* __restrict__ skip) {
  if (*skip) return;
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= f.X * f.Y * f.Z) return;
  uint32_t x, y, z;
  level_cell(f, n, x, y, z);
  smooth_prolonged_cell(f, c, x, y, z);
}

// ---- passes inside one workgroup (LDS levels, 32-bit indices) -----------------------------
// Each staged level keeps the cells of its operator (D > 0) in two lists, one
// per colour, packed x | y << 10 | z << 20.  A pass of one colour reads only
// the other colour, so the order of a list (filled by LDS atomics) changes
// nothing; the cells off the lists keep a zero correction.  Same arithmetic
// per cell as smooth_cell / smooth_prolonged_cell / restrict_cell.
typedef __attribute__((address_space(3))) uint32_t suint;
struct slist {
  suint* c0;      // colour-0 cells, from the front of the level's list segment
  suint* c1end;   // colour-1 cells, from the back: c1end[-1 - k]
  uint32_t n0, n1;
};

__device__ __forceinline__ float wg_term(bool on, float w, float e) { return on ? w * e : 0.0f; }
__device__ __forceinline__ uint32_t wg_entry(const slist& l, uint32_t colour, uint32_t k) {
  return colour == 0u ? l.c0[k] : l.c1end[-1 - (int)k];
}

template <bool PROLONG>
__device__ __forceinline__ void wg_pass(const slevel& L, const slevel& C, const slist& A, uint32_t colour, uint32_t t) {
  const uint32_t X = L.X, Y = L.Y, Z = L.Z;
  const uint32_t count = colour == 0u ? A.n0 : A.n1;
  for (uint32_t k = t; k < count; k += CS_MG_WG_THREADS) {
    const uint32_t e = wg_entry(A, colour, k);
    const uint32_t x = e & 1023u, y = (e >> 10) & 1023u, z = e >> 20;
    const uint32_t i = (z * Y + y) * X + x;
    const float d = L.D[i];
    const uint32_t sy = X, sz = X * Y;
    const bool bxm = x > 0u, bxp = x + 1u < X, bym = y > 0u, byp = y + 1u < Y, bzm = z > 0u, bzp = z + 1u < Z;
    const uint32_t xm = bxm ? i - 1u : i, xp = bxp ? i + 1u : i, ym = bym ? i - sy : i, yp = byp ? i + sy : i;
    const uint32_t zm = bzm ? i - sz : i, zp = bzp ? i + sz : i;
    const float w0 = L.Wx[xm], w1 = L.Wx[i], w2 = L.Wy[ym], w3 = L.Wy[i], w4 = L.Wz[zm], w5 = L.Wz[i];
    float e0 = L.E[xm], e1 = L.E[xp], e2 = L.E[ym], e3 = L.E[yp], e4 = L.E[zm], e5 = L.E[zp];
    if (PROLONG) {
      // the neighbours' corrections after adding CS_MG_SCALE times their parents'
      const uint32_t cX = C.X, cXY = C.X * C.Y;
      const uint32_t px = x >> 1, py = y >> 1, pz = z >> 1;
      const uint32_t pc = (pz * C.Y + py) * cX + px;
      const uint32_t pxm = ((bxm ? x - 1u : x) >> 1) + (pz * C.Y + py) * cX;
      const uint32_t pxp = ((bxp ? x + 1u : x) >> 1) + (pz * C.Y + py) * cX;
      const uint32_t pym = pc - py * cX + ((bym ? y - 1u : y) >> 1) * cX;
      const uint32_t pyp = pc - py * cX + ((byp ? y + 1u : y) >> 1) * cX;
      const uint32_t pzm = pc - pz * cXY + ((bzm ? z - 1u : z) >> 1) * cXY;
      const uint32_t pzp = pc - pz * cXY + ((bzp ? z + 1u : z) >> 1) * cXY;
      const float d0 = L.D[xm], d1 = L.D[xp], d2 = L.D[ym], d3 = L.D[yp], d4 = L.D[zm], d5 = L.D[zp];
      const float c0 = C.E[pxm], c1 = C.E[pxp], c2 = C.E[pym], c3 = C.E[pyp], c4 = C.E[pzm], c5 = C.E[pzp];
      e0 = d0 > 0.0f ? e0 + CS_MG_SCALE * c0 : e0;
      e1 = d1 > 0.0f ? e1 + CS_MG_SCALE * c1 : e1;
      e2 = d2 > 0.0f ? e2 + CS_MG_SCALE * c2 : e2;
      e3 = d3 > 0.0f ? e3 + CS_MG_SCALE * c3 : e3;
      e4 = d4 > 0.0f ? e4 + CS_MG_SCALE * c4 : e4;
      e5 = d5 > 0.0f ? e5 + CS_MG_SCALE * c5 : e5;
    }
    float sum = wg_term(bxm, w0, e0);
    sum += wg_term(bxp, w1, e1);
    sum += wg_term(bym, w2, e2);
    sum += wg_term(byp, w3, e3);
    sum += wg_term(bzm, w4, e4);
    sum += wg_term(bzp, w5, e5);
    L.E[i] = (L.R[i] + sum) / d;
  }
}

// Restriction into the coarse cells of the operator (the others' right-hand
// side is never read and their correction stays zero).
__device__ __forceinline__ void wg_restrict(const slevel& F, const slevel& C, const slist& A, uint32_t t) {
  const uint32_t total = A.n0 + A.n1;
  for (uint32_t k = t; k < total; k += CS_MG_WG_THREADS) {
    const uint32_t e = k < A.n0 ? wg_entry(A, 0u, k) : wg_entry(A, 1u, k - A.n0);
    restrict_cell(F, C, e & 1023u, (e >> 10) & 1023u, e >> 20);
  }
}

// ---- one workgroup runs a whole visit of the small levels --------------------------------
// Levels first..nl-1 (each at most CS_MG_WG_CELLS cells) are staged in LDS:
// operator, right-hand side of the first level (its parent's restriction),
// corrections and the active lists; the first level's correction goes back to
// global memory for the parent's prolongation.
struct mg_levels {
  cs_mg_level l[CS_MG_MAX_LEVELS];
};
__global__ void __launch_bounds__(CS_MG_WG_THREADS) k_mg_wg_visit(mg_levels g, int first, int nl,
                                                                  const cs_mg_op* __restrict__ ops, uint32_t nops,
                                                                  const uint32_t* __restrict__ skip) {
  if (*skip) return;
  extern __shared__ float lds_raw[];
  sfloat* const lds = (sfloat*)lds_raw;
  __shared__ slevel lv[CS_MG_MAX_LEVELS];  // descriptors indexed at run time: in LDS, not scratch
  __shared__ slist al[CS_MG_MAX_LEVELS];
  __shared__ uint32_t cnt[CS_MG_MAX_LEVELS][2];
  __shared__ cs_mg_op s_ops[CS_MG_MAX_OPS];  // the op list, so no op waits on a global load
  const uint32_t t = threadIdx.x;
  for (uint32_t o = t; o < nops; o += CS_MG_WG_THREADS) s_ops[o] = ops[o];
  if (t < (uint32_t)(2 * CS_MG_MAX_LEVELS)) cnt[t >> 1][t & 1u] = 0u;
  size_t off = 0;
  for (int j = first; j < nl; ++j) off += 6 * (size_t)(g.l[j].X * g.l[j].Y * g.l[j].Z);
  suint* const lst = (suint*)(lds + off);  // one list segment of n entries per level
  __syncthreads();
  off = 0;
  size_t loff = 0;
  for (int j = first; j < nl; ++j) {
    const cs_mg_level G = g.l[j];
    const uint32_t n = G.X * G.Y * G.Z;
    slevel L;
    L.X = G.X;
    L.Y = G.Y;
    L.Z = G.Z;
    L.mX = G.mX;
    L.mY = G.mY;
    L.HX = (G.X + 1u) >> 1;
    L.mHX = (uint32_t)(0xFFFFFFFFull / L.HX);
    L.D = lds + off;
    L.Wx = L.D + n;
    L.Wy = L.D + 2 * n;
    L.Wz = L.D + 3 * n;
    L.R = L.D + 4 * n;
    L.E = L.D + 5 * n;
    off += 6 * (size_t)n;
    suint* const seg = lst + loff;
    loff += n;
    if (t == 0) lv[j] = L;
    for (uint32_t k = t; k < n; k += CS_MG_WG_THREADS) {
      const float d = G.D[k];
      L.D[k] = d;
      L.Wx[k] = G.Wx[k];
      L.Wy[k] = G.Wy[k];
      L.Wz[k] = G.Wz[k];
      L.R[k] = j == first ? G.R[k] : 0.0f;
      L.E[k] = 0.0f;
      if (d > 0.0f) {
        uint32_t x, y, z;
        level_cell(G, k, x, y, z);
        const uint32_t c = (x + y + z) & 1u;
        const uint32_t pos = atomicAdd(&cnt[j][c], 1u);
        const uint32_t e = x | (y << 10) | (z << 20);
        if (c == 0u) seg[pos] = e;
        else seg[n - 1u - pos] = e;
      }
    }
    if (t == 0) {
      al[j].c0 = seg;
      al[j].c1end = seg + n;
    }
  }
  __syncthreads();
  if (t < (uint32_t)CS_MG_MAX_LEVELS) {
    al[t].n0 = cnt[t][0];
    al[t].n1 = cnt[t][1];
  }
  __syncthreads();
  for (uint32_t o = 0; o < nops; ++o) {
    const cs_mg_op op = s_ops[o];
    const slevel L = lv[op.level];
    if (op.kind == CS_MG_OP_SMOOTH0 || op.kind == CS_MG_OP_SMOOTH1) {
      wg_pass<false>(L, L, al[op.level], op.kind == CS_MG_OP_SMOOTH0 ? 0u : 1u, t);
    } else if (op.kind == CS_MG_OP_RESTRICT) {
      wg_restrict(L, lv[op.level + 1], al[op.level + 1], t);
    } else {  // CS_MG_OP_SMOOTH0P
      wg_pass<true>(L, lv[op.level + 1], al[op.level], 0u, t);
    }
    __syncthreads();
  }
  const slevel F = lv[first];
  const uint32_t n = F.X * F.Y * F.Z;
  for (uint32_t k = t; k < n; k += CS_MG_WG_THREADS) g.l[first].E[k] = F.E[k];
}

// ---- the two smallest levels' visit with the rows held in registers -------------------------
// When the levels from the first one small enough for a workgroup are exactly
// two (A, then the coarsest B), a visit of A is a fixed sequence and every
// operation touches only cells of the operator, so each thread keeps the same
// cells for the whole visit.  k_mg_rows writes, once per cycle, every operator
// cell's row (coordinates, diagonal, the six face weights with the minus-side
// ones read from the neighbours) into per-level, per-colour lists; the visit
// loads its rows into registers and holds only the corrections in LDS, so a
// pass is one LDS round trip for the six neighbour corrections and a barrier.
// Same per-cell arithmetic as smooth_cell, restrict_cell and prolonged:
//  - restriction: every A cell of the operator stores its residual, then each
//    B cell sums its children's in the fixed order; a child outside the
//    operator holds +0.0, which leaves the sum's bits as skipping it would
//    (the sum starts at +0.0 and an exact zero sum of floats is +0.0);
//  - prolongation as its own step (E += CS_MG_SCALE * parent's E for A's
//    operator cells) before the red pass: the same values the folded form
//    reads.
// bit f set: the neighbour across face f (-x, +x, -y, +y, -z, +z) is on the grid
__device__ __forceinline__ uint32_t mg_face_flags(uint32_t x, uint32_t y, uint32_t z, uint32_t X, uint32_t Y,
                                                  uint32_t Z) {
  return (x > 0u ? 1u : 0u) | (x + 1u < X ? 2u : 0u) | (y > 0u ? 4u : 0u) | (y + 1u < Y ? 8u : 0u) |
         (z > 0u ? 16u : 0u) | (z + 1u < Z ? 32u : 0u);
}
// Rows of A, B and (when P.X > 0) P into their colour lists, one thread per
// cell over the three levels in turn; counts A0, A1, B0, B1, P0, P1 (zero
// before the launch) reserve each workgroup's run of list positions once.  The
// order within a list is immaterial: a pass reads only the other colour, and
// every sum has its fixed order.
__global__ void __launch_bounds__(256) k_mg_rows_list(cs_mg_level A, cs_mg_level B, cs_mg_level P,
                                                      cs_mg_row* __restrict__ rows, uint32_t* __restrict__ counts,
                                                      const uint32_t* __restrict__ skip) {
  if (*skip) return;
  __shared__ uint32_t cnt[6], base[6];
  const uint32_t t = threadIdx.x;
  if (t < 6u) cnt[t] = 0u;
  __syncthreads();
  const uint32_t nA = A.X * A.Y * A.Z, nB = B.X * B.Y * B.Z, nP = P.X * P.Y * P.Z;
  uint32_t k = blockIdx.x * 256u + t;
  int lev = -1;
  if (k < nA) {
    lev = 0;
  } else if (k < nA + nB) {
    lev = 1;
    k -= nA;
  } else if (k < nA + nB + nP) {
    lev = 2;
    k -= nA + nB;
  }
  const cs_mg_level L = lev == 0 ? A : (lev == 1 ? B : P);  // by value: kernel arguments stay off the stack
  uint32_t slot = 0u, pos = 0u;
  cs_mg_row r{};
  if (lev >= 0 && L.D[k] > 0.0f) {
    uint32_t x, y, z;
    level_cell(L, k, x, y, z);
    const uint32_t c = (x + y + z) & 1u;
    const uint32_t sy = L.X, sz = L.X * L.Y;
    r.xyz = x | (y << 10) | (z << 20);
    r.d = L.D[k];
    r.w[0] = x > 0u ? L.Wx[k - 1] : 0.0f;
    r.w[1] = L.Wx[k];
    r.w[2] = y > 0u ? L.Wy[k - sy] : 0.0f;
    r.w[3] = L.Wy[k];
    r.w[4] = z > 0u ? L.Wz[k - sz] : 0.0f;
    r.w[5] = L.Wz[k];
    slot = 1u + (uint32_t)lev * 2u + c;
    pos = atomicAdd(&cnt[slot - 1u], 1u);
  }
  __syncthreads();
  if (t < 6u && cnt[t]) base[t] = atomicAdd(&counts[t], cnt[t]);
  __syncthreads();
  if (slot) {
    const uint32_t j = slot - 1u, c = j & 1u;
    const size_t off = j < 2u ? (size_t)c * nA : j < 4u ? 2 * (size_t)nA + (size_t)c * nB
                                                         : 2 * ((size_t)nA + nB) + (size_t)c * nP;
    rows[off + base[j] + pos] = r;
  }
}
// P's compact ids: colour 0's list, then colour 1's, through the dense map.
__global__ void __launch_bounds__(256) k_mg_rows_map(cs_mg_level A, cs_mg_level B, cs_mg_level P,
                                                     const cs_mg_row* __restrict__ rows,
                                                     const uint32_t* __restrict__ counts,
                                                     uint32_t* __restrict__ pmap, const uint32_t* __restrict__ skip) {
  if (*skip) return;
  const uint32_t nA = A.X * A.Y * A.Z, nB = B.X * B.Y * B.Z, nP = P.X * P.Y * P.Z;
  const uint32_t c0 = counts[4], c1 = counts[5];
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= c0 + c1) return;
  const cs_mg_row* const lP = rows + 2 * ((size_t)nA + nB);
  const uint32_t xyz = (k < c0 ? lP[k] : lP[nP + (k - c0)]).xyz;
  pmap[((xyz >> 20) * P.Y + ((xyz >> 10) & 1023u)) * P.X + (xyz & 1023u)] = k;
}
// Each P entry's six neighbours' compact ids and its own (pnb, list order; an
// off-grid neighbour or one outside the operator as CS_MG_P3_ZERO), and each A
// entry's eight children's ids in restrict_cell's order (achild; a child off
// P's grid or outside its operator as CS_MG_P3_ZERO).
__global__ void __launch_bounds__(256) k_mg_rows_nb(cs_mg_level A, cs_mg_level B, cs_mg_level P,
                                                    const cs_mg_row* __restrict__ rows,
                                                    const uint32_t* __restrict__ counts,
                                                    const uint32_t* __restrict__ pmap, uint4* __restrict__ pnb,
                                                    uint4* __restrict__ achild, const uint32_t* __restrict__ skip) {
  if (*skip) return;
  const uint32_t nA = A.X * A.Y * A.Z, nB = B.X * B.Y * B.Z, nP = P.X * P.Y * P.Z;
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  const uint32_t c0 = counts[4], c1 = counts[5], a0 = counts[0], a1 = counts[1];
  if (k < c0 + c1) {
    const uint32_t e = k < c0 ? k : nP + (k - c0);
    const cs_mg_row& r = rows[2 * ((size_t)nA + nB) + e];
    const uint32_t x = r.xyz & 1023u, y = (r.xyz >> 10) & 1023u, z = r.xyz >> 20;
    const uint32_t i = (z * P.Y + y) * P.X + x;
    const uint32_t fl = mg_face_flags(x, y, z, P.X, P.Y, P.Z);
    const uint32_t sy = P.X, sz = P.X * P.Y;
    const uint32_t ni[6] = {i - 1u, i + 1u, i - sy, i + sy, i - sz, i + sz};
    uint32_t nb[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) nb[f] = ((fl >> f) & 1u) && P.D[ni[f]] > 0.0f ? pmap[ni[f]] : CS_MG_P3_ZERO;
    pnb[e] = make_uint4(nb[0] | (nb[1] << 16), nb[2] | (nb[3] << 16), nb[4] | (nb[5] << 16), k | (i << 16));
  }
  if (k < a0 + a1) {
    const uint32_t e = k < a0 ? k : nA + (k - a0);
    const uint32_t xyz = rows[e].xyz;
    const uint32_t X = xyz & 1023u, Y = (xyz >> 10) & 1023u, Z = xyz >> 20;
    uint32_t ch[8];
    uint32_t n = 0;
    for (uint32_t dz = 0; dz < 2u; ++dz)
      for (uint32_t dy = 0; dy < 2u; ++dy)
        for (uint32_t dx = 0; dx < 2u; ++dx, ++n) {
          const uint32_t x = 2u * X + dx, y = 2u * Y + dy, z = 2u * Z + dz;
          const uint32_t ci = (z * P.Y + y) * P.X + x;
          ch[n] = x < P.X && y < P.Y && z < P.Z && P.D[ci] > 0.0f ? pmap[ci] : CS_MG_P3_ZERO;
        }
    achild[e] = make_uint4(ch[0] | (ch[1] << 16), ch[2] | (ch[3] << 16), ch[4] | (ch[5] << 16), ch[6] | (ch[7] << 16));
  }
}

// One operator cell of a thread: its row in registers.
struct vcell {
  uint32_t i;    // index in the level
  uint32_t aux;  // A: the parent's index in B; B: packed coordinates
  uint32_t fl;   // bit f set: the neighbour across face f (-x, +x, -y, +y, -z, +z) is on the grid
  float d, w0, w1, w2, w3, w4, w5, r;
  bool on;
};
__device__ __forceinline__ vcell mg_load_cell(const cs_mg_row* __restrict__ list, uint32_t count, uint32_t k,
                                              const cs_mg_level& L, const cs_mg_level* P) {
  vcell v{};
  v.on = k < count;
  if (!v.on) return v;
  const cs_mg_row row = list[k];
  const uint32_t x = row.xyz & 1023u, y = (row.xyz >> 10) & 1023u, z = row.xyz >> 20;
  v.i = (z * L.Y + y) * L.X + x;
  v.aux = P ? ((z >> 1) * P->Y + (y >> 1)) * P->X + (x >> 1) : row.xyz;
  v.fl = mg_face_flags(x, y, z, L.X, L.Y, L.Z);
  v.d = row.d;
  v.w0 = row.w[0];
  v.w1 = row.w[1];
  v.w2 = row.w[2];
  v.w3 = row.w[3];
  v.w4 = row.w[4];
  v.w5 = row.w[5];
  return v;
}
// offsum against the corrections E (same terms and order as offsum)
template <class PTR>
__device__ __forceinline__ float mg_rsum(uint32_t i, uint32_t fl, float w0, float w1, float w2, float w3, float w4,
                                         float w5, PTR E, uint32_t sy, uint32_t sz) {
  const bool bxm = fl & 1u, bxp = fl & 2u, bym = fl & 4u, byp = fl & 8u, bzm = fl & 16u, bzp = fl & 32u;
  const float e0 = E[bxm ? i - 1u : i], e1 = E[bxp ? i + 1u : i], e2 = E[bym ? i - sy : i];
  const float e3 = E[byp ? i + sy : i], e4 = E[bzm ? i - sz : i], e5 = E[bzp ? i + sz : i];
  float s = wg_term(bxm, w0, e0);
  s += wg_term(bxp, w1, e1);
  s += wg_term(bym, w2, e2);
  s += wg_term(byp, w3, e3);
  s += wg_term(bzm, w4, e4);
  s += wg_term(bzp, w5, e5);
  return s;
}
__device__ __forceinline__ float mg_vsum(const vcell& v, const sfloat* E, uint32_t sy, uint32_t sz) {
  return mg_rsum(v.i, v.fl, v.w0, v.w1, v.w2, v.w3, v.w4, v.w5, E, sy, sz);
}
__device__ __forceinline__ void mg_vsmooth(const vcell& v, sfloat* E, uint32_t sy, uint32_t sz) {
  if (!v.on) return;
  const float s = mg_vsum(v, E, sy, sz);
  E[v.i] = (v.r + s) / v.d;
}
// A's restriction into B's cell v (children dz, dy, dx ascending; a child outside
// A's operator holds +0.0 in RA), which also zeroes B's correction.
__device__ __forceinline__ void mg_vrestrict(vcell& v, const cs_mg_level& A, const sfloat* RA, sfloat* EB) {
  if (!v.on) return;
  const uint32_t X = v.aux & 1023u, Y = (v.aux >> 10) & 1023u, Z = v.aux >> 20;
  float acc = 0.0f;
  for (uint32_t dz = 0; dz < 2u; ++dz)
    for (uint32_t dy = 0; dy < 2u; ++dy)
      for (uint32_t dx = 0; dx < 2u; ++dx) {
        const uint32_t x = 2u * X + dx, y = 2u * Y + dy, z = 2u * Z + dz;
        if (x >= A.X || y >= A.Y || z >= A.Z) continue;
        acc += RA[(z * A.Y + y) * A.X + x];
      }
  v.r = acc;
  EB[v.i] = 0.0f;
}

// A visit of A (its right-hand side in a[q].r, its correction zero in EA):
// CS_MG_NU passes, then CS_MG_GAMMA times {residuals into RA, restriction into
// B, `ncoarse` passes of B, prolongation, CS_MG_NU passes}.  a[2c + s] holds
// A's colour-c list entry t + s * CS_MG_ROWS_THREADS, b[c] B's entry t.  RA
// holds +0.0 outside A's operator.
__device__ __forceinline__ void mg_visit_ab(vcell (&a)[4], vcell (&b)[2], const cs_mg_level& A,
                                            const cs_mg_level& B, sfloat* EA, sfloat* RA, sfloat* EB,
                                            uint32_t ncoarse) {
  const uint32_t ay = A.X, az = A.X * A.Y, by = B.X, bz = B.X * B.Y;
  for (uint32_t k = 0; k < CS_MG_NU; ++k) {
    mg_vsmooth(a[0], EA, ay, az);
    mg_vsmooth(a[1], EA, ay, az);
    __syncthreads();
    mg_vsmooth(a[2], EA, ay, az);
    mg_vsmooth(a[3], EA, ay, az);
    __syncthreads();
  }
  for (uint32_t g = 0; g < CS_MG_GAMMA; ++g) {
    // residuals of A's operator cells, then their sums into B's cells
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const vcell& v = a[q];
      if (v.on) {
        const float s = mg_vsum(v, EA, ay, az);
        RA[v.i] = v.r - (v.d * EA[v.i] - s);
      }
    }
    __syncthreads();
    mg_vrestrict(b[0], A, RA, EB);
    mg_vrestrict(b[1], A, RA, EB);
    __syncthreads();
    for (uint32_t k = 0; k < ncoarse; ++k) {
      mg_vsmooth(b[0], EB, by, bz);
      __syncthreads();
      mg_vsmooth(b[1], EB, by, bz);
      __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const vcell& v = a[q];
      if (v.on) EA[v.i] = EA[v.i] + CS_MG_SCALE * EB[v.aux];
    }
    __syncthreads();
    for (uint32_t k = 0; k < CS_MG_NU; ++k) {
      mg_vsmooth(a[0], EA, ay, az);
      mg_vsmooth(a[1], EA, ay, az);
      __syncthreads();
      mg_vsmooth(a[2], EA, ay, az);
      mg_vsmooth(a[3], EA, ay, az);
      __syncthreads();
    }
  }
}

__device__ __forceinline__ void mg_load_ab(vcell (&a)[4], vcell (&b)[2], const cs_mg_level& A, const cs_mg_level& B,
                                           const cs_mg_row* __restrict__ rows, const uint32_t* __restrict__ counts,
                                           uint32_t t) {
  const uint32_t nA = A.X * A.Y * A.Z, nB = B.X * B.Y * B.Z;
  const cs_mg_row* const rowsB = rows + 2 * (size_t)nA;
#pragma unroll
  for (uint32_t c = 0; c < 2u; ++c) {
#pragma unroll
    for (uint32_t s = 0; s < 2u; ++s)
      a[2 * c + s] = mg_load_cell(rows + (size_t)c * nA, counts[c], t + s * CS_MG_ROWS_THREADS, A, &B);
    b[c] = mg_load_cell(rowsB + (size_t)c * nB, counts[2 + c], t, B, nullptr);
  }
}

__global__ void __launch_bounds__(CS_MG_ROWS_THREADS) k_mg_rows_visit(cs_mg_level A, cs_mg_level B,
                                                                      const cs_mg_row* __restrict__ rows,
                                                                      const uint32_t* __restrict__ counts,
                                                                      uint32_t ncoarse,
                                                                      const uint32_t* __restrict__ skip) {
  if (*skip) return;
  __shared__ float s_ea[CS_MG_ROWS_A];
  __shared__ float s_ra[CS_MG_ROWS_A];
  __shared__ float s_eb[CS_MG_ROWS_B];
  sfloat* const EA = (sfloat*)s_ea;
  sfloat* const RA = (sfloat*)s_ra;
  sfloat* const EB = (sfloat*)s_eb;
  const uint32_t t = threadIdx.x;
  const uint32_t nA = A.X * A.Y * A.Z, nB = B.X * B.Y * B.Z;
  vcell a[4], b[2];
  mg_load_ab(a, b, A, B, rows, counts, t);
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (a[q].on) a[q].r = A.R[a[q].i];  // the parent's restriction
  for (uint32_t k = t; k < nA; k += CS_MG_ROWS_THREADS) {
    EA[k] = 0.0f;
    RA[k] = 0.0f;  // cells outside the operator restrict as +0.0
  }
  for (uint32_t k = t; k < nB; k += CS_MG_ROWS_THREADS) EB[k] = 0.0f;
  __syncthreads();
  mg_visit_ab(a, b, A, B, EA, RA, EB, ncoarse);
  for (uint32_t k = t; k < nA; k += CS_MG_ROWS_THREADS) A.E[k] = EA[k];
}

// ---- three levels: P, then A and B, all in LDS ------------------------------------------
// The level above A (P) joins the visit when its operator has at most
// CS_MG_P3_MAX cells: its corrections, residuals and right-hand side sit in
// LDS by compact id, each entry's neighbours as ids (k_mg_rows), so a pass of
// P is one LDS round trip too.  Registers hold, per thread and colour, the
// first CS_MG_ROWS_PSLOTS P entries (ids, parent in A, row, right-hand side);
// further entries read theirs from the lists each time.  A's and B's entries:
// the first per thread and colour keep only index word and parent (A) or
// coordinates (B) in registers, their rows in LDS in slot order ([field]
// [colour][thread]); A's right-hand side lives in LDS for every cell, and A's
// entries beyond the first decode their row each time.  Summation orders and
// zero terms are those of the dense passes: a neighbour outside the operator
// or off the grid contributes w * 0 = +0.0 (the weight is finite and not
// negative), as the dense pass's skipped term does.
struct pcell {
  uint32_t idi;  // compact id | dense index << 16
  uint32_t par;  // the parent's index in A
  uint32_t nb01, nb23, nb45;
  float d, w0, w1, w2, w3, w4, w5, r;
};
// The value, opaque to the optimiser: LDS addresses derived from register-held
// ids are recomputed in each pass instead of being hoisted out of the visit's
// loops.
__device__ __forceinline__ uint32_t mg_opaque(uint32_t v) {
  asm volatile("" : "+v"(v));
  return v;
}
__device__ __forceinline__ pcell mg_pcell(const cs_mg_row& row, const uint4& nb, float r, const cs_mg_level& A) {
  const uint32_t x = row.xyz & 1023u, y = (row.xyz >> 10) & 1023u, z = row.xyz >> 20;
  pcell c;
  c.idi = nb.w;
  c.par = ((z >> 1) * A.Y + (y >> 1)) * A.X + (x >> 1);
  c.nb01 = nb.x;
  c.nb23 = nb.y;
  c.nb45 = nb.z;
  c.d = row.d;
  c.w0 = row.w[0];
  c.w1 = row.w[1];
  c.w2 = row.w[2];
  c.w3 = row.w[3];
  c.w4 = row.w[4];
  c.w5 = row.w[5];
  c.r = r;
  return c;
}
// OP 0 smooth, 1 residual into RP, 2 prolongation from A, 3 write the
// correction back to P.E
template <int OP>
__device__ __forceinline__ void mg_p3_cell(const pcell& c, const cs_mg_level& P, sfloat* EP, sfloat* RP,
                                           const sfloat* EA) {
  const uint32_t idi = mg_opaque(c.idi);
  const uint32_t id = idi & 0xffffu;
  if (OP == 2) {
    EP[id] = EP[id] + CS_MG_SCALE * EA[mg_opaque(c.par)];
    return;
  }
  if (OP == 3) {
    P.E[idi >> 16] = EP[id];
    return;
  }
  const uint32_t n01 = mg_opaque(c.nb01), n23 = mg_opaque(c.nb23), n45 = mg_opaque(c.nb45);
  const float e0 = EP[n01 & 0xffffu], e1 = EP[n01 >> 16], e2 = EP[n23 & 0xffffu];
  const float e3 = EP[n23 >> 16], e4 = EP[n45 & 0xffffu], e5 = EP[n45 >> 16];
  float s = c.w0 * e0;
  s += c.w1 * e1;
  s += c.w2 * e2;
  s += c.w3 * e3;
  s += c.w4 * e4;
  s += c.w5 * e5;
  if (OP == 0) EP[id] = (c.r + s) / c.d;
  else RP[id] = c.r - (c.d * EP[id] - s);
}
template <int OP>
__device__ __forceinline__ void mg_p3_list(const pcell (&q)[CS_MG_ROWS_PSLOTS], const cs_mg_row* __restrict__ list,
                                           const uint4* __restrict__ nbl, uint32_t count, const cs_mg_level& P,
                                           const cs_mg_level& A, sfloat* EP, sfloat* RP, const sfloat* EA,
                                           uint32_t t) {
#pragma unroll
  for (int s = 0; s < CS_MG_ROWS_PSLOTS; ++s)
    if (t + (uint32_t)s * CS_MG_ROWS3_THREADS < count) mg_p3_cell<OP>(q[s], P, EP, RP, EA);
  for (uint32_t k = t + CS_MG_ROWS_PSLOTS * CS_MG_ROWS3_THREADS; k < count; k += CS_MG_ROWS3_THREADS) {
    const cs_mg_row row = list[k];
    const uint4 nb = nbl[k];
    mg_p3_cell<OP>(mg_pcell(row, nb, P.R[nb.w >> 16], A), P, EP, RP, EA);
  }
}

struct rowv {
  float d, w0, w1, w2, w3, w4, w5;
};
__device__ __forceinline__ uint32_t mg_iword(uint32_t xyz, const cs_mg_level& L) {
  const uint32_t x = xyz & 1023u, y = (xyz >> 10) & 1023u, z = xyz >> 20;
  return ((z * L.Y + y) * L.X + x) | (mg_face_flags(x, y, z, L.X, L.Y, L.Z) << 16) | 0x80000000u;
}
__device__ __forceinline__ uint32_t mg_parent(uint32_t xyz, const cs_mg_level& C) {
  const uint32_t x = xyz & 1023u, y = (xyz >> 10) & 1023u, z = xyz >> 20;
  return ((z >> 1) * C.Y + (y >> 1)) * C.X + (x >> 1);
}
__device__ __forceinline__ rowv mg_rowv(const cs_mg_row& r) {
  return rowv{r.d, r.w[0], r.w[1], r.w[2], r.w[3], r.w[4], r.w[5]};
}
__device__ __forceinline__ rowv mg_rowv_lds(const sfloat* RW, uint32_t slot, uint32_t stride) {
  return rowv{RW[slot], RW[stride + slot], RW[2 * stride + slot], RW[3 * stride + slot], RW[4 * stride + slot],
              RW[5 * stride + slot], RW[6 * stride + slot]};
}
__device__ __forceinline__ void mg_rowv_store(sfloat* RW, uint32_t slot, uint32_t stride, const rowv& v) {
  RW[slot] = v.d;
  RW[stride + slot] = v.w0;
  RW[2 * stride + slot] = v.w1;
  RW[3 * stride + slot] = v.w2;
  RW[4 * stride + slot] = v.w3;
  RW[5 * stride + slot] = v.w4;
  RW[6 * stride + slot] = v.w5;
}
template <class PTR>
__device__ __forceinline__ float mg_wsum(uint32_t iw, const rowv& w, PTR E, uint32_t sy, uint32_t sz) {
  return mg_rsum(iw & 0xffffu, (iw >> 16) & 63u, w.w0, w.w1, w.w2, w.w3, w.w4, w.w5, E, sy, sz);
}
// OP 0 smooth, 1 residual, 2 prolongation from B, 3 restriction from P (which
// zeroes A's correction), on A's cell (iw, par) with row w; ch: its children's
// compact ids in P (restrict_cell's order; CS_MG_P3_ZERO, whose RP stays +0.0,
// for a child outside P's operator or grid: adding +0.0 to a sum that starts
// at +0.0 leaves its bits as skipping the child does)
template <int OP>
__device__ __forceinline__ void mg_a3_cell(uint32_t iw, uint32_t par, const rowv& w, const uint4& ch,
                                           const cs_mg_level& A, sfloat* EA, sfloat* RAR, sfloat* RAS,
                                           const sfloat* EB, const sfloat* RP) {
  iw = mg_opaque(iw);
  const uint32_t i = iw & 0xffffu;
  if (OP == 0) {
    const float s = mg_wsum(iw, w, EA, A.X, A.X * A.Y);
    EA[i] = (RAR[i] + s) / w.d;
  } else if (OP == 1) {
    const float s = mg_wsum(iw, w, EA, A.X, A.X * A.Y);
    RAS[i] = RAR[i] - (w.d * EA[i] - s);
  } else if (OP == 2) {
    EA[i] = EA[i] + CS_MG_SCALE * EB[par];
  } else {
    const uint32_t cx = mg_opaque(ch.x), cy = mg_opaque(ch.y), cz = mg_opaque(ch.z), cw = mg_opaque(ch.w);
    const float v0 = RP[cx & 0xffffu], v1 = RP[cx >> 16], v2 = RP[cy & 0xffffu], v3 = RP[cy >> 16];
    const float v4 = RP[cz & 0xffffu], v5 = RP[cz >> 16], v6 = RP[cw & 0xffffu], v7 = RP[cw >> 16];
    float acc = 0.0f;
    acc += v0;
    acc += v1;
    acc += v2;
    acc += v3;
    acc += v4;
    acc += v5;
    acc += v6;
    acc += v7;
    RAR[i] = acc;
    EA[i] = 0.0f;
  }
}
constexpr uint32_t kA3Stride = 2u * CS_MG_ROWS3_THREADS;  // A's register-entry rows in LDS: [field][colour][thread]
constexpr uint32_t kB3Stride = CS_MG_ROWS_B;             // B's rows in LDS: [field][colour][thread < |B| / 2]
constexpr uint32_t kB3Half = CS_MG_ROWS_B / 2u;
constexpr uint32_t kP3Cap = CS_MG_P3_ZERO + 1u;           // P's compact arrays (the last id stays zero)
constexpr size_t mg_rows3_lds_words() {
  return 3 * (size_t)CS_MG_ROWS_A + CS_MG_ROWS_B + 7 * (size_t)kA3Stride + 8 * (size_t)kB3Stride + 2 * (size_t)kP3Cap;
}
static_assert(mg_rows3_lds_words() * sizeof(float) <= 160u * 1024u, "three-level visit LDS exceeds 160 KiB");
size_t mg_rows3_lds_bytes() {
  // EA, RAR, RAS; EB; A's rows; B's rows and right-hand side; EP, RP: 161,792 B of the 160 KiB
  return sizeof(float) * (3 * (size_t)CS_MG_ROWS_A + CS_MG_ROWS_B + 7 * (size_t)kA3Stride + 8 * (size_t)kB3Stride +
                          2 * (size_t)kP3Cap);
}
__global__ void __launch_bounds__(CS_MG_ROWS3_THREADS) k_mg_rows3_visit(cs_mg_level P, cs_mg_level A, cs_mg_level B,
                                                                       const cs_mg_row* __restrict__ rows,
                                                                       const uint32_t* __restrict__ counts,
                                                                       const uint4* __restrict__ pnb,
                                                                       const uint4* __restrict__ achild,
                                                                       uint32_t ncoarse,
                                                                       const uint32_t* __restrict__ skip) {
  if (*skip) return;
  extern __shared__ float lds_raw[];
  sfloat* const EA = (sfloat*)lds_raw;
  sfloat* const RAR = EA + CS_MG_ROWS_A;   // A's right-hand side
  sfloat* const RAS = RAR + CS_MG_ROWS_A;  // A's residuals (+0.0 outside its operator)
  sfloat* const EB = RAS + CS_MG_ROWS_A;
  sfloat* const AW = EB + CS_MG_ROWS_B;    // A's register-entry rows
  sfloat* const BW = AW + 7 * kA3Stride;   // B's rows, then B's right-hand side
  sfloat* const BR = BW + 7 * kB3Stride;
  sfloat* const EP = BR + kB3Stride;       // P's corrections and residuals by compact id
  sfloat* const RP = EP + kP3Cap;
  constexpr uint32_t NT = CS_MG_ROWS3_THREADS;
  const uint32_t t = threadIdx.x;
  const uint32_t nA = A.X * A.Y * A.Z, nB = B.X * B.Y * B.Z, nP = P.X * P.Y * P.Z;
  const cs_mg_row* const lA[2] = {rows, rows + nA};
  const cs_mg_row* const lB = rows + 2 * (size_t)nA;
  const cs_mg_row* const lP0 = rows + 2 * ((size_t)nA + nB);
  const cs_mg_row* const lP1 = lP0 + nP;
  const uint4* const nP0 = pnb;
  const uint4* const nP1 = pnb + nP;
  const uint32_t cA[2] = {counts[0], counts[1]}, cB[2] = {counts[2], counts[3]};
  const uint32_t cP0 = counts[4], cP1 = counts[5];
  uint32_t ia[2], pa[2], ib[2], xb[2];
  uint4 ac[2];
#pragma unroll
  for (uint32_t c = 0; c < 2u; ++c) {
    ia[c] = pa[c] = ib[c] = xb[c] = 0u;
    ac[c] = make_uint4(0u, 0u, 0u, 0u);
    if (t < cA[c]) {
      const cs_mg_row r = lA[c][t];
      ia[c] = mg_iword(r.xyz, A);
      pa[c] = mg_parent(r.xyz, B);
      ac[c] = achild[(size_t)c * nA + t];
      mg_rowv_store(AW, c * NT + t, kA3Stride, mg_rowv(r));
    }
    if (t < cB[c]) {
      const cs_mg_row r = lB[(size_t)c * nB + t];
      ib[c] = mg_iword(r.xyz, B);
      xb[c] = r.xyz;
      mg_rowv_store(BW, c * kB3Half + t, kB3Stride, mg_rowv(r));
    }
  }
  pcell q0[CS_MG_ROWS_PSLOTS], q1[CS_MG_ROWS_PSLOTS];
#pragma unroll
  for (int s = 0; s < CS_MG_ROWS_PSLOTS; ++s) {
    const uint32_t k = t + (uint32_t)s * NT;
    q0[s] = pcell{};
    q1[s] = pcell{};
    if (k < cP0) {
      const uint4 nb = nP0[k];
      q0[s] = mg_pcell(lP0[k], nb, P.R[nb.w >> 16], A);
    }
    if (k < cP1) {
      const uint4 nb = nP1[k];
      q1[s] = mg_pcell(lP1[k], nb, P.R[nb.w >> 16], A);
    }
  }
  for (uint32_t k = t; k < nA; k += NT) {
    EA[k] = 0.0f;
    RAS[k] = 0.0f;
  }
  for (uint32_t k = t; k < nB; k += NT) EB[k] = 0.0f;
  for (uint32_t k = t; k < kP3Cap; k += NT) EP[k] = 0.0f;  // P.E is zero at the visit's start
  if (t == 0u) RP[CS_MG_P3_ZERO] = 0.0f;
  __syncthreads();
  // A's colour-c list: the register entry, then the rest from the list
  auto opA = [&](auto op, uint32_t c) {
    constexpr int OP = decltype(op)::value;
    if (ia[c]) mg_a3_cell<OP>(ia[c], pa[c], mg_rowv_lds(AW, c * NT + t, kA3Stride), ac[c], A, EA, RAR, RAS, EB, RP);
    for (uint32_t k = t + NT; k < cA[c]; k += NT) {
      const cs_mg_row r = lA[c][k];
      const uint4 ch = OP == 3 ? achild[(size_t)c * nA + k] : make_uint4(0u, 0u, 0u, 0u);
      mg_a3_cell<OP>(mg_iword(r.xyz, A), mg_parent(r.xyz, B), mg_rowv(r), ch, A, EA, RAR, RAS, EB, RP);
    }
  };
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, 1>;
  using S2 = std::integral_constant<int, 2>;
  using S3 = std::integral_constant<int, 3>;
  auto passA = [&]() {
    opA(S0{}, 0u);
    __syncthreads();
    opA(S0{}, 1u);
    __syncthreads();
  };
  auto smoothB = [&](uint32_t c) {
    if (!ib[c]) return;
    const rowv w = mg_rowv_lds(BW, c * kB3Half + t, kB3Stride);
    const uint32_t iw = mg_opaque(ib[c]);
    const uint32_t i = iw & 0xffffu;
    const float s = mg_wsum(iw, w, EB, B.X, B.X * B.Y);
    EB[i] = (BR[c * kB3Half + t] + s) / w.d;
  };
  auto restrictB = [&](uint32_t c) {  // restrict_cell from A's residuals, zeroing B's correction
    if (!ib[c]) return;
    const uint32_t X = xb[c] & 1023u, Y = (xb[c] >> 10) & 1023u, Z = xb[c] >> 20;
    // the eight children's residuals loaded at once (a child off A's grid
    // reads the cell itself and adds +0.0 instead, which keeps the sum's bits)
    float v[8];
    bool in[8];
    uint32_t n = 0;
    for (uint32_t dz = 0; dz < 2u; ++dz)
      for (uint32_t dy = 0; dy < 2u; ++dy)
        for (uint32_t dx = 0; dx < 2u; ++dx, ++n) {
          const uint32_t x = 2u * X + dx, y = 2u * Y + dy, z = 2u * Z + dz;
          in[n] = x < A.X && y < A.Y && z < A.Z;
          v[n] = RAS[in[n] ? (z * A.Y + y) * A.X + x : 0u];
        }
    float acc = 0.0f;
    for (uint32_t j = 0; j < 8u; ++j) acc += in[j] ? v[j] : 0.0f;
    BR[c * kB3Half + t] = acc;
    EB[ib[c] & 0xffffu] = 0.0f;
  };
  auto passP = [&]() {
    mg_p3_list<0>(q0, lP0, nP0, cP0, P, A, EP, RP, EA, t);
    __syncthreads();
    mg_p3_list<0>(q1, lP1, nP1, cP1, P, A, EP, RP, EA, t);
    __syncthreads();
  };
  for (uint32_t k = 0; k < CS_MG_NU; ++k) passP();
  for (uint32_t g = 0; g < CS_MG_GAMMA; ++g) {
    // P's residuals, then A's right-hand side from its children (restrict_cell)
    mg_p3_list<1>(q0, lP0, nP0, cP0, P, A, EP, RP, EA, t);
    mg_p3_list<1>(q1, lP1, nP1, cP1, P, A, EP, RP, EA, t);
    __syncthreads();
    opA(S3{}, 0u);
    opA(S3{}, 1u);
    __syncthreads();
    // A's visit (the steps of mg_visit_ab)
    for (uint32_t k = 0; k < CS_MG_NU; ++k) passA();
    for (uint32_t ga = 0; ga < CS_MG_GAMMA; ++ga) {
      opA(S1{}, 0u);
      opA(S1{}, 1u);
      __syncthreads();
      restrictB(0u);
      restrictB(1u);
      __syncthreads();
      for (uint32_t k = 0; k < ncoarse; ++k) {
        smoothB(0u);
        __syncthreads();
        smoothB(1u);
        __syncthreads();
      }
      opA(S2{}, 0u);
      opA(S2{}, 1u);
      __syncthreads();
      for (uint32_t k = 0; k < CS_MG_NU; ++k) passA();
    }
    // prolongation of A's correction into P, then P's passes
    mg_p3_list<2>(q0, lP0, nP0, cP0, P, A, EP, RP, EA, t);
    mg_p3_list<2>(q1, lP1, nP1, cP1, P, A, EP, RP, EA, t);
    __syncthreads();
    for (uint32_t k = 0; k < CS_MG_NU; ++k) passP();
  }
  // P's corrections back to P.E for its parent's prolongation (every other
  // cell of P.E is already zero)
  mg_p3_list<3>(q0, lP0, nP0, cP0, P, A, EP, RP, EA, t);
  mg_p3_list<3>(q1, lP1, nP1, cP1, P, A, EP, RP, EA, t);
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
    l.mX = (uint32_t)(0xFFFFFFFFull / X);
    l.mY = (uint32_t)(0xFFFFFFFFull / Y);
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
    for (uint32_t k = 0; k < cs_mg_coarse_passes(nlevels); ++k) pass(level);
    return;
  }
  for (uint32_t k = 0; k < CS_MG_NU; ++k) pass(level);
  for (uint32_t g = 0; g < CS_MG_GAMMA; ++g) {
    out->push_back(cs_mg_op{CS_MG_OP_RESTRICT, (uint8_t)level});
    cs_mg_visit_ops(level + 1, nlevels, out);
    out->push_back(cs_mg_op{CS_MG_OP_SMOOTH0P, (uint8_t)level});
    out->push_back(cs_mg_op{CS_MG_OP_SMOOTH1, (uint8_t)level});
    for (uint32_t k = 1; k < CS_MG_NU; ++k) pass(level);
  }
}

hipError_t cs_mg_fine_setup(hipStream_t s, const cs_pblock& P, uint32_t nactive, const uint32_t* list,
                            const float* wc, const float* rhs, const int32_t* p, cs_mg_level L0, const uint32_t* skip,
                            uint32_t* row_counts) {
  const cs_params& u = P.u;
  if (!nactive) return row_counts ? cs_launch_fill_i32(s, reinterpret_cast<int32_t*>(row_counts), 6, 0) : hipSuccess;
  hipLaunchKernelGGL(k_mg_fine_setup, dim3(nactive), dim3(256), 0, s, P.d,
                     list, reinterpret_cast<const float4*>(wc), rhs,
                     p, L0, skip, row_counts);
  return hipGetLastError();
}
hipError_t cs_mg_fine_smooth(hipStream_t s, const cs_pblock& P, uint32_t nactive, const uint32_t* list,
                             cs_mg_level L0, uint32_t colour, const uint32_t* skip) {
  const cs_params& u = P.u;
  if (!nactive) return hipSuccess;
  hipLaunchKernelGGL(k_mg_fine_smooth, dim3(nactive), dim3(256), 0, s, P.d, list, L0, colour, skip);
  return hipGetLastError();
}
hipError_t cs_mg_fine_restrict(hipStream_t s, const cs_pblock& P, uint32_t nactive, const uint32_t* list,
                               cs_mg_level L0, cs_mg_level L1, const uint32_t* skip) {
  const cs_params& u = P.u;
  if (!nactive) return hipSuccess;
  hipLaunchKernelGGL(k_mg_fine_restrict, dim3(nactive), dim3(64), 0, s, P.d, list, L0, L1, skip);
  return hipGetLastError();
}
hipError_t cs_mg_fine_build(hipStream_t s, const cs_pblock& P, uint32_t nactive, const uint32_t* list,
                            cs_mg_level L0, cs_mg_level L1, const uint32_t* skip) {
  const cs_params& u = P.u;
  if (!nactive) return hipSuccess;
  hipLaunchKernelGGL(k_mg_fine_build, dim3(nactive), dim3(64), 0, s, P.d, list, L0, L1, skip);
  return hipGetLastError();
}
hipError_t cs_mg_fine_smooth_prolonged(hipStream_t s, const cs_pblock& P, uint32_t nactive, const uint32_t* list,
                                       cs_mg_level L0, cs_mg_level L1, const uint32_t* skip) {
  const cs_params& u = P.u;
  if (!nactive) return hipSuccess;
  hipLaunchKernelGGL(k_mg_fine_smooth_prolonged, dim3(nactive), dim3(256), 0, s, P.d, list, L0, L1, skip);
  return hipGetLastError();
}
hipError_t cs_mg_fine_apply(hipStream_t s, const cs_pblock& P, uint32_t nactive, const uint32_t* list,
                            cs_mg_level L0, int32_t* p, const uint32_t* skip) {
  const cs_params& u = P.u;
  if (!nactive) return hipSuccess;
  hipLaunchKernelGGL(k_mg_fine_apply, dim3(nactive), dim3(256), 0, s, P.d, list, L0, p, skip);
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
hipError_t cs_mg_smooth_prolonged(hipStream_t s, cs_mg_level f, cs_mg_level c, const uint32_t* skip) {
  hipLaunchKernelGGL(k_mg_smooth_prolonged, cells_grid(cells(f)), dim3(256), 0, s, f, c, skip);
  return hipGetLastError();
}
size_t cs_mg_wg_lds_bytes(const cs_mg_level* levels, int first, int nl) {
  size_t words = 0;  // per level: 6 arrays and the active lists, one word per cell each
  for (int j = first; j < nl; ++j) words += 7 * (size_t)cells(levels[j]);
  return words * sizeof(float);
}
hipError_t cs_mg_wg_prepare(size_t lds_bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mg_wg_visit),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
}
hipError_t cs_mg_wg_visit(hipStream_t s, const cs_mg_level* levels, int first, int nl, const cs_mg_op* ops,
                          uint32_t nops, const uint32_t* skip) {
  mg_levels lv{};
  for (int l = 0; l < CS_MG_MAX_LEVELS; ++l) lv.l[l] = levels[l];
  hipLaunchKernelGGL(k_mg_wg_visit, dim3(1), dim3(CS_MG_WG_THREADS), cs_mg_wg_lds_bytes(levels, first, nl), s, lv,
                     first, nl, ops, nops, skip);
  return hipGetLastError();
}
bool cs_mg_rows_fit(const cs_mg_level* levels, int first, int nl) {
  if (nl - first != 2) return false;
  const cs_mg_level& A = levels[first];
  const cs_mg_level& B = levels[first + 1];
  if (A.X >= 1024u || A.Y >= 1024u || A.Z >= 1024u) return false;
  // every colour list within two entries per thread (A) and one (B)
  return cells(A) <= CS_MG_ROWS_A && cells(B) <= CS_MG_ROWS_B;
}
bool cs_mg_rows3_fit(const cs_mg_level* levels, int first, int nl) {
  if (!cs_mg_rows_fit(levels, first, nl) || first < 2) return false;
  const cs_mg_level& P = levels[first - 1];
  return P.X < 1024u && P.Y < 1024u && P.Z < 1024u && cells(P) <= 65535u;
}
hipError_t cs_mg_rows(hipStream_t s, cs_mg_level A, cs_mg_level B, const cs_mg_level* P, cs_mg_row* rows,
                      uint32_t* counts, uint32_t* pmap, uint4* pnb, uint4* achild, const uint32_t* skip) {
  cs_mg_level none{};
  const cs_mg_level PP = P ? *P : none;
  const uint32_t nA = cells(A), nB = cells(B), nP = P ? cells(*P) : 0u;
  hipLaunchKernelGGL(k_mg_rows_list, cells_grid(nA + nB + nP), dim3(256), 0, s, A, B, PP, rows, counts, skip);
  if (P) {
    hipLaunchKernelGGL(k_mg_rows_map, cells_grid(nP), dim3(256), 0, s, A, B, PP, rows, counts, pmap, skip);
    hipLaunchKernelGGL(k_mg_rows_nb, cells_grid(std::max(nP, nA)), dim3(256), 0, s, A, B, PP, rows, counts, pmap, pnb,
                       achild, skip);
  }
  return hipGetLastError();
}
hipError_t cs_mg_rows_visit(hipStream_t s, cs_mg_level A, cs_mg_level B, const cs_mg_row* rows,
                            const uint32_t* counts, uint32_t coarse_passes, const uint32_t* skip) {
  hipLaunchKernelGGL(k_mg_rows_visit, dim3(1), dim3(CS_MG_ROWS_THREADS), 0, s, A, B, rows, counts, coarse_passes,
                     skip);
  return hipGetLastError();
}
hipError_t cs_mg_rows3_prepare() {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mg_rows3_visit),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)mg_rows3_lds_bytes());
}
hipError_t cs_mg_rows3_visit(hipStream_t s, cs_mg_level P, cs_mg_level A, cs_mg_level B, const cs_mg_row* rows,
                             const uint32_t* counts, const uint4* pnb, const uint4* achild, uint32_t coarse_passes,
                             const uint32_t* skip) {
  hipLaunchKernelGGL(k_mg_rows3_visit, dim3(1), dim3(CS_MG_ROWS3_THREADS), mg_rows3_lds_bytes(), s, P, A, B, rows,
                     counts, pnb, achild, coarse_passes, skip);
  return hipGetLastError();
}
