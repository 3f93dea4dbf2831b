// cs_mg.h — the multigrid start of the pressure solve (rbgs_mg_cycles).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <vector>

#include "cs_common.h"
#include "cs_solver.h"

#define CS_MG_NU 2u              // red-black smoothing passes before and after each coarse correction
#define CS_MG_GAMMA 2u           // coarse corrections per visit (W-cycle)
#define CS_MG_SCALE 1.5f         // weight of a prolonged correction
#define CS_MG_COARSE 10u         // passes per visit of the coarsest level, visited at most 4 times a cycle
#define CS_MG_COARSEST_SIDE 8u   // levels are added until no side exceeds this
#define CS_MG_MAX_LEVELS 12
#define CS_MG_WG_CELLS 4096u     // levels this small run inside one workgroup
#define CS_MG_WG_THREADS 512
#define CS_MG_MAX_OPS 4096u

// One level: a dense X*Y*Z grid of cells with the Galerkin operator (diagonal
// D, coupling to the +x/+y/+z neighbour), the right-hand side R and the
// correction E.  Level 0 is the fine grid, restricted to the free cells.
struct cs_mg_level {
  uint32_t X, Y, Z;
  uint32_t mX, mY;  // floor((2^32 - 1) / X) and / Y: division by X and Y as a multiply-high
  uint32_t bnx, bny;  // nonzero: the arrays are brick-major (8^3 bricks of 512 cells, bnx x bny
                      // bricks a layer), else x-fastest
  float* D;
  float* Wx;
  float* Wy;
  float* Wz;
  float* R;
  float* E;
};

// An operation of the single-workgroup interpreter: kind, the level it acts on.
// SMOOTH0 / SMOOTH1: a red / black pass of `level`; RESTRICT: its residual into
// level+1 (whose correction is zeroed); SMOOTH0P: level+1's correction
// prolonged into `level` together with its red pass.
enum : uint8_t { CS_MG_OP_SMOOTH0 = 0, CS_MG_OP_SMOOTH1 = 1, CS_MG_OP_RESTRICT = 2, CS_MG_OP_SMOOTH0P = 3 };
struct cs_mg_op {
  uint8_t kind;
  uint8_t level;
};

// Passes per visit of the coarsest of `nl` levels: CS_MG_COARSE while a cycle
// visits it at most 4 times (CS_MG_GAMMA^(nl-2) visits), else the same 40
// passes per cycle spread over the visits (at least 2 each).
inline uint32_t cs_mg_coarse_passes(int nl) {
  uint32_t visits = 1;
  for (int l = 2; l < nl; ++l) visits *= CS_MG_GAMMA;
  if (visits <= 4u) return CS_MG_COARSE;
  const uint32_t p = 4u * CS_MG_COARSE / visits;
  return p < 2u ? 2u : p;
}
// The level dimensions the start uses on a grid of the given size.
std::vector<cs_mg_level> cs_mg_shape(uint32_t gx, uint32_t gy, uint32_t gz);
// The operations of one W-cycle visit of `level` over levels [level, nlevels).
void cs_mg_visit_ops(int level, int nlevels, std::vector<cs_mg_op>* out);

// Fine level (over the active bricks): operator, residual b - A p and zero
// correction of the free cells; smoothing, restriction to / prolongation from
// level 1; and the final p <- clamp(p + e) in fixed point.  Every kernel returns
// at once when *skip is set.
// cs_mg_fine_setup also zeroes row_counts (cs_mg_rows's six counts) when given,
// skipped cycle or not.
hipError_t cs_mg_fine_setup(hipStream_t s, const cs_pblock& P, uint32_t nactive, const uint32_t* list,
                            const float* wc, const float* rhs, const int32_t* p, cs_mg_level L0, const uint32_t* skip,
                            uint32_t* row_counts);
hipError_t cs_mg_fine_smooth(hipStream_t s, const cs_pblock& P, uint32_t nactive, const uint32_t* list,
                             cs_mg_level L0, uint32_t colour, const uint32_t* skip);
hipError_t cs_mg_fine_restrict(hipStream_t s, const cs_pblock& P, uint32_t nactive, const uint32_t* list,
                               cs_mg_level L0, cs_mg_level L1, const uint32_t* skip);
hipError_t cs_mg_fine_build(hipStream_t s, const cs_pblock& P, uint32_t nactive, const uint32_t* list,
                            cs_mg_level L0, cs_mg_level L1, const uint32_t* skip);
hipError_t cs_mg_fine_smooth_prolonged(hipStream_t s, const cs_pblock& P, uint32_t nactive, const uint32_t* list,
                                       cs_mg_level L0, cs_mg_level L1, const uint32_t* skip);
hipError_t cs_mg_fine_apply(hipStream_t s, const cs_pblock& P, uint32_t nactive, const uint32_t* list,
                            cs_mg_level L0, int32_t* p, const uint32_t* skip);
// Dense levels.
hipError_t cs_mg_smooth(hipStream_t s, cs_mg_level L, uint32_t colour, const uint32_t* skip);
hipError_t cs_mg_restrict(hipStream_t s, cs_mg_level f, cs_mg_level c, const uint32_t* skip);
hipError_t cs_mg_build(hipStream_t s, cs_mg_level f, cs_mg_level c, const uint32_t* skip);
// The coarse correction prolonged into f together with f's red pass.
hipError_t cs_mg_smooth_prolonged(hipStream_t s, cs_mg_level f, cs_mg_level c, const uint32_t* skip);
// The whole visit of level `first` in one workgroup, levels first..nl-1 staged
// in LDS: `ops` (device) as cs_mg_visit_ops lists them.
size_t cs_mg_wg_lds_bytes(const cs_mg_level* levels, int first, int nl);
hipError_t cs_mg_wg_prepare(size_t lds_bytes);
hipError_t cs_mg_wg_visit(hipStream_t s, const cs_mg_level* levels, int first, int nl, const cs_mg_op* ops,
                          uint32_t nops, const uint32_t* skip);
// The same visit with each operator cell's row held in registers, for the
// usual case of exactly two small levels (A = levels[first], B the coarsest):
// k_mg_rows lists the rows once per cycle (after the Galerkin builds), by level
// and colour: A's colour c at rows[c * |A|], B's at rows[2|A| + c * |B|], and
// with a third level P = levels[first - 1] P's at rows[2(|A| + |B|) + c * |P|];
// the counts A0, A1, B0, B1, P0, P1 in `counts`.  Same results as
// cs_mg_wg_visit, bit for bit.
#define CS_MG_ROWS_THREADS 1024
#define CS_MG_ROWS_A 4096u  // A's cells at most: two list entries per thread and colour
#define CS_MG_ROWS_B 512u   // B's cells at most: one list entry per thread and colour
#define CS_MG_ROWS3_THREADS 512  // threads of the three-level visit
#define CS_MG_ROWS_PSLOTS 6      // P's list entries per thread and colour held in registers (more are read per pass)
#define CS_MG_P3_ZERO 8191u      // P's compact id that stays zero (off-grid and non-operator neighbours)
#define CS_MG_P3_MAX 8191u       // P's operator cells at most for the three-level visit
struct cs_mg_row {
  uint32_t xyz;  // x | y << 10 | z << 20
  float d;
  float w[6];    // face weights -x, +x, -y, +y, -z, +z (minus sides read from the neighbours)
};
bool cs_mg_rows_fit(const cs_mg_level* levels, int first, int nl);
// The three-level visit's level shapes fit: a level P above A (not the fine
// level) of at most 65,535 cells with sides below 1024.  Its operator must also
// have at most CS_MG_P3_MAX cells, which the caller bounds from the active
// bricks before choosing it.
bool cs_mg_rows3_fit(const cs_mg_level* levels, int first, int nl);
// pmap (|P| words), pnb (2|P| entries) and achild (2|A| entries): P's compact
// ids, P's entries' neighbour ids and A's entries' children ids, written when
// P is given.  counts must be zero before the call (the cycle's
// cs_mg_fine_setup zeroes them).
hipError_t cs_mg_rows(hipStream_t s, cs_mg_level A, cs_mg_level B, const cs_mg_level* P, cs_mg_row* rows,
                      uint32_t* counts, uint32_t* pmap, uint4* pnb, uint4* achild, const uint32_t* skip);
hipError_t cs_mg_rows_visit(hipStream_t s, cs_mg_level A, cs_mg_level B, const cs_mg_row* rows,
                            const uint32_t* counts, uint32_t coarse_passes, const uint32_t* skip);
// A whole visit of P (A's and B's visits inside) in one workgroup, all three
// levels in LDS; P's corrections written back to P.E at the end.
hipError_t cs_mg_rows3_prepare();  // once, before the first cs_mg_rows3_visit
hipError_t cs_mg_rows3_visit(hipStream_t s, cs_mg_level P, cs_mg_level A, cs_mg_level B, const cs_mg_row* rows,
                             const uint32_t* counts, const uint4* pnb, const uint4* achild, uint32_t coarse_passes,
                             const uint32_t* skip);
// flags, set by each residual test of the start (cs_launch_residual_gate):
// flags[0] = skip the next cycle, flags[1] = the start before any cycle was
// within the tolerance.  Every MG kernel reads flags[0].
