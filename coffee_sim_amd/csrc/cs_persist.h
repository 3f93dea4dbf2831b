// cs_persist.h — the persistent red-black Gauss-Seidel solve (cs_rbgs_persist.hip):
// one launch of co-resident workgroups, one active brick each, that runs a
// whole solve with neighbour-to-neighbour synchronisation.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "cs_common.h"
#include "cs_solver.h"

#define CS_PERSIST_THREADS 256
#define CS_PERSIST_FLAG_STRIDE 16u  // u32 words between two workgroups' flags (64 B)
#define CS_PERSIST_FACE_WORDS (6u * 2u * 32u)  // face-buffer words per brick
#define CS_PERSIST_MAX_BRICKS 1024              // active bricks one persistent solve takes

struct cs_persist_args {
  const cs_dparams* dp;      // the substep's parameters, epoch base and halt word
  const uint32_t* list;      // active bricks, ascending (cs_launch_brick_scan)
  const int32_t* slot;       // per brick: its position in `list` (its run), or -1
  uint32_t* xcc;             // [max_runs]: 1 + the XCC id each run's workgroup runs on
  const float4* wc;          // face weights and diagonal per cell (cs_launch_face_weights)
  const float* rhs;          // c = dx^2 * rhs per cell
  int32_t* p;                // the pressure: the start on entry, the solution on exit
  uint32_t* metrics;         // clamp fires are added to CS_M_P_CLAMP
  cs_solve_state* st;        // active_bricks in; iterations, residuals, done, faults out
  uint64_t* fb;              // [nbricks][6 faces][2 colours][32]: surface cells as
                             // (epoch << 32 | fixed-point pressure)
  uint32_t* pflags;          // [max_runs * stride]: residual checks published per brick
  double* partials;          // [2][max_runs][2]: (rr, bb) per brick and check parity
  uint32_t max_runs;         // capacity of pflags and partials (the brick count)
  uint32_t fixed_pairs;      // 0: iterate to `tol`; else exactly this many red+black pairs
  uint32_t every;            // iterations between residual tests (tolerance mode)
  uint32_t cap;              // iteration cap (tolerance mode)
  double tol;
  uint64_t spin_limit;       // 100 MHz ticks a wait may last before the solve gives up
  uint32_t poll_delay;       // s_sleep(1) steps (64 clocks each) before a halo word's first poll
  const uint32_t* min_gate;  // when non-null and *min_gate == 0 (a multigrid start ran), the
  uint32_t min_iters;        //   tolerance stops the solve only from min_iters iterations on
};

// Clears the slot map, check flags, XCD table and failure state, then maps each
// listed brick to its slot (reads the active count from `st`).
hipError_t cs_persist_prepare(hipStream_t s, const cs_params& u, const uint32_t* list, cs_solve_state* st,
                              int32_t* slot, uint32_t* pflags, uint32_t* xcc);
// Epochs a solve of at most `iters` red+black pairs uses (the next solve's base moves by this).
inline uint32_t cs_persist_epochs(uint32_t iters) { return 2u * iters + 2u; }
// Workgroups launched for `nactive` bricks (a multiple of 8, one per XCD slice).
uint32_t cs_persist_groups(uint32_t nactive);
// Workgroups of the persistent kernel that fit on the device at once.
hipError_t cs_persist_max_groups(int device, uint32_t* out);
// With `cooperative` the launch fails instead of starting workgroups that
// would wait for ones not yet resident.
hipError_t cs_persist_launch(hipStream_t s, const cs_persist_args& a, uint32_t groups, bool cooperative);
