// cs_solver.h — launchers of the brick-sparse red-black Gauss-Seidel solve.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "cs_common.h"

// Device-resident state of one pressure solve.
struct cs_solve_state {
  uint32_t done;           // residual reached the tolerance
  uint32_t iterations;     // red+black pairs executed
  uint32_t launches;       // sweep launches executed (parity says which buffer is current)
  uint32_t active_bricks;  // bricks holding a fluid cell
  uint32_t pending_fires;   // clamp fires of a residual-computing launch, committed by the deciding one
  float rescale;           // Galerkin factor of the one-time rescale (rbgs_rescale_after)
  double residual;         // latest relative residual
  double initial_residual; // before the first sweep
  uint32_t abort;          // persistent solve: a workgroup gave up waiting, the others stop
  uint32_t faults;         // persistent solve: workgroups that gave up
  uint64_t verdict;        // persistent solve: (residual test number << 32) | stop
};

hipError_t cs_launch_brick_scan(hipStream_t s, const cs_pblock& P, const uint32_t* flags, uint32_t* list,
                                cs_solve_state* st);
// The state into rec[d->slot] (slots below n; nothing once the block is halted)
// and the halt word into *halt; rec and halt may be pinned host memory.
hipError_t cs_launch_solve_record(hipStream_t s, const cs_dparams* d, const cs_solve_state* st, cs_solve_state* rec,
                                  uint32_t* halt, uint32_t n);
// Face weights and diagonal of every cell, float4 per cell in `wc` [4 * nc].
hipError_t cs_launch_face_weights(hipStream_t s, const cs_pblock& P, const float* kf, float* wc);
// One launch = `halo` half sweeps (colour color0, color0^1, ...) over every active brick;
// (halo, threads) in {(1,256), (2,256), (2,512), (4,256), (4,512)}.
hipError_t cs_launch_rbgs_bricks(hipStream_t s, const cs_params& u, uint32_t nactive, int halo, int threads,
                                 uint32_t color0, const uint32_t* list, const float* wc, const float* rhs,
                                 const int32_t* pin, int32_t* pout, uint32_t* metrics, const cs_solve_state* st);
// Convergence test folded into the multi-brick sweeps (never both in one launch).
#define CS_RES_COMPUTE 1u  // sum the residual of the launch's input pressure into partials[groups]
#define CS_RES_DECIDE 2u   // reduce the previous launch's partials, advance the state, stop if converged

// Four half sweeps per launch with `groups` workgroups of 1024 threads (one per compute unit),
// each walking a run of the active-brick list and prefetching the next brick while it sweeps.
// `sweep_table` holds CS_SWEEP_TABLE_WORDS codes from cs_build_sweep_table.
#define CS_MULTI_HALO 4
#define CS_MULTI_THREADS 1024
#define CS_SWEEP_TABLE_WORDS (2 * CS_MULTI_HALO * CS_MULTI_THREADS)
void cs_build_sweep_table(uint32_t* out /* CS_SWEEP_TABLE_WORDS */);
hipError_t cs_launch_rbgs_multi(hipStream_t s, const cs_params& u, uint32_t nactive, uint32_t groups,
                                uint32_t color0, const uint32_t* list, const float* wc, const float* rhs,
                                const int32_t* pin, int32_t* pout, uint32_t* metrics, cs_solve_state* st,
                                double* partials, uint32_t res_mode, uint32_t chunk_iters, uint32_t chunk_launches,
                                double tol, const uint32_t* sweep_table);
// mode 0: initial residual (always runs); mode 1: after `iters` iterations in `launches` launches.
// Two launches: per-brick partials into `partials` [2 * nactive], then a one-workgroup reduce.
hipError_t cs_launch_residual_bricks(hipStream_t s, const cs_pblock& P, uint32_t nactive, const uint32_t* list,
                                     const float* wc, const float* rhs, const int32_t* p, double* partials,
                                     cs_solve_state* st, uint32_t mode, uint32_t iters, uint32_t launches,
                                     double tol);
// The initial residual (mode 0) followed, in the reduce's launch, by the
// multigrid start's gate on it: flags[0] skip the next cycle, flags[1] the
// start was within `tol` (set by the first test), flags[2] cycles run.  After
// the last test st->done is cleared unless flags[1]: after a multigrid start
// the iteration runs at least one chunk before its first test.
hipError_t cs_launch_residual_gate(hipStream_t s, const cs_pblock& P, uint32_t nactive, const uint32_t* list,
                                   const float* wc, const float* rhs, const int32_t* p, double* partials,
                                   cs_solve_state* st, double tol, uint32_t* flags, bool first, bool last);
// The one-time rescale p <- alpha p (alpha = <b,p>/<p,Ap> over the residual's rows) of the
// fluid cells of `p`, skipped when the state says done; alpha is left in st->rescale.
hipError_t cs_launch_rescale(hipStream_t s, const cs_params& u, uint32_t nactive, const uint32_t* list,
                             const float* wc, const float* rhs, const float* kf, int32_t* p, double* partials,
                             cs_solve_state* st, uint32_t* metrics);
// Brick cells of `b` into `a`: when the state's launch count is odd (`by_state`) or always.
hipError_t cs_launch_select_final(hipStream_t s, const cs_params& u, uint32_t nactive, const uint32_t* list,
                                  int32_t* a, const int32_t* b, const cs_solve_state* st, bool by_state);
