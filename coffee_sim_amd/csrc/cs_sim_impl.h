// cs_sim_impl.h — the simulation handle's internals, shared by the
// single-scene substep (cs_sim.cpp) and the batched one (cs_batch.cpp).
#pragma once

#include <hip/hip_runtime_api.h>

#include <vector>

#include "cs_ctx.h"
#include "cs_host.h"
#include "cs_mg.h"
#include "cs_persist.h"
#include "cs_solver.h"
#include "cs_unet_host.h"

enum Pass { P_P2G = 0, P_GRID, P_CLASSIFY, P_SOLVE, P_POST, P_G2P, P_END, P_COUNT };

#define CS_PARAM_RING 64     // pinned host copies of the parameter block in flight at most
#define CS_SOLVE_RECORDS 64  // recorded substeps replayed between two waits of the host

struct cs_sim {
  cs_ctx* ctx = nullptr;
  cs_settings s;
  uint32_t gx = 0, gy = 0, gz = 0;
  size_t nc = 0;
  uint32_t nbricks = 0;
  // device state
  float* part = nullptr;     // 8 floats per particle
  float* aff = nullptr;      // 12 floats per particle
  int32_t* lanes = nullptr;  // [6][nc]
  float* gv = nullptr;       // float4 per cell
  float* gv2 = nullptr;      // float4 per cell (after viscosity)
  float* sdf = nullptr;
  uint8_t* cls = nullptr;
  int32_t* bed_id = nullptr;
  cs_bed_cell* bed = nullptr;
  uint32_t num_bed_cells = 0;
  // a bed of particles (reference mode): their own arrays, re-stamped every substep
  uint32_t num_bed = 0;
  float* bed_part = nullptr;        // 8 floats per bed particle
  float* bed_aff = nullptr;         // 12 floats per bed particle
  float* bed_ext = nullptr;         // BedExtract, 8 floats per bed particle
  int32_t* bed_lookup = nullptr;    // per cell: highest stamping bed particle, or -1
  uint32_t* bed_counter = nullptr;  // table slots taken this substep
  int32_t* bed_dw = nullptr;        // per bed particle: [4][num_bed] water credit, then reaction impulse x/y/z (fixed point)
  float* bed_snap = nullptr;        // per bed particle: position and radius before bed_dynamics moves it
  uint8_t* alias4 = nullptr;        // per cell: the viscosity scratch's z lane reads as the solid kind
  std::vector<float> init_bed_ext;  // the scene's BedExtracts, restored by reset
  std::vector<float> init_bed_part; // the scene's bed particles and affine rows, restored by reset
  std::vector<float> init_bed_aff;
  int8_t* kind = nullptr;
  float* kf = nullptr;
  int32_t* pres = nullptr;
  float* rhs = nullptr;
  int32_t* pack = nullptr;
  float* tgt = nullptr;      // classify's target divergence per fluid cell
  int32_t* prev_p = nullptr;
  float* warm = nullptr;
  float* render = nullptr;   // RenderParticle per particle slot, packed on demand
  bool has_warm = false;
  uint32_t* metrics = nullptr;
  double* res_partials = nullptr;
  double* res_out = nullptr;
  // brick-sparse solve
  int32_t* pres2 = nullptr;          // second pressure buffer
  float* wc = nullptr;               // float4 per cell: face weights and diagonal
  uint32_t* brick_flag = nullptr;    // [nbricks]
  uint32_t* brick_list = nullptr;    // [nbricks]
  double* brick_partials = nullptr;  // [2 * nbricks]
  cs_solve_state* dstate = nullptr;
  cs_solve_state* h_state = nullptr; // pinned
  uint32_t* h_done = nullptr;        // pinned, one slot per in-flight batch
  hipEvent_t batch_ev[2] = {};
  std::vector<hipEvent_t> chunk_ev;  // start/end pairs for sweep timing
  bool dense_solver = false;         // one-thread-per-cell reference layout (CS_SOLVER=dense)
  int halo = 4;                      // half sweeps per launch (CS_RBGS_HALO = 1, 2 or 4)
  int threads = 512;                 // threads per brick workgroup (CS_RBGS_THREADS = 256 or 512)
  bool multi = true;                 // halo 4: one prefetching workgroup per compute unit (CS_RBGS_MULTI)
  uint32_t multi_groups = 256;       // workgroups of the multi-brick launch: the compute-unit count
  uint32_t batch_chunks = 4;         // residual checks enqueued ahead of the host
  bool fuse_residual = true;         // residual test inside the multi-brick sweeps (CS_RBGS_FUSE)
  double* group_partials = nullptr;  // [2 * 1024]: per-workgroup sums of the fused test
  uint32_t* sweep_table = nullptr;   // per-thread cell lists of the multi-brick sweeps
  bool use_graphs = true;            // replay the steady-state batch as a hipGraph (CS_GRAPHS)
  hipGraph_t batch_graph = nullptr;
  hipGraphExec_t batch_exec = nullptr;
  uint32_t batch_launches_per_chunk = 0;
  // persistent solve (cs_rbgs_persist.hip): the whole solve in one launch
  bool persist = true;               // CS_RBGS_PERSIST (off when another layout is asked for)
  uint32_t persist_max_groups = 0;   // co-resident workgroups of the persistent kernel
  uint64_t* persist_fb = nullptr;    // [nbricks * CS_PERSIST_FACE_WORDS] face buffers
  uint32_t persist_epoch = 1;        // base of the next solve's epochs
  uint32_t* persist_pflags = nullptr;
  int32_t* persist_slot = nullptr;   // [nbricks] brick -> run
  uint32_t* persist_xcc = nullptr;   // [nbricks] run -> 1 + XCC id
  double* persist_partials = nullptr;  // [2 * nbricks * 2]
  uint64_t persist_solves = 0;       // solves run by the persistent kernel
  uint32_t persist_poll_delay = 8;   // CS_PERSIST_POLL_DELAY: about 0.25 us, measured best of 0-24
  // multigrid start (cs_mg.hip), allocated on first use
  std::vector<cs_mg_level> mg_levels;  // level 0 = the fine grid
  float* mg_buf = nullptr;             // every level's D, Wx, Wy, Wz, R, E
  size_t mg_floats = 0;
  size_t mg_n0 = 0;                    // cells of the fine level's arrays (brick-major: whole bricks)
  bool mg_brick_major = true;          // CS_MG_BRICKS: the fine level's arrays brick-major (0: x-fastest)
  uint32_t* mg_flags = nullptr;        // [0] skip the next cycle, [1] no cycle ran (start within the tolerance),
                                       // [2] cycles run in the current start
  cs_mg_op* mg_ops = nullptr;          // one visit of level mg_wg_first, run by one workgroup
  uint32_t mg_nops = 0;
  int mg_wg_first = 0;
  uint32_t nactive = 0;                // active bricks of the current substep's solve (brick_list)
  int mg_rows_allowed = 2;             // CS_MG_ROWS: 0 the operation-list visit, 1 at most the two-level
                                       // register visit, 2 also the three-level one (cross-checks)
  bool mg_rows = false;                // the small levels' visit keeps its rows in registers (cs_mg_rows_visit)
  bool mg_rows3_shape = false;         // the level above them can join it (cs_mg_rows3_visit) ...
  bool mg_rows3 = false;               // ... and does in the current start (its operator fits for sure)
  cs_mg_row* mg_row_buf = nullptr;     // their row lists, rebuilt every cycle
  uint32_t* mg_row_counts = nullptr;   // [6]
  uint32_t* mg_pmap = nullptr;         // the three-level visit's top level: dense index -> compact id
  uint4* mg_pnb = nullptr;             // and each list entry's neighbour ids
  uint4* mg_achild = nullptr;          // the level below it: each list entry's children's compact ids
  hipEvent_t mg_ev[2] = {nullptr, nullptr};
  bool mg_timed = false;               // mg_ev holds an unharvested start
  // the device-resident parameter block (cs_dparams) and the pinned ring its host
  // part is uploaded from, one slot per substep
  cs_dparams* dparams = nullptr;
  cs_dparams* h_pring = nullptr;
  uint32_t pring_next = 0;
  // recorded substeps: the whole substep as one graph, replayed (step_recorded)
  bool substep_graphs = true;        // CS_SUBSTEP_GRAPHS (and CS_GRAPHS)
  uint32_t substep_cap = 0;          // CS_SUBSTEP_CAP: a smaller recorded brick capacity (0: the limit)
  hipGraph_t sg_graph = nullptr;
  hipGraphExec_t sg_exec = nullptr;
  uint32_t sg_cap = 0;               // the recorded graph's active-brick capacity ...
  uint32_t sg_emit = 0;              // ... and emission width
  cs_solve_state* h_records = nullptr;  // pinned [CS_SOLVE_RECORDS]: each replayed substep's final solve
                                        // state, written by the device
  uint32_t* h_halt = nullptr;           // pinned: the block's halt word, written by the device
  uint64_t recorded_batches = 0;     // host waits of step_recorded
  uint64_t recorded_substeps = 0;    // substeps run as replays to the end
  uint64_t recorded_halts = 0;       // replays halted at the brick scan (finished launch by launch)
  // tile-binned P2G scratch
  bool p2g_tiled = true;             // CS_P2G=atomic selects the one-thread-per-particle layout
  bool g2p_tiled = true;             // G2P over P2G's tile bins, footprints in LDS (CS_G2P=global: per particle)
  uint32_t* tile_count = nullptr;    // [nbricks]
  uint32_t* tile_start = nullptr;    // [nbricks]
  uint32_t* tile_chunk = nullptr;    // [nbricks + 1] the chunked tile pass's work-item prefix
  uint2* p2g_key = nullptr;          // [max_particles]
  uint32_t* p2g_sorted = nullptr;    // [max_particles]
  cs_unet* warm_net = nullptr;       // network of CS_WARM_LEARNED
  // host state
  double* h_res = nullptr;    // pinned
  std::vector<float> init_particles, init_affine;
  uint32_t init_count = 0;
  cs::Inflow inflow;
  uint32_t num_water = 0;
  float total_time = 0.0f;
  float total_emitted_mass = 0.0f;
  uint32_t total_dropped = 0;
  cs_solve_stats stats;
  bool final_residual_pending = false;
  bool pending_from_state = false;  // lazily read residual lives in h_state (brick solve)
  bool timing = false;
  hipEvent_t ev[P_COUNT] = {};
  float pass_us[7] = {};
  // per-kernel timing (cs_kernel_timing): [p2g start, end, g2p start, end, rbgs start, end]
  hipEvent_t kev[6] = {};
  cs_kernel_timing kt = {};
};

// Sets the message cs_last_error returns and returns `code`.
int cs_fail(int code, const char* msg);
// CS_ERR_DEVICE with the HIP error's text.
int cs_hip_fail(hipError_t e, const char* what);
