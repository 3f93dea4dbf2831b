// cs_kernels.h — host-side launchers for the gfx950 kernels in cs_kernels.hip.
// Pointers are device pointers; every launcher enqueues on `stream` and
// returns the launch status.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#include "cs_common.h"

hipError_t cs_launch_fill_i32(hipStream_t stream, int32_t* p, size_t n, int32_t v);
hipError_t cs_launch_copy_i32(hipStream_t stream, const int32_t* src, int32_t* dst, size_t n);
// The same, as no-ops once the substep's parameter block is halted (cs_dparams).
hipError_t cs_launch_fill_i32_gated(hipStream_t stream, const cs_dparams* d, int32_t* p, size_t n, int32_t v);
hipError_t cs_launch_copy_i32_gated(hipStream_t stream, const cs_dparams* d, const int32_t* src, int32_t* dst,
                                    size_t n);
hipError_t cs_launch_p2g(hipStream_t stream, const cs_pblock& P, const float* part, const float* aff,
                         int32_t* lanes, uint32_t* metrics);
// Tile-binned P2G (cs_p2g.hip): scratch tile_count / tile_start [nbricks], tile_chunk
// [nbricks + 1] (the chunk prefix of the tile pass), key [max_particles] (tile, rank),
// sorted [max_particles].
hipError_t cs_launch_p2g_tiled(hipStream_t stream, const cs_pblock& P, const float* part, const float* aff,
                               int32_t* lanes, uint32_t* metrics, uint32_t* tile_count, uint32_t* tile_start,
                               uint32_t* tile_chunk, uint2* key, uint32_t* sorted);
hipError_t cs_launch_grid(hipStream_t stream, const cs_pblock& P, const int32_t* lanes, float* gvel,
                          const float* sdf);
hipError_t cs_launch_classify(hipStream_t stream, const cs_pblock& P, const float* gvel, const uint8_t* cls,
                              const int32_t* lanes, const int32_t* bed_id, const cs_bed_cell* bed, int8_t* kind,
                              float* kf, int32_t* pres, float* rhs, int32_t* pack, float* tgt, uint32_t* metrics,
                              const float* sdf, uint32_t* brick_flag);
hipError_t cs_launch_warm_previous(hipStream_t stream, const cs_pblock& P, const float* kf, const int32_t* prev,
                                   int32_t* pres);
hipError_t cs_launch_warm_external(hipStream_t stream, const cs_params& u, const float* kf, const float* warm,
                                   int32_t* pres, uint32_t* metrics);
hipError_t cs_launch_rbgs_pair(hipStream_t stream, const cs_params& u, const float* kf, const float* rhs,
                               int32_t* pres, uint32_t* metrics);
hipError_t cs_launch_residual(hipStream_t stream, const cs_params& u, const float* kf, const float* rhs,
                              const int32_t* pres, double* partials, uint32_t max_blocks, double* out);
hipError_t cs_launch_project(hipStream_t stream, const cs_pblock& P, float* gvel, const float* kf,
                             const int32_t* pres, const int32_t* bed_id, const cs_bed_cell* bed, const float* sdf,
                             const cs_grains& grains);
// The divergence of the projected velocity over the fluid cells (metric slots 5
// and 6), from classify's per-cell target `tgt`.
hipError_t cs_launch_post_div(hipStream_t stream, const cs_params& u, const float* gvel, const float* kf,
                              const uint8_t* cls, const float* sdf, const float* tgt, uint32_t* metrics);
// The same over the `nactive` bricks of `list` (every fluid cell lies in an active brick).
hipError_t cs_launch_post_div_bricks(hipStream_t stream, const cs_pblock& P, uint32_t nactive, const uint32_t* list,
                                     const float* gvel, const float* kf, const uint8_t* cls, const float* sdf,
                                     const float* tgt, uint32_t* metrics);
hipError_t cs_launch_pack(hipStream_t stream, const cs_pblock& P, float* gvel, const float* kf, const int32_t* pack,
                          const float* sdf);
// alias4 (may be null): per cell 1 when the viscosity scratch's z lane holds
// CELL_SOLID, the lane bed_dynamics reads as the kind (shader.rs:158-164, 238-259).
hipError_t cs_launch_visc(hipStream_t stream, const cs_pblock& P, const float* gin, float* gout, const uint8_t* cls,
                          const float* sdf, uint8_t* alias4);
hipError_t cs_launch_g2p(hipStream_t stream, const cs_pblock& P, float* part, float* aff, const float* gvel,
                         const int32_t* bed_id, const cs_bed_cell* bed, const float* sdf, const cs_grains& grains);
// The same over the tile bins of this substep's cs_launch_p2g_tiled (its
// tile_count / tile_start / tile_chunk / key / sorted, the particles unmoved
// since): each tile's footprint staged in LDS, same results as cs_launch_g2p.
hipError_t cs_launch_g2p_tiled(hipStream_t stream, const cs_pblock& P, float* part, float* aff, const float* gvel,
                               const int32_t* bed_id, const cs_bed_cell* bed, const float* sdf, const cs_grains& grains,
                               const uint32_t* tile_count, const uint32_t* tile_start, const uint32_t* tile_chunk,
                               const uint2* key, const uint32_t* sorted);
// prepare_render (shader.rs:2242-2276): 8 floats per particle into `out`; particles
// below num_bed are bed particles, with their BedExtracts in bed_ext (may be null).
hipError_t cs_launch_pack_render(hipStream_t stream, const cs_params& u, const float* part, const float* aff,
                                 uint32_t num_bed, const float* bed_ext, float* out);
hipError_t cs_launch_expf(hipStream_t stream, const float* in, float* out, uint32_t n);
hipError_t cs_launch_sincosf(hipStream_t stream, const float* in, float* sn, float* cs, uint32_t n);
// The scene's SDF (sdf_res^3 floats) and class mask (one byte per cell) from the
// settings' obstacles, as cs_generate_sdf / cs_generate_sdf_class make them on the host.
hipError_t cs_launch_gen_sdf(hipStream_t stream, const cs_settings& s, float* out);
hipError_t cs_launch_gen_sdf_class(hipStream_t stream, const cs_settings& s, const float* sdf, uint8_t* out);
// Reference-mode bed: clear the lookup, stamp it from the nbed bed particles
// (bed_lookup_clear / bed_lookup_scatter) and build the table of the stamped cells
// (at most `capacity`), mapping each cell to its entry in bed_id.
hipError_t cs_launch_bed_lookup(hipStream_t st, const cs_pblock& P, uint32_t nbed, const float* bed_part,
                                const float* bed_aff, const float* bed_ext, uint32_t capacity, int32_t* lookup,
                                uint32_t* counter, int32_t* bed_id, cs_bed_cell* table);
// bed_coupling over the water particles, then extraction_advect over the nbed bed
// particles (shader.rs:1885-1987); dw holds one i32 water credit per bed particle.
hipError_t cs_launch_bed_coupling(hipStream_t st, const cs_pblock& P, float* part, const float* aff,
                                  const int32_t* lookup, uint32_t nbed, float* bed_ext, int32_t* dw);
// bed_dynamics (shader.rs:1991-2238) over the nbed grains: snapshots their
// positions into `snap` (4 floats each), then moves them; reads the final grid
// velocity, the solved pressure and the viscosity pass's alias4 flags.
hipError_t cs_launch_bed_dynamics(hipStream_t st, const cs_pblock& P, uint32_t nbed, float* bed_part, float* bed_aff,
                                  float* bed_ext, int32_t* dw, float* snap, const int32_t* lookup, const float* gvel,
                                  const int32_t* pres, const uint8_t* alias4, const float* sdf);
// The particles of one substep's emission (d->ef.count of them at slots
// d->ef.first..) and their zero affine records; `threads` >= that count.
hipError_t cs_launch_emit(hipStream_t stream, const cs_dparams* d, uint32_t threads, float* part, float* aff);

// Blocks used by the residual reduction (size of the partials buffer / 2).
#define CS_RESIDUAL_BLOCKS 1024u
// Features of the learned warm start ([nc][4]) and, with `target`, its training
// target, loss mask and fluid-cell count.
hipError_t cs_launch_unet_features(hipStream_t stream, const cs_params& u, const float* kf, const float* rhs,
                                   const int32_t* pres, const int32_t* bed_id, const cs_bed_cell* bed, float rhs_scale,
                                   float inv_p_scale, float* f4, float* target, uint8_t* mask, uint32_t* count);
