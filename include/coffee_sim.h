/*
 * coffee_sim.h — C ABI of the MI355X-native MLS-MPM substep.
 *
 * Drop-in boundary for the GPU solver of craigxchen/coffee-sim.  The reference
 * owns its solver through `MpmSim3D` (crates/sim-wasm/src/mpm_3d/mod.rs:179-193)
 * and talks to the device through one wgpu bind group
 * (crates/sim-wasm/src/mpm_3d/pipelines.rs:31-141).  Every entry point below
 * replaces one method of `MpmSim3D` (or one buffer write/readback the reference
 * performs through `queue.write_buffer` / `map_async`); the citation on each
 * declaration names it.
 *
 * Conventions
 *  - Every function returns int: CS_OK (0) or a negative CS_ERR_* code.  Nothing
 *    aborts; `cs_last_error()` returns a message for the last failure on the
 *    calling thread.
 *  - Handles own their device memory.  Host pointers are borrowed for the call;
 *    copies are synchronous.
 *  - One `cs_sim` is externally single-threaded (the reference's `&mut self`).
 *  - Byte layouts of particles (32 B), affine (48 B) and the grid lanes are the
 *    reference's (state.rs:90-118; shader.rs:24-33) so a Rust host can
 *    `bytemuck`-cast exactly as inflow.rs:175-217 does today.
 *  - Axes are the reference's: y is up, cell = iz*gx*gy + iy*gx + ix
 *    (shader.rs:148-150).
 */
#ifndef COFFEE_SIM_H
#define COFFEE_SIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CS_ABI_VERSION 1

/* ---- error codes ------------------------------------------------------- */
#define CS_OK 0
#define CS_ERR_INVALID -1  /* bad argument / inconsistent settings          */
#define CS_ERR_DEVICE -2   /* HIP runtime error or no device                 */
#define CS_ERR_NOMEM -3    /* device or host allocation failed               */
#define CS_ERR_RANGE -4    /* offset/count outside the allocated capacity    */
#define CS_ERR_STATE -5    /* call not valid in the current state            */

#define CS_MAX_OBSTACLES 4
#define CS_METRICS_SLOTS 8

/* ---- scene description -------------------------------------------------- */

/* Obstacle kinds, mpm/mod.rs:71-85 (`Obstacle`). */
#define CS_OBSTACLE_NONE 0
#define CS_OBSTACLE_TRUNCATED_CONE 1
#define CS_OBSTACLE_CYLINDER 2

typedef struct cs_obstacle {
  int32_t kind;
  float center[3];
  float top_radius; /* cone */
  float bot_radius; /* cone */
  float radius;     /* cylinder */
  float top_y;
  float bot_y;
} cs_obstacle;

/* Spout / kettle model, mirrors `SpoutSettings` (mpm/inflow.rs:9-25). */
typedef struct cs_spout {
  float origin[3];
  float direction[3];
  float target[3];
  float nozzle_radius;
  float stem_radius;
  float activation_angle_deg;
  float full_flow_angle_deg;
  float head_at_activation;
  float head_at_full_angle;
  float discharge_coeff;
  float volume_to_ml;
  float max_flow_rate_ml_s;
  float gentle_exit_speed;
  float max_exit_speed;
  float stem_length;
} cs_spout;

/* Emitter modes. */
#define CS_EMIT_FLOW_MODEL 0  /* reference: kettle angle -> head -> flow (inflow.rs:109-293) */
#define CS_EMIT_FIXED_COUNT 1 /* fixed particles/substep at a given speed and jet radius    */

/* Pressure solve modes. */
#define CS_RBGS_FIXED_PAIRS 0 /* reference: `pressure_rbgs_pairs` red+black pairs (mod.rs:352-357) */
#define CS_RBGS_TOLERANCE 1   /* iterate until the relative residual <= tolerance            */

/* Warm starts for the pressure solve (the reference always starts at 0, shader.rs:1187). */
#define CS_WARM_ZERO 0
#define CS_WARM_PREVIOUS 1 /* previous substep's converged pressure                        */
#define CS_WARM_EXTERNAL 2 /* pressure supplied by cs_sim_set_warm_start (learned predictor) */
#define CS_WARM_LEARNED 3  /* pressure predicted on the device by an attached cs_unet          */

/* Analytic contact fallbacks that the reference hard-codes in WGSL
 * (shader.rs:744-823, 1778-1779).  Reference mode plugs in its constants;
 * config scenes describe their own cone and disable the cup. */
typedef struct cs_geometry {
  uint32_t cone_enabled;
  float cone_bot_y, cone_top_y;       /* shader.rs:751-752: -3, 3             */
  float cone_bot_radius;              /* dripper_outlet_radius (0.42)         */
  float cone_top_radius;              /* outlet + 4.2634315 (shader.rs:101)   */
  uint32_t cup_enabled;
  float cup_top_y;                    /* -3.5 (shader.rs:805)                  */
  float cup_radius;                   /* 3.0  (shader.rs:806)                  */
  float cup_bottom_y;                 /* -8.0 (shader.rs:107)                  */
  /* The paper filter that holds a bed of particles (resolve_scene_obstacles with
   * is_bed, shader.rs:765-801; FilterConfig::default, filter.rs:14-26): bed
   * grains stay inside the paper cone and rest no lower than its apex floor
   * (coffee_filter_floor_y, shader.rs:136-146).  Water passes through. */
  uint32_t filter_enabled;
  float filter_center_y;              /* -0.35                                  */
  float filter_top_y;                 /* 2.75 above the centre                  */
  float filter_bot_y;                 /* -3.02 below the centre                 */
  float filter_top_radius;            /* 4.10                                   */
  float filter_thickness;             /* 0.08                                   */
} cs_geometry;

/* Superset of `MpmSettings` (mpm/mod.rs:87-104).  Fields after `initial_kettle_angle_deg`
 * are extensions; cs_settings_reference() fills them with the reference's values. */
typedef struct cs_settings {
  float bounds_size[3];
  uint32_t grid_dims[3];
  uint32_t max_particles;
  uint32_t substeps;
  float gravity;
  float bulk_modulus;
  float viscosity;
  float render_radius;
  uint32_t pressure_rbgs_pairs;
  uint32_t use_sdf_cache;
  uint32_t num_obstacles;
  cs_obstacle obstacles[CS_MAX_OBSTACLES];
  cs_spout spout;
  float initial_kettle_angle_deg;
  /* ---- extensions ---- */
  uint32_t sdf_res;            /* 128 (state.rs:20)                               */
  float particle_vol_factor;   /* p_vol = dx^3 * factor; 0.25 (mod.rs:576)         */
  float wall_thickness;        /* OBSTACLE_WALL_THICKNESS 0.4 (mod.rs:33)          */
  float water_viscosity_m2_s;  /* Darcy nu, 1e-6 (brew_config.rs:125)              */
  float min_bed_permeability;  /* 1e-12 (brew_config.rs:126)                       */
  cs_geometry geometry;
  uint32_t emitter_mode;       /* CS_EMIT_*                                        */
  uint32_t emit_per_substep;   /* fixed-count emitter                              */
  uint32_t emit_total;         /* stop after this many particles (0 = unlimited)   */
  float emit_speed;            /* scene units / s                                  */
  float emit_jet_radius;       /* scene units                                      */
  uint32_t rbgs_mode;          /* CS_RBGS_*                                        */
  float rbgs_tolerance;        /* relative residual target (1e-4)                  */
  uint32_t rbgs_check_every;   /* iterations between residual checks (8)           */
  uint32_t rbgs_max_iters;     /* iteration cap (20000)                            */
  uint32_t warm_start;         /* CS_WARM_*                                        */
  uint32_t has_bed_field;      /* a per-cell porous bed is supplied with the scene */
  float warm_fallback_residual; /* an external or learned start whose relative residual
                                 * exceeds this restarts from zero (1: no worse than the
                                 * zero start's residual)                          */
  uint32_t rbgs_rescale_after;  /* residual mode, brick solver: after this many iterations
                                 * the pressure is rescaled once, p <- alpha p with the
                                 * Galerkin factor alpha = <b,p>/<p,Ap> (rounded to a
                                 * multiple of 2^-20), then the solve goes on.  0 = off
                                 * (the reference's plain iteration); otherwise a
                                 * multiple of 4 and of rbgs_check_every, with at least
                                 * rbgs_check_every iterations left under rbgs_max_iters.
                                 * The iterations before the rescale count toward
                                 * cs_solve_stats.iterations.  No reference counterpart. */
  uint32_t rbgs_mg_cycles;      /* residual mode, brick solver: before the iteration, this
                                 * many multigrid W-cycles correct the start (the cells at
                                 * the pressure clamp held fixed; 2x2x2 Galerkin coarsening
                                 * down to 8^3; 2 pre- and post-smoothing red-black passes
                                 * per level; corrections scaled by 1.5; 10 passes on the
                                 * coarsest grid).  The iteration then runs to the tolerance
                                 * as before and counts only its own passes.  0 = off.
                                 * Exclusive with rbgs_rescale_after.  No reference
                                 * counterpart. */
  uint32_t emit_sample_period;  /* 0: the golden-angle disk's sample index is the running
                                 * emission count (inflow.rs:195-199).  Otherwise each
                                 * substep's samples start from that count modulo this period:
                                 * the index is converted to f32 and multiplied by the golden
                                 * angle, and past a few million samples the products lose the
                                 * bits that spread the disk, so emitted particles coincide.
                                 * No reference counterpart. */
} cs_settings;

/* Per-cell porous bed ("field mode").  The reference stamps bed-particle ids into
 * `bed_lookup` and reads porosity / permeability / compaction / saturation from
 * `bed_extract` and the stamping particle's height (shader.rs:419-666, 2291-2335);
 * a field bed gives every bed cell its own record with the same meaning. */
typedef struct cs_bed_record {
  float porosity;     /* be.bed.y                                   */
  float permeability; /* be.bed.z (m^2)                             */
  float compaction;   /* be.bed.w                                   */
  float saturation;   /* be.extract.w                               */
  float sample_y;     /* height of the bed sample (particles[id].pos.y) */
  float velocity[3];  /* bed matrix velocity (particles[id].vel)    */
} cs_bed_record;

/* Host-side scene arrays handed to cs_sim_create.  Any pointer may be NULL:
 * the SDF and class mask are then generated from the obstacles exactly as
 * MpmBuffers::new does (state.rs:178-182). */
typedef struct cs_scene {
  const float* sdf;          /* sdf_res^3, x fastest                         */
  const uint8_t* sdf_class;  /* gx*gy*gz, 1 = solid                           */
  const int32_t* bed_lookup; /* gx*gy*gz, -1 or index into bed_records        */
  const cs_bed_record* bed_records;
  uint32_t num_bed_records;
  const float* particles;    /* 8 floats per particle: pos.xyz, J, vel.xyz, m  */
  const float* affine;       /* 12 floats per particle: C columns, .w phase    */
  uint32_t num_particles;    /* initial water particles                       */
  /* A bed of particles ("reference mode", exclusive with has_bed_field): the
   * reference's particle buffer holds these first, then the water (mod.rs:239-249).
   * Here they sit in their own arrays; p2g and g2p skip them either way
   * (shader.rs:967, 1693).  Every substep clears and re-stamps the cell lookup
   * from them (bed_lookup_clear / bed_lookup_scatter, shader.rs:2291-2335) and
   * rebuilds the bed cells' records from their BedExtracts.  After G2P,
   * bed_coupling and extraction_advect move water into the bed's extracts and
   * bed_dynamics moves the grains (mod.rs:391-409); the water's reaction on
   * them is gathered in project_pressure and G2P (shader.rs:458-522).
   * cs_sim_reset restores the particles, affine rows and extracts given here. */
  const float* bed_particles; /* 8 floats per bed particle, as `particles`     */
  const float* bed_affine;    /* 12 floats per bed particle, .w of col0 = phase */
  const float* bed_extracts;  /* 8 floats per bed particle: BedExtract bed.xyzw
                               * (pore water, porosity, permeability, compaction)
                               * and extract.xyzw (extractable, dissolved, temp,
                               * saturation), shader.rs:35-38                   */
  uint32_t num_bed_particles;
} cs_scene;

/* Solver statistics of the most recent substep. */
typedef struct cs_solve_stats {
  uint32_t iterations;        /* red+black pairs executed                    */
  uint32_t fluid_cells;       /* cells in the pressure system                */
  double initial_residual;    /* before the first sweep                      */
  double final_residual;      /* at exit (NaN when never measured)           */
  uint32_t converged;         /* 1 when final_residual <= tolerance          */
  uint32_t fell_back;         /* external warm start rejected (residual grew) */
} cs_solve_stats;

/* ---- settings presets (host only, no device needed) ---------------------- */

/* MpmSettings::default_v60 / benchmark_free_stream / benchmark_center_pour
 * (mpm/mod.rs:107-161) plus reference values for every extension field. */
#define CS_PRESET_DEFAULT_V60 0
#define CS_PRESET_FREE_STREAM 1
#define CS_PRESET_CENTER_POUR 2
int cs_settings_preset(int preset, cs_settings* out);

/* FilterConfig (filter.rs:3-26): the paper filter's frustum, y relative to center.y. */
typedef struct cs_filter_config {
  float center[3];
  float top_y, bot_y, top_radius, bot_radius, thickness, hole_radius;
} cs_filter_config;
/* BedConfig (bed.rs:7-38): the granular bed's frustum, y relative to center.y. */
typedef struct cs_bed_config {
  float center[3];
  float top_y, bot_y, top_radius, bot_radius;
  uint32_t num_particles;      /* target count (DEFAULT_BREW.bed_particle_samples) */
  float initial_porosity;
  float initial_permeability;  /* m^2, Kozeny-Carman of the default grind */
  float extractable_mass;
} cs_bed_config;
/* FilterConfig::default (filter.rs:14-26) and BedConfig::default (bed.rs:20-38). */
void cs_filter_config_default(cs_filter_config* out);
void cs_bed_config_default(cs_bed_config* out);
/* BedConfig::seated_in_filter (bed.rs:41-81), the bed of MpmSettings::default_v60. */
void cs_bed_config_seated_in_filter(const cs_filter_config* filter, cs_bed_config* out);
/* init_bed_particles (bed.rs:107-206) with build_cell_lookup (bed.rs:208-288).
 * *count receives the bed particle count (at most config->num_particles); the
 * arrays, when not NULL, receive 8 (particle), 12 (affine, phase 1) and 8
 * (BedExtract) floats per particle, up to `capacity` particles, and the initial
 * gx*gy*gz cell lookup.  Returns CS_ERR_RANGE when *count exceeds a non-zero
 * `capacity` with arrays given.  Host only. */
int cs_bed_init(const cs_bed_config* config, const uint32_t grid_dims[3], const float bounds_size[3],
                uint32_t capacity, float* particles, float* affine, float* extracts, int32_t* cell_lookup,
                uint32_t* count);

/* Seeded synthetic workloads (no reference counterpart; the reference has no RNG).
 *   CS_WORKLOAD_V60_SMALL: 64^3, dx 2 mm, 32,768-particle jittered cylinder over a flat bed.
 *   CS_WORKLOAD_V60_WORKING: 128^3, dx 1 mm, 1,048,576 particles (70% pre-wetted slab,
 *   30% poured at 2,000 particles/substep), value-noise porosity, Kozeny-Carman permeability.
 *   CS_WORKLOAD_V60_LARGE: 256^3, dx 0.5 mm, 8,388,608 particles (70% pre-wetted slab, 30% poured
 *   at 16,000 particles/substep), porosity noise at a 16-cell wavelength and +-0.08 amplitude,
 *   2.5e-4 s substeps.
 * `scale` (>= 1) divides the grid side and particle counts for quick tests (1 = full size).
 * The arrays are allocated by the library; release them with cs_workload_free. */
#define CS_WORKLOAD_V60_SMALL 0
#define CS_WORKLOAD_V60_WORKING 1
#define CS_WORKLOAD_V60_LARGE 2
typedef struct cs_workload {
  cs_settings settings;
  cs_scene scene;
  float frame_dt;          /* seconds per cs_sim_step_frame call               */
  uint32_t substeps_total; /* run length of the workload                       */
  uint32_t total_particles;/* initial + emitted particles                      */
  void* _owner;            /* library-owned storage                            */
} cs_workload;
int cs_workload_build(int workload, uint64_t seed, uint32_t scale, cs_workload* out);
/* A pour scene of the working kind with its own size (no reference counterpart):
 * a cubic grid of `side` cells over the same 128 mm domain (side a multiple of 16,
 * 32..256), cone apex side/8 and bed 3*side/16 cells up, porosity noise as in
 * V60_WORKING (8-cell lattice, +-0.05), 70% of `particles` as the pre-wetted slab
 * and the rest poured `pour_per_substep` at a time from a spout tilted
 * `kettle_angle_deg` below horizontal; RBGS to 1e-4 every 8 iterations. */
typedef struct cs_pour_params {
  uint32_t side;
  uint32_t particles;
  uint32_t pour_per_substep;
  float kettle_angle_deg;
  uint64_t seed;
} cs_pour_params;
int cs_workload_build_pour(const cs_pour_params* p, cs_workload* out);
int cs_workload_free(cs_workload* w);

/* Scene generators exposed for verification (state.rs:250-344). */
int cs_generate_sdf(const cs_settings* s, float* out_sdf /* sdf_res^3 */);
int cs_generate_sdf_class(const cs_settings* s, const float* sdf, uint8_t* out_class /* gx*gy*gz */);

/* ---- context ------------------------------------------------------------- */
typedef struct cs_ctx cs_ctx;
typedef struct cs_sim cs_sim;

/* Device + stream.  Replaces the wgpu Device/Queue pair that MpmSim3D borrows
 * per call (mod.rs:196, mod.rs:264). */
int cs_ctx_create(int device, cs_ctx** out);
int cs_ctx_destroy(cs_ctx* ctx);
const char* cs_last_error(void);
int cs_device_count(int* out);

/* ---- simulation ---------------------------------------------------------- */

/* MpmSim3D::new (mod.rs:196-220): buffers (state.rs:85-201), SDF + class
 * textures, bed lookup, initial particles. */
int cs_sim_create(cs_ctx* ctx, const cs_settings* s, const cs_scene* scene, cs_sim** out);
int cs_sim_destroy(cs_sim* sim);

/* MpmSim3D::reset (mod.rs:432-446): counters, inflow state, initial particles. */
int cs_sim_reset(cs_sim* sim);

/* MpmSim3D::step_frame (mod.rs:264-430): dt clamped to 1/30, split into
 * `substeps`, each substep = emit -> uniforms -> clear -> P2G -> grid update ->
 * boundary -> classify -> pressure -> project -> packing -> viscosity -> G2P.
 * The reference records a substep's passes once and submits them as one
 * command buffer (mod.rs:305-416) after writing the substep's uniforms
 * (mod.rs:569-670).  Here, when the settings allow it (brick solver with the
 * persistent solve, zero or previous-step start, no kernel timing), the whole
 * substep is recorded once per simulation as a HIP graph whose kernels read a
 * device-resident parameter block; each substep is one parameter copy and one
 * graph launch, and the host waits once per frame (per 64 substeps).  A
 * substep whose active bricks exceed the recorded launch widths stops at its
 * brick scan and is finished launch by launch, with the same results. */
int cs_sim_step_frame(cs_sim* sim, float dt);

/* Substeps run as graph replays to the end, and replays that stopped at the
 * brick scan and were finished launch by launch, since the handle was created
 * (the recorded form of step_frame's single submit, mod.rs:305-416). */
int cs_sim_get_recorded_counts(cs_sim* sim, uint64_t* replayed, uint64_t* halted);

/* Blocks until every pass enqueued by this handle has finished (the reference's
 * device.poll(wait_indefinitely), physics_tests.rs:103). */
int cs_sim_synchronize(cs_sim* sim);

/* Setters/getters, mod.rs:448-538. */
int cs_sim_set_kettle_angle(cs_sim* sim, float angle_deg);
int cs_sim_set_spout_position(cs_sim* sim, float x, float y, float z);
int cs_sim_set_spout_target(cs_sim* sim, float x, float y, float z);
int cs_sim_get_counts(cs_sim* sim, uint32_t* num_water, uint32_t* num_bed, uint32_t* max_particles);
int cs_sim_get_time(cs_sim* sim, float* total_time, float* total_emitted_mass, uint32_t* total_dropped);
int cs_sim_get_inflow(cs_sim* sim, float* kettle_angle, float* flow_rate_ml_s, float* exit_speed);

/* Particle append contract (inflow.rs:221-235): writes `count` particles at
 * slot `offset` (particle8 = 8 floats, affine12 = 12 floats per particle). */
int cs_sim_write_particles(cs_sim* sim, uint32_t offset, uint32_t count,
                           const float* particle8, const float* affine12);
int cs_sim_set_particle_count(cs_sim* sim, uint32_t num_water);

/* Readbacks (the reference's staging copy + map_async, physics_tests.rs:61-139).
 * Sizes are in bytes and must match the buffer extent requested. */
int cs_sim_read_particles(cs_sim* sim, uint32_t offset, uint32_t count, float* particle8);
int cs_sim_read_affine(cs_sim* sim, uint32_t offset, uint32_t count, float* affine12);
/* The 6 i32 lanes (mass, mom xyz, rest vol, current vol) as accumulated by the
 * most recent P2G, lane-major [6][cells] (state.rs:106-111). */
int cs_sim_read_p2g_lanes(cs_sim* sim, int32_t* lanes6);
/* grid_vel (v.xyz, mass) after the last pass of the most recent substep. */
int cs_sim_read_grid_vel(cs_sim* sim, float* vel4);
/* Pressure (i32 fixed point, 2^-18) and cell kind (0 air .. 4 solid) of the most recent solve. */
int cs_sim_read_pressure(cs_sim* sim, int32_t* pressure_fp, int32_t* kind);
/* The scene's SDF (sdf_res^3 floats, x fastest) and solid-class mask (one byte per
 * cell) as the simulation holds them: the scene's arrays, or the ones generated
 * on the device from the settings when the scene gave none (generate_sdf_data
 * and generate_sdf_class_data, state.rs:206-344).  Either pointer may be NULL. */
int cs_sim_read_sdf(cs_sim* sim, float* sdf, uint8_t* sdf_class);
/* The bed lookup of the most recent substep (gx*gy*gz: the highest stamping bed
 * particle id, or -1), as bed_lookup_scatter leaves it (shader.rs:2304-2335).
 * A simulation without bed particles reads all -1. */
int cs_sim_read_bed_lookup(cs_sim* sim, int32_t* lookup);
/* The bed particles' BedExtracts (8 floats each) as bed_coupling and
 * extraction_advect (shader.rs:1885-1987) leave them after the most recent
 * substep: pore water and saturation grow as the bed absorbs water, extractable
 * mass moves to dissolved.  Nothing is written for a simulation without bed
 * particles. */
int cs_sim_read_bed_extracts(cs_sim* sim, float* extracts8);
/* The bed particles (8 floats each, as cs_scene.bed_particles) and their
 * affine rows (12 floats each: .w of column 0 is the phase, 1 anchored or 2
 * suspended; column 1 xyz is the anchored rest position) as bed_dynamics
 * (shader.rs:1991-2238) leaves them after the most recent substep.  Either
 * pointer may be null; nothing is written for a simulation without bed
 * particles. */
int cs_sim_read_bed_particles(cs_sim* sim, float* particles8, float* affine12);
/* Metric slots (state.rs:22-28). */
int cs_sim_read_metrics(cs_sim* sim, uint32_t* slots8);
/* prepare_render (shader.rs:2242-2276, run by step_frame at mod.rs:410-412 and
 * exposed through render_buffer, mod.rs:500-502): the RenderParticle layout
 * (shader.rs:40-43), 8 floats per particle, (pos.xyz, colour) and (radius, 0,
 * 0, 0), for every particle slot in use.  Packed on demand from the current
 * particle state; `render8` holds particle_count x 8 floats, a bed of
 * particles first (as the reference's buffer holds it), then the water. */
int cs_sim_pack_render(cs_sim* sim, float* render8);
/* cs_sim_pack_render with no device allocation: no counterpart in the reference;
 * for callers that cannot allocate (code under stream capture, tight frame
 * loops, hosts that forbid allocation).  The vertices are staged in a fixed
 * device array inside the library (at most 262144 particles, bed plus water,
 * 8 MiB per GPU, held until the process ends; CS_ERR_RANGE beyond that).  The
 * array is shared by every call on a GPU: calls that overlap on one GPU race. */
int cs_sim_pack_render_noalloc(cs_sim* sim, float* render8);
int cs_sim_get_solve_stats(cs_sim* sim, cs_solve_stats* out);

/* Warm start for the next pressure solve (CS_WARM_EXTERNAL): float pressure per cell. */
int cs_sim_set_warm_start(cs_sim* sim, const float* pressure);

/* Checkpoint / restore of the persistent state (particles, affine, inflow, counters). */
int cs_sim_save(cs_sim* sim, void* buffer, size_t capacity, size_t* written);
int cs_sim_load(cs_sim* sim, const void* buffer, size_t size);

/* Device-side timing of the passes of the most recent substep, in microseconds,
 * measured with HIP events on the solver's stream: [p2g, grid, classify,
 * pressure, project+packing+viscosity, g2p, total]. */
int cs_sim_get_pass_times(cs_sim* sim, float* us7);
int cs_sim_enable_timing(cs_sim* sim, int enable);

/* Device time of the three hot kernels summed since timing was enabled,
 * measured with HIP events on the solver's stream around their launches. */
typedef struct cs_kernel_timing {
  double p2g_us;               /* P2G launches                                  */
  uint64_t p2g_launches;
  uint64_t p2g_particles;      /* particles scattered                           */
  double g2p_us;               /* G2P launches                                  */
  uint64_t g2p_launches;
  uint64_t g2p_particles;
  double rbgs_us;              /* red-black sweep launches (residual checks excluded) */
  uint64_t rbgs_launches;
  uint64_t rbgs_iterations;    /* red+black pairs                               */
  uint64_t rbgs_cell_updates;  /* fluid-cell updates (iterations x fluid cells)  */
  uint64_t rbgs_cells_swept;   /* cells the sweep launches visited              */
  double mg_us;                /* multigrid starts (rbgs_mg_cycles), whole     */
  uint64_t mg_starts;          /* solves that ran the multigrid start section   */
  uint64_t mg_cycles;          /* W-cycles those starts ran (the rest returned
                                * at once once the start met the tolerance)     */
} cs_kernel_timing;
int cs_sim_get_kernel_timing(cs_sim* sim, cs_kernel_timing* out);
int cs_sim_reset_kernel_timing(cs_sim* sim);

/* Verification entry point (no reference counterpart): evaluates the solver's
 * device exp (used by the J update, shader.rs:1853) on `n` inputs. */
int cs_device_expf(cs_ctx* ctx, const float* in, float* out, uint32_t n);
/* Verification entry point (no reference counterpart): the emitter's sin and cos
 * (theta.sin() / theta.cos() of inflow.rs:205-207, evaluated as described at
 * cs_sincosf in the kernels) on `n` inputs. */
int cs_device_sincosf(cs_ctx* ctx, const float* in, float* sn, float* cs, uint32_t n);

/* ---- learned warm start (no reference counterpart) ------------------------------
 * A 3-D U-Net (3 levels, 16/32/64 channels, 3x3x3 kernels, bf16 weights and
 * activations, fp32 accumulation) that predicts the converged pressure of a solve
 * from four features per cell: rhs * rhs_scale, liquid fill (0 when not fluid),
 * solid mask, bed porosity (0 outside the bed).  It replaces the zero pressure
 * the reference starts every solve from (shader.rs:1187); RBGS then runs to the
 * tolerance from the prediction and falls back to a zero start when the
 * prediction's relative residual exceeds the settings' warm_fallback_residual,
 * so the plain solve still decides the result.
 * The grid side must be a multiple of 32.  Parameters are one fp32 vector:
 * for each convolution e1a, e1b, e2a, e2b, ba, bb, d2a, d2b, d1a, d1b the
 * weights [27][cin][cout] (tap = (dz+1)*9 + (dy+1)*3 + (dx+1)) then the bias
 * [cout]; last the head's 16 weights and bias.  Host arrays: features
 * [side^3][4], predictions and targets [side^3] in units of pressure_scale,
 * masks [side^3] (1 = counted in the loss).
 * Training pairs live on the device: cs_unet_capture appends the last solve
 * of a simulation (its features and converged pressure; its mask counts the
 * fluid cells whose pressure is not held at the clamp). */
typedef struct cs_unet cs_unet;
int cs_unet_create(cs_ctx* ctx, uint32_t side, uint64_t seed, cs_unet** out);
int cs_unet_destroy(cs_unet* net);
int cs_unet_param_count(const cs_unet* net, uint32_t* count);
int cs_unet_get_params(cs_unet* net, float* out);
int cs_unet_set_params(cs_unet* net, const float* in);
int cs_unet_set_scales(cs_unet* net, float rhs_scale, float pressure_scale);
int cs_unet_get_scales(const cs_unet* net, float* rhs_scale, float* pressure_scale);
int cs_unet_forward(cs_unet* net, const float* features4, float* pred);
/* Loss (masked mean squared error against `target`) and its gradient for every parameter. */
int cs_unet_gradient(cs_unet* net, const float* features4, const float* target, const uint8_t* mask, float* loss,
                     float* grad);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on one sample. */
int cs_unet_train_step(cs_unet* net, const float* features4, const float* target, const uint8_t* mask, float lr,
                       float* loss);
int cs_unet_capture(cs_unet* net, cs_sim* sim);
int cs_unet_dataset_size(const cs_unet* net, uint32_t* count);
int cs_unet_dataset_clear(cs_unet* net);
int cs_unet_read_sample(cs_unet* net, uint32_t index, float* features4, float* target, uint8_t* mask);
/* One pass over the captured samples in a seeded random order, one Adam step each. */
int cs_unet_train_epoch(cs_unet* net, float lr, uint64_t seed, float* mean_loss);
/* Device time of the network on its last input (cs_unet_forward / _gradient /
 * _train_step), measured with HIP events on the context's stream over `reps`
 * repetitions: one forward pass, and one forward + backward pass (loss against
 * the last target and mask; the parameters are not changed). */
int cs_unet_benchmark(cs_unet* net, uint32_t reps, double* forward_us, double* gradient_us);
/* The network a simulation uses when its settings ask for CS_WARM_LEARNED (NULL detaches). */
int cs_sim_set_warm_net(cs_sim* sim, cs_unet* net);
/* The warm-start pressure (float per cell) the last CS_WARM_EXTERNAL / CS_WARM_LEARNED
 * solve started from, before it was injected into the fluid cells. */
int cs_sim_read_warm_start(cs_sim* sim, float* pressure);

/* ---- batched scenes (no reference counterpart) -----------------------------------
 * Advances many independent simulations together, one launch per pass over
 * every scene: each workgroup of a batched launch runs one block of one scene's
 * own launch.  The pressure solves share one launch per four half sweeps over
 * the active bricks of the scenes still solving, and every scene stops at its
 * own tolerance.  Each scene's state after a batched step equals, bit for bit,
 * the state of stepping it alone with the residual test as a pass of its own
 * (CS_RBGS_FUSE=0).  Requirements: one context; the brick solver; the same
 * rbgs_mode, substeps, and rbgs_check_every and rbgs_max_iters (tolerance mode,
 * both even) or pressure_rbgs_pairs (fixed mode, even); warm_start ZERO or
 * PREVIOUS; at most 1024 scenes and 65535 bricks per scene.  A simulation in a
 * batch is not stepped on its own until the batch is destroyed. */
typedef struct cs_batch cs_batch;
typedef struct cs_batch_config {
  uint32_t sweep_groups;     /* workgroups of the persistent sweep launch (0: one per compute unit) */
  uint32_t residual_groups;  /* workgroups of the persistent residual launch (0: 4 per compute unit) */
  uint32_t check_batches;    /* residual tests enqueued before the host reads a verdict (0: 4)     */
  uint32_t use_graphs;       /* replay the fixed launch sequences as hipGraphs                      */
} cs_batch_config;
int cs_batch_default_config(cs_batch_config* out);
int cs_batch_create(cs_ctx* ctx, cs_sim* const* sims, uint32_t count, const cs_batch_config* config,
                    cs_batch** out);
int cs_batch_destroy(cs_batch* batch);
/* P2G layout of one scene: 1 = tile-binned (default), 0 = one thread per particle. */
int cs_batch_set_scene_p2g(cs_batch* batch, uint32_t scene, int tiled);
int cs_batch_step_frame(cs_batch* batch, float dt);
typedef struct cs_batch_stats {
  uint64_t substeps;           /* batched substeps                                */
  uint64_t particle_substeps;  /* sum over substeps of the scenes' particle counts */
  uint64_t sweep_launches;     /* batched sweep launches (four half sweeps each)   */
  uint64_t residual_checks;    /* batched residual tests                           */
  uint64_t max_iterations;     /* largest per-scene iteration count of a substep, summed */
  double substep_us;           /* host wall time inside cs_batch_step_frame       */
  double solve_us;             /* device time of the solve loops (sweeps and residual tests),
                                  HIP events on the batch's stream              */
  uint64_t cell_updates;       /* sum over scenes and substeps of iterations x fluid cells */
} cs_batch_stats;
int cs_batch_get_stats(cs_batch* batch, cs_batch_stats* out);
int cs_batch_reset_stats(cs_batch* batch);

#ifdef __cplusplus
}
#endif
#endif /* COFFEE_SIM_H */
