/*
 * orc.h — CPU oracle for the MLS-MPM substep.  TEST INFRASTRUCTURE ONLY.
 *
 * This is a single-threaded C restatement of the reference solver of
 * craigxchen/coffee-sim: the WGSL kernels of crates/sim-wasm/src/mpm_3d/shader.rs
 * (P2G, grid update, boundary projection, classify, red-black Gauss-Seidel,
 * projection, packing, viscosity, G2P), the pass order of MpmSim3D::step_frame
 * (mod.rs:264-430), the uniform values of write_uniforms (mod.rs:569-670), the
 * scene generators of state.rs:206-344, the emitter of inflow.rs:109-293 and,
 * for a bed of particles, bed_lookup_clear / bed_lookup_scatter, bed_coupling
 * and extraction_advect (shader.rs:1885-1987, 2291-2335).
 * Every function cites the reference lines it follows.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * load this library, and only as the checker.  The product (coffee_sim_amd/)
 * never links, loads or calls it.
 *
 * Numeric rules shared with the product (so results can be compared bitwise):
 * IEEE binary32, no FMA contraction, correctly rounded / and sqrt, NaN-avoiding
 * min/max with ties returning the first operand, saturating float->i32 with
 * NaN -> 0, WGSL mix(a,b,t) = a*(1-t) + b*t, exp from Cody-Waite reduction and
 * a degree-7 Taylor polynomial, and the emitter's sin/cos from a Cody-Waite
 * reduction by pi/2 in double and Taylor series, rounded once to float.
 */
#ifndef ORC_H
#define ORC_H
#include "../include/coffee_sim.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

/* presets and generators */
int orc_settings_preset(int preset, cs_settings* out);
void orc_generate_sdf(const cs_settings* s, float* out);
void orc_generate_sdf_class(const cs_settings* s, const float* sdf, uint8_t* out);

/* inflow model, inflow.rs:244-293 */
float orc_effective_head_from_angle(float angle_deg, const cs_spout* sp);
float orc_exit_speed_cap_from_angle(float angle_deg, const cs_spout* sp);
float orc_pour_fraction_from_angle(float angle_deg, const cs_spout* sp);
float orc_exit_speed_from_head(float head, float max_exit_speed);
float orc_flow_rate_from_speed(float exit_speed, float nozzle_radius, float discharge_coeff,
                               float volume_to_ml, float max_flow_rate_ml_s);
void orc_spout_aim_at(cs_spout* sp, float x, float y, float z);
void orc_spout_translate_origin_to(cs_spout* sp, float x, float y, float z);
float orc_kozeny_carman_permeability(float grind_diameter_um, float porosity);

/* math */
float orc_expf(float x);
void orc_sincosf(float x, float* sn, float* cs);
float orc_smoothstep(float lo, float hi, float x);
int32_t orc_f2i(float x);

/* simulation */
orc_sim* orc_create(const cs_settings* s, const cs_scene* scene);
void orc_destroy(orc_sim* sim);
int orc_reset(orc_sim* sim);
int orc_step_frame(orc_sim* sim, float dt);
int orc_set_kettle_angle(orc_sim* sim, float angle_deg);
int orc_set_spout_position(orc_sim* sim, float x, float y, float z);
int orc_set_spout_target(orc_sim* sim, float x, float y, float z);
int orc_set_warm_start(orc_sim* sim, const float* pressure);
int orc_write_particles(orc_sim* sim, uint32_t offset, uint32_t count, const float* p8, const float* a12);
int orc_set_particle_count(orc_sim* sim, uint32_t num_water);
void orc_get_counts(orc_sim* sim, uint32_t* num_water, uint32_t* num_bed, uint32_t* max_particles);
void orc_get_time(orc_sim* sim, float* total_time, float* total_emitted_mass, uint32_t* total_dropped);
void orc_get_inflow(orc_sim* sim, float* angle, float* flow_rate, float* exit_speed);
void orc_read_particles(orc_sim* sim, uint32_t offset, uint32_t count, float* p8);
void orc_read_affine(orc_sim* sim, uint32_t offset, uint32_t count, float* a12);
void orc_read_p2g_lanes(orc_sim* sim, int32_t* lanes6);
void orc_read_grid_vel(orc_sim* sim, float* vel4);
void orc_read_pressure(orc_sim* sim, int32_t* p, int32_t* kind);
/* The bed lookup: per cell the highest stamping bed particle (reference mode),
 * the scene's lookup (field mode) or -1. */
void orc_read_bed_lookup(orc_sim* sim, int32_t* lookup);
/* The bed particles' BedExtracts, 8 floats each (reference mode). */
void orc_read_bed_extracts(orc_sim* sim, float* ext);
/* bed grains (8 floats) and affine rows (12 floats) after bed_dynamics; either may be null */
void orc_read_bed_particles(orc_sim* sim, float* part, float* aff);
void orc_read_metrics(orc_sim* sim, uint32_t* slots8);
/* prepare_render (shader.rs:2242-2276): RenderParticle, 8 floats per particle in use. */
void orc_pack_render(orc_sim* sim, float* render8);
void orc_get_solve_stats(orc_sim* sim, cs_solve_stats* out);
/* relative residual of the current pressure (definition shared with the product) */
double orc_residual(orc_sim* sim);
/* Galerkin factor <p,b>/<p,Ap> of the current pressure over the residual's rows,
 * unrounded (NaN when <p,Ap> <= 0); the solve rounds it as rbgs_rescale_after says. */
double orc_rescale_factor(orc_sim* sim);
/* Overwrite the pressure lane (fixed point, one value per cell), for tests. */
void orc_debug_set_pressure(orc_sim* sim, const int32_t* p);

/* Dense pressure-system pieces after classify, for solver unit tests: per fluid
 * cell the 6 face weights (order -x,+x,-y,+y,-z,+z), their running sum and the
 * right-hand side c = dx*dx*rhs.  Non-fluid cells get zeros. */
void orc_read_system(orc_sim* sim, float* w6, float* count, float* c);
/* Run a single pass of the substep in isolation for unit tests (debug hooks). */
int orc_debug_step_until(orc_sim* sim, float sub_dt, int stop_after_pass);
/* Pass ids for orc_debug_step_until. */
#define ORC_PASS_P2G 1
#define ORC_PASS_GRID 2
#define ORC_PASS_CLASSIFY 3
#define ORC_PASS_RBGS 4
#define ORC_PASS_PROJECT 5
#define ORC_PASS_PACKING 6
#define ORC_PASS_VISCOSITY 7
#define ORC_PASS_G2P 8
#define ORC_PASS_BOUNDARY 9
#define ORC_PASS_POST_DIVERGENCE 10
#define ORC_PASS_BED_STAMP 11    /* bed_lookup_clear + bed_lookup_scatter and the bed cells' records */
#define ORC_PASS_BED_DYNAMICS 12 /* bed_dynamics alone */
#define ORC_PASS_MG 13           /* the multigrid start (rbgs_mg_cycles W-cycles) on the current pressure */
/* One pass on the current state (uniforms of a substep of sub_dt, nothing
 * cleared): ORC_PASS_RBGS runs one red+black pair, PACKING and VISCOSITY their
 * prepare + apply, PROJECT the projection alone. */
int orc_debug_run_pass(orc_sim* sim, float sub_dt, int pass);
/* The pressure a CS_WARM_PREVIOUS solve starts from (fixed point, one value per
 * cell), e.g. a GPU run's last solve, so the oracle continues that run. */
void orc_set_previous_pressure(orc_sim* sim, const int32_t* p);
/* Overwrite / read the working grid state, for the pass tests: grid_vel (v.xyz,
 * mass per cell) and the six i32 lanes as the passes see them (after grid
 * update: pressure, divergence, packing pressure, kind, rest and current volume). */
void orc_debug_set_grid_vel(orc_sim* sim, const float* vel4);
void orc_debug_set_grid_lanes(orc_sim* sim, const int32_t* lanes6);
void orc_debug_read_grid_lanes(orc_sim* sim, int32_t* lanes6);
void orc_debug_clear_metrics(orc_sim* sim);
/* rbgs_mg_cycles of the oracle's own settings (the ORC_PASS_MG pass and later solves) */
void orc_set_mg_cycles(orc_sim* sim, uint32_t cycles);
/* The bed particles' bed_delta lanes, [4][n_bed] i32: water credit, then the
 * reaction impulse x, y, z (fixed point). */
void orc_debug_read_bed_delta(orc_sim* sim, int32_t* delta4);
void orc_debug_set_bed_delta(orc_sim* sim, const int32_t* delta4);
/* The multigrid start's operator on `level` (0: the fine level over the free
 * cells) for the current pressure and kinds: diagonal and +x/+y/+z couplings,
 * one value per cell of that level's grid.  Returns the level count, 0 when
 * `level` does not exist. */
int orc_debug_mg_operator(orc_sim* sim, float sub_dt, int level, float* D, float* Wx, float* Wy, float* Wz);

#ifdef __cplusplus
}
#endif
#endif
