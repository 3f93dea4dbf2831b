/* orc_internal.h — shared state and numeric helpers of the CPU oracle (test
 * infrastructure; see orc.h). */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

/* ---- constants (reference values) -------------------------------------- */
/* state.rs:13, state.rs:18, state.rs:28 */
#define O_FP_SCALE 262144.0f
#define O_FP_VALUE_LIMIT 8192.0f
#define O_METRICS_DIV_FP_SCALE 1024.0f
#define O_SDF_NO_CONSTRAINT 999.0f
/* shader.rs:174-178 */
#define O_CELL_AIR 0
#define O_CELL_SURFACE 1
#define O_CELL_INTERIOR 2
#define O_CELL_BED 3
#define O_CELL_SOLID 4
/* shader.rs:67-71 */
#define O_M_MAX_ABS_DIV 0
#define O_M_FLUID_CELLS 1
#define O_M_DIV_CLAMP 2
#define O_M_P_CLAMP 3
#define O_M_OVERFLOW 4
/* slots the reference leaves free, used by this build: the projection's
 * outcome, max|div| and max|div - target| over the fluid cells after
 * project_pressure + boundary_project, and max|div - target| before it */
#define O_M_POST_MAX_ABS_DIV 5
#define O_M_POST_MAX_ABS_RES 6
#define O_M_PRE_MAX_ABS_RES 7

/* units.rs:13-27 evaluated in binary32 like Rust's const evaluation. */
float orc_sim_units_per_meter(void);
float orc_max_velocity(void);
float orc_gravity_sim_units(void);

/* ---- numeric helpers ------------------------------------------------------ */
static inline float o_min(float a, float b) { return (a != a) ? b : ((b < a) ? b : a); }
static inline float o_max(float a, float b) { return (a != a) ? b : ((b > a) ? b : a); }
static inline float o_clamp(float x, float lo, float hi) { return o_min(o_max(x, lo), hi); }
static inline float o_mix(float a, float b, float t) { return a * (1.0f - t) + b * t; }
static inline int32_t o_f2i(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return INT32_MAX;
  if (x <= -2147483648.0f) return INT32_MIN;
  return (int32_t)x;
}
static inline uint32_t o_f2u(float x) {
  if (x != x || x <= 0.0f) return 0u;
  if (x >= 4294967296.0f) return 0xffffffffu;
  return (uint32_t)x;
}
static inline int32_t o_wrap_add(int32_t a, int32_t b) {
  return (int32_t)((uint32_t)a + (uint32_t)b);
}
static inline float o_smooth(float lo, float hi, float x) {
  float t = o_clamp((x - lo) / (hi - lo), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}

typedef struct { float x, y, z; } v3;
static inline v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 vdiv(v3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float vlen(v3 a) { return sqrtf(vdot(a, a)); }

/* ---- per-substep parameter block (superset of MpmUniforms, state.rs:31-56) ---- */
typedef struct orc_params {
  uint32_t gx, gy, gz, nc;
  uint32_t num_water, num_bed_particles, max_particles, use_sdf_cache;
  float dt, gravity, dx, inv_dx;
  float origin[3], bmax[3];
  float bulk_K, viscosity, nominal_mass, p_vol;
  float fp_scale, inv_fp_scale, vel_cap, outlet_radius;
  float sdf_res, friction, restitution, contact_offset;
  float water_nu, min_k;
  float total_time, pore_capacity_scale, pore_overfill_alpha;
  float div_clamp, pressure_clamp, metrics_div_fp_scale;
  float j_alpha, j_expand_alpha, max_rest_fraction, surface_void_scale;
  float wall_half;
  cs_geometry geo;
  uint32_t num_bed_records;
} orc_params;

typedef struct orc_inflow {
  float kettle_angle_deg, flow_rate, exit_speed, accumulator;
  uint32_t emission_sequence;
  uint32_t emitted_total; /* fixed-count emitter progress */
} orc_inflow;

struct orc_sim {
  cs_settings s;
  uint32_t gx, gy, gz;
  size_t nc;
  uint32_t sdf_res;
  float* sdf;
  uint8_t* cls;
  int32_t* bed_lookup;
  cs_bed_record* bed;
  uint32_t num_bed;
  /* a bed of particles (reference mode): the reference's leading particles,
   * kept apart; S->bed then holds one record per bed particle */
  uint32_t bed_n;
  float* bed_part; /* 8 per bed particle */
  float* bed_aff;  /* 12 per bed particle */
  float* bed_ext;  /* BedExtract, 8 per bed particle */
  float* bed_ext0; /* the scene's BedExtracts, restored by reset */
  float* bed_part0; /* the scene's bed particles and affine state, restored by reset */
  float* bed_aff0;
  float* bed_snap; /* bed positions before bed_dynamics (3 per bed particle) */
  int32_t* bed_dw; /* bed_delta (shader.rs:451-453): [4][bed_n] lanes water, impulse x, y, z */
  float* particles; /* 8 per particle */
  float* affine;    /* 12 per particle */
  float* init_particles;
  float* init_affine;
  uint32_t init_count;
  int32_t* grid;    /* 6 lanes */
  int32_t* p2g_lanes;
  float* grid_vel;  /* 4 per cell */
  int32_t* prev_p;  /* previous solve, CS_WARM_PREVIOUS */
  int32_t* saved_p; /* pressure and kind lanes at the end of the last solve */
  int32_t* saved_kind;
  float* tgt;       /* classify's target divergence per fluid cell (post-projection metric) */
  float* warm;      /* CS_WARM_EXTERNAL */
  int has_warm;
  uint32_t metrics[8];
  cs_solve_stats stats;
  orc_inflow inflow;
  uint32_t num_water;
  float total_time, total_emitted_mass;
  uint32_t total_dropped;
  orc_params u;
};

/* host pieces shared between orc_host.c and orc_step.c */
void orc_write_uniforms(orc_sim* sim, float dt, orc_params* u);
void orc_emit(orc_sim* sim, float dt);

#endif
