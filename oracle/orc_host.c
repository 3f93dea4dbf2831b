/* orc_host.c — CPU oracle, host half: unit constants, presets, SDF and class
 * generation, the inflow emitter and the uniform block.  TEST INFRASTRUCTURE
 * ONLY (see orc.h). */
#include <stdio.h>

#include "orc_internal.h"

/* ---- units.rs:10-34 (binary32 constant evaluation) ------------------------ */
float orc_sim_units_per_meter(void) {
  volatile float top = 2.75f, bot = -3.02f, h = 0.10f;
  float height_units = top - bot;      /* units.rs:13 */
  return height_units / h;             /* units.rs:17-18 */
}
float orc_max_velocity(void) { return 1.5f * orc_sim_units_per_meter(); } /* units.rs:26-27 */
float orc_gravity_sim_units(void) {
  return -9.80665f * orc_sim_units_per_meter(); /* units.rs:23-24 */
}
static float meters_per_sim_unit(void) { return 1.0f / orc_sim_units_per_meter(); }
static float ml_per_sim_unit_cubed(void) {
  float m = meters_per_sim_unit();
  return m * m * m * 1000000.0f; /* units.rs:20-21 */
}

/* brew_config.rs:160-164 */
float orc_kozeny_carman_permeability(float grind_diameter_um, float porosity) {
  float d_m = grind_diameter_um * 1.0e-6f;
  float solid = 1.0f - porosity;
  return d_m * d_m * porosity * porosity * porosity / (180.0f * solid * solid);
}

/* ---- sim-core Vec3 (crates/sim-core/src/math.rs:21-92) --------------------- */
static float vlen_rs(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }
static void vnormalize_rs(float* v) {
  float len = vlen_rs(v[0], v[1], v[2]);
  if (len <= 1.0e-6f) { v[0] = v[1] = v[2] = 0.0f; return; }      /* math.rs:41-48 */
  if (fabsf(len) <= 1.0e-6f) { v[0] = v[1] = v[2] = 0.0f; return; } /* math.rs:83-92 */
  v[0] /= len; v[1] /= len; v[2] /= len;
}

/* SpoutSettings::aim_at / translate_origin_to (inflow.rs:50-66) */
void orc_spout_aim_at(cs_spout* sp, float x, float y, float z) {
  sp->target[0] = x; sp->target[1] = y; sp->target[2] = z;
  float d[3] = {x - sp->origin[0], y - sp->origin[1], z - sp->origin[2]};
  if (vlen_rs(d[0], d[1], d[2]) > 1e-5f) {
    vnormalize_rs(d);
    sp->direction[0] = d[0]; sp->direction[1] = d[1]; sp->direction[2] = d[2];
  }
}
void orc_spout_translate_origin_to(cs_spout* sp, float x, float y, float z) {
  float d[3] = {x - sp->origin[0], y - sp->origin[1], z - sp->origin[2]};
  sp->origin[0] = x; sp->origin[1] = y; sp->origin[2] = z;
  sp->target[0] = sp->target[0] + d[0];
  sp->target[1] = sp->target[1] + d[1];
  sp->target[2] = sp->target[2] + d[2];
}

/* SpoutSettings::default (inflow.rs:27-47) */
static void spout_default(cs_spout* sp) {
  memset(sp, 0, sizeof(*sp));
  sp->origin[0] = -3.4f; sp->origin[1] = 7.3f; sp->origin[2] = 0.9f;
  float d[3] = {0.36f, -0.92f, -0.12f};
  vnormalize_rs(d);
  memcpy(sp->direction, d, sizeof(d));
  sp->target[0] = -0.7f; sp->target[1] = 0.4f; sp->target[2] = 0.0f;
  sp->nozzle_radius = 0.18f;
  sp->stem_radius = 0.24f;
  sp->activation_angle_deg = 8.0f;
  sp->full_flow_angle_deg = 56.0f;
  sp->head_at_activation = 0.04f;
  sp->head_at_full_angle = 24.0f;
  sp->discharge_coeff = 0.92f;
  sp->volume_to_ml = ml_per_sim_unit_cubed();
  sp->max_flow_rate_ml_s = 12.0f;
  sp->gentle_exit_speed = 0.12f * orc_sim_units_per_meter();
  sp->max_exit_speed = 0.45f * orc_sim_units_per_meter();
  sp->stem_length = 1.9f;
}

/* MpmSettings presets (mod.rs:107-177) with reference values for the
 * extension fields. */
int orc_settings_preset(int preset, cs_settings* s) {
  if (!s || preset < 0 || preset > 2) return CS_ERR_INVALID;
  memset(s, 0, sizeof(*s));
  s->bounds_size[0] = 14.0f; s->bounds_size[1] = 20.0f; s->bounds_size[2] = 14.0f;
  uint32_t gx = 80u;
  float dx = s->bounds_size[0] / (float)gx;
  uint32_t gy = (uint32_t)ceilf(s->bounds_size[1] / dx);
  s->grid_dims[0] = gx; s->grid_dims[1] = gy; s->grid_dims[2] = 80u;
  s->max_particles = 220000u;
  s->substeps = 10u;
  s->gravity = orc_gravity_sim_units();
  s->bulk_modulus = 900.0f;
  s->viscosity = 1.2f;
  s->render_radius = dx * 0.7f;
  s->pressure_rbgs_pairs = 40u;
  s->use_sdf_cache = 1u;
  /* v60_support_cone (mod.rs:163-177) against FilterConfig::default (filter.rs:211-223) */
  float f_top = 2.75f, f_bot = -3.02f, f_top_r = 4.10f, f_bot_r = 0.0f;
  float filter_height = o_max(f_top - f_bot, 1e-6f);
  float filter_slope = (f_top_r - f_bot_r) / filter_height;
  float cone_top = 3.0f, cone_bot = -3.0f, outlet = 0.42f;
  s->num_obstacles = 2;
  s->obstacles[0].kind = CS_OBSTACLE_TRUNCATED_CONE;
  s->obstacles[0].top_radius = outlet + filter_slope * (cone_top - cone_bot);
  s->obstacles[0].bot_radius = outlet;
  s->obstacles[0].top_y = cone_top;
  s->obstacles[0].bot_y = cone_bot;
  s->obstacles[1].kind = CS_OBSTACLE_CYLINDER;
  s->obstacles[1].radius = 3.0f;
  s->obstacles[1].top_y = -3.5f;
  s->obstacles[1].bot_y = -8.0f;
  spout_default(&s->spout);
  s->initial_kettle_angle_deg = 9.0f;
  /* extensions */
  s->sdf_res = 128u;
  s->particle_vol_factor = 0.25f;
  s->wall_thickness = 0.4f;
  s->water_viscosity_m2_s = 1.0e-6f;
  s->min_bed_permeability = 1.0e-12f;
  s->geometry.cone_enabled = 1u;
  s->geometry.cone_bot_y = -3.0f;
  s->geometry.cone_top_y = 3.0f;
  s->geometry.cone_bot_radius = outlet;
  s->geometry.cone_top_radius = outlet + 4.2634315f;
  s->geometry.cup_enabled = 1u;
  s->geometry.cup_top_y = -3.5f;
  s->geometry.cup_radius = 3.0f;
  s->geometry.cup_bottom_y = -8.0f;
  s->geometry.filter_enabled = 1u;
  s->geometry.filter_center_y = -0.35f;
  s->geometry.filter_top_y = 2.75f;
  s->geometry.filter_bot_y = -3.02f;
  s->geometry.filter_top_radius = 4.10f;
  s->geometry.filter_thickness = 0.08f;
  s->emitter_mode = CS_EMIT_FLOW_MODEL;
  s->rbgs_mode = CS_RBGS_FIXED_PAIRS;
  s->rbgs_tolerance = 1.0e-4f;
  s->rbgs_check_every = 8u;
  s->rbgs_max_iters = 20000u;
  s->warm_start = CS_WARM_ZERO;
  s->has_bed_field = 0u;
  s->warm_fallback_residual = 1.0f;
  s->rbgs_rescale_after = 0u;
  s->rbgs_mg_cycles = 0u;
  if (preset == CS_PRESET_FREE_STREAM) {           /* mod.rs:145-152 */
    s->spout.origin[0] = 0.0f; s->spout.origin[1] = 6.8f; s->spout.origin[2] = 0.0f;
    orc_spout_aim_at(&s->spout, 0.0f, -6.8f, 0.0f);
  } else if (preset == CS_PRESET_CENTER_POUR) {    /* mod.rs:154-160 */
    s->spout.origin[0] = 0.0f; s->spout.origin[1] = 7.1f; s->spout.origin[2] = 0.0f;
    orc_spout_aim_at(&s->spout, 0.0f, 0.4f, 0.0f);
  }
  return CS_OK;
}

/* ---- SDF generation (state.rs:206-276) ------------------------------------- */
static float sdf_interior(const cs_obstacle* o, float px, float py, float pz) {
  if (o->kind == CS_OBSTACLE_TRUNCATED_CONE) {
    float local_y = py - o->center[1];
    if (local_y > o->top_y || local_y < o->bot_y) return O_SDF_NO_CONSTRAINT;
    float height = o->top_y - o->bot_y;
    if (height < 1e-10f) return O_SDF_NO_CONSTRAINT;
    float t = (local_y - o->bot_y) / height;
    float cone_r = o->bot_radius + (o->top_radius - o->bot_radius) * t;
    float dx = px - o->center[0], dz = pz - o->center[2];
    float r = sqrtf(dx * dx + dz * dz);
    return cone_r - r;
  }
  if (o->kind == CS_OBSTACLE_CYLINDER) {
    float local_y = py - o->center[1];
    if (local_y > o->top_y) return O_SDF_NO_CONSTRAINT;
    float dx = px - o->center[0], dz = pz - o->center[2];
    float r = sqrtf(dx * dx + dz * dz);
    float radial = o->radius - r;
    float floor_d = local_y - o->bot_y;
    return fminf(radial, floor_d);
  }
  return O_SDF_NO_CONSTRAINT;
}

void orc_generate_sdf(const cs_settings* s, float* data) {
  uint32_t n = s->sdf_res;
  size_t total = (size_t)n * n * n;
  for (size_t i = 0; i < total; ++i) data[i] = O_SDF_NO_CONSTRAINT;
  float bx = s->bounds_size[0], by = s->bounds_size[1], bz = s->bounds_size[2];
  float half = s->wall_thickness * 0.5f;
  for (uint32_t oi = 0; oi < s->num_obstacles; ++oi) {
    const cs_obstacle* o = &s->obstacles[oi];
    for (uint32_t iz = 0; iz < n; ++iz)
      for (uint32_t iy = 0; iy < n; ++iy)
        for (uint32_t ix = 0; ix < n; ++ix) {
          float px = -bx * 0.5f + ((float)ix + 0.5f) * bx / (float)n;
          float py = -by * 0.5f + ((float)iy + 0.5f) * by / (float)n;
          float pz = -bz * 0.5f + ((float)iz + 0.5f) * bz / (float)n;
          float sd = sdf_interior(o, px, py, pz) - half;
          size_t idx = (size_t)iz * n * n + (size_t)iy * n + ix;
          if (data[idx] >= O_SDF_NO_CONSTRAINT - 1.0f) data[idx] = sd;
          else data[idx] = fmaxf(data[idx], sd);
        }
  }
}

/* Rust `f32 as i32` and `f32::fract` (x - trunc(x)) used by the CPU sampler. */
static int32_t rs_as_i32(float x) { return o_f2i(x); }
static float rs_fract(float x) { return x - truncf(x); }

static float load_texel(const float* d, uint32_t res, int32_t x, int32_t y, int32_t z) {
  int32_t mx = (int32_t)res - 1;
  x = x < 0 ? 0 : (x > mx ? mx : x);
  y = y < 0 ? 0 : (y > mx ? mx : y);
  z = z < 0 ? 0 : (z > mx ? mx : z);
  return d[(size_t)z * res * res + (size_t)y * res + (size_t)x];
}

/* sample_sdf_from_data (state.rs:287-320): lerp form c + (d - c) * f. */
static float sample_sdf_cpu(const cs_settings* s, const float* d, float px, float py, float pz) {
  float res = (float)s->sdf_res;
  float bx = s->bounds_size[0], by = s->bounds_size[1], bz = s->bounds_size[2];
  float ux = (px + bx * 0.5f) / bx * res - 0.5f;
  float uy = (py + by * 0.5f) / by * res - 0.5f;
  float uz = (pz + bz * 0.5f) / bz * res - 0.5f;
  int32_t b0 = rs_as_i32(floorf(ux)), b1 = rs_as_i32(floorf(uy)), b2 = rs_as_i32(floorf(uz));
  float fx = rs_fract(ux), fy = rs_fract(uy), fz = rs_fract(uz);
  uint32_t r = s->sdf_res;
  float c000 = load_texel(d, r, b0, b1, b2);
  float c100 = load_texel(d, r, b0 + 1, b1, b2);
  float c010 = load_texel(d, r, b0, b1 + 1, b2);
  float c110 = load_texel(d, r, b0 + 1, b1 + 1, b2);
  float c001 = load_texel(d, r, b0, b1, b2 + 1);
  float c101 = load_texel(d, r, b0 + 1, b1, b2 + 1);
  float c011 = load_texel(d, r, b0, b1 + 1, b2 + 1);
  float c111 = load_texel(d, r, b0 + 1, b1 + 1, b2 + 1);
  float c00 = c000 + (c100 - c000) * fx;
  float c10 = c010 + (c110 - c010) * fx;
  float c01 = c001 + (c101 - c001) * fx;
  float c11 = c011 + (c111 - c011) * fx;
  float c0 = c00 + (c10 - c00) * fy;
  float c1 = c01 + (c11 - c01) * fy;
  return c0 + (c1 - c0) * fz;
}

/* generate_sdf_class_data (state.rs:322-344) */
void orc_generate_sdf_class(const cs_settings* s, const float* sdf, uint8_t* out) {
  uint32_t gx = s->grid_dims[0], gy = s->grid_dims[1], gz = s->grid_dims[2];
  float bx = s->bounds_size[0], by = s->bounds_size[1], bz = s->bounds_size[2];
  float dx = bx / (float)gx;
  float ox = -bx * 0.5f, oy = -by * 0.5f, oz = -bz * 0.5f;
  for (uint32_t iz = 0; iz < gz; ++iz)
    for (uint32_t iy = 0; iy < gy; ++iy)
      for (uint32_t ix = 0; ix < gx; ++ix) {
        float cx = ox + ((float)ix + 0.5f) * dx;
        float cy = oy + ((float)iy + 0.5f) * dx;
        float cz = oz + ((float)iz + 0.5f) * dx;
        size_t idx = (size_t)iz * gx * gy + (size_t)iy * gx + ix;
        out[idx] = (uint8_t)(sample_sdf_cpu(s, sdf, cx, cy, cz) < 0.0f);
      }
}

/* ---- inflow model (inflow.rs:244-293) ------------------------------------- */
float orc_pour_fraction_from_angle(float angle, const cs_spout* sp) {
  if (angle <= sp->activation_angle_deg) return 0.0f;
  if (angle >= sp->full_flow_angle_deg) return 1.0f;
  float t = (angle - sp->activation_angle_deg) / (sp->full_flow_angle_deg - sp->activation_angle_deg);
  return t * t * (3.0f - 2.0f * t);
}
float orc_effective_head_from_angle(float angle, const cs_spout* sp) {
  if (angle <= sp->activation_angle_deg) return 0.0f;
  if (angle >= sp->full_flow_angle_deg) return sp->head_at_full_angle;
  float s = orc_pour_fraction_from_angle(angle, sp);
  return sp->head_at_activation + s * (sp->head_at_full_angle - sp->head_at_activation);
}
float orc_exit_speed_cap_from_angle(float angle, const cs_spout* sp) {
  float s = orc_pour_fraction_from_angle(angle, sp);
  return sp->gentle_exit_speed + s * (sp->max_exit_speed - sp->gentle_exit_speed);
}
float orc_exit_speed_from_head(float head, float max_exit_speed) {
  if (head < 1e-6f) return 0.0f;
  float speed = sqrtf(2.0f * 9.80665f * orc_sim_units_per_meter() * head);
  return fminf(speed, max_exit_speed);
}
float orc_flow_rate_from_speed(float v, float nozzle_radius, float cd, float vol_to_ml, float max_flow) {
  if (v < 1e-6f) return 0.0f;
  float area = 3.14159265358979323846f * nozzle_radius * nozzle_radius;
  float q = cd * area * v;
  return fminf(q * vol_to_ml, max_flow);
}

/* InflowState::update (inflow.rs:109-120) */
static void inflow_update(orc_inflow* in, const cs_spout* sp) {
  float head = orc_effective_head_from_angle(in->kettle_angle_deg, sp);
  float cap = orc_exit_speed_cap_from_angle(in->kettle_angle_deg, sp);
  in->exit_speed = orc_exit_speed_from_head(head, cap);
  in->flow_rate = orc_flow_rate_from_speed(in->exit_speed, sp->nozzle_radius, sp->discharge_coeff,
                                           sp->volume_to_ml, sp->max_flow_rate_ml_s);
}

static uint32_t rs_as_u32(float x) { return o_f2u(x); }

static void cross(const float* a, const float* b, float* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

/* Golden-angle disk emission (inflow.rs:165-219) of `count` particles at slot
 * `slot0`, shared by both emitter modes. */
static void emit_disk(orc_sim* S, uint32_t slot0, uint32_t count, float dt, float speed,
                      float jet_radius, float mass) {
  const cs_spout* sp = &S->s.spout;
  const float* dir = sp->direction;
  float up[3];
  if (fabsf(dir[1]) < 0.9f) { up[0] = 0.0f; up[1] = 1.0f; up[2] = 0.0f; }
  else { up[0] = 1.0f; up[1] = 0.0f; up[2] = 0.0f; }
  float t1[3], t2[3];
  cross(dir, up, t1); vnormalize_rs(t1);
  cross(dir, t1, t2); vnormalize_rs(t2);
  float vel[3] = {dir[0] * speed, dir[1] * speed, dir[2] * speed};
  float eo[3] = {sp->origin[0] + dir[0] * 0.18f, sp->origin[1] + dir[1] * 0.18f,
                 sp->origin[2] + dir[2] * 0.18f};
  const float golden = 2.399963f, radial_irr = 0.7548777f;
  /* emit_sample_period (no reference counterpart): the substep's samples start
   * from the running count modulo the period */
  const uint32_t period = S->s.emit_sample_period;
  const uint32_t seq0 = period ? S->inflow.emission_sequence % period : S->inflow.emission_sequence;
  for (uint32_t i = 0; i < count; ++i) {
    uint32_t sample = seq0 + i;
    float sf = (float)sample;
    float radial_u = rs_fract((sf + 0.5f) * radial_irr);
    float r = jet_radius * sqrtf(radial_u);
    float theta = sf * golden;
    float age = (((float)i + 0.5f) / (float)count) * dt;
    float ct, st;
    orc_sincosf(theta, &st, &ct);
    float off[3];
    for (int c = 0; c < 3; ++c) off[c] = t1[c] * r * ct + t2[c] * r * st;
    float* p = &S->particles[(size_t)(slot0 + i) * 8];
    for (int c = 0; c < 3; ++c) p[c] = eo[c] + off[c] - vel[c] * age;
    p[3] = 1.0f;
    p[4] = vel[0]; p[5] = vel[1]; p[6] = vel[2];
    p[7] = mass;
    memset(&S->affine[(size_t)(slot0 + i) * 12], 0, 12 * sizeof(float));
  }
  S->inflow.emission_sequence += count;
}

/* InflowState::emit_particles (inflow.rs:122-241) plus the fixed-count mode. */
void orc_emit(orc_sim* S, float dt) {
  float mass = 80.0f / 320.0f; /* MASS_UNITS_PER_ML / PARTICLES_PER_ML (mod.rs:278) */
  uint32_t total = S->num_water + S->bed_n; /* current_water + current_bed (inflow.rs:155) */
  uint32_t avail = S->s.max_particles > total ? S->s.max_particles - total : 0u;
  if (S->s.emitter_mode == CS_EMIT_FIXED_COUNT) {
    uint32_t want = S->s.emit_per_substep;
    if (S->s.emit_total) {
      uint32_t left = S->s.emit_total > S->inflow.emitted_total ? S->s.emit_total - S->inflow.emitted_total : 0u;
      if (want > left) want = left;
    }
    uint32_t count = want < avail ? want : avail;
    S->total_dropped += want - count;
    if (count == 0) return;
    emit_disk(S, S->num_water, count, dt, S->s.emit_speed, S->s.emit_jet_radius, mass); /* the water array */
    S->inflow.emitted_total += count;
    S->num_water += count;
    S->total_emitted_mass += (float)count * mass;
    return;
  }
  inflow_update(&S->inflow, &S->s.spout);
  if (S->inflow.flow_rate < 1e-6f) { S->inflow.accumulator = 0.0f; return; }
  float pps = S->inflow.flow_rate * 320.0f;
  S->inflow.accumulator += pps * dt;
  uint32_t requested = rs_as_u32(S->inflow.accumulator);
  if (requested == 0) return;
  S->inflow.accumulator -= (float)requested;
  uint32_t count = requested < avail ? requested : avail;
  if (count == 0) { S->total_dropped += requested; return; }
  emit_disk(S, S->num_water, count, dt, S->inflow.exit_speed, S->s.spout.nozzle_radius * 0.58f, mass);
  S->num_water += count;
  S->total_emitted_mass += (float)count * mass;
  S->total_dropped += requested - count;
}

/* ---- write_uniforms (mod.rs:569-670) -------------------------------------- */
void orc_write_uniforms(orc_sim* S, float dt, orc_params* u) {
  memset(u, 0, sizeof(*u));
  const cs_settings* s = &S->s;
  u->gx = S->gx; u->gy = S->gy; u->gz = S->gz; u->nc = (uint32_t)S->nc;
  float dx = s->bounds_size[0] / (float)S->gx;
  float inv_dx = 1.0f / dx;
  u->num_water = S->num_water;
  u->num_bed_particles = 0u;
  u->max_particles = s->max_particles;
  u->use_sdf_cache = s->use_sdf_cache;
  u->dt = dt; u->gravity = s->gravity; u->dx = dx; u->inv_dx = inv_dx;
  for (int c = 0; c < 3; ++c) {
    u->origin[c] = -s->bounds_size[c] * 0.5f;
    u->bmax[c] = s->bounds_size[c] * 0.5f;
  }
  u->bulk_K = s->bulk_modulus;
  u->viscosity = s->viscosity;
  u->nominal_mass = 80.0f / 320.0f;
  u->p_vol = dx * dx * dx * s->particle_vol_factor;
  u->fp_scale = O_FP_SCALE;
  u->inv_fp_scale = 1.0f / O_FP_SCALE;
  u->vel_cap = orc_max_velocity();
  u->outlet_radius = s->geometry.cone_bot_radius;
  u->sdf_res = (float)s->sdf_res;
  u->friction = 0.3f; u->restitution = 0.0f; u->contact_offset = 0.05f;
  u->water_nu = s->water_viscosity_m2_s;
  u->min_k = s->min_bed_permeability;
  u->total_time = S->total_time;
  u->pore_capacity_scale = 1.0f;
  u->pore_overfill_alpha = 18.0f;
  float bound = 4.0f * orc_max_velocity() * inv_dx;
  u->div_clamp = fminf(bound, O_FP_VALUE_LIMIT - 1.0f);
  u->pressure_clamp = O_FP_VALUE_LIMIT - 1.0f;
  u->metrics_div_fp_scale = O_METRICS_DIV_FP_SCALE;
  u->j_alpha = 32.0f; u->j_expand_alpha = 2.0f; u->max_rest_fraction = 1.20f;
  u->surface_void_scale = 1.0f;
  u->wall_half = s->wall_thickness * 0.5f;
  u->geo = s->geometry;
  u->num_bed_records = S->num_bed;
}

/* ---- create / reset -------------------------------------------------------- */
orc_sim* orc_create(const cs_settings* s, const cs_scene* scene) {
  if (!s || s->grid_dims[0] == 0 || s->grid_dims[1] == 0 || s->grid_dims[2] == 0) return NULL;
  orc_sim* S = (orc_sim*)calloc(1, sizeof(orc_sim));
  if (!S) return NULL;
  S->s = *s;
  S->gx = s->grid_dims[0]; S->gy = s->grid_dims[1]; S->gz = s->grid_dims[2];
  S->nc = (size_t)S->gx * S->gy * S->gz;
  S->sdf_res = s->sdf_res;
  size_t ns = (size_t)s->sdf_res * s->sdf_res * s->sdf_res;
  S->sdf = (float*)malloc(ns * sizeof(float));
  S->cls = (uint8_t*)malloc(S->nc);
  S->bed_lookup = (int32_t*)malloc(S->nc * sizeof(int32_t));
  S->particles = (float*)calloc((size_t)s->max_particles * 8, sizeof(float));
  S->affine = (float*)calloc((size_t)s->max_particles * 12, sizeof(float));
  S->grid = (int32_t*)calloc(S->nc * 6, sizeof(int32_t));
  S->p2g_lanes = (int32_t*)calloc(S->nc * 6, sizeof(int32_t));
  S->grid_vel = (float*)calloc(S->nc * 4, sizeof(float));
  S->prev_p = (int32_t*)calloc(S->nc, sizeof(int32_t));
  S->warm = (float*)calloc(S->nc, sizeof(float));
  S->saved_p = (int32_t*)calloc(S->nc, sizeof(int32_t));
  S->saved_kind = (int32_t*)calloc(S->nc, sizeof(int32_t));
  S->tgt = (float*)calloc(S->nc, sizeof(float));
  if (!S->tgt || !S->saved_p || !S->saved_kind ||!S->sdf || !S->cls || !S->bed_lookup || !S->particles || !S->affine || !S->grid ||
      !S->p2g_lanes || !S->grid_vel || !S->prev_p || !S->warm) {
    orc_destroy(S);
    return NULL;
  }
  if (scene && scene->sdf) memcpy(S->sdf, scene->sdf, ns * sizeof(float));
  else orc_generate_sdf(s, S->sdf);
  if (scene && scene->sdf_class) memcpy(S->cls, scene->sdf_class, S->nc);
  else orc_generate_sdf_class(s, S->sdf, S->cls);
  for (size_t i = 0; i < S->nc; ++i) S->bed_lookup[i] = -1;
  if (s->has_bed_field && scene && scene->bed_lookup && scene->bed_records) {
    memcpy(S->bed_lookup, scene->bed_lookup, S->nc * sizeof(int32_t));
    S->num_bed = scene->num_bed_records;
    S->bed = (cs_bed_record*)malloc((size_t)S->num_bed * sizeof(cs_bed_record) + 1);
    memcpy(S->bed, scene->bed_records, (size_t)S->num_bed * sizeof(cs_bed_record));
  } else if (scene && scene->num_bed_particles && scene->bed_particles && scene->bed_affine && scene->bed_extracts) {
    uint32_t n = scene->num_bed_particles;
    S->bed_n = n;
    S->num_bed = n;
    S->bed = (cs_bed_record*)calloc((size_t)n + 1, sizeof(cs_bed_record));
    S->bed_part = (float*)malloc((size_t)n * 8 * sizeof(float));
    S->bed_aff = (float*)malloc((size_t)n * 12 * sizeof(float));
    S->bed_ext = (float*)malloc((size_t)n * 8 * sizeof(float));
    memcpy(S->bed_part, scene->bed_particles, (size_t)n * 8 * sizeof(float));
    memcpy(S->bed_aff, scene->bed_affine, (size_t)n * 12 * sizeof(float));
    memcpy(S->bed_ext, scene->bed_extracts, (size_t)n * 8 * sizeof(float));
    S->bed_ext0 = (float*)malloc((size_t)n * 8 * sizeof(float));
    memcpy(S->bed_ext0, scene->bed_extracts, (size_t)n * 8 * sizeof(float));
    S->bed_part0 = (float*)malloc((size_t)n * 8 * sizeof(float));
    S->bed_aff0 = (float*)malloc((size_t)n * 12 * sizeof(float));
    memcpy(S->bed_part0, scene->bed_particles, (size_t)n * 8 * sizeof(float));
    memcpy(S->bed_aff0, scene->bed_affine, (size_t)n * 12 * sizeof(float));
    S->bed_snap = (float*)malloc((size_t)n * 3 * sizeof(float));
    S->bed_dw = (int32_t*)calloc((size_t)n * 4, sizeof(int32_t)); /* bed_delta lanes: water, impulse xyz */
  }
  if (scene && scene->particles && scene->num_particles) {
    uint32_t n = scene->num_particles;
    if (n > s->max_particles) n = s->max_particles;
    S->init_count = n;
    S->init_particles = (float*)malloc((size_t)n * 8 * sizeof(float));
    S->init_affine = (float*)calloc((size_t)n * 12, sizeof(float));
    memcpy(S->init_particles, scene->particles, (size_t)n * 8 * sizeof(float));
    if (scene->affine) memcpy(S->init_affine, scene->affine, (size_t)n * 12 * sizeof(float));
  }
  orc_reset(S);
  return S;
}

void orc_destroy(orc_sim* S) {
  if (!S) return;
  free(S->sdf); free(S->cls); free(S->bed_lookup); free(S->bed); free(S->bed_part); free(S->bed_aff); free(S->bed_ext); free(S->bed_ext0); free(S->bed_dw);
  free(S->bed_part0); free(S->bed_aff0); free(S->bed_snap);
  free(S->particles); free(S->affine); free(S->init_particles); free(S->init_affine);
  free(S->grid); free(S->p2g_lanes); free(S->grid_vel); free(S->prev_p); free(S->warm);
  free(S->saved_p); free(S->saved_kind); free(S->tgt);
  free(S);
}

/* MpmSim3D::reset (mod.rs:432-446) */
int orc_reset(orc_sim* S) {
  S->num_water = 0;
  S->total_time = 0.0f;
  S->total_emitted_mass = 0.0f;
  S->total_dropped = 0;
  memset(&S->inflow, 0, sizeof(S->inflow));
  S->inflow.kettle_angle_deg = S->s.initial_kettle_angle_deg;
  memset(S->particles, 0, (size_t)S->s.max_particles * 8 * sizeof(float));
  memset(S->affine, 0, (size_t)S->s.max_particles * 12 * sizeof(float));
  if (S->init_count) {
    memcpy(S->particles, S->init_particles, (size_t)S->init_count * 8 * sizeof(float));
    memcpy(S->affine, S->init_affine, (size_t)S->init_count * 12 * sizeof(float));
    S->num_water = S->init_count;
  }
  memset(S->prev_p, 0, S->nc * sizeof(int32_t));
  if (S->bed_n) {
    /* the bed as the scene gave it (the reference re-runs init_bed, mod.rs:432-446) */
    memcpy(S->bed_part, S->bed_part0, (size_t)S->bed_n * 8 * sizeof(float));
    memcpy(S->bed_aff, S->bed_aff0, (size_t)S->bed_n * 12 * sizeof(float));
    memcpy(S->bed_ext, S->bed_ext0, (size_t)S->bed_n * 8 * sizeof(float));
    memset(S->bed_dw, 0, (size_t)S->bed_n * 4 * sizeof(int32_t));
  }
  S->has_warm = 0;
  memset(&S->stats, 0, sizeof(S->stats));
  memset(S->metrics, 0, sizeof(S->metrics));
  return CS_OK;
}

int orc_set_kettle_angle(orc_sim* S, float a) { S->inflow.kettle_angle_deg = a; return CS_OK; }
int orc_set_spout_position(orc_sim* S, float x, float y, float z) {
  orc_spout_translate_origin_to(&S->s.spout, x, y, z);
  return CS_OK;
}
int orc_set_spout_target(orc_sim* S, float x, float y, float z) {
  orc_spout_aim_at(&S->s.spout, x, y, z);
  return CS_OK;
}
int orc_set_warm_start(orc_sim* S, const float* p) {
  memcpy(S->warm, p, S->nc * sizeof(float));
  S->has_warm = 1;
  return CS_OK;
}
int orc_write_particles(orc_sim* S, uint32_t off, uint32_t n, const float* p8, const float* a12) {
  if ((uint64_t)off + n > S->s.max_particles) return CS_ERR_RANGE;
  if (p8) memcpy(&S->particles[(size_t)off * 8], p8, (size_t)n * 8 * sizeof(float));
  if (a12) memcpy(&S->affine[(size_t)off * 12], a12, (size_t)n * 12 * sizeof(float));
  return CS_OK;
}
int orc_set_particle_count(orc_sim* S, uint32_t n) {
  if (n > S->s.max_particles) return CS_ERR_RANGE;
  S->num_water = n;
  return CS_OK;
}
void orc_get_counts(orc_sim* S, uint32_t* w, uint32_t* b, uint32_t* m) {
  if (w) *w = S->num_water;
  if (b) *b = S->bed_n;
  if (m) *m = S->s.max_particles;
}
void orc_get_time(orc_sim* S, float* t, float* m, uint32_t* d) {
  if (t) *t = S->total_time;
  if (m) *m = S->total_emitted_mass;
  if (d) *d = S->total_dropped;
}
void orc_get_inflow(orc_sim* S, float* a, float* f, float* e) {
  if (a) *a = S->inflow.kettle_angle_deg;
  if (f) *f = S->inflow.flow_rate;
  if (e) *e = S->inflow.exit_speed;
}
void orc_read_particles(orc_sim* S, uint32_t off, uint32_t n, float* p8) {
  memcpy(p8, &S->particles[(size_t)off * 8], (size_t)n * 8 * sizeof(float));
}
void orc_read_affine(orc_sim* S, uint32_t off, uint32_t n, float* a12) {
  memcpy(a12, &S->affine[(size_t)off * 12], (size_t)n * 12 * sizeof(float));
}
void orc_read_p2g_lanes(orc_sim* S, int32_t* l) { memcpy(l, S->p2g_lanes, S->nc * 6 * sizeof(int32_t)); }
void orc_read_grid_vel(orc_sim* S, float* v) { memcpy(v, S->grid_vel, S->nc * 4 * sizeof(float)); }
void orc_read_metrics(orc_sim* S, uint32_t* m) { memcpy(m, S->metrics, sizeof(S->metrics)); }

/* prepare_render (shader.rs:2242-2276) on the current particle state: an
 * inactive particle (mass <= inactive_mass_threshold, shader.rs:116) is parked
 * at y = -1e6 with colour -999; water colours by speed / 10 clamped to
 * [0, 2]; other phases by -1 - saturation, where a bed particle
 * (pid < num_bed) also takes the bed radius.  Radii are dx times the
 * DEFAULT_BREW sample radii (brew_config.rs:50-51, shader.rs:127-128). */
void orc_pack_render(orc_sim* S, float* out) {
  orc_params u;
  orc_write_uniforms(S, 1.0e-4f, &u);  /* dt does not enter the packing */
  /* the reference's particle order: a bed of particles first, then the water */
  const uint32_t nb = S->bed_n;
  const uint32_t n = nb + u.num_water + u.num_bed_particles;
  for (uint32_t pid = 0; pid < n; ++pid) {
    const int bed = pid < nb;
    const float* p = bed ? &S->bed_part[(size_t)pid * 8] : &S->particles[(size_t)(pid - nb) * 8];
    const float phase = bed ? S->bed_aff[(size_t)pid * 12 + 3] : S->affine[(size_t)(pid - nb) * 12 + 3];
    float* o = &out[(size_t)pid * 8];
    if (p[7] <= u.nominal_mass * 0.10f) {
      const float park[8] = {0.0f, -1e6f, 0.0f, -999.0f, 0.0f, 0.0f, 0.0f, 0.0f};
      memcpy(o, park, sizeof(park));
      continue;
    }
    float color_t, radius = u.dx * 0.18f;
    if (phase < 0.5f) {
      const float speed = sqrtf(p[4] * p[4] + p[5] * p[5] + p[6] * p[6]);
      color_t = o_clamp(speed / 10.0f, 0.0f, 2.0f);
    } else {
      float sat = 0.0f;  /* bed_extract[pid].extract.w; a field bed has no extraction state */
      if (bed) {
        sat = S->bed_ext[(size_t)pid * 8 + 7];
        radius = u.dx * 0.62f;
      } else if (pid < u.num_bed_particles) {
        radius = u.dx * 0.62f;
      }
      color_t = -1.0f - sat;
    }
    o[0] = p[0]; o[1] = p[1]; o[2] = p[2]; o[3] = color_t;
    o[4] = radius; o[5] = 0.0f; o[6] = 0.0f; o[7] = 0.0f;
  }
}
void orc_get_solve_stats(orc_sim* S, cs_solve_stats* o) { *o = S->stats; }

/* ---- exp from Cody-Waite range reduction + degree-7 Taylor polynomial ------- */
float orc_expf(float x) {
  if (x != x) return x;
  if (x > 88.0f) x = 88.0f;
  if (x < -87.0f) x = -87.0f;
  const float log2e = 1.44269502f;
  const float ln2_hi = 0.693145751953125f; /* 16 significant bits: k*ln2_hi exact */
  const float ln2_lo = 1.42860677e-06f;    /* ln2 - ln2_hi */
  float kf = floorf(x * log2e + 0.5f);
  float r = (x - kf * ln2_hi) - kf * ln2_lo;
  float p = 1.0f / 5040.0f;
  p = p * r + 1.0f / 720.0f;
  p = p * r + 1.0f / 120.0f;
  p = p * r + 1.0f / 24.0f;
  p = p * r + 1.0f / 6.0f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  int32_t k = (int32_t)kf;
  union { uint32_t u; float f; } sc;
  sc.u = (uint32_t)(k + 127) << 23;
  return p * sc.f;
}
/* ---- sin and cos for the emitter's golden angle (inflow.rs:205-207) ----------
 * The reference calls f32::sin / f32::cos (a libm).  The oracle and the product
 * share one evaluation instead, so host and device agree bit for bit: reduce
 * x = k pi/2 + r in double with pi/2 split in three (22 + 22 significant bits and
 * a double remainder, so k*c1 and k*c2 are exact for |k| < 2^31), sum the Taylor
 * series of sin r and cos r (odd terms through r^17, even through r^18) by
 * Horner's rule, pick the quadrant k mod 4 and round each once to float. */
static const double kSinC[8] = {
  0x1.952c77030ad4ap-49,  -0x1.ae7f3e733b81fp-41, 0x1.6124613a86d09p-33, -0x1.ae64567f544e4p-26,
  0x1.71de3a556c734p-19,  -0x1.a01a01a01a01ap-13, 0x1.1111111111111p-7,  -0x1.5555555555555p-3};
static const double kCosC[9] = {
  -0x1.6827863b97d97p-53, 0x1.ae7f3e733b81fp-45,  -0x1.93974a8c07c9dp-37, 0x1.1eed8eff8d898p-29,
  -0x1.27e4fb7789f5cp-22, 0x1.a01a01a01a01ap-16,  -0x1.6c16c16c16c17p-10, 0x1.5555555555555p-5, -0.5};
void orc_sincosf(float x, float* sn, float* cs) {
  const double xd = (double)x;
  const double k = rint(xd * 0x1.45f306dc9c883p-1); /* nearest multiple of pi/2 */
  double r = xd - k * 0x1.921fb8p+0;
  r = r - k * -0x1.5dde98p-23;
  r = r - k * 0x1.8469898cc5170p-48;
  const double r2 = r * r;
  double ps = kSinC[0], pc = kCosC[0];
  for (int i = 1; i < 8; ++i) ps = ps * r2 + kSinC[i];
  for (int i = 1; i < 9; ++i) pc = pc * r2 + kCosC[i];
  const double sr = r + (r * r2) * ps, cr = 1.0 + r2 * pc;
  const double sv[4] = {sr, cr, -sr, -cr}, cv[4] = {cr, -sr, -cr, sr};
  const int q = (int)((long long)k & 3);
  *sn = (float)sv[q];
  *cs = (float)cv[q];
}
float orc_smoothstep(float lo, float hi, float x) { return o_smooth(lo, hi, x); }
int32_t orc_f2i(float x) { return o_f2i(x); }
