// cs_scene.cpp — CPU-side scene construction for the MLS-MPM substep:
// units, settings presets, SDF and class-mask generation, the inflow emitter,
// the parameter block and the static bed table.  Follows the reference's host
// code (crates/sim-wasm/src/mpm_3d/{mod,state,inflow,units,brew_config}.rs) with
// binary32 arithmetic in the same operation order, so the values handed to the
// kernels are the reference's bit for bit.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "cs_host.h"
#include "cs_bed.h"
#include "cs_sdf.h"

namespace cs {

// ---- units.rs:10-34 ---------------------------------------------------------------
float sim_units_per_meter() {
  volatile float top = 2.75f, bot = -3.02f, h = 0.10f;
  float height_units = top - bot;
  return height_units / h;
}
float max_velocity() { return 1.5f * sim_units_per_meter(); }
float gravity_sim_units() { return -9.80665f * sim_units_per_meter(); }
float ml_per_sim_unit_cubed() {
  float m = 1.0f / sim_units_per_meter();
  return m * m * m * 1000000.0f;
}
// brew_config.rs:160-164
float kozeny_carman_permeability(float grind_diameter_um, float porosity) {
  float d_m = grind_diameter_um * 1.0e-6f;
  float solid = 1.0f - porosity;
  return d_m * d_m * porosity * porosity * porosity / (180.0f * solid * solid);
}

// ---- sim-core Vec3 semantics (crates/sim-core/src/math.rs:37-92) -------------------
static float rs_len(const float* v) { return sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
static void rs_normalize(float* v) {
  float len = rs_len(v);
  if (len <= 1.0e-6f || fabsf(len) <= 1.0e-6f) {
    v[0] = v[1] = v[2] = 0.0f;
    return;
  }
  v[0] /= len;
  v[1] /= len;
  v[2] /= len;
}
static void rs_cross(const float* a, const float* b, float* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// SpoutSettings::aim_at / translate_origin_to (inflow.rs:50-66)
void spout_aim_at(cs_spout* sp, float x, float y, float z) {
  sp->target[0] = x;
  sp->target[1] = y;
  sp->target[2] = z;
  float d[3] = {x - sp->origin[0], y - sp->origin[1], z - sp->origin[2]};
  if (rs_len(d) > 1e-5f) {
    rs_normalize(d);
    memcpy(sp->direction, d, sizeof(d));
  }
}
void spout_translate_origin_to(cs_spout* sp, float x, float y, float z) {
  float d[3] = {x - sp->origin[0], y - sp->origin[1], z - sp->origin[2]};
  sp->origin[0] = x;
  sp->origin[1] = y;
  sp->origin[2] = z;
  for (int c = 0; c < 3; ++c) sp->target[c] = sp->target[c] + d[c];
}

// ---- presets: MpmSettings::{default_v60, benchmark_free_stream,
//      benchmark_center_pour} (mod.rs:106-177), SpoutSettings::default (inflow.rs:27-47)
int settings_preset(int preset, cs_settings* s) {
  if (!s || preset < CS_PRESET_DEFAULT_V60 || preset > CS_PRESET_CENTER_POUR) return CS_ERR_INVALID;
  memset(s, 0, sizeof(*s));
  s->bounds_size[0] = 14.0f;
  s->bounds_size[1] = 20.0f;
  s->bounds_size[2] = 14.0f;
  const uint32_t gx = 80u;
  const float dx = s->bounds_size[0] / (float)gx;
  s->grid_dims[0] = gx;
  s->grid_dims[1] = (uint32_t)ceilf(s->bounds_size[1] / dx);
  s->grid_dims[2] = 80u;
  s->max_particles = 220000u;
  s->substeps = 10u;
  s->gravity = gravity_sim_units();
  s->bulk_modulus = 900.0f;
  s->viscosity = 1.2f;  // DEFAULT_BREW.water_viscosity (brew_config.rs:124)
  s->render_radius = dx * 0.7f;
  s->pressure_rbgs_pairs = 40u;
  s->use_sdf_cache = 1u;
  // v60_support_cone against FilterConfig::default (filter.rs:211-223)
  const float filter_height = cs_max(2.75f - -3.02f, 1e-6f);
  const float slope = (4.10f - 0.0f) / filter_height;
  const float outlet = 0.42f;  // DEFAULT_BREW.dripper_outlet_radius
  s->num_obstacles = 2u;
  s->obstacles[0].kind = CS_OBSTACLE_TRUNCATED_CONE;
  s->obstacles[0].top_radius = outlet + slope * (3.0f - -3.0f);
  s->obstacles[0].bot_radius = outlet;
  s->obstacles[0].top_y = 3.0f;
  s->obstacles[0].bot_y = -3.0f;
  s->obstacles[1].kind = CS_OBSTACLE_CYLINDER;
  s->obstacles[1].radius = 3.0f;
  s->obstacles[1].top_y = -3.5f;
  s->obstacles[1].bot_y = -8.0f;
  cs_spout& sp = s->spout;
  sp.origin[0] = -3.4f;
  sp.origin[1] = 7.3f;
  sp.origin[2] = 0.9f;
  float dir[3] = {0.36f, -0.92f, -0.12f};
  rs_normalize(dir);
  memcpy(sp.direction, dir, sizeof(dir));
  sp.target[0] = -0.7f;
  sp.target[1] = 0.4f;
  sp.target[2] = 0.0f;
  sp.nozzle_radius = 0.18f;
  sp.stem_radius = 0.24f;
  sp.activation_angle_deg = 8.0f;
  sp.full_flow_angle_deg = 56.0f;
  sp.head_at_activation = 0.04f;
  sp.head_at_full_angle = 24.0f;
  sp.discharge_coeff = 0.92f;
  sp.volume_to_ml = ml_per_sim_unit_cubed();
  sp.max_flow_rate_ml_s = 12.0f;
  sp.gentle_exit_speed = 0.12f * sim_units_per_meter();
  sp.max_exit_speed = 0.45f * sim_units_per_meter();
  sp.stem_length = 1.9f;
  s->initial_kettle_angle_deg = 9.0f;
  // extensions at their reference values
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
  if (preset == CS_PRESET_FREE_STREAM) {
    sp.origin[0] = 0.0f;
    sp.origin[1] = 6.8f;
    sp.origin[2] = 0.0f;
    spout_aim_at(&sp, 0.0f, -6.8f, 0.0f);
  } else if (preset == CS_PRESET_CENTER_POUR) {
    sp.origin[0] = 0.0f;
    sp.origin[1] = 7.1f;
    sp.origin[2] = 0.0f;
    spout_aim_at(&sp, 0.0f, 0.4f, 0.0f);
  }
  return CS_OK;
}

int validate_settings(const cs_settings& s, std::string* why) {
  auto fail = [&](const char* m) {
    if (why) *why = m;
    return CS_ERR_INVALID;
  };
  for (int c = 0; c < 3; ++c) {
    if (s.grid_dims[c] == 0u || s.grid_dims[c] > 1024u) return fail("grid dims must be in [1, 1024]");
    if (!(s.bounds_size[c] > 0.0f)) return fail("bounds_size must be positive");
  }
  if ((uint64_t)s.grid_dims[0] * s.grid_dims[1] * s.grid_dims[2] > (1ull << 30)) return fail("grid too large");
  if (s.sdf_res < 2u || s.sdf_res > 1024u) return fail("sdf_res must be in [2, 1024]");
  if (s.max_particles == 0u || s.max_particles > (1u << 28)) return fail("max_particles out of range");
  if (s.num_obstacles > CS_MAX_OBSTACLES) return fail("too many obstacles");
  if (s.rbgs_mode > CS_RBGS_TOLERANCE) return fail("unknown rbgs_mode");
  if (s.warm_start > CS_WARM_LEARNED) return fail("unknown warm_start");
  if (!(s.warm_fallback_residual > 0.0f)) return fail("warm_fallback_residual must be positive");
  if (s.rbgs_rescale_after) {
    const uint32_t every = s.rbgs_check_every ? s.rbgs_check_every : 8u;
    if (s.rbgs_mode != CS_RBGS_TOLERANCE) return fail("rbgs_rescale_after needs the residual-driven rbgs_mode");
    if (s.rbgs_rescale_after % 4u != 0u || s.rbgs_rescale_after % every != 0u)
      return fail("rbgs_rescale_after must be a multiple of 4 and of rbgs_check_every");
    if ((uint64_t)s.rbgs_rescale_after + every > s.rbgs_max_iters)
      return fail("rbgs_rescale_after leaves fewer than rbgs_check_every iterations under rbgs_max_iters");
  }
  if (s.rbgs_mg_cycles) {
    if (s.rbgs_mode != CS_RBGS_TOLERANCE) return fail("rbgs_mg_cycles needs the residual-driven rbgs_mode");
    if (s.rbgs_rescale_after) return fail("rbgs_mg_cycles and rbgs_rescale_after exclude each other");
    if (s.rbgs_mg_cycles > 64u) return fail("rbgs_mg_cycles must be at most 64");
  }
  if (s.emitter_mode > CS_EMIT_FIXED_COUNT) return fail("unknown emitter_mode");
  return CS_OK;
}

// ---- SDF generation (state.rs:206-344) -----------------------------------------------
// The per-voxel and per-cell rules are in cs_sdf.h, shared with the device generator.
void generate_sdf(const cs_settings& s, float* data) {
  const uint32_t n = s.sdf_res;
  for (uint32_t iz = 0; iz < n; ++iz)
    for (uint32_t iy = 0; iy < n; ++iy)
      for (uint32_t ix = 0; ix < n; ++ix) data[((size_t)iz * n + iy) * n + ix] = cs_sdf_voxel(s, ix, iy, iz);
}

void generate_sdf_class(const cs_settings& s, const float* sdf, uint8_t* out) {
  const uint32_t gx = s.grid_dims[0], gy = s.grid_dims[1], gz = s.grid_dims[2];
  for (uint32_t iz = 0; iz < gz; ++iz)
    for (uint32_t iy = 0; iy < gy; ++iy)
      for (uint32_t ix = 0; ix < gx; ++ix) out[((size_t)iz * gy + iy) * gx + ix] = cs_sdf_class_cell(s, sdf, ix, iy, iz);
}

// ---- inflow model (inflow.rs:244-293) -----------------------------------------------
float pour_fraction_from_angle(float a, const cs_spout& sp) {
  if (a <= sp.activation_angle_deg) return 0.0f;
  if (a >= sp.full_flow_angle_deg) return 1.0f;
  float t = (a - sp.activation_angle_deg) / (sp.full_flow_angle_deg - sp.activation_angle_deg);
  return t * t * (3.0f - 2.0f * t);
}
float effective_head_from_angle(float a, const cs_spout& sp) {
  if (a <= sp.activation_angle_deg) return 0.0f;
  if (a >= sp.full_flow_angle_deg) return sp.head_at_full_angle;
  float s = pour_fraction_from_angle(a, sp);
  return sp.head_at_activation + s * (sp.head_at_full_angle - sp.head_at_activation);
}
float exit_speed_cap_from_angle(float a, const cs_spout& sp) {
  float s = pour_fraction_from_angle(a, sp);
  return sp.gentle_exit_speed + s * (sp.max_exit_speed - sp.gentle_exit_speed);
}
float exit_speed_from_head(float head, float max_exit_speed) {
  if (head < 1e-6f) return 0.0f;
  float speed = sqrtf(2.0f * 9.80665f * sim_units_per_meter() * head);
  return fminf(speed, max_exit_speed);
}
float flow_rate_from_speed(float v, float nozzle_radius, float cd, float vol_to_ml, float max_flow) {
  if (v < 1e-6f) return 0.0f;
  float area = 3.14159265358979323846f * nozzle_radius * nozzle_radius;
  float q = cd * area * v;
  return fminf(q * vol_to_ml, max_flow);
}

// Golden-angle disk (inflow.rs:165-219).
// The golden-angle disk frame of one substep's emission (inflow.rs:165-193):
// tangents of the spout direction, the contracted emit origin and the velocity.
static void disk_frame(const cs_spout& sp, Inflow& in, uint32_t count, float dt, float speed, float jet_radius,
                       float mass, uint32_t period, cs_emit_frame* f) {
  const float* dir = sp.direction;
  float up[3] = {0.0f, 1.0f, 0.0f};
  if (!(fabsf(dir[1]) < 0.9f)) {
    up[0] = 1.0f;
    up[1] = 0.0f;
  }
  rs_cross(dir, up, f->t1);
  rs_normalize(f->t1);
  rs_cross(dir, f->t1, f->t2);
  rs_normalize(f->t2);
  for (int c = 0; c < 3; ++c) {
    f->vel[c] = dir[c] * speed;
    f->origin[c] = sp.origin[c] + dir[c] * 0.18f;
  }
  f->jet_radius = jet_radius;
  f->dt = dt;
  f->mass = mass;
  f->seq0 = period ? in.emission_sequence % period : in.emission_sequence;
  f->count = count;
  in.emission_sequence += count;
}

uint32_t emit_plan(const cs_settings& s, Inflow& in, float dt, uint32_t current, cs_emit_frame* f,
                   uint32_t* dropped) {
  const float mass = 80.0f / 320.0f;  // MASS_UNITS_PER_ML / PARTICLES_PER_ML (mod.rs:278)
  uint32_t avail = s.max_particles > current ? s.max_particles - current : 0u;
  memset(f, 0, sizeof(*f));
  f->first = current;
  if (s.emitter_mode == CS_EMIT_FIXED_COUNT) {
    uint32_t want = s.emit_per_substep;
    if (s.emit_total) {
      uint32_t left = s.emit_total > in.emitted_total ? s.emit_total - in.emitted_total : 0u;
      want = std::min(want, left);
    }
    uint32_t count = std::min(want, avail);
    if (dropped) *dropped += want - count;
    if (count == 0) return 0;
    disk_frame(s.spout, in, count, dt, s.emit_speed, s.emit_jet_radius, mass, s.emit_sample_period, f);
    in.emitted_total += count;
    return count;
  }
  // InflowState::update (inflow.rs:109-120)
  float head = effective_head_from_angle(in.kettle_angle_deg, s.spout);
  float cap = exit_speed_cap_from_angle(in.kettle_angle_deg, s.spout);
  in.exit_speed = exit_speed_from_head(head, cap);
  in.flow_rate = flow_rate_from_speed(in.exit_speed, s.spout.nozzle_radius, s.spout.discharge_coeff,
                                      s.spout.volume_to_ml, s.spout.max_flow_rate_ml_s);
  if (in.flow_rate < 1e-6f) {
    in.accumulator = 0.0f;
    return 0;
  }
  in.accumulator += in.flow_rate * 320.0f * dt;
  uint32_t requested = cs_f2u(in.accumulator);
  if (requested == 0) return 0;
  in.accumulator -= (float)requested;
  uint32_t count = std::min(requested, avail);
  if (dropped) *dropped += requested - count;
  if (count == 0) return 0;
  disk_frame(s.spout, in, count, dt, in.exit_speed, s.spout.nozzle_radius * 0.58f, mass, s.emit_sample_period, f);
  return count;
}

uint32_t emit(const cs_settings& s, Inflow& in, float dt, uint32_t current, std::vector<float>& out,
              uint32_t* dropped) {
  cs_emit_frame f;
  const uint32_t count = emit_plan(s, in, dt, current, &f, dropped);
  const size_t base = out.size();
  out.resize(base + (size_t)count * 8);
  for (uint32_t i = 0; i < count; ++i) cs_emit_particle(f, i, &out[base + (size_t)i * 8]);
  return count;
}

// ---- write_uniforms (mod.rs:569-670) -----------------------------------------------------
void make_params(const cs_settings& s, uint32_t num_particles, uint32_t num_bed_cells, float dt, cs_params* u) {
  memset(u, 0, sizeof(*u));
  u->gx = s.grid_dims[0];
  u->gy = s.grid_dims[1];
  u->gz = s.grid_dims[2];
  u->nc = u->gx * u->gy * u->gz;
  u->nbx = (u->gx + CS_BRICK - 1) / CS_BRICK;
  u->nby = (u->gy + CS_BRICK - 1) / CS_BRICK;
  u->nbz = (u->gz + CS_BRICK - 1) / CS_BRICK;
  u->nbricks = u->nbx * u->nby * u->nbz;
  u->num_particles = num_particles;
  u->num_bed_cells = num_bed_cells;
  u->max_particles = s.max_particles;
  u->use_sdf_cache = s.use_sdf_cache;
  const float dx = s.bounds_size[0] / (float)u->gx;
  const float inv_dx = 1.0f / dx;
  u->dt = dt;
  u->gravity = s.gravity;
  u->dx = dx;
  u->inv_dx = inv_dx;
  for (int c = 0; c < 3; ++c) {
    u->origin[c] = -s.bounds_size[c] * 0.5f;
    u->bmax[c] = s.bounds_size[c] * 0.5f;
  }
  u->bulk_K = s.bulk_modulus;
  u->viscosity = s.viscosity;
  u->nominal_mass = 80.0f / 320.0f;
  u->p_vol = dx * dx * dx * s.particle_vol_factor;
  u->fp_scale = CS_FP_SCALE;
  u->inv_fp_scale = 1.0f / CS_FP_SCALE;
  u->vel_cap = max_velocity();
  u->sdf_res = s.sdf_res;
  u->sdf_res_f = (float)s.sdf_res;
  u->friction = 0.3f;
  u->restitution = 0.0f;
  u->contact_offset = 0.05f;
  u->water_nu = s.water_viscosity_m2_s;
  u->min_k = s.min_bed_permeability;
  float bound = 4.0f * max_velocity() * inv_dx;
  u->div_clamp = fminf(bound, CS_FP_VALUE_LIMIT - 1.0f);
  u->pressure_clamp = CS_FP_VALUE_LIMIT - 1.0f;
  u->metrics_div_fp_scale = CS_METRICS_DIV_FP_SCALE;
  u->j_alpha = 32.0f;
  u->j_expand_alpha = 2.0f;
  u->max_rest_fraction = 1.20f;
  u->wall_half = s.wall_thickness * 0.5f;
  u->geo = s.geometry;
}

// ---- static bed table (field mode of shader.rs:380-449, 621-666) -------------------------
bool build_bed_table(const cs_settings& s, const int32_t* lookup, const cs_bed_record* rec, uint32_t nrec,
                     std::vector<int32_t>& cell_index, std::vector<cs_bed_cell>& table) {
  const uint32_t gx = s.grid_dims[0], gy = s.grid_dims[1], gz = s.grid_dims[2];
  const size_t nc = (size_t)gx * gy * gz;
  cell_index.assign(nc, -1);
  table.clear();
  if (!lookup || !rec || nrec == 0) return true;
  for (size_t i = 0; i < nc; ++i)
    if (lookup[i] >= (int32_t)nrec || lookup[i] < -1) return false;
  cs_params u;
  make_params(s, 0, 0, 0.0f, &u);
  auto at = [&](int x, int y, int z) -> int32_t {
    if (x < 0 || y < 0 || z < 0 || x >= (int)gx || y >= (int)gy || z >= (int)gz) return -2;
    return lookup[((size_t)z * gy + y) * gx + x];
  };
  auto vel = [&](int32_t id, float* v) {
    v[0] = rec[id].velocity[0];
    v[1] = rec[id].velocity[1];
    v[2] = rec[id].velocity[2];
  };
  for (uint32_t iz = 0; iz < gz; ++iz)
    for (uint32_t iy = 0; iy < gy; ++iy)
      for (uint32_t ix = 0; ix < gx; ++ix) {
        const size_t c = ((size_t)iz * gy + iy) * gx + ix;
        const int32_t id = lookup[c];
        if (id < 0) continue;
        cs_bed_cell t;
        memset(&t, 0, sizeof(t));
        cs_bed_cell_build(u, gx, gy, gz, (int)ix, (int)iy, (int)iz, rec[id], at, vel, &t);
        cell_index[c] = (int32_t)table.size();
        table.push_back(t);
      }
  return true;
}

}  // namespace cs
