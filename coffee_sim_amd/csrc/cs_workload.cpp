// cs_workload.cpp — seeded synthetic pour-over workloads.
//
// The reference has no RNG, seed or noise and no scene at these sizes; these
// builders produce the inputs (settings, SDF, class mask, porous bed field and
// initial particles) that both this library and the CPU oracle consume, so
// the two run on identical data.  Config axes are z-up; the solver is y-up
// (gravity acts on v.y, shader.rs:1074), so config z maps to internal y.
//
//   V60_SMALL   64^3, dx 2 mm, cone apex 8 cells up, flat bed 12 cells deep
//               (porosity 0.4, k = 1e-9 m^2), 32,768 particles in a jittered
//               cylinder at 8 per cell moving at 0.5 m/s down, 40 RBGS pairs.
//   V60_WORKING 128^3, dx 1 mm, apex 16 cells up, bed 24 cells deep with
//               porosity 0.4 +- 0.05 from seeded value noise (8-cell lattice)
//               and Kozeny-Carman permeability (k(0.4) = 1e-9 m^2);
//               734,004 particles in a pre-wetted slab filling the cone from
//               the apex, 314,572 more poured at 2,000 per substep; RBGS to
//               a 1e-4 relative residual.
//   V60_LARGE   256^3, dx 0.5 mm, apex 32 cells up, bed 48 cells deep with
//               porosity 0.4 +- 0.08 from seeded value noise (16-cell
//               lattice); 5,872,026 particles in the pre-wetted slab,
//               2,516,582 more poured at 16,000 per substep from the 45 degree
//               spout (8,388,608 in all); substep 2.5e-4 s, half the working
//               scene's with half its cell, so the jet crosses the same
//               fraction of a cell per substep; RBGS to 1e-4.  The emitter's
//               disk samples restart every 65,536 (emit_sample_period): the
//               reference's running f32 sample index loses the disk's spread
//               long before 2.5M emitted particles.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <vector>

#include "cs_host.h"

namespace {

// SplitMix64 (Steele, Lea, Flood 2014) as a counter-based generator.
inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
inline float u01(uint64_t seed, uint64_t stream, uint64_t index) {
  uint64_t h = splitmix64(seed ^ splitmix64(stream * 0x632BE59BD9B4E019ull + index));
  return (float)(h >> 40) * (1.0f / 16777216.0f);
}

// Trilinear value noise in [-1, 1] on a lattice of `wavelength` cells.
float value_noise(uint64_t seed, float x, float y, float z, float wavelength) {
  float gx = x / wavelength, gy = y / wavelength, gz = z / wavelength;
  float fx0 = floorf(gx), fy0 = floorf(gy), fz0 = floorf(gz);
  int ix = (int)fx0, iy = (int)fy0, iz = (int)fz0;
  float tx = gx - fx0, ty = gy - fy0, tz = gz - fz0;
  auto lat = [&](int a, int b, int c) {
    uint64_t key = ((uint64_t)(uint32_t)(a + 4096) << 42) ^ ((uint64_t)(uint32_t)(b + 4096) << 21) ^
                   (uint64_t)(uint32_t)(c + 4096);
    return 2.0f * u01(seed, 7, key) - 1.0f;
  };
  float v = 0.0f;
  for (int c = 0; c < 8; ++c) {
    int a = c & 1, b = (c >> 1) & 1, d = (c >> 2) & 1;
    float w = (a ? tx : 1.0f - tx) * (b ? ty : 1.0f - ty) * (d ? tz : 1.0f - tz);
    v += w * lat(ix + a, iy + b, iz + d);
  }
  return v;
}

// Substep of the working-size workload.  The pressure lanes hold i32 fixed
// point clamped at +-8191 (state.rs:15-18, shader.rs:202-209) and the
// projection needs p ~ v dx / dt to stop a flow of speed v within a step; at
// dx = 1 mm that stays representable up to about the speed cap only for
// dt >= ~5e-4 s, which also keeps the CFL number at the cap below one cell
// (86.55 u/s x 5e-4 s = 0.75 dx).
constexpr float kWorkingDt = 5.0e-4f;

struct Owner {
  std::vector<float> sdf;
  std::vector<uint8_t> cls;
  std::vector<int32_t> lookup;
  std::vector<cs_bed_record> records;
  std::vector<float> particles;
  std::vector<float> affine;
};

// Places jittered particles in cell (ix,iy,iz), one per 2x2x2 sub-cell.  With
// fill < 1 (water in a porous bed's pores) the cell gets floor(8*fill + u)
// particles in a random subset of its sub-cells, u uniform in [0,1).
void seed_cell(std::vector<float>& out, uint64_t seed, const cs_params& u, uint32_t ix, uint32_t iy, uint32_t iz,
               const float vel[3], uint32_t limit, float fill = 1.0f) {
  uint64_t cell_key = ((uint64_t)iz * 65536u + iy) * 65536u + ix;
  int order[8] = {0, 1, 2, 3, 4, 5, 6, 7};
  int count = 8;
  if (fill < 1.0f) {
    count = (int)floorf(8.0f * fill + u01(seed, 11, cell_key));
    for (int i = 7; i > 0; --i) {  // Fisher-Yates
      int j = (int)(u01(seed, 12 + (uint64_t)i, cell_key) * (float)(i + 1));
      if (j > i) j = i;
      int t = order[i];
      order[i] = order[j];
      order[j] = t;
    }
  }
  for (int q = 0; q < count && out.size() / 8 < limit; ++q) {
    const int s = order[q];
    uint64_t pid = out.size() / 8;
    float sub[3] = {(float)(s & 1), (float)((s >> 1) & 1), (float)((s >> 2) & 1)};
    uint32_t cell[3] = {ix, iy, iz};
    float p[8];
    for (int c = 0; c < 3; ++c) {
      float j = u01(seed, 1 + c, pid);
      p[c] = u.origin[c] + ((float)cell[c] + (sub[c] + j) * 0.5f) * u.dx;
    }
    p[3] = 1.0f;
    p[4] = vel[0];
    p[5] = vel[1];
    p[6] = vel[2];
    p[7] = 80.0f / 320.0f;
    out.insert(out.end(), p, p + 8);
  }
}

// Everything that differs between the generated scenes.
struct Spec {
  bool working;            // pre-wetted slab + pour (else the small cylinder scene)
  uint32_t n;              // grid side
  float dx_m;              // cell size in metres
  uint32_t apex_cells;     // cone apex height
  uint32_t bed_cells;      // bed depth above the apex
  uint32_t n_total;        // particles in the run
  uint32_t n_initial;      // particles at the start
  uint32_t emit_per_substep;
  float noise_wavelength;  // porosity noise lattice, cells
  float noise_amp;         // porosity noise amplitude
  float kettle_deg;        // spout tilt (informational)
  float aim_dx, aim_dy;    // spout direction in the x-y plane
  uint32_t radius_cells;   // small scene: initial cylinder radius
  uint64_t seed;
  float frame_dt;
  uint32_t substeps_total;
  uint32_t sdf_res;        // SDF texture side
  uint32_t emit_sample_period = 0u;  // cs_settings.emit_sample_period
};

int build(const Spec& sp, cs_workload* out);

}  // namespace

extern "C" int cs_workload_build(int workload, uint64_t seed, uint32_t scale, cs_workload* out) {
  if (!out || (workload != CS_WORKLOAD_V60_SMALL && workload != CS_WORKLOAD_V60_WORKING &&
               workload != CS_WORKLOAD_V60_LARGE))
    return CS_ERR_INVALID;
  if (scale == 0 || scale > 8) return CS_ERR_INVALID;
  if (workload == CS_WORKLOAD_V60_LARGE) {
    const uint32_t s3 = scale * scale * scale;
    Spec sp;
    sp.working = true;
    sp.n = 256u / scale;
    sp.dx_m = 0.0005f * (float)scale;
    sp.apex_cells = 32u / scale;
    sp.bed_cells = 48u / scale;
    sp.n_total = 8388608u / s3;
    sp.n_initial = 5872026u / s3;
    sp.emit_per_substep = std::max(1u, 16000u / s3);
    sp.noise_wavelength = 16.0f / (float)scale;
    sp.noise_amp = 0.08f;
    sp.kettle_deg = 45.0f;
    sp.aim_dx = 1.0f;
    sp.aim_dy = -1.0f;
    sp.radius_cells = 1u;
    sp.seed = seed;
    sp.frame_dt = kWorkingDt * 0.5f * (float)scale;
    sp.substeps_total = 500u;
    sp.sdf_res = sp.n;
    sp.emit_sample_period = 65536u;
    return build(sp, out);
  }
  const bool working = workload == CS_WORKLOAD_V60_WORKING;
  const uint32_t s3 = scale * scale * scale;
  Spec sp;
  sp.working = working;
  sp.n = (working ? 128u : 64u) / scale;
  sp.dx_m = (working ? 0.001f : 0.002f) * (float)scale;
  sp.apex_cells = (working ? 16u : 8u) / scale;
  sp.bed_cells = (working ? 24u : 12u) / scale;
  sp.n_total = (working ? 1048576u : 32768u) / s3;
  sp.n_initial = working ? 734004u / s3 : sp.n_total;
  sp.emit_per_substep = std::max(1u, 2000u / s3);
  sp.noise_wavelength = 8.0f / (float)scale;
  sp.noise_amp = 0.05f;
  sp.kettle_deg = 45.0f;
  sp.aim_dx = 1.0f;  // 45 degrees below horizontal
  sp.aim_dy = -1.0f;
  sp.radius_cells = std::max(1u, 4u / scale);
  sp.seed = seed;
  sp.frame_dt = working ? kWorkingDt : 1.0e-4f;
  sp.substeps_total = working ? 1000u : 200u;
  sp.sdf_res = 128u;
  return build(sp, out);
}

extern "C" int cs_workload_build_pour(const cs_pour_params* p, cs_workload* out) {
  if (!p || !out) return CS_ERR_INVALID;
  if (p->side < 32 || p->side > 256 || p->side % 16 != 0) return CS_ERR_INVALID;
  if (p->particles == 0 || p->pour_per_substep == 0) return CS_ERR_INVALID;
  if (!(p->kettle_angle_deg > 0.0f && p->kettle_angle_deg < 90.0f)) return CS_ERR_INVALID;
  Spec sp;
  sp.working = true;
  sp.n = p->side;
  sp.dx_m = 0.128f / (float)p->side;  // the working scene's 128 mm domain
  sp.apex_cells = p->side / 8u;
  sp.bed_cells = 3u * p->side / 16u;
  sp.n_total = p->particles;
  sp.n_initial = (uint32_t)((double)p->particles * 0.7 + 0.5);
  sp.emit_per_substep = p->pour_per_substep;
  sp.noise_wavelength = 8.0f;
  sp.noise_amp = 0.05f;
  sp.kettle_deg = p->kettle_angle_deg;
  const float a = p->kettle_angle_deg * 3.14159265f / 180.0f;
  sp.aim_dx = cosf(a);
  sp.aim_dy = -sinf(a);
  sp.radius_cells = 1u;
  sp.seed = p->seed;
  sp.frame_dt = kWorkingDt;
  sp.substeps_total = 500u;
  sp.sdf_res = 128u;
  return build(sp, out);
}

namespace {

int build(const Spec& sp, cs_workload* out) {
  memset(out, 0, sizeof(*out));
  const bool working = sp.working;
  const uint64_t seed = sp.seed;
  const float supm = cs::sim_units_per_meter();
  const uint32_t n = sp.n;
  const float dx_m = sp.dx_m;
  const float dx = dx_m * supm;
  const uint32_t apex_cells = sp.apex_cells;
  const uint32_t bed_cells = sp.bed_cells;
  const uint32_t n_total = sp.n_total;
  const uint32_t n_initial = sp.n_initial;
  const uint32_t n_pour = n_total - n_initial;

  cs_settings& s = out->settings;
  cs::settings_preset(CS_PRESET_FREE_STREAM, &s);
  const float side = (float)n * dx;
  for (int c = 0; c < 3; ++c) {
    s.bounds_size[c] = side;
    s.grid_dims[c] = n;
  }
  s.max_particles = n_total;
  s.substeps = 1u;
  s.viscosity = 1.0e-6f * supm * supm;  // nu = 1e-6 m^2/s in scene units
  s.render_radius = dx * 0.7f;
  s.particle_vol_factor = 0.125f;        // 8 particles per cell at rest
  s.wall_thickness = 2.0f * dx;          // SDF zero level one cell inside the cone
  s.geometry.cone_enabled = 0u;          // the reference's analytic V60/cup fallbacks
  s.geometry.cup_enabled = 0u;           // describe its own scene, not this one
  s.geometry.filter_enabled = 0u;
  s.sdf_res = sp.sdf_res;
  s.emit_sample_period = sp.emit_sample_period;
  const float half = side * 0.5f;
  const float apex_y = -half + (float)apex_cells * dx;
  const float tan60 = 1.7320508f;
  s.num_obstacles = 1u;
  memset(s.obstacles, 0, sizeof(s.obstacles));
  s.obstacles[0].kind = CS_OBSTACLE_TRUNCATED_CONE;
  s.obstacles[0].bot_y = apex_y;
  s.obstacles[0].bot_radius = 0.0f;
  s.obstacles[0].top_y = half;
  s.obstacles[0].top_radius = (half - apex_y) * tan60;
  s.has_bed_field = 1u;
  s.initial_kettle_angle_deg = working ? sp.kettle_deg : 0.0f;
  s.emitter_mode = CS_EMIT_FIXED_COUNT;
  s.emit_per_substep = 0u;
  s.emit_total = 0u;
  if (working) {
    s.rbgs_mode = CS_RBGS_TOLERANCE;
    s.rbgs_tolerance = 1.0e-4f;
    s.rbgs_check_every = 8u;
    s.rbgs_max_iters = 20000u;
    // 2,000 particles per substep (scaled with the particle count) from a
    // spout tilted 45 degrees, at half the speed cap; the jet radius keeps
    // the poured column at 8 particles per cell: N = 8 pi R^2 (v dt / dx).
    s.emit_per_substep = sp.emit_per_substep;
    s.emit_total = n_pour;
    s.emit_speed = 0.5f * cs::max_velocity();
    const float dt = sp.frame_dt;
    float cells = sqrtf((float)s.emit_per_substep / (8.0f * 3.14159265f * (s.emit_speed * dt / dx)));
    s.emit_jet_radius = cells * dx;
    s.spout.origin[0] = -0.30f * side;
    s.spout.origin[1] = 0.30f * side;
    s.spout.origin[2] = 0.0f;
    cs::spout_aim_at(&s.spout, s.spout.origin[0] + sp.aim_dx, s.spout.origin[1] + sp.aim_dy, 0.0f);
  } else {
    s.rbgs_mode = CS_RBGS_FIXED_PAIRS;
    s.pressure_rbgs_pairs = 40u;
  }

  std::unique_ptr<Owner> own(new Owner());
  const size_t nc = (size_t)n * n * n;
  own->sdf.resize((size_t)s.sdf_res * s.sdf_res * s.sdf_res);
  cs::generate_sdf(s, own->sdf.data());
  own->cls.resize(nc);
  cs::generate_sdf_class(s, own->sdf.data(), own->cls.data());
  cs_params u;
  cs::make_params(s, 0, 0, 1.0e-4f, &u);

  // porous bed: open cells of the cone from the apex up `bed_cells` layers
  own->lookup.assign(nc, -1);
  const float bed_top = apex_y + (float)bed_cells * dx;
  const float k0 = 1.0e-9f;
  const float d_um = 1.0e6f * sqrtf(k0 * 180.0f * 0.36f / 0.064f);  // k(0.4) = 1e-9 m^2
  for (uint32_t iz = 0; iz < n; ++iz)
    for (uint32_t iy = apex_cells; iy < apex_cells + bed_cells && iy < n; ++iy)
      for (uint32_t ix = 0; ix < n; ++ix) {
        size_t c = ((size_t)iz * n + iy) * n + ix;
        if (own->cls[c]) continue;
        cs_bed_record r;
        memset(&r, 0, sizeof(r));
        float ccy = u.origin[1] + ((float)iy + 0.5f) * dx;
        if (working) {
          float nz = value_noise(seed + 1, (float)ix + 0.5f, (float)iy + 0.5f, (float)iz + 0.5f, sp.noise_wavelength);
          r.porosity = 0.4f + sp.noise_amp * nz;
          r.permeability = cs::kozeny_carman_permeability(d_um, r.porosity);
        } else {
          r.porosity = 0.4f;
          r.permeability = k0;
        }
        // height of the bed sample a particle lattice would stamp on this
        // cell: one cell above inside the bed, the top layer's own centre at
        // the surface
        r.sample_y = cs_min(ccy + dx, bed_top - 0.5f * dx);
        own->lookup[c] = (int32_t)own->records.size();
        own->records.push_back(r);
      }

  // initial particles
  const float vel_small[3] = {0.0f, -0.5f * supm, 0.0f};
  const float vel_rest[3] = {0.0f, 0.0f, 0.0f};
  own->particles.reserve((size_t)n_initial * 8);
  if (working) {
    // pre-wetted slab filling the open cone from the apex upward: bed cells
    // hold water only in their pores (porosity x 8 particles, the pore
    // capacity of shader.rs:654-658), open cells above the bed 8 per cell
    for (uint32_t iy = apex_cells; iy < n && own->particles.size() / 8 < n_initial; ++iy)
      for (uint32_t iz = 0; iz < n && own->particles.size() / 8 < n_initial; ++iz)
        for (uint32_t ix = 0; ix < n && own->particles.size() / 8 < n_initial; ++ix) {
          size_t c = ((size_t)iz * n + iy) * n + ix;
          if (own->cls[c]) continue;
          float fill = own->lookup[c] >= 0 ? own->records[own->lookup[c]].porosity : 1.0f;
          seed_cell(own->particles, seed, u, ix, iy, iz, vel_rest, n_initial, fill);
        }
  } else {
    // jittered cylinder above the bed centre; radius 4 cells as specified,
    // widened to the smallest radius whose column fits below the top margin
    const uint32_t y0 = apex_cells + bed_cells + 1u;
    const uint32_t top = n - 3u;
    uint32_t r_cells = sp.radius_cells;
    for (;; ++r_cells) {
      uint32_t per_layer = 0;
      float rr = (float)r_cells;
      for (uint32_t iz = 0; iz < n; ++iz)
        for (uint32_t ix = 0; ix < n; ++ix) {
          float x = (float)ix + 0.5f - 0.5f * (float)n, z = (float)iz + 0.5f - 0.5f * (float)n;
          if (x * x + z * z <= rr * rr) ++per_layer;
        }
      uint32_t layers = (n_initial / 8 + per_layer - 1) / std::max(per_layer, 1u);
      if (y0 + layers <= top || r_cells >= n / 2) break;
    }
    const float rr = (float)r_cells;
    for (uint32_t iy = y0; iy < n && own->particles.size() / 8 < n_initial; ++iy)
      for (uint32_t iz = 0; iz < n && own->particles.size() / 8 < n_initial; ++iz)
        for (uint32_t ix = 0; ix < n && own->particles.size() / 8 < n_initial; ++ix) {
          float x = (float)ix + 0.5f - 0.5f * (float)n, z = (float)iz + 0.5f - 0.5f * (float)n;
          if (x * x + z * z > rr * rr) continue;
          if (own->cls[((size_t)iz * n + iy) * n + ix]) continue;
          seed_cell(own->particles, seed, u, ix, iy, iz, vel_small, n_initial);
        }
  }
  uint32_t placed = (uint32_t)(own->particles.size() / 8);
  own->affine.assign((size_t)placed * 12, 0.0f);

  cs_scene& sc = out->scene;
  sc.sdf = own->sdf.data();
  sc.sdf_class = own->cls.data();
  sc.bed_lookup = own->lookup.data();
  sc.bed_records = own->records.data();
  sc.num_bed_records = (uint32_t)own->records.size();
  sc.particles = own->particles.data();
  sc.affine = own->affine.data();
  sc.num_particles = placed;
  out->frame_dt = sp.frame_dt;
  out->substeps_total = sp.substeps_total;
  out->total_particles = placed + n_pour;
  out->_owner = own.release();
  return CS_OK;
}

}  // namespace

extern "C" int cs_workload_free(cs_workload* w) {
  if (!w) return CS_ERR_INVALID;
  delete static_cast<Owner*>(w->_owner);
  memset(w, 0, sizeof(*w));
  return CS_OK;
}
