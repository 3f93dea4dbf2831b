// cs_bedinit.cpp — the granular bed's initial particles (bed.rs:7-288): the
// default bed and filter, the bed seated in the filter, the lattice of bed
// particles with its downsampling to the target count, and the initial cell
// lookup.  Host only; the result is scene input for cs_sim_create.
#include <math.h>
#include <string.h>
#include <vector>
#include "cs_sim_impl.h"

namespace {

// f32::clamp for ordered bounds (bed.rs calls it only with lo <= hi)
float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }

// order_bounds (bed.rs:83-90)
void order_bounds(float lo, float hi, float* out_lo, float* out_hi) {
  if (lo <= hi) {
    *out_lo = lo;
    *out_hi = hi;
  } else {
    const float mid = (lo + hi) * 0.5f;
    *out_lo = mid;
    *out_hi = mid;
  }
}

// FilterConfig::radius_at_y / inner_radius_at_y (filter.rs:29-37)
float filter_radius_at_y(const cs_filter_config& f, float y) {
  const float height = fmaxf(f.top_y - f.bot_y, 1e-6f);
  const float t = clampf((y - f.bot_y) / height, 0.0f, 1.0f);
  return f.bot_radius + (f.top_radius - f.bot_radius) * t;
}
float filter_inner_radius_at_y(const cs_filter_config& f, float y) {
  return fmaxf(filter_radius_at_y(f, y) - f.thickness, f.hole_radius);
}

// kozeny_carman_permeability_m2 (brew_config.rs:92-96)
float kozeny_carman_m2(float grind_um, float porosity) {
  const float d_m = grind_um * 1.0e-6f;
  const float solid = 1.0f - porosity;
  return d_m * d_m * porosity * porosity * porosity / (180.0f * solid * solid);
}

// Rust's `f as i32`: saturating, NaN to 0
int32_t f2i(float v) {
  if (v != v) return 0;
  if (v >= 2147483648.0f) return INT32_MAX;
  if (v < -2147483648.0f) return INT32_MIN;
  return (int32_t)v;
}

}  // namespace

extern "C" {

void cs_filter_config_default(cs_filter_config* out) {
  if (!out) return;
  memset(out, 0, sizeof(*out));
  out->center[1] = -0.35f;
  out->top_y = 2.75f;
  out->bot_y = -3.02f;
  out->top_radius = 4.10f;
  out->bot_radius = 0.0f;
  out->thickness = 0.08f;
  out->hole_radius = 0.0f;
}

void cs_bed_config_default(cs_bed_config* out) {
  if (!out) return;
  memset(out, 0, sizeof(*out));
  out->center[1] = -0.35f;
  out->top_y = 0.45f;
  out->bot_y = -2.70f;
  out->top_radius = 2.65f;
  out->bot_radius = 0.18f;
  out->num_particles = 12000u;                              // DEFAULT_BREW.bed_particle_samples
  out->initial_porosity = 0.40f;                            // DEFAULT_BREW.bed_porosity
  out->initial_permeability = kozeny_carman_m2(450.0f, 0.40f);  // grind_diameter_um 450
  out->extractable_mass = 0.15f;
}

void cs_bed_config_seated_in_filter(const cs_filter_config* filter, cs_bed_config* out) {
  if (!filter || !out) return;
  const cs_filter_config& f = *filter;
  cs_bed_config b;
  cs_bed_config_default(&b);
  b.center[0] = f.center[0];
  b.center[1] = f.center[1];
  b.center[2] = f.center[2];
  const float filter_bot_abs = f.center[1] + f.bot_y;
  const float filter_top_abs = f.center[1] + f.top_y;
  float lo, hi;
  order_bounds(filter_bot_abs + 0.6f, filter_top_abs - 0.35f, &lo, &hi);
  const float top_abs = clampf(b.center[1] + b.top_y, lo, hi);
  order_bounds(filter_bot_abs + 0.4f, top_abs - 1.4f, &lo, &hi);
  const float bot_abs = clampf(b.center[1] + b.bot_y, lo, hi);
  b.top_y = top_abs - b.center[1];
  b.bot_y = bot_abs - b.center[1];
  const float top_local = top_abs - f.center[1];
  const float bot_local = bot_abs - f.center[1];
  order_bounds(f.hole_radius + 0.8f, f.top_radius - f.thickness - 0.1f, &lo, &hi);
  b.top_radius = clampf(filter_inner_radius_at_y(f, top_local) - 0.18f, lo, hi);
  order_bounds(f.hole_radius + 0.14f, b.top_radius - 0.25f, &lo, &hi);
  b.bot_radius = clampf(filter_inner_radius_at_y(f, bot_local) - 0.12f, lo, hi);
  *out = b;
}

int cs_bed_init(const cs_bed_config* config, const uint32_t grid_dims[3], const float bounds_size[3],
                uint32_t capacity, float* particles, float* affine, float* extracts, int32_t* cell_lookup,
                uint32_t* count) {
  if (!config || !grid_dims || !bounds_size || !count) return cs_fail(CS_ERR_INVALID, "cs_bed_init: null argument");
  const cs_bed_config& c = *config;
  const uint32_t gx = grid_dims[0], gy = grid_dims[1], gz = grid_dims[2];
  const size_t nc = (size_t)gx * gy * gz;
  *count = 0;
  if (cell_lookup)
    for (size_t i = 0; i < nc; ++i) cell_lookup[i] = -1;
  const float height = c.top_y - c.bot_y;
  if (height < 1e-6f) return CS_OK;
  const float pi = 3.14159265358979323846f;
  const float avg_radius = (c.top_radius + c.bot_radius) * 0.5f;
  const float q = c.bot_radius / avg_radius;
  const float volume = pi * avg_radius * avg_radius * height / 3.0f * (1.0f + q + q * q);
  const float spacing = cbrtf(volume / (float)(c.num_particles > 0u ? c.num_particles : 1u));
  const int32_t nx = f2i(ceilf((c.top_radius * 2.0f) / spacing));
  const int32_t ny = f2i(ceilf(height / spacing));
  const int32_t nz = nx;
  if (nx <= 0 || ny <= 0) return CS_OK;
  // lattice points in generation order; `key` is the lattice cell of each point
  std::vector<float> part, aff, ext;
  std::vector<int32_t> key;
  for (int32_t iy = 0; iy < ny; ++iy) {
    const float base_y = c.center[1] + c.bot_y + ((float)iy + 0.5f) * spacing;
    const float t = (base_y - (c.center[1] + c.bot_y)) / height;
    const float max_r = c.bot_radius + (c.top_radius - c.bot_radius) * t;
    for (int32_t ix = 0; ix < nx; ++ix) {
      const float x = c.center[0] - max_r + ((float)ix + 0.5f) * spacing;
      for (int32_t iz = 0; iz < nz; ++iz) {
        const float z = c.center[2] - max_r + ((float)iz + 0.5f) * spacing;
        const float dx = x - c.center[0];
        const float dz = z - c.center[2];
        const float r = sqrtf(dx * dx + dz * dz);
        if (r > max_r) continue;
        // the rounded bottom plug (bed.rs:154-159)
        const float cap_t = clampf(t / 0.16f, 0.0f, 1.0f);
        const float radial_t = max_r > 1e-6f ? r / max_r : 0.0f;
        const float one_m = 1.0f - cap_t;
        const float edge_lift = spacing * 0.75f * (one_m * one_m) * (radial_t * radial_t);
        const float y = base_y + edge_lift;
        key.push_back((iy * nx + ix) * nz + iz);
        const float p8[8] = {x, y, z, 1.0f, 0.0f, 0.0f, 0.0f, 1.0f};
        const float a12[12] = {0.0f, 0.0f, 0.0f, 1.0f, x, y, z, 0.0f, y, 0.0f, 0.0f, 0.0f};
        const float e8[8] = {0.0f, c.initial_porosity, c.initial_permeability, 0.0f,
                             c.extractable_mass, 0.0f, 93.0f, 0.0f};
        part.insert(part.end(), p8, p8 + 8);
        aff.insert(aff.end(), a12, a12 + 12);
        ext.insert(ext.end(), e8, e8 + 8);
      }
    }
  }
  // lattice cell -> particle index (-1: no particle)
  std::vector<int32_t> lattice((size_t)nx * ny * nz, -1);
  size_t n = key.size();
  std::vector<size_t> src(n);
  for (size_t i = 0; i < n; ++i) src[i] = i;
  const size_t target = c.num_particles;
  if (n > target) {
    // evenly strided downsampling (bed.rs:184-206)
    src.resize(target);
    for (size_t o = 0; o < target; ++o) {
      const float pick = floorf((((float)o + 0.5f) * (float)n) / (float)target);
      const float capped = fminf(pick, (float)(n - 1));
      src[o] = (size_t)capped;
    }
    n = target;
  }
  for (size_t o = 0; o < n; ++o) lattice[key[src[o]]] = (int32_t)o;
  *count = (uint32_t)n;
  if ((particles || affine || extracts) && capacity && n > capacity)
    return cs_fail(CS_ERR_RANGE, "cs_bed_init: more bed particles than capacity");
  for (size_t o = 0; o < n; ++o) {
    if (particles) memcpy(particles + o * 8, &part[src[o] * 8], 8 * sizeof(float));
    if (affine) memcpy(affine + o * 12, &aff[src[o] * 12], 12 * sizeof(float));
    if (extracts) memcpy(extracts + o * 8, &ext[src[o] * 8], 8 * sizeof(float));
  }
  if (!cell_lookup) return CS_OK;
  // build_cell_lookup (bed.rs:208-288): the nearest lattice particle among the
  // 3x3x3 lattice cells around each bed cell's guess
  const float ox = -bounds_size[0] * 0.5f, oy = -bounds_size[1] * 0.5f, oz = -bounds_size[2] * 0.5f;
  const float cdx = bounds_size[0] / (float)gx;
  const float bed_bottom = c.center[1] + c.bot_y;
  for (uint32_t iz = 0; iz < gz; ++iz)
    for (uint32_t iy = 0; iy < gy; ++iy)
      for (uint32_t ix = 0; ix < gx; ++ix) {
        const float px = ox + ((float)ix + 0.5f) * cdx;
        const float py = oy + ((float)iy + 0.5f) * cdx;
        const float pz = oz + ((float)iz + 0.5f) * cdx;
        if (py < bed_bottom || py > c.center[1] + c.top_y) continue;
        const float t = clampf((py - bed_bottom) / height, 0.0f, 1.0f);
        const float max_r = c.bot_radius + (c.top_radius - c.bot_radius) * t;
        const float dxr = px - c.center[0];
        const float dzr = pz - c.center[2];
        if (sqrtf(dxr * dxr + dzr * dzr) > max_r) continue;
        const int32_t iy_guess = f2i(roundf(((py - bed_bottom) / spacing) - 0.5f));
        const int32_t ix_guess = f2i(roundf(((px - (c.center[0] - max_r)) / spacing) - 0.5f));
        const int32_t iz_guess = f2i(roundf(((pz - (c.center[2] - max_r)) / spacing) - 0.5f));
        int32_t best = -1;
        float best_d2 = INFINITY;
        for (int dy = -1; dy <= 1; ++dy)
          for (int dxi = -1; dxi <= 1; ++dxi)
            for (int dzi = -1; dzi <= 1; ++dzi) {
              const int32_t kx = std::min(std::max(ix_guess + dxi, 0), nx - 1);
              const int32_t ky = std::min(std::max(iy_guess + dy, 0), ny - 1);
              const int32_t kz = std::min(std::max(iz_guess + dzi, 0), nz - 1);
              const int32_t pidx = lattice[((size_t)ky * nx + kx) * nz + kz];
              if (pidx < 0) continue;
              const float ly = bed_bottom + ((float)ky + 0.5f) * spacing;
              const float ly_t = clampf((ly - bed_bottom) / height, 0.0f, 1.0f);
              const float ly_r = c.bot_radius + (c.top_radius - c.bot_radius) * ly_t;
              const float lx = c.center[0] - ly_r + ((float)kx + 0.5f) * spacing;
              const float lz = c.center[2] - ly_r + ((float)kz + 0.5f) * spacing;
              const float ddx = px - lx, ddy = py - ly, ddz = pz - lz;
              const float d2 = ddx * ddx + ddy * ddy + ddz * ddz;
              if (d2 < best_d2) {
                best_d2 = d2;
                best = pidx;
              }
            }
        cell_lookup[((size_t)iz * gy + iy) * gx + ix] = best;
      }
  return CS_OK;
}

}  // extern "C"
