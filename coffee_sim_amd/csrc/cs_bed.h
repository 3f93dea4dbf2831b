// cs_bed.h — one cell of the porous bed's table (the per-cell factors of
// shader.rs:380-449 and 621-666), shared by the host builder (a bed field handed
// over with the scene, built once) and the device builder (a bed of particles,
// re-stamped every substep) so both give the same bits.
#pragma once
#include "cs_common.h"

// kozeny_porosity_factor (shader.rs:429-433)
CS_HD float cs_bed_kozeny_factor(float phi_in) {
  const float phi = cs_clamp(phi_in, 0.08f, 0.82f);
  const float solid = cs_max(1.0f - phi, 1e-4f);
  return (phi * phi * phi) / (solid * solid);
}
// bed_compacted_porosity (shader.rs:435-440)
CS_HD float cs_bed_compacted_porosity(float porosity, float compaction) {
  const float base = cs_clamp(porosity, 0.18f, 0.70f);
  const float comp = cs_clamp(compaction, 0.0f, 2.25f);
  return cs_clamp(base - comp * 0.12f, 0.16f, base);
}

// bed_matrix_velocity_cell (shader.rs:380-407): the mean record velocity over the
// 3x3x3 block around (x, y, z).  at(x, y, z) is the record id of a cell (< 0 for
// none or off the grid); vel(id, v) loads that record's velocity.
template <class At, class Vel>
CS_HD void cs_bed_matrix_velocity(int x, int y, int z, const At& at, const Vel& vel, float* out) {
  float sx = 0.0f, sy = 0.0f, sz = 0.0f, w = 0.0f;
  for (int di = -1; di <= 1; ++di)
    for (int dj = -1; dj <= 1; ++dj)
      for (int dk = -1; dk <= 1; ++dk) {
        const int32_t id = at(x + di, y + dj, z + dk);
        if (id < 0) continue;
        float v[3];
        vel(id, v);
        sx = sx + v[0];
        sy = sy + v[1];
        sz = sz + v[2];
        w += 1.0f;
      }
  if (w <= 0.0f) {
    out[0] = out[1] = out[2] = 0.0f;
  } else {
    out[0] = sx / w;
    out[1] = sy / w;
    out[2] = sz / w;
  }
}

// The table entry of cell (ix, iy, iz), whose record is b.
template <class At, class Vel>
CS_HD void cs_bed_cell_build(const cs_params& u, uint32_t gx, uint32_t gy, uint32_t gz, int ix, int iy, int iz,
                             const cs_bed_record& b, const At& at, const Vel& vel, cs_bed_cell* t) {
  const float dx = u.dx, inv_dx = u.inv_dx;
  // bed_compacted_permeability (shader.rs:442-449)
  const float base_p = cs_clamp(b.porosity, 0.18f, 0.70f);
  const float compacted = cs_bed_compacted_porosity(b.porosity, b.compaction);
  const float ratio = cs_bed_kozeny_factor(compacted) / cs_max(cs_bed_kozeny_factor(base_p), 1e-8f);
  t->k_record = cs_max(b.permeability * ratio, u.min_k);
  t->v_record[0] = b.velocity[0];
  t->v_record[1] = b.velocity[1];
  t->v_record[2] = b.velocity[2];
  // bed_pore_projection_target_divergence factors (shader.rs:628-662)
  const float ccy = u.origin[1] + ((float)iy + 0.5f) * dx;
  const float base = compacted;
  const float sat = cs_clamp(b.saturation, 0.0f, 1.0f);
  const float comp = cs_clamp(b.compaction, 0.0f, 2.25f);
  const float surface_t = cs_smoothstep(-0.25f, 0.75f, (ccy - b.sample_y) * inv_dx);
  const float open = cs_clamp((1.0f - 0.65f * sat) * (1.0f - 0.12f * comp), 0.20f, 1.0f);
  const float opening = surface_t * cs_clamp(1.0f, 0.0f, 1.0f) * open;  // bed_surface_void_scale = 1
  const float surf_por = cs_min(base + 0.35f, 0.82f);
  const float eff = cs_mix(base, surf_por, opening);
  const float occ = cs_mix(1.0f, 0.35f, opening);
  const float satc = cs_mix(1.0f, 0.32f, sat);
  const float compc = cs_mix(1.0f, 0.58f, cs_clamp(comp / 1.5f, 0.0f, 1.0f));
  const float cv = dx * dx * dx;
  t->pore_capacity = eff * cv * cs_max(1.0f, 0.0f) * satc * compc;  // bed_pore_capacity_scale = 1
  t->overfill_resp = 18.0f * occ * cs_mix(1.0f, 1.45f, sat);      // bed_pore_overfill_alpha = 18
  t->porosity = cs_clamp(compacted, 0.08f, 1.0f);
  t->solid_fraction = 1.0f - t->porosity;
  // bed_velocity_divergence (shader.rs:419-427) over bed_velocity_load_cell
  // (shader.rs:409-417: zero off the grid)
  auto bvl = [&](int x, int y, int z, float* o) {
    if (x < 0 || y < 0 || z < 0 || x >= (int)gx || y >= (int)gy || z >= (int)gz) {
      o[0] = o[1] = o[2] = 0.0f;
      return;
    }
    cs_bed_matrix_velocity(x, y, z, at, vel, o);
  };
  float a[3];
  bvl(ix - 1, iy, iz, a);
  const float vxm = a[0];
  bvl(ix + 1, iy, iz, a);
  const float vxp = a[0];
  bvl(ix, iy - 1, iz, a);
  const float vym = a[1];
  bvl(ix, iy + 1, iz, a);
  const float vyp = a[1];
  bvl(ix, iy, iz - 1, a);
  const float vzm = a[2];
  bvl(ix, iy, iz + 1, a);
  const float vzp = a[2];
  t->solid_div = 0.5f * inv_dx * ((vxp - vxm) + (vyp - vym) + (vzp - vzm));
  cs_bed_matrix_velocity(ix, iy, iz, at, vel, t->bmv);
}

// The record of bed particle `id` (a BedExtract, shader.rs:35-38, and the
// particle): porosity, permeability and compaction are bed.yzw, saturation is
// extract.w, the sample height is the particle's y and the velocity its velocity.
CS_HD cs_bed_record cs_bed_particle_record(const float* ext8, const float* part8) {
  cs_bed_record r;
  r.porosity = ext8[1];
  r.permeability = ext8[2];
  r.compaction = ext8[3];
  r.saturation = ext8[7];
  r.sample_y = part8[1];
  r.velocity[0] = part8[4];
  r.velocity[1] = part8[5];
  r.velocity[2] = part8[6];
  return r;
}
