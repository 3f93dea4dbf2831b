/* orc_step.c — CPU oracle, device half: the WGSL compute kernels of
 * crates/sim-wasm/src/mpm_3d/shader.rs restated pass by pass, the substep order
 * of MpmSim3D::step_frame (mod.rs:264-430) and the residual-driven pressure
 * solve.  TEST INFRASTRUCTURE ONLY (see orc.h). */
#include <stdio.h>

#include "orc_internal.h"

typedef struct { int32_t x, y, z; } i3;

static inline size_t cidx(const orc_params* u, uint32_t ix, uint32_t iy, uint32_t iz) {
  return (size_t)iz * u->gx * u->gy + (size_t)iy * u->gx + ix; /* shader.rs:148-150 */
}
static inline int on_grid(const orc_params* u, int32_t x, int32_t y, int32_t z) {
  return !(x < 0 || y < 0 || z < 0) && !((uint32_t)x >= u->gx || (uint32_t)y >= u->gy || (uint32_t)z >= u->gz);
}
static inline void decompose(const orc_params* u, size_t idx, uint32_t* ix, uint32_t* iy, uint32_t* iz) {
  size_t plane = (size_t)u->gx * u->gy;
  *iz = (uint32_t)(idx / plane);
  size_t rem = idx % plane;
  *iy = (uint32_t)(rem / u->gx);
  *ix = (uint32_t)(rem % u->gx);
}

/* Built with -fopenmp (liborc_omp.so, the CPU timing baseline) the cell and
 * particle loops run in parallel: every pass writes only its own cell or
 * particle except the P2G scatter and the metric counters, which then use
 * atomics (integer addition, so the lanes and counts are the same).  Without
 * OpenMP (liborc.so, the checker) the same code runs serially. */
#ifdef _OPENMP
#define ORC_PARALLEL_FOR _Pragma("omp parallel for schedule(static)")
#define ORC_COUNT(p) __atomic_fetch_add((p), 1u, __ATOMIC_RELAXED)
#define ORC_LANE_ADD(p, v) __atomic_fetch_add((uint32_t*)(p), (uint32_t)(v), __ATOMIC_RELAXED)
static inline void orc_max_u32(uint32_t* p, uint32_t v) {
  uint32_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
  while (v > cur && !__atomic_compare_exchange_n(p, &cur, v, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
  }
}
#else
#define ORC_PARALLEL_FOR
#define ORC_COUNT(p) (*(p) += 1u)
#define ORC_LANE_ADD(p, v) (*(p) = o_wrap_add(*(p), (v)))
static inline void orc_max_u32(uint32_t* p, uint32_t v) {
  if (v > *p) *p = v;
}
#endif

/* lane helpers (shader.rs:152-164) */
#define LANE(S, l, c) ((S)->grid[(size_t)(l) * (S)->nc + (c)])
static inline int32_t kind_load(orc_sim* S, size_t c) { return LANE(S, 3, c); }
static inline float pressure_load(orc_sim* S, const orc_params* u, size_t c) {
  return (float)LANE(S, 0, c) * u->inv_fp_scale;
}
/* shader.rs:202-209 */
static inline void pressure_store(orc_sim* S, const orc_params* u, size_t c, float v) {
  float lim = u->pressure_clamp;
  float cl = o_clamp(v, -lim, lim);
  if (cl != v) ORC_COUNT(&S->metrics[O_M_P_CLAMP]);
  LANE(S, 0, c) = o_f2i(cl * u->fp_scale);
}
static inline float packing_load(orc_sim* S, const orc_params* u, size_t c) {
  return (float)LANE(S, 2, c) * u->inv_fp_scale;
}
/* shader.rs:215-222 */
static inline void packing_store(orc_sim* S, const orc_params* u, size_t c, float v) {
  float cl = o_clamp(v, 0.0f, u->pressure_clamp);
  if (cl != v) ORC_COUNT(&S->metrics[O_M_P_CLAMP]);
  LANE(S, 2, c) = o_f2i(cl * u->fp_scale);
}
static inline float div_load(orc_sim* S, const orc_params* u, size_t c) {
  return (float)LANE(S, 1, c) * u->inv_fp_scale;
}
/* shader.rs:352-359 */
static inline void div_store(orc_sim* S, const orc_params* u, size_t c, float v) {
  float lim = u->div_clamp;
  float cl = o_clamp(v, -lim, lim);
  if (cl != v) ORC_COUNT(&S->metrics[O_M_DIV_CLAMP]);
  LANE(S, 1, c) = o_f2i(cl * u->fp_scale);
}
static inline float rest_vol(orc_sim* S, const orc_params* u, size_t c) {
  return (float)LANE(S, 4, c) * u->inv_fp_scale;
}
static inline float cur_vol(orc_sim* S, const orc_params* u, size_t c) {
  return (float)LANE(S, 5, c) * u->inv_fp_scale;
}
/* shader.rs:274-283 */
static inline void vscratch_store(orc_sim* S, const orc_params* u, size_t c, v3 v) {
  float speed = vlen(v);
  if (speed > u->vel_cap) v = vmul(v, u->vel_cap / speed);
  LANE(S, 1, c) = o_f2i(v.x * u->fp_scale);
  LANE(S, 2, c) = o_f2i(v.y * u->fp_scale);
  LANE(S, 3, c) = o_f2i(v.z * u->fp_scale);
}
static inline v3 vscratch_load(orc_sim* S, const orc_params* u, size_t c) {
  return V((float)LANE(S, 1, c) * u->inv_fp_scale, (float)LANE(S, 2, c) * u->inv_fp_scale,
           (float)LANE(S, 3, c) * u->inv_fp_scale);
}

static inline int is_fluid(int32_t k) { return k == O_CELL_INTERIOR || k == O_CELL_SURFACE || k == O_CELL_BED; }
static inline int is_solid(int32_t k) { return k == O_CELL_SOLID; }

/* shader.rs:293-304 */
static float fill_fraction(orc_sim* S, const orc_params* u, size_t c, int32_t kind) {
  if (kind == O_CELL_INTERIOR || kind == O_CELL_BED) return 1.0f;
  if (kind != O_CELL_SURFACE) return 0.0f;
  float cv = o_max(u->dx * u->dx * u->dx, 1e-8f);
  float dep = o_max(rest_vol(S, u, c), cur_vol(S, u, c)) / cv;
  return o_clamp(dep, 0.0f, 1.0f);
}
/* shader.rs:306-323 */
static float face_weight(orc_sim* S, const orc_params* u, int32_t self_kind, float self_fill, size_t nc, int32_t nk) {
  if (is_solid(nk)) return 0.0f;
  if (self_kind == O_CELL_BED || nk == O_CELL_BED) return 1.0f;
  if (!is_fluid(nk)) return self_fill;
  return o_min(self_fill, fill_fraction(S, u, nc, nk));
}

/* ---- SDF sampling (shader.rs:673-715) -------------------------------------- */
static float load_sdf_texel(orc_sim* S, const orc_params* u, int32_t x, int32_t y, int32_t z) {
  int32_t mx = o_f2i(u->sdf_res) - 1;
  x = x < 0 ? 0 : (x > mx ? mx : x);
  y = y < 0 ? 0 : (y > mx ? mx : y);
  z = z < 0 ? 0 : (z > mx ? mx : z);
  size_t r = S->sdf_res;
  return S->sdf[(size_t)z * r * r + (size_t)y * r + (size_t)x];
}
static float sample_sdf(orc_sim* S, const orc_params* u, v3 p) {
  float res = u->sdf_res;
  float bsx = u->bmax[0] * 2.0f, bsy = u->bmax[1] * 2.0f, bsz = u->bmax[2] * 2.0f;
  float ux = (p.x + u->bmax[0]) / bsx * res - 0.5f;
  float uy = (p.y + u->bmax[1]) / bsy * res - 0.5f;
  float uz = (p.z + u->bmax[2]) / bsz * res - 0.5f;
  float flx = floorf(ux), fly = floorf(uy), flz = floorf(uz);
  int32_t bx = o_f2i(flx), by = o_f2i(fly), bz = o_f2i(flz);
  float fx = ux - flx, fy = uy - fly, fz = uz - flz; /* WGSL fract */
  float c000 = load_sdf_texel(S, u, bx, by, bz);
  float c100 = load_sdf_texel(S, u, bx + 1, by, bz);
  float c010 = load_sdf_texel(S, u, bx, by + 1, bz);
  float c110 = load_sdf_texel(S, u, bx + 1, by + 1, bz);
  float c001 = load_sdf_texel(S, u, bx, by, bz + 1);
  float c101 = load_sdf_texel(S, u, bx + 1, by, bz + 1);
  float c011 = load_sdf_texel(S, u, bx, by + 1, bz + 1);
  float c111 = load_sdf_texel(S, u, bx + 1, by + 1, bz + 1);
  float c00 = o_mix(c000, c100, fx);
  float c10 = o_mix(c010, c110, fx);
  float c01 = o_mix(c001, c101, fx);
  float c11 = o_mix(c011, c111, fx);
  float c0 = o_mix(c00, c10, fy);
  float c1 = o_mix(c01, c11, fy);
  return o_mix(c0, c1, fz);
}
static v3 sdf_gradient(orc_sim* S, const orc_params* u, v3 p) {
  float e = u->dx;
  float gxv = sample_sdf(S, u, V(p.x + e, p.y + 0.0f, p.z + 0.0f)) - sample_sdf(S, u, V(p.x - e, p.y - 0.0f, p.z - 0.0f));
  float gyv = sample_sdf(S, u, V(p.x + 0.0f, p.y + e, p.z + 0.0f)) - sample_sdf(S, u, V(p.x - 0.0f, p.y - e, p.z - 0.0f));
  float gzv = sample_sdf(S, u, V(p.x + 0.0f, p.y + 0.0f, p.z + e)) - sample_sdf(S, u, V(p.x - 0.0f, p.y - 0.0f, p.z - e));
  v3 g = V(gxv, gyv, gzv);
  float len = vlen(g);
  if (len < 1e-8f) return V(0.0f, 0.0f, 0.0f);
  return vdiv(g, len);
}
static int class_solid(orc_sim* S, const orc_params* u, int32_t x, int32_t y, int32_t z) {
  return S->cls[cidx(u, (uint32_t)x, (uint32_t)y, (uint32_t)z)] != 0;
}

/* ---- contact (shader.rs:717-903) ------------------------------------------ */
typedef struct { v3 pos, vel; } contact_t;
static contact_t radial_barrier(const orc_params* u, v3 pos, v3 vel, float cx, float cz, float max_r) {
  float rx = pos.x - cx, rz = pos.z - cz;
  float r = sqrtf(rx * rx + rz * rz);
  if (r > max_r && r > 1e-6f) {
    float ox = rx / r, oz = rz / r;
    pos.x = cx + ox * max_r;
    pos.z = cz + oz * max_r;
    float vn = vel.x * ox + vel.z * oz;
    if (vn > 0.0f) {
      float tx = vel.x - ox * vn, tz = vel.z - oz * vn;
      vel.x = tx * (1.0f - u->friction * 0.35f);
      vel.z = tz * (1.0f - u->friction * 0.35f);
    }
  }
  contact_t c = {pos, vel};
  return c;
}
/* bed_particle_radius (shader.rs:128): dx * bed_sample_radius_dx (brew_config.rs:51) */
static float bed_particle_radius(const orc_params* u) { return u->dx * 0.62f; }
/* coffee_filter_floor_y (shader.rs:136-146) with the filter of the geometry */
static float coffee_filter_floor_y(const orc_params* u) {
  const cs_geometry* g = &u->geo;
  float fb = g->filter_center_y + g->filter_bot_y, ft = g->filter_center_y + g->filter_top_y;
  float bco = o_max(u->contact_offset, bed_particle_radius(u));
  float fh = o_max(ft - fb, 1e-6f);
  float slope = g->filter_top_radius / fh;
  return fb + (bco + g->filter_thickness) / o_max(slope, 1e-6f);
}
/* resolve_scene_obstacles (shader.rs:744-823): the cone, then for bed grains the
 * paper filter and its apex floor, then the cup */
static contact_t scene_obstacles(const orc_params* u, v3 pos, v3 vel, int is_bed) {
  const cs_geometry* g = &u->geo;
  if (g->cone_enabled && pos.y <= g->cone_top_y && pos.y >= g->cone_bot_y) {
    float t = o_clamp((pos.y - g->cone_bot_y) / (g->cone_top_y - g->cone_bot_y), 0.0f, 1.0f);
    float cone_r = o_mix(g->cone_bot_radius, g->cone_top_radius, t) - u->contact_offset;
    contact_t c = radial_barrier(u, pos, vel, 0.0f, 0.0f, cone_r);
    pos = c.pos; vel = c.vel;
  }
  if (is_bed && g->filter_enabled) {
    float ft = g->filter_center_y + g->filter_top_y, fb = g->filter_center_y + g->filter_bot_y;
    float bco = o_max(u->contact_offset, bed_particle_radius(u));
    float floor_y = coffee_filter_floor_y(u);
    if (pos.y <= ft && pos.y >= fb) {
      float t = o_clamp((pos.y - fb) / (ft - fb), 0.0f, 1.0f);
      float outer_r = o_mix(0.0f, g->filter_top_radius, t);
      float fr = o_max(outer_r - g->filter_thickness, 0.0f) - bco;
      contact_t c = radial_barrier(u, pos, vel, 0.0f, 0.0f, o_max(fr, 0.0f));
      pos = c.pos; vel = c.vel;
    }
    if (pos.y <= floor_y) {
      pos.y = floor_y;
      pos.x = 0.0f;
      pos.z = 0.0f;
      if (vel.y < 0.0f) {
        vel.y = 0.0f;
        vel.x *= 1.0f - u->friction * 0.55f;
        vel.z *= 1.0f - u->friction * 0.55f;
      }
      vel.x *= 1.0f - u->friction * 0.55f;
      vel.z *= 1.0f - u->friction * 0.55f;
    }
  }
  if (g->cup_enabled && pos.y <= g->cup_top_y) {
    float cup_r = g->cup_radius - u->wall_half - u->contact_offset;
    contact_t c = radial_barrier(u, pos, vel, 0.0f, 0.0f, cup_r);
    pos = c.pos; vel = c.vel;
    float floor_y = g->cup_bottom_y + u->wall_half + u->contact_offset;
    if (pos.y < floor_y) {
      pos.y = floor_y;
      if (vel.y < 0.0f) {
        vel.y = 0.0f;
        vel.x *= 1.0f - u->friction * 0.55f;
        vel.z *= 1.0f - u->friction * 0.55f;
      }
    }
  }
  contact_t c = {pos, vel};
  return c;
}
/* resolve_sdf_contact (shader.rs:879-903) */
static contact_t sdf_contact_as(orc_sim* S, const orc_params* u, v3 pos, v3 vel, int is_bed) {
  float sv = sample_sdf(S, u, pos);
  if (sv < u->contact_offset) {
    v3 n = sdf_gradient(S, u, pos);
    if (vlen(n) > 1e-6f) {
      pos = vadd(pos, vmul(n, u->contact_offset - sv));
      float vn = vdot(vel, n);
      if (vn < 0.0f) {
        vel = vsub(vel, vmul(vmul(n, vn), 1.0f + u->restitution));
        v3 vt = vsub(vel, vmul(n, vdot(vel, n)));
        float vtl = vlen(vt);
        if (vtl > 1e-6f) {
          float fi = o_min(u->friction * fabsf(vn), vtl);
          vel = vsub(vel, vmul(vt, fi / vtl));
        }
      }
    }
  }
  return scene_obstacles(u, pos, vel, is_bed);
}
static contact_t sdf_contact(orc_sim* S, const orc_params* u, v3 pos, v3 vel) {
  return sdf_contact_as(S, u, pos, vel, 0);
}

/* ---- bed record access (field mode of shader.rs:361-449) ------------------- */
static int bed_valid(orc_sim* S, int32_t id) { return id >= 0 && (uint32_t)id < S->num_bed; }
static float kozeny_factor(float phi_in) { /* shader.rs:429-433 */
  float phi = o_clamp(phi_in, 0.08f, 0.82f);
  float solid = o_max(1.0f - phi, 1e-4f);
  return (phi * phi * phi) / (solid * solid);
}
static float bed_compacted_porosity(orc_sim* S, int32_t id) { /* shader.rs:435-440 */
  const cs_bed_record* b = &S->bed[id];
  float base = o_clamp(b->porosity, 0.18f, 0.70f);
  float comp = o_clamp(b->compaction, 0.0f, 2.25f);
  return o_clamp(base - comp * 0.12f, 0.16f, base);
}
static float bed_compacted_permeability(orc_sim* S, const orc_params* u, int32_t id) { /* shader.rs:442-449 */
  const cs_bed_record* b = &S->bed[id];
  float base = o_clamp(b->porosity, 0.18f, 0.70f);
  float compacted = bed_compacted_porosity(S, id);
  float ratio = kozeny_factor(compacted) / o_max(kozeny_factor(base), 1e-8f);
  return o_max(b->permeability * ratio, u->min_k);
}
static v3 bed_matrix_velocity_cell(orc_sim* S, const orc_params* u, i3 cell) { /* shader.rs:380-407 */
  v3 sum = V(0.0f, 0.0f, 0.0f);
  float wsum = 0.0f;
  for (int di = -1; di <= 1; ++di)
    for (int dj = -1; dj <= 1; ++dj)
      for (int dk = -1; dk <= 1; ++dk) {
        int32_t x = cell.x + di, y = cell.y + dj, z = cell.z + dk;
        if (!on_grid(u, x, y, z)) continue;
        int32_t id = S->bed_lookup[cidx(u, (uint32_t)x, (uint32_t)y, (uint32_t)z)];
        if (!bed_valid(S, id)) continue;
        sum = vadd(sum, V(S->bed[id].velocity[0], S->bed[id].velocity[1], S->bed[id].velocity[2]));
        wsum += 1.0f;
      }
  if (wsum <= 0.0f) return V(0.0f, 0.0f, 0.0f);
  return vdiv(sum, wsum);
}
static v3 bed_velocity_load_cell(orc_sim* S, const orc_params* u, i3 c) { /* shader.rs:409-417 */
  if (!on_grid(u, c.x, c.y, c.z)) return V(0.0f, 0.0f, 0.0f);
  return bed_matrix_velocity_cell(S, u, c);
}
static float bed_velocity_divergence(orc_sim* S, const orc_params* u, i3 c) { /* shader.rs:419-427 */
  i3 a;
  a = c; a.x -= 1; float vxm = bed_velocity_load_cell(S, u, a).x;
  a = c; a.x += 1; float vxp = bed_velocity_load_cell(S, u, a).x;
  a = c; a.y -= 1; float vym = bed_velocity_load_cell(S, u, a).y;
  a = c; a.y += 1; float vyp = bed_velocity_load_cell(S, u, a).y;
  a = c; a.z -= 1; float vzm = bed_velocity_load_cell(S, u, a).z;
  a = c; a.z += 1; float vzp = bed_velocity_load_cell(S, u, a).z;
  return 0.5f * u->inv_dx * ((vxp - vxm) + (vyp - vym) + (vzp - vzm));
}

/* ---- bed reaction impulses (shader.rs:458-515) ------------------------------
 * The momentum a bed removes from the water (grid Darcy in project_pressure,
 * particle drag in G2P) is handed back to the bed grains through bed_delta
 * lanes 1-3 (i32 fixed point, wrapping adds), read and cleared by
 * bed_dynamics.  Only a bed of particles has grains to receive it. */
#define O_BED_REACTION_ALPHA 0.04f      /* shader.rs:76 */
#define O_BED_REACTION_IMPULSE_CAP 0.012f /* shader.rs:77 */
static int bed_solid_id(orc_sim* S, int32_t id) { /* is_valid_bed_solid_idx, shader.rs:365-370 */
  return id >= 0 && (uint32_t)id < S->bed_n && !(S->bed_aff[12 * (size_t)id + 3] < 0.5f);
}
static void bed_impulse_add(orc_sim* S, const orc_params* u, int32_t bid, v3 imp) { /* shader.rs:462-477 */
  float cap = u->nominal_mass * u->vel_cap * O_BED_REACTION_IMPULSE_CAP;
  float cx = o_clamp(imp.x, -cap, cap), cy = o_clamp(imp.y, -cap, cap), cz = o_clamp(imp.z, -cap, cap);
  ORC_LANE_ADD(&S->bed_dw[1u * S->bed_n + (uint32_t)bid], o_f2i(cx * u->fp_scale));
  ORC_LANE_ADD(&S->bed_dw[2u * S->bed_n + (uint32_t)bid], o_f2i(cy * u->fp_scale));
  ORC_LANE_ADD(&S->bed_dw[3u * S->bed_n + (uint32_t)bid], o_f2i(cz * u->fp_scale));
}
/* bed_impulse_delta_add_neighborhood (shader.rs:479-515): split over the grains
 * stamped in the 3x3x3 cells around `cell` (one share per stamped cell) */
static void bed_impulse_neighborhood(orc_sim* S, const orc_params* u, i3 cell, v3 imp) {
  float count = 0.0f;
  for (int di = -1; di <= 1; ++di)
    for (int dj = -1; dj <= 1; ++dj)
      for (int dk = -1; dk <= 1; ++dk) {
        int32_t x = cell.x + di, y = cell.y + dj, z = cell.z + dk;
        if (!on_grid(u, x, y, z)) continue;
        if (bed_solid_id(S, S->bed_lookup[cidx(u, (uint32_t)x, (uint32_t)y, (uint32_t)z)])) count += 1.0f;
      }
  if (count <= 0.0f) return;
  v3 share = V(imp.x / count, imp.y / count, imp.z / count);
  for (int di = -1; di <= 1; ++di)
    for (int dj = -1; dj <= 1; ++dj)
      for (int dk = -1; dk <= 1; ++dk) {
        int32_t x = cell.x + di, y = cell.y + dj, z = cell.z + dk;
        if (!on_grid(u, x, y, z)) continue;
        int32_t id = S->bed_lookup[cidx(u, (uint32_t)x, (uint32_t)y, (uint32_t)z)];
        if (bed_solid_id(S, id)) bed_impulse_add(S, u, id, share);
      }
}
static i3 world_to_cell(const orc_params* u, v3 p) { /* shader.rs:668-671 */
  i3 c = {o_f2i(floorf((p.x - u->origin[0]) * u->inv_dx)), o_f2i(floorf((p.y - u->origin[1]) * u->inv_dx)),
          o_f2i(floorf((p.z - u->origin[2]) * u->inv_dx))};
  return c;
}

/* shader.rs:597-619 */
static float volume_target_div(const orc_params* u, float rest, float cur) {
  float cv = u->dx * u->dx * u->dx;
  float eps = o_max(cv * 1e-5f, u->p_vol * 1e-3f);
  if (rest <= eps) return 0.0f;
  float j = cur / rest;
  float comp = o_clamp(1.0f - j, 0.0f, 0.75f);
  float exp_ = o_clamp(j - 1.0f, 0.0f, 0.75f);
  float target = comp * u->j_alpha - exp_ * u->j_expand_alpha;
  float packed = o_max(rest, cur) / o_max(cv, 1e-8f);
  float over = o_max(packed - u->max_rest_fraction, 0.0f);
  target = o_max(target, over * u->j_alpha);
  return target;
}
/* shader.rs:621-666 with the bed sample height taken from the bed record */
static float bed_pore_target_div(orc_sim* S, const orc_params* u, int32_t id, v3 cc, float rest, float cur) {
  float target = volume_target_div(u, rest, cur);
  float cv = u->dx * u->dx * u->dx;
  const cs_bed_record* b = &S->bed[id];
  float base = bed_compacted_porosity(S, id);
  float sat = o_clamp(b->saturation, 0.0f, 1.0f);
  float comp = o_clamp(b->compaction, 0.0f, 2.25f);
  float bed_y = b->sample_y;
  float surface_t = o_smooth(-0.25f, 0.75f, (cc.y - bed_y) * u->inv_dx);
  float open = o_clamp((1.0f - 0.65f * sat) * (1.0f - 0.12f * comp), 0.20f, 1.0f);
  float opening = surface_t * o_clamp(u->surface_void_scale, 0.0f, 1.0f) * open;
  float surf_por = o_min(base + 0.35f, 0.82f);
  float eff = o_mix(base, surf_por, opening);
  float occ = o_mix(1.0f, 0.35f, opening);
  float satc = o_mix(1.0f, 0.32f, sat);
  float compc = o_mix(1.0f, 0.58f, o_clamp(comp / 1.5f, 0.0f, 1.0f));
  float cap = eff * cv * o_max(u->pore_capacity_scale, 0.0f) * satc * compc;
  float overfill = o_max(cur - cap, 0.0f) / o_max(cv, 1e-8f);
  float resp = u->pore_overfill_alpha * occ * o_mix(1.0f, 1.45f, sat);
  target = o_max(target, overfill * resp);
  return target;
}

/* ---- P2G (shader.rs:955-1052) --------------------------------------------- */
static void k_p2g(orc_sim* S, const orc_params* u) {
  uint32_t np = u->num_water + u->num_bed_particles;
  float thr = u->nominal_mass * 0.10f;
  ORC_PARALLEL_FOR
  for (uint32_t pid = 0; pid < np; ++pid) {
    const float* p = &S->particles[(size_t)pid * 8];
    const float* a = &S->affine[(size_t)pid * 12];
    v3 xp = V(p[0], p[1], p[2]), vp = V(p[4], p[5], p[6]);
    float J = p[3], mass_p = p[7], phase = a[3];
    if (!(phase < 0.5f) || mass_p <= thr) continue;
    v3 gp = vmul(vsub(xp, V(u->origin[0], u->origin[1], u->origin[2])), u->inv_dx);
    int32_t bx = o_f2i(floorf(gp.x - 0.5f)), by = o_f2i(floorf(gp.y - 0.5f)), bz = o_f2i(floorf(gp.z - 0.5f));
    v3 fx = V(gp.x - (float)bx, gp.y - (float)by, gp.z - (float)bz);
    float wx[3], wy[3], wz[3];
    wx[0] = 0.5f * (1.5f - fx.x) * (1.5f - fx.x);
    wx[1] = 0.75f - (fx.x - 1.0f) * (fx.x - 1.0f);
    wx[2] = 0.5f * (fx.x - 0.5f) * (fx.x - 0.5f);
    wy[0] = 0.5f * (1.5f - fx.y) * (1.5f - fx.y);
    wy[1] = 0.75f - (fx.y - 1.0f) * (fx.y - 1.0f);
    wy[2] = 0.5f * (fx.y - 0.5f) * (fx.y - 0.5f);
    wz[0] = 0.5f * (1.5f - fx.z) * (1.5f - fx.z);
    wz[1] = 0.75f - (fx.z - 1.0f) * (fx.z - 1.0f);
    wz[2] = 0.5f * (fx.z - 0.5f) * (fx.z - 0.5f);
    v3 ac0 = V(mass_p * a[0], mass_p * a[1], mass_p * a[2]);
    v3 ac1 = V(mass_p * a[4], mass_p * a[5], mass_p * a[6]);
    v3 ac2 = V(mass_p * a[8], mass_p * a[9], mass_p * a[10]);
    float fp = u->fp_scale, cdx = u->dx;
    float rpv = u->p_vol * mass_p / o_max(u->nominal_mass, 1e-6f);
    float cpv = rpv * o_clamp(J, 0.40f, 2.00f);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          int32_t cx = bx + i, cy = by + j, cz = bz + k;
          if (!on_grid(u, cx, cy, cz)) continue;
          float w = wx[i] * wy[j] * wz[k];
          v3 dpos = vmul(vsub(V((float)i, (float)j, (float)k), fx), cdx);
          float mc = w * mass_p, rvc = w * rpv, cvc = w * cpv;
          v3 inner = vadd(vmul(vp, mass_p), V(vdot(ac0, dpos), vdot(ac1, dpos), vdot(ac2, dpos)));
          v3 mom = vmul(inner, w);
          size_t ci = cidx(u, (uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
          float lim = 1.0e9f;
          float mfp = mc * fp, rfp = rvc * fp, cfp = cvc * fp;
          float mxf = mom.x * fp, myf = mom.y * fp, mzf = mom.z * fp;
          if (fabsf(mfp) > lim || fabsf(mxf) > lim || fabsf(myf) > lim || fabsf(mzf) > lim ||
              fabsf(rfp) > lim || fabsf(cfp) > lim)
            ORC_COUNT(&S->metrics[O_M_OVERFLOW]);
          ORC_LANE_ADD(&LANE(S, 0, ci), o_f2i(mfp));
          ORC_LANE_ADD(&LANE(S, 1, ci), o_f2i(mxf));
          ORC_LANE_ADD(&LANE(S, 2, ci), o_f2i(myf));
          ORC_LANE_ADD(&LANE(S, 3, ci), o_f2i(mzf));
          ORC_LANE_ADD(&LANE(S, 4, ci), o_f2i(rfp));
          ORC_LANE_ADD(&LANE(S, 5, ci), o_f2i(cfp));
        }
  }
}

/* ---- grid_update (shader.rs:1056-1082) ------------------------------------ */
static void k_grid_update(orc_sim* S, const orc_params* u) {
  ORC_PARALLEL_FOR
  for (size_t idx = 0; idx < S->nc; ++idx) {
    float inv_fp = u->inv_fp_scale;
    float mass = (float)LANE(S, 0, idx) * inv_fp;
    if (mass < 1e-6f) continue;
    v3 v = V((float)LANE(S, 1, idx) * inv_fp / mass, (float)LANE(S, 2, idx) * inv_fp / mass,
             (float)LANE(S, 3, idx) * inv_fp / mass);
    v.y += u->gravity * u->dt;
    float speed = vlen(v);
    if (speed > u->vel_cap) v = vmul(v, u->vel_cap / speed);
    float* g = &S->grid_vel[idx * 4];
    g[0] = v.x; g[1] = v.y; g[2] = v.z; g[3] = mass;
  }
}

/* ---- boundary_project (shader.rs:1630-1679) -------------------------------- */
static void k_boundary(orc_sim* S, const orc_params* u) {
  ORC_PARALLEL_FOR
  for (size_t idx = 0; idx < S->nc; ++idx) {
    float* g = &S->grid_vel[idx * 4];
    if (g[3] < 1e-6f) continue;
    v3 v = V(g[0], g[1], g[2]);
    uint32_t ix, iy, iz;
    decompose(u, idx, &ix, &iy, &iz);
    v3 cp = V(u->origin[0] + ((float)ix + 0.5f) * u->dx, u->origin[1] + ((float)iy + 0.5f) * u->dx,
              u->origin[2] + ((float)iz + 0.5f) * u->dx);
    float sv = sample_sdf(S, u, cp);
    if (sv < u->contact_offset) {
      v3 n = sdf_gradient(S, u, cp);
      float vn = vdot(v, n);
      if (vn < 0.0f) {
        v = vsub(v, vmul(vmul(n, vn), 1.0f + u->restitution));
        v3 vt = vsub(v, vmul(n, vdot(v, n)));
        float vtl = vlen(vt);
        if (vtl > 1e-6f) {
          float fi = o_min(u->friction * fabsf(vn), vtl);
          v = vsub(v, vmul(vt, fi / vtl));
        }
      }
    }
    float margin = 2.0f * u->dx;
    float bminx = u->origin[0] + margin, bminy = u->origin[1] + margin, bminz = u->origin[2] + margin;
    float bmaxx = u->bmax[0] - margin, bmaxy = u->bmax[1] - margin, bmaxz = u->bmax[2] - margin;
    if (cp.x < bminx && v.x < 0.0f) v.x = 0.0f;
    if (cp.x > bmaxx && v.x > 0.0f) v.x = 0.0f;
    if (cp.y < bminy && v.y < 0.0f) v.y = 0.0f;
    if (cp.y > bmaxy && v.y > 0.0f) v.y = 0.0f;
    if (cp.z < bminz && v.z < 0.0f) v.z = 0.0f;
    if (cp.z > bmaxz && v.z > 0.0f) v.z = 0.0f;
    g[0] = v.x; g[1] = v.y; g[2] = v.z;
  }
}

/* ---- classify_cells (shader.rs:1181-1367) ---------------------------------- */
static const int32_t OFF6[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};

/* The central-difference divergence of grid_vel at a fluid cell, with the
 * ghost-mirror rule on solid and off-grid faces (shader.rs:1278-1335). */
static float cell_divergence(orc_sim* S, const orc_params* u, uint32_t ix, uint32_t iy, uint32_t iz, v3 cc) {
  size_t idx = cidx(u, ix, iy, iz);
  const float* sv = &S->grid_vel[idx * 4];
  float dxv = u->dx;
  float vxm = -sv[0], vxp = -sv[0], vym = -sv[1], vyp = -sv[1], vzm = -sv[2], vzp = -sv[2];
#define OPEN(cond_sdf_pos, cx, cy, cz) \
  (u->use_sdf_cache ? !class_solid(S, u, (cx), (cy), (cz)) : (sample_sdf(S, u, cond_sdf_pos) >= 0.0f))
  if (ix > 0u && OPEN(V(cc.x + -dxv, cc.y + 0.0f, cc.z + 0.0f), (int32_t)ix - 1, (int32_t)iy, (int32_t)iz))
    vxm = S->grid_vel[cidx(u, ix - 1u, iy, iz) * 4 + 0];
  if (ix + 1u < u->gx && OPEN(V(cc.x + dxv, cc.y + 0.0f, cc.z + 0.0f), (int32_t)ix + 1, (int32_t)iy, (int32_t)iz))
    vxp = S->grid_vel[cidx(u, ix + 1u, iy, iz) * 4 + 0];
  if (iy > 0u && OPEN(V(cc.x + 0.0f, cc.y + -dxv, cc.z + 0.0f), (int32_t)ix, (int32_t)iy - 1, (int32_t)iz))
    vym = S->grid_vel[cidx(u, ix, iy - 1u, iz) * 4 + 1];
  if (iy + 1u < u->gy && OPEN(V(cc.x + 0.0f, cc.y + dxv, cc.z + 0.0f), (int32_t)ix, (int32_t)iy + 1, (int32_t)iz))
    vyp = S->grid_vel[cidx(u, ix, iy + 1u, iz) * 4 + 1];
  if (iz > 0u && OPEN(V(cc.x + 0.0f, cc.y + 0.0f, cc.z + -dxv), (int32_t)ix, (int32_t)iy, (int32_t)iz - 1))
    vzm = S->grid_vel[cidx(u, ix, iy, iz - 1u) * 4 + 2];
  if (iz + 1u < u->gz && OPEN(V(cc.x + 0.0f, cc.y + 0.0f, cc.z + dxv), (int32_t)ix, (int32_t)iy, (int32_t)iz + 1))
    vzp = S->grid_vel[cidx(u, ix, iy, iz + 1u) * 4 + 2];
#undef OPEN
  float div = 0.5f * u->inv_dx * ((vxp - vxm) + (vyp - vym) + (vzp - vzm));
  return div;
}

static void k_classify(orc_sim* S, const orc_params* u) {
  float occ = u->nominal_mass * 0.1f;
  ORC_PARALLEL_FOR
  for (size_t idx = 0; idx < S->nc; ++idx) {
    float mass = S->grid_vel[idx * 4 + 3];
    pressure_store(S, u, idx, 0.0f);
    uint32_t ix, iy, iz;
    decompose(u, idx, &ix, &iy, &iz);
    v3 cc = V(u->origin[0] + ((float)ix + 0.5f) * u->dx, u->origin[1] + ((float)iy + 0.5f) * u->dx,
              u->origin[2] + ((float)iz + 0.5f) * u->dx);
    int self_solid = u->use_sdf_cache ? class_solid(S, u, (int32_t)ix, (int32_t)iy, (int32_t)iz)
                                      : (sample_sdf(S, u, cc) < 0.0f);
    if (self_solid) {
      LANE(S, 3, idx) = O_CELL_SOLID;
      div_store(S, u, idx, 0.0f);
      continue;
    }
    if (mass <= occ) {
      LANE(S, 3, idx) = O_CELL_AIR;
      div_store(S, u, idx, 0.0f);
      continue;
    }
    if (S->bed_lookup[idx] >= 0) {
      LANE(S, 3, idx) = O_CELL_BED;
    } else {
      int air = 0;
      for (int n = 0; n < 6; ++n) {
        int32_t x = (int32_t)ix + OFF6[n][0], y = (int32_t)iy + OFF6[n][1], z = (int32_t)iz + OFF6[n][2];
        if (!on_grid(u, x, y, z)) { air = 1; break; }
        if (S->grid_vel[cidx(u, (uint32_t)x, (uint32_t)y, (uint32_t)z) * 4 + 3] <= occ) { air = 1; break; }
      }
      LANE(S, 3, idx) = air ? O_CELL_SURFACE : O_CELL_INTERIOR;
    }
    int32_t kind = kind_load(S, idx);
    if (!is_fluid(kind)) { div_store(S, u, idx, 0.0f); continue; }
    float div = cell_divergence(S, u, ix, iy, iz, cc);
    float rest = rest_vol(S, u, idx), cur = cur_vol(S, u, idx);
    float target = volume_target_div(u, rest, cur);
    if (kind == O_CELL_BED) {
      int32_t id = S->bed_lookup[idx];
      if (id >= 0 && (uint32_t)id < S->num_bed) {
        float pore = bed_pore_target_div(S, u, id, cc, rest, cur);
        float por = o_clamp(bed_compacted_porosity(S, id), 0.08f, 1.0f);
        float solid_fraction = 1.0f - por;
        i3 c3 = {(int32_t)ix, (int32_t)iy, (int32_t)iz};
        float solid_div = bed_velocity_divergence(S, u, c3);
        target = (pore - solid_fraction * solid_div) / por;
      }
    }
    div_store(S, u, idx, div - target);
    S->tgt[idx] = target;
    uint32_t fpr = o_f2u(o_clamp(fabsf(div - target) * u->metrics_div_fp_scale, 0.0f, 2147483648.0f));
    orc_max_u32(&S->metrics[O_M_PRE_MAX_ABS_RES], fpr);
    float abs_div = fabsf(div);
    uint32_t fpd = o_f2u(o_clamp(abs_div * u->metrics_div_fp_scale, 0.0f, 2147483648.0f));
    orc_max_u32(&S->metrics[O_M_MAX_ABS_DIV], fpd);
    ORC_COUNT(&S->metrics[O_M_FLUID_CELLS]);
  }
}

/* Face weights, their sum and the right-hand side of one fluid cell, exactly as
 * pressure_update (shader.rs:1371-1439) forms them. */
static void cell_system(orc_sim* S, const orc_params* u, size_t idx, int32_t kind, float* w6, size_t* nb6,
                        int* nf6, float* count, float* c) {
  uint32_t ix, iy, iz;
  decompose(u, idx, &ix, &iy, &iz);
  float self_fill = fill_fraction(S, u, idx, kind);
  float cnt = 0.0f;
  for (int n = 0; n < 6; ++n) {
    w6[n] = 0.0f; nb6[n] = idx; nf6[n] = 0;
    int32_t x = (int32_t)ix + OFF6[n][0], y = (int32_t)iy + OFF6[n][1], z = (int32_t)iz + OFF6[n][2];
    if (!on_grid(u, x, y, z)) continue;
    size_t ni = cidx(u, (uint32_t)x, (uint32_t)y, (uint32_t)z);
    int32_t nk = kind_load(S, ni);
    if (is_solid(nk)) continue;
    float fw = face_weight(S, u, kind, self_fill, ni, nk);
    if (fw <= 0.0f) continue;
    cnt += fw;
    w6[n] = fw; nb6[n] = ni; nf6[n] = is_fluid(nk);
  }
  *count = cnt;
  float rhs = div_load(S, u, idx) * self_fill / o_max(u->dt, 1e-6f);
  *c = u->dx * u->dx * rhs;
}

/* One row of the pressure system.  Kinds, fills and the divergence do not
 * change during the solve, so the rows are formed once per solve; the
 * arithmetic of each update is unchanged. */
typedef struct {
  size_t idx;
  float w[6];
  size_t nb[6];
  int nf[6];
  float count, c;
} sys_row;

static sys_row* g_rows[2] = {NULL, NULL};
static size_t g_nrows[2] = {0, 0};
static size_t g_cap[2] = {0, 0};

/* Splits the fluid cells by colour (ix+iy+iz)&1 in ascending index order. */
static void build_rows(orc_sim* S, const orc_params* u) {
  g_nrows[0] = g_nrows[1] = 0;
  for (size_t idx = 0; idx < S->nc; ++idx) {
    int32_t kind = kind_load(S, idx);
    if (!is_fluid(kind)) continue;
    uint32_t ix, iy, iz;
    decompose(u, idx, &ix, &iy, &iz);
    int col = (int)((ix + iy + iz) & 1u);
    if (g_nrows[col] == g_cap[col]) {
      g_cap[col] = g_cap[col] ? g_cap[col] * 2 : 4096;
      g_rows[col] = (sys_row*)realloc(g_rows[col], g_cap[col] * sizeof(sys_row));
    }
    sys_row* r = &g_rows[col][g_nrows[col]++];
    r->idx = idx;
    cell_system(S, u, idx, kind, r->w, r->nb, r->nf, &r->count, &r->c);
  }
}

/* pressure_update (shader.rs:1371-1439) over one colour.  A cell of one
 * colour reads only cells of the other colour, so the visiting order within a
 * colour does not change any result; non-fluid cells keep the 0 written by
 * classify. */
static void k_rbgs_half(orc_sim* S, const orc_params* u, uint32_t parity) {
  sys_row* rows = g_rows[parity];
  size_t n = g_nrows[parity];
  ORC_PARALLEL_FOR
  for (size_t i = 0; i < n; ++i) {
    const sys_row* r = &rows[i];
    float sum = 0.0f;
    for (int k = 0; k < 6; ++k)
      if (r->w[k] > 0.0f && r->nf[k]) sum += r->w[k] * pressure_load(S, u, r->nb[k]);
    if (r->count <= 0.0f) { pressure_store(S, u, r->idx, 0.0f); continue; }
    float p_new = (sum - r->c) / r->count;
    pressure_store(S, u, r->idx, p_new);
  }
}

/* Relative residual r = ||b - A p||_2 / ||b||_2 over fluid cells with a
 * non-empty stencil, where the row of cell i reads count_i p_i - sum_n w_n p_n = -c_i
 * (the fixed point of pressure_update); summed in ascending cell order.
 * A cell whose pressure sits at the clamp (|p| = pressure_clamp, the bound
 * pressure_update enforces through pressure_store, shader.rs:202-209) is left out: there the bound
 * replaces the row, so the clamped iteration's fixed point leaves that row
 * unsatisfied and it would hold the residual above any tolerance. */
/* One row's contribution to the residual sums (nothing for a row the clamp holds). */
static inline void row_residual(orc_sim* S, const orc_params* u, const sys_row* r, double* rr, double* bb) {
  if (r->count <= 0.0f) return;
  if (fabsf(pressure_load(S, u, r->idx)) == u->pressure_clamp) return;
  double ap = (double)r->count * (double)pressure_load(S, u, r->idx);
  for (int k = 0; k < 6; ++k)
    if (r->w[k] > 0.0f && r->nf[k]) ap -= (double)r->w[k] * (double)pressure_load(S, u, r->nb[k]);
  double b = -(double)r->c;
  *rr += (b - ap) * (b - ap);
  *bb += b * b;
}

static double residual_rows(orc_sim* S, const orc_params* u) {
  double rr = 0.0, bb = 0.0;
#ifdef _OPENMP
  /* the timing build sums each colour's rows in parallel (another order) */
  for (int col = 0; col < 2; ++col) {
    const sys_row* rows = g_rows[col];
    const size_t n = g_nrows[col];
#pragma omp parallel for reduction(+ : rr, bb) schedule(static)
    for (size_t i = 0; i < n; ++i) row_residual(S, u, &rows[i], &rr, &bb);
  }
#else
  size_t i0 = 0, i1 = 0;
  while (i0 < g_nrows[0] || i1 < g_nrows[1]) {
    const sys_row* r;
    if (i1 >= g_nrows[1] || (i0 < g_nrows[0] && g_rows[0][i0].idx < g_rows[1][i1].idx)) r = &g_rows[0][i0++];
    else r = &g_rows[1][i1++];
    row_residual(S, u, r, &rr, &bb);
  }
#endif
  if (bb > 0.0) return sqrt(rr / bb);
  return rr > 0.0 ? INFINITY : 0.0;
}

double orc_residual(orc_sim* S) {
  build_rows(S, &S->u);
  return residual_rows(S, &S->u);
}

/* One-time Galerkin rescale (rbgs_rescale_after; no reference counterpart).
 * Over the residual's rows, <p, A p> and <p, b>; the factor alpha = <p,b>/<p,Ap>
 * minimises the energy norm of the error of alpha p.  It is rounded to a
 * multiple of 2^-20 and kept in [1/16, 16] (1 for degenerate sums). */
static void energy_rows(orc_sim* S, const orc_params* u, double* pap_out, double* pb_out) {
  double pap = 0.0, pb = 0.0;
  for (int col = 0; col < 2; ++col) {
    const sys_row* rows = g_rows[col];
    const size_t n = g_nrows[col];
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : pap, pb) schedule(static)
#endif
    for (size_t i = 0; i < n; ++i) {
      const sys_row* r = &rows[i];
      if (r->count <= 0.0f) continue;
      const float pf = pressure_load(S, u, r->idx);
      if (fabsf(pf) == u->pressure_clamp) continue;
      double ap = (double)r->count * (double)pf;
      for (int k = 0; k < 6; ++k)
        if (r->w[k] > 0.0f && r->nf[k]) ap -= (double)r->w[k] * (double)pressure_load(S, u, r->nb[k]);
      pap += (double)pf * ap;
      pb += (double)pf * -(double)r->c;
    }
  }
  *pap_out = pap;
  *pb_out = pb;
}
static float rescale_factor(double pap, double pb) {
  if (!(pap > 0.0)) return 1.0f;
  const double a = pb / pap;
  if (!(a >= 0.0625 && a <= 16.0)) return a > 16.0 ? 16.0f : (a < 0.0625 ? 0.0625f : 1.0f);
  return (float)(rint(a * 1048576.0) * (1.0 / 1048576.0));
}
static void rescale_pressure(orc_sim* S, const orc_params* u) {
  double pap, pb;
  energy_rows(S, u, &pap, &pb);
  const float a = rescale_factor(pap, pb);
  for (size_t i = 0; i < S->nc; ++i)
    if (is_fluid(kind_load(S, i))) pressure_store(S, u, i, a * pressure_load(S, u, i));
}
double orc_rescale_factor(orc_sim* S) {
  double pap, pb;
  build_rows(S, &S->u);
  energy_rows(S, &S->u, &pap, &pb);
  return pap > 0.0 ? pb / pap : NAN;
}
void orc_debug_set_pressure(orc_sim* S, const int32_t* p) {
  for (size_t i = 0; i < S->nc; ++i) LANE(S, 0, i) = p[i];
}

void orc_read_system(orc_sim* S, float* w6o, float* cnto, float* co) {
  const orc_params* u = &S->u;
  for (size_t idx = 0; idx < S->nc; ++idx) {
    int32_t kind = kind_load(S, idx);
    for (int n = 0; n < 6; ++n) w6o[idx * 6 + n] = 0.0f;
    cnto[idx] = 0.0f; co[idx] = 0.0f;
    if (!is_fluid(kind)) continue;
    float w6[6], count, c;
    size_t nb6[6];
    int nf6[6];
    cell_system(S, u, idx, kind, w6, nb6, nf6, &count, &c);
    for (int n = 0; n < 6; ++n) w6o[idx * 6 + n] = w6[n];
    cnto[idx] = count; co[idx] = c;
  }
}

void orc_read_bed_extracts(orc_sim* S, float* ext) {
  if (S->bed_n) memcpy(ext, S->bed_ext, (size_t)S->bed_n * 8 * sizeof(float));
}
void orc_read_bed_particles(orc_sim* S, float* part, float* aff) {
  if (!S->bed_n) return;
  if (part) memcpy(part, S->bed_part, (size_t)S->bed_n * 8 * sizeof(float));
  if (aff) memcpy(aff, S->bed_aff, (size_t)S->bed_n * 12 * sizeof(float));
}
void orc_read_bed_lookup(orc_sim* S, int32_t* lookup) {
  memcpy(lookup, S->bed_lookup, S->nc * sizeof(int32_t));
}
void orc_read_pressure(orc_sim* S, int32_t* p, int32_t* kind) {
  /* lane 0 keeps the pressure to the end of the substep; the kind lane is
   * overwritten by the viscosity scratch (shader.rs:274-283), so the copy
   * taken at the end of the solve is returned */
  memcpy(p, S->saved_p, S->nc * sizeof(int32_t));
  if (kind) memcpy(kind, S->saved_kind, S->nc * sizeof(int32_t));
}

/* ---- multigrid start (rbgs_mg_cycles; no reference counterpart) --------------
 * Before the red-black iteration, W-cycles of a truncated aggregation multigrid
 * correct the start.  The free cells are the fluid cells with a non-empty row
 * whose pressure is inside the clamp; a cell at the clamp keeps its pressure
 * (its correction is held at zero, as a Dirichlet cell).  Level 0 is the fine
 * grid's error equation A e = b - A p over the free cells; level l+1 has one
 * cell per 2x2x2 block of level l (the grid halved, rounded up) and the
 * Galerkin operator of piecewise-constant aggregation: a coarse cell's diagonal
 * is its children's diagonals less twice their internal couplings, a coarse
 * face weight is the sum of the fine couplings across that face.  Levels are
 * added until no side exceeds 8.  A visit of level l (W-cycle, gamma 2):
 * MG_NU red-black smoothing passes, then twice { restrict the residual to
 * level l+1, zero its correction, visit it, add MG_SCALE times the prolonged
 * (piecewise-constant) correction, MG_NU passes }; the coarsest level gets
 * mg_coarse_passes(nl) passes.  Every per-cell sum runs in a fixed order (faces -x, +x,
 * -y, +y, -z, +z; children dz, dy, dx ascending), so the device reproduces it
 * bit for bit; a cycle ends with p <- clamp(p + e) stored in fixed point. */
#define MG_NU 2u
#define MG_GAMMA 2u
#define MG_SCALE 1.5f
#define MG_COARSE 10u /* per visit of the coarsest level while a cycle visits it at most 4 times */
#define MG_COARSEST_SIDE 8u
#define MG_MAX_LEVELS 12

typedef struct {
  uint32_t X, Y, Z;
  size_t n;
  float *D, *Wx, *Wy, *Wz, *R, *E;
} mg_level;

static inline size_t mg_idx(const mg_level* L, uint32_t x, uint32_t y, uint32_t z) {
  return ((size_t)z * L->Y + y) * L->X + x;
}
/* the six neighbour terms of cell (x,y,z) against values v, in face order */
static float mg_offsum(const mg_level* L, const float* v, uint32_t x, uint32_t y, uint32_t z) {
  const size_t i = mg_idx(L, x, y, z);
  float t0 = x > 0 ? L->Wx[i - 1] * v[i - 1] : 0.0f;
  float t1 = x + 1 < L->X ? L->Wx[i] * v[i + 1] : 0.0f;
  float t2 = y > 0 ? L->Wy[i - L->X] * v[i - L->X] : 0.0f;
  float t3 = y + 1 < L->Y ? L->Wy[i] * v[i + L->X] : 0.0f;
  const size_t sl = (size_t)L->X * L->Y;
  float t4 = z > 0 ? L->Wz[i - sl] * v[i - sl] : 0.0f;
  float t5 = z + 1 < L->Z ? L->Wz[i] * v[i + sl] : 0.0f;
  float s = t0;
  s += t1;
  s += t2;
  s += t3;
  s += t4;
  s += t5;
  return s;
}
static void mg_smooth(mg_level* L, uint32_t col) {
  for (uint32_t z = 0; z < L->Z; ++z)
    for (uint32_t y = 0; y < L->Y; ++y)
      for (uint32_t x = 0; x < L->X; ++x) {
        if (((x + y + z) & 1u) != col) continue;
        const size_t i = mg_idx(L, x, y, z);
        if (!(L->D[i] > 0.0f)) continue;
        L->E[i] = (L->R[i] + mg_offsum(L, L->E, x, y, z)) / L->D[i];
      }
}
/* children of coarse cell (X,Y,Z) in a level of (fx,fy,fz) cells */
#define MG_CHILDREN(FX, FY, FZ, CX, CY, CZ, BODY)                                              \
  for (uint32_t dz_ = 0; dz_ < 2u; ++dz_)                                                      \
    for (uint32_t dy_ = 0; dy_ < 2u; ++dy_)                                                    \
      for (uint32_t dx_ = 0; dx_ < 2u; ++dx_) {                                                \
        const uint32_t x = 2u * (CX) + dx_, y = 2u * (CY) + dy_, z = 2u * (CZ) + dz_;         \
        if (x >= (FX) || y >= (FY) || z >= (FZ)) continue;                                     \
        BODY                                                                                   \
      }
/* residual of level f restricted to level c (children summed in order) */
static void mg_restrict(const mg_level* f, mg_level* c) {
  for (uint32_t Z = 0; Z < c->Z; ++Z)
    for (uint32_t Y = 0; Y < c->Y; ++Y)
      for (uint32_t X = 0; X < c->X; ++X) {
        float acc = 0.0f;
        MG_CHILDREN(f->X, f->Y, f->Z, X, Y, Z, {
          const size_t i = mg_idx(f, x, y, z);
          if (f->D[i] > 0.0f) {
            const float s = mg_offsum(f, f->E, x, y, z);
            acc += f->R[i] - (f->D[i] * f->E[i] - s);
          }
        })
        c->R[mg_idx(c, X, Y, Z)] = acc;
      }
}
/* Galerkin operator of level c from level f */
static void mg_build(const mg_level* f, mg_level* c) {
  for (uint32_t Z = 0; Z < c->Z; ++Z)
    for (uint32_t Y = 0; Y < c->Y; ++Y)
      for (uint32_t X = 0; X < c->X; ++X) {
        float d = 0.0f, wx = 0.0f, wy = 0.0f, wz = 0.0f;
        MG_CHILDREN(f->X, f->Y, f->Z, X, Y, Z, { d += f->D[mg_idx(f, x, y, z)]; })
        MG_CHILDREN(f->X, f->Y, f->Z, X, Y, Z, {
          const size_t i = mg_idx(f, x, y, z);
          if (dx_ == 0u && x + 1 < f->X) d -= 2.0f * f->Wx[i];
          if (dy_ == 0u && y + 1 < f->Y) d -= 2.0f * f->Wy[i];
          if (dz_ == 0u && z + 1 < f->Z) d -= 2.0f * f->Wz[i];
          if (dx_ == 1u) wx += f->Wx[i];
          if (dy_ == 1u) wy += f->Wy[i];
          if (dz_ == 1u) wz += f->Wz[i];
        })
        const size_t I = mg_idx(c, X, Y, Z);
        c->D[I] = d > 0.0f ? d : 0.0f;
        c->Wx[I] = wx;
        c->Wy[I] = wy;
        c->Wz[I] = wz;
      }
}
static void mg_prolong(mg_level* f, const mg_level* c) {
  for (uint32_t z = 0; z < f->Z; ++z)
    for (uint32_t y = 0; y < f->Y; ++y)
      for (uint32_t x = 0; x < f->X; ++x) {
        const size_t i = mg_idx(f, x, y, z);
        if (f->D[i] > 0.0f) f->E[i] = f->E[i] + MG_SCALE * c->E[mg_idx(c, x >> 1, y >> 1, z >> 1)];
      }
}
/* Passes per visit of the coarsest of nl levels: MG_COARSE while a cycle visits
 * it at most 4 times (MG_GAMMA^(nl-2) visits), else the same 40 passes per
 * cycle spread over the visits (at least 2 each). */
static uint32_t mg_coarse_passes(int nl) {
  uint32_t visits = 1;
  for (int l = 2; l < nl; ++l) visits *= MG_GAMMA;
  if (visits <= 4u) return MG_COARSE;
  const uint32_t p = 4u * MG_COARSE / visits;
  return p < 2u ? 2u : p;
}
static void mg_visit(mg_level* L, int l, int nl) {
  if (l == nl - 1) {
    for (uint32_t k = 0; k < mg_coarse_passes(nl); ++k) { mg_smooth(&L[l], 0u); mg_smooth(&L[l], 1u); }
    return;
  }
  for (uint32_t k = 0; k < MG_NU; ++k) { mg_smooth(&L[l], 0u); mg_smooth(&L[l], 1u); }
  for (uint32_t g = 0; g < MG_GAMMA; ++g) {
    mg_restrict(&L[l], &L[l + 1]);
    memset(L[l + 1].E, 0, L[l + 1].n * sizeof(float));
    mg_visit(L, l + 1, nl);
    mg_prolong(&L[l], &L[l + 1]);
    for (uint32_t k = 0; k < MG_NU; ++k) { mg_smooth(&L[l], 0u); mg_smooth(&L[l], 1u); }
  }
}
/* Level 0 over the free cells: D = the row's count, face weights toward free
 * neighbours, R = b - A p, E = 0.  The pressure sum is pressure_update's. */
static void mg_fine_level(orc_sim* S, const orc_params* u, mg_level* L0, uint8_t* fluid) {
  const float lim = u->pressure_clamp;
  memset(L0->D, 0, S->nc * sizeof(float));
  memset(L0->Wx, 0, S->nc * sizeof(float));
  memset(L0->Wy, 0, S->nc * sizeof(float));
  memset(L0->Wz, 0, S->nc * sizeof(float));
  memset(L0->R, 0, S->nc * sizeof(float));
  memset(L0->E, 0, S->nc * sizeof(float));
  uint8_t* fr = (uint8_t*)calloc(S->nc, 1);
  float* sums = (float*)calloc(S->nc, sizeof(float));
  for (int col = 0; col < 2; ++col)
    for (size_t k = 0; k < g_nrows[col]; ++k) {
      const sys_row* r = &g_rows[col][k];
      float t[6];
      for (int n = 0; n < 6; ++n) t[n] = (r->w[n] > 0.0f && r->nf[n]) ? r->w[n] * pressure_load(S, u, r->nb[n]) : 0.0f;
      float sum = t[0];
      sum += t[1];
      sum += t[2];
      sum += t[3];
      sum += t[4];
      sum += t[5];
      sums[r->idx] = sum;
      if (!(r->count > 0.0f)) continue;
      const float p = pressure_load(S, u, r->idx);
      if (fabsf(p) < lim) {
        fr[r->idx] = 1u;
      } else {
        /* at the clamp: free when pressure_update would move it back inside */
        const float v = (sum - r->c) / r->count;
        fr[r->idx] = (p > 0.0f ? v < lim : v > -lim) ? 1u : 0u;
      }
    }
  for (int col = 0; col < 2; ++col)
    for (size_t k = 0; k < g_nrows[col]; ++k) {
      const sys_row* r = &g_rows[col][k];
      const size_t i = r->idx;
      fluid[i] = 1u;
      if (!fr[i]) continue;
      /* +x, +y, +z couplings to free neighbours (faces 1, 3, 5) */
      if (r->nf[1] && fr[r->nb[1]]) L0->Wx[i] = r->w[1];
      if (r->nf[3] && fr[r->nb[3]]) L0->Wy[i] = r->w[3];
      if (r->nf[5] && fr[r->nb[5]]) L0->Wz[i] = r->w[5];
      L0->D[i] = r->count;
      L0->R[i] = (sums[i] - r->c) - r->count * pressure_load(S, u, i);
    }
  free(sums);
  free(fr);
}
static int mg_alloc(orc_sim* S, mg_level* L) {
  int nl = 0;
  uint32_t X = S->gx, Y = S->gy, Z = S->gz;
  for (;;) {
    mg_level* l = &L[nl++];
    l->X = X; l->Y = Y; l->Z = Z;
    l->n = (size_t)X * Y * Z;
    float* buf = (float*)calloc(6 * l->n, sizeof(float));
    l->D = buf; l->Wx = buf + l->n; l->Wy = buf + 2 * l->n; l->Wz = buf + 3 * l->n;
    l->R = buf + 4 * l->n; l->E = buf + 5 * l->n;
    if ((X <= MG_COARSEST_SIDE && Y <= MG_COARSEST_SIDE && Z <= MG_COARSEST_SIDE) || nl == MG_MAX_LEVELS) break;
    X = (X + 1u) / 2u; Y = (Y + 1u) / 2u; Z = (Z + 1u) / 2u;
  }
  return nl;
}
static double residual_rows(orc_sim* S, const orc_params* u);
/* Up to `cycles` W-cycles, each followed by the residual test; returns the last
 * residual (r when no cycle runs). */
static double mg_start(orc_sim* S, const orc_params* u, uint32_t cycles, double r) {
  mg_level L[MG_MAX_LEVELS];
  const int nl = mg_alloc(S, L);
  uint8_t* fluid = (uint8_t*)calloc(S->nc, 1);
  const float lim = u->pressure_clamp;
  for (uint32_t c = 0; c < cycles && !(r <= (double)S->s.rbgs_tolerance); ++c) {
    mg_fine_level(S, u, &L[0], fluid);
    for (int l = 0; l + 1 < nl; ++l) mg_build(&L[l], &L[l + 1]);
    if (nl > 1) mg_visit(L, 0, nl);
    else for (uint32_t k = 0; k < mg_coarse_passes(nl); ++k) { mg_smooth(&L[0], 0u); mg_smooth(&L[0], 1u); }
    for (size_t i = 0; i < S->nc; ++i) {
      if (!(L[0].D[i] > 0.0f)) continue;
      const float pn = pressure_load(S, u, i) + L[0].E[i];
      LANE(S, 0, i) = o_f2i(o_clamp(pn, -lim, lim) * u->fp_scale);
    }
    r = residual_rows(S, u);
  }
  free(fluid);
  for (int l = 0; l < nl; ++l) free(L[l].D);
  return r;
}
/* The multigrid operator of `level` (0 = the fine level over the free cells)
 * for the current pressure and kinds; returns the level count, or 0 when
 * `level` is beyond the last. */
int orc_debug_mg_operator(orc_sim* S, float sub_dt, int level, float* D, float* Wx, float* Wy, float* Wz) {
  orc_write_uniforms(S, sub_dt, &S->u);
  const orc_params* u = &S->u;
  build_rows(S, u);
  mg_level L[MG_MAX_LEVELS];
  const int nl = mg_alloc(S, L);
  uint8_t* fluid = (uint8_t*)calloc(S->nc, 1);
  mg_fine_level(S, u, &L[0], fluid);
  for (int l = 0; l + 1 < nl && l < level; ++l) mg_build(&L[l], &L[l + 1]);
  int ok = level >= 0 && level < nl;
  if (ok) {
    const mg_level* l = &L[level];
    memcpy(D, l->D, l->n * sizeof(float));
    memcpy(Wx, l->Wx, l->n * sizeof(float));
    memcpy(Wy, l->Wy, l->n * sizeof(float));
    memcpy(Wz, l->Wz, l->n * sizeof(float));
  }
  free(fluid);
  for (int l = 0; l < nl; ++l) free(L[l].D);
  return ok ? nl : 0;
}

/* ---- pressure solve driver ------------------------------------------------- */
static void solve_pressure(orc_sim* S, const orc_params* u) {
  cs_solve_stats* st = &S->stats;
  memset(st, 0, sizeof(*st));
  st->final_residual = NAN;
  uint32_t fluid = 0;
  for (size_t i = 0; i < S->nc; ++i) fluid += is_fluid(kind_load(S, i)) ? 1u : 0u;
  st->fluid_cells = fluid;
  /* warm start injected after classify zeroed the pressure (shader.rs:1187) */
  int warmed = 0;
  if (S->s.warm_start == CS_WARM_PREVIOUS) {
    for (size_t i = 0; i < S->nc; ++i)
      if (is_fluid(kind_load(S, i))) LANE(S, 0, i) = S->prev_p[i];
    warmed = 1;
  } else if (S->s.warm_start == CS_WARM_EXTERNAL && S->has_warm) {
    for (size_t i = 0; i < S->nc; ++i)
      if (is_fluid(kind_load(S, i))) pressure_store(S, u, i, S->warm[i]);
    warmed = 1;
  }
  build_rows(S, u);
  if (S->s.rbgs_mode == CS_RBGS_FIXED_PAIRS) {
    for (uint32_t k = 0; k < S->s.pressure_rbgs_pairs; ++k) {
      k_rbgs_half(S, u, 0u);
      k_rbgs_half(S, u, 1u);
    }
    st->iterations = S->s.pressure_rbgs_pairs;
    st->final_residual = residual_rows(S, u);
    st->converged = st->final_residual <= (double)S->s.rbgs_tolerance;
  } else {
    double r = residual_rows(S, u);
    if (warmed && S->s.warm_start == CS_WARM_EXTERNAL && r > (double)S->s.warm_fallback_residual) {
      /* the predicted start is worse than the zero start: fall back */
      for (size_t i = 0; i < S->nc; ++i) LANE(S, 0, i) = 0;
      st->fell_back = 1;
      r = residual_rows(S, u);
    }
    uint32_t it = 0, every = S->s.rbgs_check_every ? S->s.rbgs_check_every : 8u;
    /* the multigrid start; after it the first residual test comes after
     * `every` iterations, so the clamp's projected updates always run */
    uint32_t min_it = 0;
    if (S->s.rbgs_mg_cycles && !(r <= (double)S->s.rbgs_tolerance)) {
      r = mg_start(S, u, S->s.rbgs_mg_cycles, r);
      min_it = every;
    }
    st->initial_residual = r;
    if (S->s.rbgs_rescale_after && !(r <= (double)S->s.rbgs_tolerance)) {
      /* untested iterations, then the rescale; the next test is at it = rescale_after + every */
      for (uint32_t k = 0; k < S->s.rbgs_rescale_after; ++k) {
        k_rbgs_half(S, u, 0u);
        k_rbgs_half(S, u, 1u);
      }
      it = S->s.rbgs_rescale_after;
      rescale_pressure(S, u);
    }
    while ((it < min_it || !(r <= (double)S->s.rbgs_tolerance)) && it < S->s.rbgs_max_iters) {
      k_rbgs_half(S, u, 0u);
      k_rbgs_half(S, u, 1u);
      ++it;
      if (it % every == 0u || it >= S->s.rbgs_max_iters) r = residual_rows(S, u);
    }
    st->iterations = it;
    st->final_residual = r;
    st->converged = r <= (double)S->s.rbgs_tolerance;
  }
  for (size_t i = 0; i < S->nc; ++i) {
    S->prev_p[i] = LANE(S, 0, i);
    S->saved_p[i] = LANE(S, 0, i);
    S->saved_kind[i] = LANE(S, 3, i);
  }
}

/* ---- project_pressure (shader.rs:1457-1543) -------------------------------- */
static float pressure_weighted_or_mirror(orc_sim* S, const orc_params* u, i3 c, float mirror, int32_t sk, float sf) {
  if (!on_grid(u, c.x, c.y, c.z)) return mirror;
  size_t ci = cidx(u, (uint32_t)c.x, (uint32_t)c.y, (uint32_t)c.z);
  int32_t k = kind_load(S, ci);
  if (is_solid(k)) return mirror;
  float np = 0.0f;
  if (is_fluid(k)) np = pressure_load(S, u, ci);
  float fw = face_weight(S, u, sk, sf, ci, k);
  return o_mix(mirror, np, fw);
}
static void k_project(orc_sim* S, const orc_params* u) {
  ORC_PARALLEL_FOR
  for (size_t idx = 0; idx < S->nc; ++idx) {
    float* g = &S->grid_vel[idx * 4];
    if (g[3] < 1e-6f) continue;
    int32_t kind = kind_load(S, idx);
    if (!is_fluid(kind)) continue;
    uint32_t ix, iy, iz;
    decompose(u, idx, &ix, &iy, &iz);
    float ph = pressure_load(S, u, idx);
    float sf = fill_fraction(S, u, idx, kind);
    i3 c = {(int32_t)ix, (int32_t)iy, (int32_t)iz}, a;
    a = c; a.x -= 1; float pxm = pressure_weighted_or_mirror(S, u, a, ph, kind, sf);
    a = c; a.x += 1; float pxp = pressure_weighted_or_mirror(S, u, a, ph, kind, sf);
    a = c; a.y -= 1; float pym = pressure_weighted_or_mirror(S, u, a, ph, kind, sf);
    a = c; a.y += 1; float pyp = pressure_weighted_or_mirror(S, u, a, ph, kind, sf);
    a = c; a.z -= 1; float pzm = pressure_weighted_or_mirror(S, u, a, ph, kind, sf);
    a = c; a.z += 1; float pzp = pressure_weighted_or_mirror(S, u, a, ph, kind, sf);
    float h = 0.5f * u->inv_dx;
    v3 grad = V(h * (pxp - pxm), h * (pyp - pym), h * (pzp - pzm));
    v3 v = vsub(V(g[0], g[1], g[2]), vmul(grad, u->dt));
    if (kind == O_CELL_BED) {
      int32_t id = S->bed_lookup[idx];
      if (id >= 0 && (uint32_t)id < S->num_bed) {
        float k = bed_compacted_permeability(S, u, id);
        float rate = u->water_nu / k;
        float damping = 1.0f / (1.0f + o_max(rate * u->dt, 0.0f));
        v3 bv = bed_matrix_velocity_cell(S, u, c);
        v3 before = v;
        v = vadd(bv, vmul(vsub(v, bv), damping));
        /* the reaction on the grains (shader.rs:1533-1535); a field bed has none */
        if (S->bed_n) {
          v3 wi = vmul(vsub(v, before), g[3]);
          bed_impulse_neighborhood(S, u, c, vmul(V(-wi.x, -wi.y, -wi.z), O_BED_REACTION_ALPHA));
        }
      }
    }
    float speed = vlen(v);
    if (speed > u->vel_cap) v = vmul(v, u->vel_cap / speed);
    g[0] = v.x; g[1] = v.y; g[2] = v.z;
  }
}

/* ---- post-projection divergence (no reference pass) ------------------------
 * The A7 divergence stencil (shader.rs:1278-1335) evaluated on grid_vel after
 * project_pressure and its boundary_project, over the cells classify made
 * fluid: max|div| (encoded like shader.rs:1363-1365) and max|div - target|
 * (the quantity the projection drives to zero) into the free metric slots. */
static void k_post_divergence(orc_sim* S, const orc_params* u) {
  ORC_PARALLEL_FOR
  for (size_t idx = 0; idx < S->nc; ++idx) {
    if (!is_fluid(kind_load(S, idx))) continue;
    uint32_t ix, iy, iz;
    decompose(u, idx, &ix, &iy, &iz);
    v3 cc = V(u->origin[0] + ((float)ix + 0.5f) * u->dx, u->origin[1] + ((float)iy + 0.5f) * u->dx,
              u->origin[2] + ((float)iz + 0.5f) * u->dx);
    float div = cell_divergence(S, u, ix, iy, iz, cc);
    uint32_t fd = o_f2u(o_clamp(fabsf(div) * u->metrics_div_fp_scale, 0.0f, 2147483648.0f));
    uint32_t fr = o_f2u(o_clamp(fabsf(div - S->tgt[idx]) * u->metrics_div_fp_scale, 0.0f, 2147483648.0f));
    orc_max_u32(&S->metrics[O_M_POST_MAX_ABS_DIV], fd);
    orc_max_u32(&S->metrics[O_M_POST_MAX_ABS_RES], fr);
  }
}

/* ---- packing (shader.rs:1547-1626) ----------------------------------------- */
static float packing_or_mirror(orc_sim* S, const orc_params* u, i3 c, float mirror) {
  if (!on_grid(u, c.x, c.y, c.z)) return mirror;
  size_t ci = cidx(u, (uint32_t)c.x, (uint32_t)c.y, (uint32_t)c.z);
  int32_t k = kind_load(S, ci);
  if (is_solid(k)) return mirror;
  if (!is_fluid(k)) return 0.0f;
  return packing_load(S, u, ci);
}
static void k_packing_prepare(orc_sim* S, const orc_params* u) {
  ORC_PARALLEL_FOR
  for (size_t idx = 0; idx < S->nc; ++idx) {
    int32_t kind = kind_load(S, idx);
    if (!is_fluid(kind) || kind == O_CELL_BED) { packing_store(S, u, idx, 0.0f); continue; }
    float cv = u->dx * u->dx * u->dx;
    float packed = o_max(rest_vol(S, u, idx), cur_vol(S, u, idx)) / o_max(cv, 1e-8f);
    float tgt = o_min(u->max_rest_fraction, 1.0f);
    float over = o_max(packed - tgt, 0.0f);
    packing_store(S, u, idx, u->bulk_K * over);
  }
}
static void k_packing_apply(orc_sim* S, const orc_params* u) {
  float occ = u->nominal_mass * 0.1f;
  ORC_PARALLEL_FOR
  for (size_t idx = 0; idx < S->nc; ++idx) {
    float* g = &S->grid_vel[idx * 4];
    if (g[3] <= occ) continue;
    int32_t kind = kind_load(S, idx);
    if (!is_fluid(kind)) continue;
    uint32_t ix, iy, iz;
    decompose(u, idx, &ix, &iy, &iz);
    i3 c = {(int32_t)ix, (int32_t)iy, (int32_t)iz}, a;
    float ph = packing_load(S, u, idx);
    a = c; a.x -= 1; float pxm = packing_or_mirror(S, u, a, ph);
    a = c; a.x += 1; float pxp = packing_or_mirror(S, u, a, ph);
    a = c; a.y -= 1; float pym = packing_or_mirror(S, u, a, ph);
    a = c; a.y += 1; float pyp = packing_or_mirror(S, u, a, ph);
    a = c; a.z -= 1; float pzm = packing_or_mirror(S, u, a, ph);
    a = c; a.z += 1; float pzp = packing_or_mirror(S, u, a, ph);
    float h = 0.5f * u->inv_dx;
    v3 grad = V(h * (pxp - pxm), h * (pyp - pym), h * (pzp - pzm));
    v3 v = vsub(V(g[0], g[1], g[2]), vmul(grad, u->dt));
    float speed = vlen(v);
    if (speed > u->vel_cap) v = vmul(v, u->vel_cap / speed);
    g[0] = v.x; g[1] = v.y; g[2] = v.z;
  }
}

/* ---- viscosity (shader.rs:1086-1177) ----------------------------------------- */
static void k_viscosity_prepare(orc_sim* S, const orc_params* u) {
  float occ = u->nominal_mass * 0.1f, support = u->nominal_mass * 2.0f;
  ORC_PARALLEL_FOR
  for (size_t idx = 0; idx < S->nc; ++idx) {
    const float* gv = &S->grid_vel[idx * 4];
    v3 vh = V(gv[0], gv[1], gv[2]);
    if (gv[3] <= occ || u->viscosity <= 0.0f) { vscratch_store(S, u, idx, vh); continue; }
    uint32_t ix, iy, iz;
    decompose(u, idx, &ix, &iy, &iz);
    if (gv[3] <= support) { vscratch_store(S, u, idx, vh); continue; }
    float alpha = o_clamp(u->viscosity * u->dt / o_max(u->dx * u->dx, 1e-6f), 0.0f, 0.14f);
    v3 nsum = V(0.0f, 0.0f, 0.0f);
    float wsum = 0.0f;
    uint32_t nfl = 0;
    for (int n = 0; n < 6; ++n) {
      int32_t x = (int32_t)ix + OFF6[n][0], y = (int32_t)iy + OFF6[n][1], z = (int32_t)iz + OFF6[n][2];
      if (!on_grid(u, x, y, z)) { wsum += 1.0f; continue; }
      if (class_solid(S, u, x, y, z)) { wsum += 1.0f; continue; }
      const float* ng = &S->grid_vel[cidx(u, (uint32_t)x, (uint32_t)y, (uint32_t)z) * 4];
      if (ng[3] <= occ) continue;
      float nw = o_clamp(ng[3] / o_max(o_max(gv[3], u->nominal_mass), 1e-6f), 0.0f, 1.0f);
      nsum = vadd(nsum, vmul(V(ng[0], ng[1], ng[2]), nw));
      wsum += nw;
      nfl += 1u;
    }
    if (nfl < 3u || wsum <= 1e-6f) { vscratch_store(S, u, idx, vh); continue; }
    v3 avg = vdiv(nsum, wsum);
    float blend = o_clamp(alpha * wsum, 0.0f, 0.65f);
    vscratch_store(S, u, idx, V(o_mix(vh.x, avg.x, blend), o_mix(vh.y, avg.y, blend), o_mix(vh.z, avg.z, blend)));
  }
}
static void k_viscosity_apply(orc_sim* S, const orc_params* u) {
  float occ = u->nominal_mass * 0.1f;
  ORC_PARALLEL_FOR
  for (size_t idx = 0; idx < S->nc; ++idx) {
    float* g = &S->grid_vel[idx * 4];
    float mass = g[3];
    if (mass <= occ) { g[0] = g[1] = g[2] = g[3] = 0.0f; continue; }
    v3 v = vscratch_load(S, u, idx);
    g[0] = v.x; g[1] = v.y; g[2] = v.z; g[3] = mass;
  }
}

/* ---- G2P (shader.rs:1683-1881) --------------------------------------------- */
static void k_g2p(orc_sim* S, const orc_params* u) {
  uint32_t np = u->num_water + u->num_bed_particles;
  float thr = u->nominal_mass * 0.10f;
  ORC_PARALLEL_FOR
  for (uint32_t pid = 0; pid < np; ++pid) {
    float* p = &S->particles[(size_t)pid * 8];
    float* a = &S->affine[(size_t)pid * 12];
    v3 xp = V(p[0], p[1], p[2]);
    float J_old = p[3], mass_p = p[7], phase = a[3];
    if (!(phase < 0.5f)) continue;
    if (mass_p <= thr) { p[7] = 0.0f; continue; }
    v3 gp = vmul(vsub(xp, V(u->origin[0], u->origin[1], u->origin[2])), u->inv_dx);
    int32_t bx = o_f2i(floorf(gp.x - 0.5f)), by = o_f2i(floorf(gp.y - 0.5f)), bz = o_f2i(floorf(gp.z - 0.5f));
    v3 fx = V(gp.x - (float)bx, gp.y - (float)by, gp.z - (float)bz);
    float wx[3], wy[3], wz[3];
    wx[0] = 0.5f * (1.5f - fx.x) * (1.5f - fx.x);
    wx[1] = 0.75f - (fx.x - 1.0f) * (fx.x - 1.0f);
    wx[2] = 0.5f * (fx.x - 0.5f) * (fx.x - 0.5f);
    wy[0] = 0.5f * (1.5f - fx.y) * (1.5f - fx.y);
    wy[1] = 0.75f - (fx.y - 1.0f) * (fx.y - 1.0f);
    wy[2] = 0.5f * (fx.y - 0.5f) * (fx.y - 0.5f);
    wz[0] = 0.5f * (1.5f - fx.z) * (1.5f - fx.z);
    wz[1] = 0.75f - (fx.z - 1.0f) * (fx.z - 1.0f);
    wz[2] = 0.5f * (fx.z - 0.5f) * (fx.z - 0.5f);
    v3 nv = V(0, 0, 0), C0 = V(0, 0, 0), C1 = V(0, 0, 0), C2 = V(0, 0, 0);
    float supported = 0.0f, local_mass = 0.0f, bed_overlap = 0.0f, bed_perm_sum = 0.0f;
    v3 bed_vsum = V(0, 0, 0);
    float B = 4.0f * u->inv_dx * u->inv_dx;
    float cdx = u->dx;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          int32_t cx = bx + i, cy = by + j, cz = bz + k;
          if (!on_grid(u, cx, cy, cz)) continue;
          float w = wx[i] * wy[j] * wz[k];
          size_t ci = cidx(u, (uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
          const float* g = &S->grid_vel[ci * 4];
          v3 gv = V(g[0], g[1], g[2]);
          v3 dpos = vmul(vsub(V((float)i, (float)j, (float)k), fx), cdx);
          float gm = g[3];
          if (gm > 1e-6f) {
            nv = vadd(nv, vmul(gv, w));
            float wB = w * B;
            C0 = vadd(C0, vmul(vmul(gv, wB), dpos.x));
            C1 = vadd(C1, vmul(vmul(gv, wB), dpos.y));
            C2 = vadd(C2, vmul(vmul(gv, wB), dpos.z));
            supported += w;
            local_mass += w * gm;
          }
          int32_t id = S->bed_lookup[ci];
          if (S->num_bed > 0u && bed_valid(S, id)) {
            bed_overlap += w;
            bed_vsum = vadd(bed_vsum, vmul(V(S->bed[id].velocity[0], S->bed[id].velocity[1], S->bed[id].velocity[2]), w));
            bed_perm_sum += w * bed_compacted_permeability(S, u, id);
          }
        }
    float support_ratio = o_clamp(supported, 0.0f, 1.0f);
    float j_support = support_ratio;
    if (supported > 1e-6f) {
      float inv = 1.0f / supported;
      nv = vmul(nv, inv); C0 = vmul(C0, inv); C1 = vmul(C1, inv); C2 = vmul(C2, inv);
    }
    int in_cup = 0;
    if (u->geo.cup_enabled) {
      float lim = u->geo.cup_radius + u->contact_offset;
      in_cup = xp.y < u->geo.cup_top_y && (xp.x * xp.x + xp.z * xp.z) < lim * lim;
    }
    float dense_ratio = o_clamp(local_mass / o_max(u->nominal_mass * 4.0f, 1e-6f), 0.0f, 1.0f);
    float porous = o_clamp(bed_overlap, 0.0f, 1.0f);
    v3 pbv = V(0, 0, 0);
    float pk = u->min_k;
    if (bed_overlap > 1e-6f) {
      float ib = 1.0f / bed_overlap;
      pbv = vmul(bed_vsum, ib);
      pk = o_max(bed_perm_sum * ib, u->min_k);
    }
    if (support_ratio < 0.999f) {
      v3 ball = V(p[4], p[5] + u->gravity * u->dt, p[6]);
      float fgain = in_cup ? 1.15f : 1.35f;
      float fcap = in_cup ? 0.95f : 0.995f;
      float gain = o_mix(fgain, 0.18f, porous);
      float cap = o_mix(fcap, 0.16f, porous);
      float damp = o_mix(1.0f, 0.12f, dense_ratio);
      float pres = o_clamp((1.0f - support_ratio) * gain * damp, 0.0f, cap);
      nv = V(o_mix(nv.x, ball.x, pres), o_mix(nv.y, ball.y, pres), o_mix(nv.z, ball.z, pres));
      float ad = 1.0f - pres * 0.75f;
      C0 = vmul(C0, ad); C1 = vmul(C1, ad); C2 = vmul(C2, ad);
    }
    int airborne = porous <= 0.05f && sample_sdf(S, u, xp) > u->contact_offset * 2.0f;
    if (airborne) {
      float dense_mass = u->nominal_mass * 4.0f;
      float dr = o_clamp(local_mass / o_max(dense_mass, 1e-6f), 0.0f, 1.0f);
      j_support = o_min(j_support, dr);
      v3 ball = V(p[4], p[5] + u->gravity * u->dt, p[6]);
      float gain = in_cup ? 0.72f : 0.98f;
      float cap = in_cup ? 0.88f : 0.98f;
      float pres = o_clamp((1.0f - dr) * gain, 0.0f, cap);
      nv = V(o_mix(nv.x, ball.x, pres), o_mix(nv.y, ball.y, pres), o_mix(nv.z, ball.z, pres));
      float ad = 1.0f - pres * 0.65f;
      C0 = vmul(C0, ad); C1 = vmul(C1, ad); C2 = vmul(C2, ad);
    }
    if (porous > 1e-4f) {
      v3 rel = vsub(nv, pbv);
      float rate = u->water_nu / pk;
      float dd = 1.0f / (1.0f + o_max(rate * u->dt * porous, 0.0f));
      float rs = vlen(rel);
      float idd = 1.0f / (1.0f + o_max(rs * u->dt * u->inv_dx * porous * 0.35f, 0.0f));
      v3 before = nv;
      nv = vadd(pbv, vmul(vmul(rel, dd), idd));
      /* the reaction on the grains (shader.rs:1837-1843); a field bed has none */
      if (S->bed_n) {
        v3 wi = vmul(vsub(nv, before), mass_p);
        bed_impulse_neighborhood(S, u, world_to_cell(u, xp), vmul(V(-wi.x, -wi.y, -wi.z), O_BED_REACTION_ALPHA));
      }
      float ad = o_mix(1.0f, dd * idd, porous);
      C0 = vmul(C0, ad); C1 = vmul(C1, ad); C2 = vmul(C2, ad);
      j_support = o_max(j_support, porous);
    }
    float tr = C0.x + C1.y + C2.z;
    float Jc = o_clamp(J_old > 0.0f ? J_old : 1.0f, 0.40f, 2.00f);
    float Ja = o_clamp(Jc * orc_expf(o_clamp(u->dt * tr, -0.35f, 0.35f)), 0.40f, 2.00f);
    float Jb = o_smooth(0.35f, 0.85f, o_clamp(j_support, 0.0f, 1.0f));
    float Jn = o_clamp(o_mix(Jc, Ja, Jb), 0.40f, 2.00f);
    v3 np_ = vadd(xp, vmul(nv, u->dt));
    v3 mid = V(o_mix(xp.x, np_.x, 0.5f), o_mix(xp.y, np_.y, 0.5f), o_mix(xp.z, np_.z, 0.5f));
    contact_t c = sdf_contact(S, u, mid, nv);
    nv = c.vel;
    c = sdf_contact(S, u, np_, nv);
    np_ = c.pos; nv = c.vel;
    float margin = u->dx * 0.5f;
    np_.x = o_clamp(np_.x, u->origin[0] + margin, u->bmax[0] - margin);
    np_.y = o_clamp(np_.y, u->origin[1] + margin, u->bmax[1] - margin);
    np_.z = o_clamp(np_.z, u->origin[2] + margin, u->bmax[2] - margin);
    p[0] = np_.x; p[1] = np_.y; p[2] = np_.z; p[3] = Jn;
    p[4] = nv.x; p[5] = nv.y; p[6] = nv.z; p[7] = mass_p;
    a[0] = C0.x; a[1] = C0.y; a[2] = C0.z; a[3] = phase;
    a[4] = C1.x; a[5] = C1.y; a[6] = C1.z; a[7] = 0.0f;
    a[8] = C2.x; a[9] = C2.y; a[10] = C2.z; a[11] = 0.0f;
  }
}

/* ---- bed of particles (reference mode) ------------------------------------- */
/* bed_lookup_clear and bed_lookup_scatter (shader.rs:2291-2335): every bed
 * particle that is not water stamps its index into the 3x3x3 cells around
 * world_to_cell (shader.rs:668-671); the highest index wins (atomicMax).  Then
 * each particle's record is read from its BedExtract (shader.rs:35-38: porosity,
 * permeability, compaction in bed.yzw, saturation in extract.w) and from the
 * particle itself (height, velocity), as the cell readers of shader.rs:361-449
 * and 621-666 take them. */
static void bed_restamp(orc_sim* S, const orc_params* u) {
  for (size_t i = 0; i < S->nc; ++i) S->bed_lookup[i] = -1;
  for (uint32_t bid = 0; bid < S->bed_n; ++bid) {
    const float* p = S->bed_part + 8 * (size_t)bid;
    if (S->bed_aff[12 * (size_t)bid + 3] < 0.5f) continue; /* is_water_phase */
    int c[3];
    for (int k = 0; k < 3; ++k) c[k] = (int)floorf((p[k] - u->origin[k]) * u->inv_dx);
    for (int di = -1; di <= 1; ++di)
      for (int dj = -1; dj <= 1; ++dj)
        for (int dk = -1; dk <= 1; ++dk) {
          int x = c[0] + di, y = c[1] + dj, z = c[2] + dk;
          if (x < 0 || y < 0 || z < 0 || x >= (int)S->gx || y >= (int)S->gy || z >= (int)S->gz) continue;
          int32_t* slot = &S->bed_lookup[cidx(u, (uint32_t)x, (uint32_t)y, (uint32_t)z)];
          if ((int32_t)bid > *slot) *slot = (int32_t)bid;
        }
  }
  for (uint32_t bid = 0; bid < S->bed_n; ++bid) {
    const float* p = S->bed_part + 8 * (size_t)bid;
    const float* e = S->bed_ext + 8 * (size_t)bid;
    cs_bed_record* r = &S->bed[bid];
    r->porosity = e[1];
    r->permeability = e[2];
    r->compaction = e[3];
    r->saturation = e[7];
    r->sample_y = p[1];
    r->velocity[0] = p[4];
    r->velocity[1] = p[5];
    r->velocity[2] = p[6];
  }
}

/* bed_water_delta_add (shader.rs:455-460) */
static void bed_water_add(orc_sim* S, const orc_params* u, int32_t idx, float amount) {
  if (amount <= 0.0f) return;
  S->bed_dw[idx] += o_f2i(amount * u->fp_scale);
}

/* bed_coupling (shader.rs:1885-1957) with deposit_absorbed_bed_water
 * (shader.rs:524-576).  The filter-paper branch (shader.rs:1904-1911) runs
 * only for filter_absorption_rate > 0, which DEFAULT_BREW sets to 0
 * (brew_config.rs:68); it is left out. */
static void bed_coupling(orc_sim* S, const orc_params* u) {
  const uint32_t nbed = S->bed_n;
  const float max_sat = 42.0f * 80.0f / (float)nbed; /* mod.rs:577-581 */
  const float inactive = u->nominal_mass * 0.10f;
  for (uint32_t pid = 0; pid < u->num_water; ++pid) {
    float* p = S->particles + 8 * (size_t)pid;
    if (!(S->affine[12 * (size_t)pid + 3] < 0.5f)) continue;
    int c[3];
    for (int k = 0; k < 3; ++k) c[k] = (int)floorf((p[k] - u->origin[k]) * u->inv_dx);
    if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= (int)S->gx || c[1] >= (int)S->gy || c[2] >= (int)S->gz) continue;
    const float mass_p = p[7];
    const int32_t home = S->bed_lookup[cidx(u, (uint32_t)c[0], (uint32_t)c[1], (uint32_t)c[2])];
    if (home < 0 || (uint32_t)home >= nbed) continue;
    const float* be = S->bed_ext + 8 * (size_t)home;
    const float saturation = be[7];
    const float capacity = o_max(max_sat - be[0], 0.0f);
    if (capacity <= 1e-6f) continue;
    const float abs_rate = 1.6f * (1.0f - saturation) * u->dt; /* bed_absorption_rate, brew_config.rs:59 */
    const float speed = sqrtf(p[4] * p[4] + p[5] * p[5] + p[6] * p[6]);
    float absorbed = o_min(o_min(mass_p * o_clamp(abs_rate, 0.0f, 0.25f), mass_p * 0.5f), capacity);
    const float rest = mass_p - absorbed;
    if (rest > 0.0f && rest <= u->nominal_mass * 0.22f) {
      absorbed = o_min(mass_p, capacity);
    } else if (saturation > 0.55f && speed < 1.35f) {
      const float almost = o_min(mass_p, capacity);
      if (mass_p - almost <= u->nominal_mass * 0.35f) absorbed = almost;
    }
    if (mass_p - absorbed <= inactive && absorbed < mass_p) absorbed = o_max(mass_p - inactive * 1.05f, 0.0f);
    if (absorbed <= 1e-6f) continue;
    const float remaining = mass_p - absorbed;
    if (remaining <= inactive) {
      p[4] = p[5] = p[6] = p[7] = 0.0f;
    } else {
      p[7] = remaining;
    }
    const float neighbor_share = absorbed * 0.35f;
    bed_water_add(S, u, home, absorbed - neighbor_share);
    static const int off[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    int32_t ids[6];
    uint32_t n = 0;
    for (int k = 0; k < 6; ++k) {
      int x = c[0] + off[k][0], y = c[1] + off[k][1], z = c[2] + off[k][2];
      if (x < 0 || y < 0 || z < 0 || x >= (int)S->gx || y >= (int)S->gy || z >= (int)S->gz) continue;
      int32_t id = S->bed_lookup[cidx(u, (uint32_t)x, (uint32_t)y, (uint32_t)z)];
      if (id < 0 || id == home || (uint32_t)id >= nbed) continue;
      int dup = 0;
      for (uint32_t j = 0; j < n; ++j) if (ids[j] == id) dup = 1;
      if (!dup) ids[n++] = id;
    }
    if (n == 0) {
      bed_water_add(S, u, home, neighbor_share);
      continue;
    }
    const float each = neighbor_share / (float)n;
    for (uint32_t j = 0; j < n; ++j) bed_water_add(S, u, ids[j], each);
  }
}

/* extraction_advect (shader.rs:1961-1987): the water credit into pore water and
 * saturation, then extractable to dissolved mass at extraction_params.x = 0.01
 * (mod.rs:645). */
static void extraction_advect(orc_sim* S, const orc_params* u) {
  const float max_sat = 42.0f * 80.0f / (float)S->bed_n;
  for (uint32_t bid = 0; bid < S->bed_n; ++bid) {
    float* be = S->bed_ext + 8 * (size_t)bid;
    const float absorbed = (float)S->bed_dw[bid] * u->inv_fp_scale;
    S->bed_dw[bid] = 0;
    if (absorbed > 0.0f) {
      be[0] += absorbed;
      be[7] = o_clamp(be[0] / o_max(max_sat, 1e-6f), 0.0f, 1.0f);
    }
    const float sat = be[7];
    if (sat > 0.01f) {
      const float flux = 0.01f * be[4] * sat * u->dt;
      be[4] = o_max(be[4] - flux, 0.0f);
      be[5] += flux;
    }
  }
}

/* pressure_gradient_at_cell (shader.rs:238-259) as bed_dynamics sees it: lane 0
 * still holds the solved pressure, but lane 3, which the reference reads as the
 * cell kind, holds viscosity_prepare's velocity scratch z by then (shader.rs:
 * 158-164, 274-283), so a neighbour counts as solid (mirror) exactly when that
 * scratch is 4 in fixed point.  The oracle keeps the reference's lane layout,
 * so reading the lanes reproduces this. */
static float grad_pressure_or_mirror(orc_sim* S, const orc_params* u, i3 c, float mirror) {
  if (!on_grid(u, c.x, c.y, c.z)) return mirror;
  size_t ci = cidx(u, (uint32_t)c.x, (uint32_t)c.y, (uint32_t)c.z);
  if (is_solid(kind_load(S, ci))) return mirror;
  return pressure_load(S, u, ci);
}
static v3 pressure_gradient_at_cell(orc_sim* S, const orc_params* u, i3 c) {
  if (!on_grid(u, c.x, c.y, c.z)) return V(0.0f, 0.0f, 0.0f);
  float ph = pressure_load(S, u, cidx(u, (uint32_t)c.x, (uint32_t)c.y, (uint32_t)c.z));
  i3 a;
  a = c; a.x -= 1; float pxm = grad_pressure_or_mirror(S, u, a, ph);
  a = c; a.x += 1; float pxp = grad_pressure_or_mirror(S, u, a, ph);
  a = c; a.y -= 1; float pym = grad_pressure_or_mirror(S, u, a, ph);
  a = c; a.y += 1; float pyp = grad_pressure_or_mirror(S, u, a, ph);
  a = c; a.z -= 1; float pzm = grad_pressure_or_mirror(S, u, a, ph);
  a = c; a.z += 1; float pzp = grad_pressure_or_mirror(S, u, a, ph);
  float h = 0.5f * u->inv_dx;
  return V(h * (pxp - pxm), h * (pyp - pym), h * (pzp - pzm));
}

/* resolve_coffee_particle_packing (shader.rs:825-877): push a grain out of its
 * nearest stamped neighbour closer than 1.78 grain radii.  The reference reads
 * the neighbours' positions while bed_dynamics moves them, so what it sees
 * depends on thread scheduling; here every grain reads the positions from
 * before the pass (`snap`), which the device pass reproduces in any order. */
static contact_t coffee_packing(orc_sim* S, const orc_params* u, v3 pos, v3 vel, uint32_t self, const float* snap) {
  v3 cd = V(0.0f, 0.0f, 0.0f);
  float closest = 1e9f;
  int found = 0;
  i3 cc = world_to_cell(u, pos);
  for (int di = -1; di <= 1; ++di)
    for (int dj = -1; dj <= 1; ++dj)
      for (int dk = -1; dk <= 1; ++dk) {
        int32_t x = cc.x + di, y = cc.y + dj, z = cc.z + dk;
        if (!on_grid(u, x, y, z)) continue;
        int32_t o = S->bed_lookup[cidx(u, (uint32_t)x, (uint32_t)y, (uint32_t)z)];
        if (o < 0 || (uint32_t)o == self || (uint32_t)o >= S->bed_n) continue;
        v3 d = vsub(pos, V(snap[3 * (size_t)o], snap[3 * (size_t)o + 1], snap[3 * (size_t)o + 2]));
        float dist = vlen(d);
        if (dist < closest) { closest = dist; cd = d; found = 1; }
      }
  float min_sep = bed_particle_radius(u) * 1.78f;
  if (found && closest < min_sep) {
    v3 n = V(0.0f, 1.0f, 0.0f);
    if (closest > 1e-5f) n = vdiv(cd, closest);
    float pen = min_sep - closest;
    pos = vadd(pos, vmul(n, o_min(pen * 0.65f, u->dx * 0.22f)));
    float vn = vdot(vel, n);
    if (vn < 0.0f) vel = vsub(vel, vmul(n, vn));
  }
  contact_t c = {pos, vel};
  return c;
}

/* bed_dynamics (shader.rs:1991-2238): each grain takes its reaction impulse,
 * then either drifts with the water (suspended phase 2) or moves as an
 * overdamped, plastically compacting porous matrix around its rest position
 * (anchored phase 1), with packing against its neighbours and the paper
 * filter's contact.  bed_damping 8.5, bed_compaction_rate 5.5, bed_impact 8.0
 * (mod.rs:644-649, brew_config.rs:69-70). */
static void bed_dynamics(orc_sim* S, const orc_params* u) {
  const uint32_t nbed = S->bed_n;
  const float bed_damping = 8.5f, bed_compaction = 5.5f, bed_impact = 8.0f;
  const float floor_y = coffee_filter_floor_y(u);
  const float dx = u->dx, dt = u->dt;
  for (uint32_t b = 0; b < nbed; ++b) {
    S->bed_snap[3 * (size_t)b] = S->bed_part[8 * (size_t)b];
    S->bed_snap[3 * (size_t)b + 1] = S->bed_part[8 * (size_t)b + 1];
    S->bed_snap[3 * (size_t)b + 2] = S->bed_part[8 * (size_t)b + 2];
  }
  ORC_PARALLEL_FOR
  for (uint32_t bid = 0; bid < nbed; ++bid) {
    float* p = S->bed_part + 8 * (size_t)bid;
    float* a = S->bed_aff + 12 * (size_t)bid;
    float* be = S->bed_ext + 8 * (size_t)bid;
    float phase = a[3];
    v3 rest = V(a[4], a[5], a[6]);
    v3 pos = V(p[0], p[1], p[2]);
    const float mass_p = p[7];
    /* bed_impulse_delta_exchange (shader.rs:517-522) */
    v3 imp = V((float)S->bed_dw[1u * nbed + bid] * u->inv_fp_scale, (float)S->bed_dw[2u * nbed + bid] * u->inv_fp_scale,
               (float)S->bed_dw[3u * nbed + bid] * u->inv_fp_scale);
    S->bed_dw[1u * nbed + bid] = S->bed_dw[2u * nbed + bid] = S->bed_dw[3u * nbed + bid] = 0;
    v3 brv = vdiv(imp, o_max(mass_p, u->nominal_mass * 0.25f));
    const float max_rs = o_min(u->vel_cap * 0.012f, dx / o_max(dt, 1e-6f) * 0.02f);
    const float rs = vlen(brv);
    if (rs > max_rs && rs > 1e-6f) brv = vmul(brv, max_rs / rs);
    /* the water around the grain: quadratic B-spline gather of grid_vel */
    v3 gp = vmul(vsub(pos, V(u->origin[0], u->origin[1], u->origin[2])), u->inv_dx);
    int32_t bx = o_f2i(floorf(gp.x - 0.5f)), by = o_f2i(floorf(gp.y - 0.5f)), bz = o_f2i(floorf(gp.z - 0.5f));
    v3 fx = V(gp.x - (float)bx, gp.y - (float)by, gp.z - (float)bz);
    float wx[3], wy[3], wz[3];
    wx[0] = 0.5f * (1.5f - fx.x) * (1.5f - fx.x);
    wx[1] = 0.75f - (fx.x - 1.0f) * (fx.x - 1.0f);
    wx[2] = 0.5f * (fx.x - 0.5f) * (fx.x - 0.5f);
    wy[0] = 0.5f * (1.5f - fx.y) * (1.5f - fx.y);
    wy[1] = 0.75f - (fx.y - 1.0f) * (fx.y - 1.0f);
    wy[2] = 0.5f * (fx.y - 0.5f) * (fx.y - 0.5f);
    wz[0] = 0.5f * (1.5f - fx.z) * (1.5f - fx.z);
    wz[1] = 0.75f - (fx.z - 1.0f) * (fx.z - 1.0f);
    wz[2] = 0.5f * (fx.z - 0.5f) * (fx.z - 0.5f);
    v3 water_v = V(0.0f, 0.0f, 0.0f);
    float water_mass = 0.0f, asw = 0.0f, lasw = 0.0f;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          int32_t cx = bx + i, cy = by + j, cz = bz + k;
          if (!on_grid(u, cx, cy, cz)) continue;
          float w = wx[i] * wy[j] * wz[k];
          size_t ci = cidx(u, (uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
          const float* g = &S->grid_vel[ci * 4];
          water_v = vadd(water_v, vmul(V(g[0], g[1], g[2]), w));
          water_mass += w * g[3];
          int32_t sb = S->bed_lookup[ci];
          if (sb >= 0 && (uint32_t)sb != bid) {
            asw += w;
            float ccy = u->origin[1] + ((float)cy + 0.5f) * dx;
            if (ccy <= pos.y + dx * 0.25f) lasw += w;
          }
        }
    if (phase >= 1.5f) { /* suspended grain (shader.rs:2065-2105) */
      float ws = o_clamp(water_mass / o_max(u->nominal_mass * 2.0f, 1e-6f), 0.0f, 1.0f);
      float drag = o_clamp((8.0f + 18.0f * ws) * dt, 0.0f, 1.0f);
      v3 unsup = vadd(vadd(V(p[4], p[5], p[6]), brv), V(0.0f, u->gravity * dt * 0.85f, 0.0f));
      float settle = o_mix(1.10f, 0.32f, ws);
      v3 carried = vadd(water_v, V(0.0f, -settle, 0.0f));
      v3 sv = V(o_mix(unsup.x, carried.x, drag), o_mix(unsup.y, carried.y, drag), o_mix(unsup.z, carried.z, drag));
      float sp = vlen(sv);
      if (sp > u->vel_cap) sv = vmul(sv, u->vel_cap / sp);
      v3 spos = vadd(pos, vmul(sv, dt));
      contact_t c = sdf_contact_as(S, u, spos, sv, 1);
      spos = c.pos; sv = c.vel;
      p[0] = spos.x; p[1] = spos.y; p[2] = spos.z;
      p[4] = sv.x; p[5] = sv.y; p[6] = sv.z; p[7] = mass_p;
      int anchored = lasw > 0.06f || asw > 0.20f;
      int dense = asw > 0.42f;
      int redeposit = (spos.y <= floor_y + dx * 0.35f && ws < 0.35f) ||
                      (anchored && sv.y <= 0.20f && (ws < 0.55f || dense));
      if (redeposit) phase = 1.0f;
      a[3] = phase;
      a[4] = spos.x; a[5] = spos.y; a[6] = spos.z; a[7] = 0.0f;
      be[3] *= 0.985f;
      continue;
    }
    /* anchored grain (shader.rs:2107-2237) */
    const float sat = be[7];
    const float wl = o_clamp(water_mass / o_max(u->nominal_mass * 1.5f, 1e-6f), 0.0f, 1.0f);
    const float wet = 1.0f + sat * 0.40f;
    const float mob = o_clamp((0.30f + wl * 0.62f) * wet, 0.0f, 1.0f);
    const float sf = o_clamp((rest.y + 3.0f) / 3.5f, 0.12f, 1.0f);
    const float damping = o_clamp(1.0f - bed_damping * dt, 0.0f, 1.0f);
    v3 vel = vadd(V(p[4], p[5], p[6]), brv);
    const float sr = o_clamp(1.0f - asw / 0.55f, 0.0f, 1.0f);
    const float ed = o_clamp((2.0f + 5.0f * sat) * wl * sr * dt, 0.0f, 0.08f);
    v3 target = vadd(water_v, V(0.0f, -o_mix(0.95f, 0.35f, wl), 0.0f));
    vel = V(o_mix(vel.x, target.x, ed), o_mix(vel.y, target.y, ed), o_mix(vel.z, target.z, ed));
    const float ls = sqrtf(water_v.x * water_v.x + water_v.z * water_v.z);
    const float ldx = ls > 1e-6f ? water_v.x / ls : 0.0f, ldz = ls > 1e-6f ? water_v.z / ls : 0.0f;
    const float lss = o_min(ls, fabsf(o_min(water_v.y, 0.0f)) * 0.18f);
    v3 impact = V(ldx * lss * 0.08f, o_min(water_v.y, 0.0f) * 0.62f, ldz * lss * 0.08f);
    vel = vadd(vel, vmul(vmul(vmul(vmul(impact, bed_impact), mob), sf), dt));
    v3 pg = pressure_gradient_at_cell(S, u, world_to_cell(u, pos));
    const float sfx = -pg.x, sfz = -pg.z;
    const float sps = sqrtf(sfx * sfx + sfz * sfz);
    const float sdx = sps > 1e-6f ? sfx / sps : 0.0f, sdz = sps > 1e-6f ? sfz / sps : 0.0f;
    const float drive = o_min(sps, u->vel_cap / o_max(dt, 1e-6f));
    const float pl = o_clamp(water_mass * (0.25f + sat), 0.0f, 1.0f);
    const float slx = sdx * drive * pl * 0.08f, slz = sdz * drive * pl * 0.08f;
    vel.x += slx * bed_impact * mob * sf * dt;
    vel.z += slz * bed_impact * mob * sf * dt;
    vel = vmul(vel, damping);
    v3 np = vadd(pos, vmul(vel, dt));
    v3 off = vsub(np, rest);
    const float ll = sqrtf(off.x * off.x + off.z * off.z);
    const float lpt = dx * 0.22f;
    if (ll > lpt && ll > 1e-6f) {
      const float ox = off.x / ll, oz = off.z / ll;
      const float excess = ll - lpt;
      const float lp = o_clamp((0.04f + sat * 0.35f + mob * 0.45f + pl * 0.25f) * sf * bed_compaction * dt, 0.0f, 0.18f);
      rest.x += ox * excess * lp;
      rest.z += oz * excess * lp;
      off = vsub(np, rest);
    }
    const float cll = sqrtf(off.x * off.x + off.z * off.z);
    const float ml = dx * 0.9f * sf;
    if (cll > ml && cll > 1e-6f) {
      const float ox = off.x / cll, oz = off.z / cll;
      off.x = ox * ml;
      off.z = oz * ml;
      vel.x *= 0.4f;
      vel.z *= 0.4f;
    }
    off.y = o_clamp(off.y, -dx * (1.75f * sf + 0.2f), dx * 0.04f);
    np = vadd(rest, off);
    contact_t pc = coffee_packing(S, u, np, vel, bid, S->bed_snap);
    np = pc.pos; vel = pc.vel;
    const float comp = o_max(rest.y - np.y, 0.0f);
    const float pth = dx * 0.18f;
    if (comp > pth) {
      const float ex = comp - pth;
      const float plast = o_clamp((0.05f + sat * 0.35f + mob * 0.55f) * sf * bed_compaction * dt, 0.0f, 0.18f);
      rest.y -= ex * plast;
    }
    const int detach = sat > 0.82f && pl > 0.65f && wl > 0.65f && sr > 0.35f &&
                       (comp > dx * 0.65f || cll > dx * 0.65f);
    const int entrained = sat > 0.58f && wl > 0.74f && sr > 0.55f && vlen(water_v) > 0.95f &&
                          np.y > floor_y + dx * 0.70f;
    const int isolated = sat > 0.55f && water_mass < u->nominal_mass * 0.75f && asw < 0.04f && lasw < 0.03f &&
                         np.y > floor_y + dx * 0.50f;
    if (detach || entrained || isolated) {
      phase = 2.0f;
      rest = np;
    }
    contact_t c = sdf_contact_as(S, u, np, vel, 1);
    np = c.pos; vel = c.vel;
    p[0] = np.x; p[1] = np.y; p[2] = np.z;
    p[4] = vel.x; p[5] = vel.y; p[6] = vel.z; p[7] = mass_p;
    a[3] = phase;
    a[4] = rest.x; a[5] = rest.y; a[6] = rest.z; a[7] = 0.0f;
    const float gc = o_max((rest.y - np.y) / o_max(dx, 1e-6f), 0.0f);
    be[3] = o_max(be[3] * 0.995f, gc);
  }
}

/* ---- one substep in step_frame order (mod.rs:271-418) ---------------------- */
static int substep(orc_sim* S, float sub_dt, int stop_after, int emit) {
  if (emit) orc_emit(S, sub_dt);
  orc_write_uniforms(S, sub_dt, &S->u);
  orc_params* u = &S->u;
  memset(S->grid, 0, S->nc * 6 * sizeof(int32_t));
  memset(S->grid_vel, 0, S->nc * 4 * sizeof(float));
  memset(S->metrics, 0, sizeof(S->metrics));
  if (S->bed_n) bed_restamp(S, u); /* mod.rs:322-331 */
  k_p2g(S, u);
  memcpy(S->p2g_lanes, S->grid, S->nc * 6 * sizeof(int32_t));
  if (stop_after == ORC_PASS_P2G) return 0;
  k_grid_update(S, u);
  k_boundary(S, u);
  if (stop_after == ORC_PASS_GRID) return 0;
  k_classify(S, u);
  if (stop_after == ORC_PASS_CLASSIFY) return 0;
  solve_pressure(S, u);
  if (stop_after == ORC_PASS_RBGS) return 0;
  k_project(S, u);
  k_boundary(S, u);
  k_post_divergence(S, u);
  if (stop_after == ORC_PASS_PROJECT) return 0;
  k_packing_prepare(S, u);
  k_packing_apply(S, u);
  k_boundary(S, u);
  if (stop_after == ORC_PASS_PACKING) return 0;
  k_viscosity_prepare(S, u);
  k_viscosity_apply(S, u);
  k_boundary(S, u);
  if (stop_after == ORC_PASS_VISCOSITY) return 0;
  k_g2p(S, u);
  if (stop_after == ORC_PASS_G2P) return 0;
  if (S->bed_n) { /* mod.rs:391-406 */
    bed_coupling(S, u);
    extraction_advect(S, u);
    bed_dynamics(S, u);
  }
  S->total_time += sub_dt;
  return 0;
}

int orc_debug_step_until(orc_sim* S, float sub_dt, int stop_after) {
  return substep(S, sub_dt, stop_after, 0);
}

/* One pass on the current state, with the uniforms of a substep of sub_dt and
 * nothing cleared, for the closed-form tests of the passes. */
int orc_debug_run_pass(orc_sim* S, float sub_dt, int pass) {
  orc_write_uniforms(S, sub_dt, &S->u);
  orc_params* u = &S->u;
  switch (pass) {
    case ORC_PASS_P2G: k_p2g(S, u); break;
    case ORC_PASS_GRID: k_grid_update(S, u); break;
    case ORC_PASS_BOUNDARY: k_boundary(S, u); break;
    case ORC_PASS_CLASSIFY: k_classify(S, u); break;
    case ORC_PASS_RBGS: build_rows(S, u); k_rbgs_half(S, u, 0u); k_rbgs_half(S, u, 1u); break;
    case ORC_PASS_PROJECT: k_project(S, u); break;
    case ORC_PASS_PACKING: k_packing_prepare(S, u); k_packing_apply(S, u); break;
    case ORC_PASS_VISCOSITY: k_viscosity_prepare(S, u); k_viscosity_apply(S, u); break;
    case ORC_PASS_G2P: k_g2p(S, u); break;
    case ORC_PASS_POST_DIVERGENCE: k_post_divergence(S, u); break;
    case ORC_PASS_BED_STAMP: if (S->bed_n) bed_restamp(S, u); break;
    case ORC_PASS_BED_DYNAMICS: if (S->bed_n) bed_dynamics(S, u); break;
    case ORC_PASS_MG: build_rows(S, u); (void)mg_start(S, u, S->s.rbgs_mg_cycles, INFINITY); break;
    default: return CS_ERR_INVALID;
  }
  return CS_OK;
}
void orc_set_previous_pressure(orc_sim* S, const int32_t* p) { memcpy(S->prev_p, p, S->nc * sizeof(int32_t)); }
void orc_debug_set_grid_vel(orc_sim* S, const float* vel4) { memcpy(S->grid_vel, vel4, S->nc * 4 * sizeof(float)); }
void orc_debug_set_grid_lanes(orc_sim* S, const int32_t* lanes6) { memcpy(S->grid, lanes6, S->nc * 6 * sizeof(int32_t)); }
void orc_debug_read_grid_lanes(orc_sim* S, int32_t* lanes6) { memcpy(lanes6, S->grid, S->nc * 6 * sizeof(int32_t)); }
void orc_debug_clear_metrics(orc_sim* S) { memset(S->metrics, 0, sizeof(S->metrics)); }
void orc_set_mg_cycles(orc_sim* S, uint32_t cycles) { S->s.rbgs_mg_cycles = cycles; }
void orc_debug_read_bed_delta(orc_sim* S, int32_t* d) {
  if (S->bed_n) memcpy(d, S->bed_dw, (size_t)S->bed_n * 4 * sizeof(int32_t));
}
void orc_debug_set_bed_delta(orc_sim* S, const int32_t* d) {
  if (S->bed_n) memcpy(S->bed_dw, d, (size_t)S->bed_n * 4 * sizeof(int32_t));
}

/* MpmSim3D::step_frame (mod.rs:264-430) */
int orc_step_frame(orc_sim* S, float dt) {
  dt = o_min(dt, 1.0f / 30.0f);
  uint32_t n = S->s.substeps > 1u ? S->s.substeps : 1u;
  float sub_dt = dt / (float)n;
  for (uint32_t i = 0; i < n; ++i) substep(S, sub_dt, 0, 1);
  return CS_OK;
}
