// cs_device.h — device helpers shared by the kernel translation units:
// grid indexing, SDF sampling and contact (shader.rs:148-150, 673-903), the
// packed kind/fill word and the volume target of the pressure right-hand side.
#pragma once

#include "cs_common.h"

// A single-scene kernel's parameters, read from the substep's device block.  A
// halted block makes the kernel return before it reads or writes anything else.
#define CS_PARAMS_FROM(dp) \
  if ((dp)->halt) return;  \
  const cs_params u = (dp)->u

namespace {

// ---- tile-binned particle passes (P2G in cs_p2g.hip, G2P in cs_kernels.hip) -------------
// A tile is one CS_BRICK^3 brick; its footprint, the nodes its particles' 3x3x3
// stencils reach, is (CS_BRICK + 2)^3 nodes from the brick's first cell; a work
// item of the tile passes is at most kChunk particles of one tile.
constexpr int TB = CS_BRICK;               // tile side in cells
constexpr int FP = TB + 2;                 // footprint side in nodes
constexpr int FP3 = FP * FP * FP;          // 1000 nodes
constexpr uint32_t kNoTile = 0xffffffffu;  // the tile of a particle P2G skips
constexpr uint32_t kChunk = 1024;          // particles per work item

__device__ __forceinline__ uint32_t cell_index(const cs_params& u, uint32_t ix, uint32_t iy, uint32_t iz) {
  return iz * u.gx * u.gy + iy * u.gx + ix;  // shader.rs:148-150
}
__device__ __forceinline__ bool on_grid(const cs_params& u, int x, int y, int z) {
  return !(x < 0 || y < 0 || z < 0) && !((uint32_t)x >= u.gx || (uint32_t)y >= u.gy || (uint32_t)z >= u.gz);
}
__device__ __forceinline__ void decompose(const cs_params& u, uint32_t idx, uint32_t& ix, uint32_t& iy, uint32_t& iz) {
  uint32_t plane = u.gx * u.gy;
  iz = idx / plane;
  uint32_t rem = idx % plane;
  iy = rem / u.gx;
  ix = rem % u.gx;
}
__device__ __forceinline__ cs_f3 cell_center(const cs_params& u, uint32_t ix, uint32_t iy, uint32_t iz) {
  return f3(u.origin[0] + ((float)ix + 0.5f) * u.dx, u.origin[1] + ((float)iy + 0.5f) * u.dx,
            u.origin[2] + ((float)iz + 0.5f) * u.dx);
}

// ---- SDF (shader.rs:673-715) ------------------------------------------------
__device__ __forceinline__ float sdf_texel(const cs_params& u, const float* __restrict__ sdf, int x, int y, int z) {
  int mx = (int)u.sdf_res - 1;
  x = x < 0 ? 0 : (x > mx ? mx : x);
  y = y < 0 ? 0 : (y > mx ? mx : y);
  z = z < 0 ? 0 : (z > mx ? mx : z);
  return sdf[((size_t)z * u.sdf_res + (size_t)y) * u.sdf_res + (size_t)x];
}
__device__ float sample_sdf(const cs_params& u, const float* __restrict__ sdf, cs_f3 p) {
  float res = u.sdf_res_f;
  float ux = (p.x + u.bmax[0]) / (u.bmax[0] * 2.0f) * res - 0.5f;
  float uy = (p.y + u.bmax[1]) / (u.bmax[1] * 2.0f) * res - 0.5f;
  float uz = (p.z + u.bmax[2]) / (u.bmax[2] * 2.0f) * res - 0.5f;
  float flx = floorf(ux), fly = floorf(uy), flz = floorf(uz);
  int bx = cs_f2i(flx), by = cs_f2i(fly), bz = cs_f2i(flz);
  float fx = ux - flx, fy = uy - fly, fz = uz - flz;
  float c000 = sdf_texel(u, sdf, bx, by, bz);
  float c100 = sdf_texel(u, sdf, bx + 1, by, bz);
  float c010 = sdf_texel(u, sdf, bx, by + 1, bz);
  float c110 = sdf_texel(u, sdf, bx + 1, by + 1, bz);
  float c001 = sdf_texel(u, sdf, bx, by, bz + 1);
  float c101 = sdf_texel(u, sdf, bx + 1, by, bz + 1);
  float c011 = sdf_texel(u, sdf, bx, by + 1, bz + 1);
  float c111 = sdf_texel(u, sdf, bx + 1, by + 1, bz + 1);
  float c00 = cs_mix(c000, c100, fx);
  float c10 = cs_mix(c010, c110, fx);
  float c01 = cs_mix(c001, c101, fx);
  float c11 = cs_mix(c011, c111, fx);
  float c0 = cs_mix(c00, c10, fy);
  float c1 = cs_mix(c01, c11, fy);
  return cs_mix(c0, c1, fz);
}
__device__ cs_f3 sdf_gradient(const cs_params& u, const float* __restrict__ sdf, cs_f3 p) {
  float e = u.dx;
  float gxv = sample_sdf(u, sdf, f3(p.x + e, p.y + 0.0f, p.z + 0.0f)) - sample_sdf(u, sdf, f3(p.x - e, p.y - 0.0f, p.z - 0.0f));
  float gyv = sample_sdf(u, sdf, f3(p.x + 0.0f, p.y + e, p.z + 0.0f)) - sample_sdf(u, sdf, f3(p.x - 0.0f, p.y - e, p.z - 0.0f));
  float gzv = sample_sdf(u, sdf, f3(p.x + 0.0f, p.y + 0.0f, p.z + e)) - sample_sdf(u, sdf, f3(p.x - 0.0f, p.y - 0.0f, p.z - e));
  cs_f3 g = f3(gxv, gyv, gzv);
  float len = len3(g);
  if (len < 1e-8f) return f3(0.0f, 0.0f, 0.0f);
  return g / len;
}

// boundary_project body (shader.rs:1638-1678) for a cell with mass >= 1e-6.
__device__ cs_f3 boundary_apply(const cs_params& u, const float* __restrict__ sdf, cs_f3 v, cs_f3 cp) {
  float sv = sample_sdf(u, sdf, cp);
  if (sv < u.contact_offset) {
    cs_f3 n = sdf_gradient(u, sdf, cp);
    float vn = dot3(v, n);
    if (vn < 0.0f) {
      v = v - (n * vn) * (1.0f + u.restitution);
      cs_f3 vt = v - n * dot3(v, n);
      float vtl = len3(vt);
      if (vtl > 1e-6f) {
        float fi = cs_min(u.friction * fabsf(vn), vtl);
        v = v - vt * (fi / vtl);
      }
    }
  }
  float margin = 2.0f * u.dx;
  float bminx = u.origin[0] + margin, bminy = u.origin[1] + margin, bminz = u.origin[2] + margin;
  float bmaxx = u.bmax[0] - margin, bmaxy = u.bmax[1] - margin, bmaxz = u.bmax[2] - margin;
  if (cp.x < bminx && v.x < 0.0f) v.x = 0.0f;
  if (cp.x > bmaxx && v.x > 0.0f) v.x = 0.0f;
  if (cp.y < bminy && v.y < 0.0f) v.y = 0.0f;
  if (cp.y > bmaxy && v.y > 0.0f) v.y = 0.0f;
  if (cp.z < bminz && v.z < 0.0f) v.z = 0.0f;
  if (cp.z > bmaxz && v.z > 0.0f) v.z = 0.0f;
  return v;
}

__device__ __forceinline__ cs_f3 cap_speed(const cs_params& u, cs_f3 v) {
  float speed = len3(v);
  if (speed > u.vel_cap) v = v * (u.vel_cap / speed);
  return v;
}

// ---- contact for particles (shader.rs:717-903) ---------------------------------
struct contact {
  cs_f3 pos, vel;
};
__device__ contact radial_barrier(const cs_params& u, cs_f3 pos, cs_f3 vel, float max_r) {
  float rx = pos.x - 0.0f, rz = pos.z - 0.0f;
  float r = sqrtf(rx * rx + rz * rz);
  if (r > max_r && r > 1e-6f) {
    float ox = rx / r, oz = rz / r;
    pos.x = 0.0f + ox * max_r;
    pos.z = 0.0f + oz * max_r;
    float vn = vel.x * ox + vel.z * oz;
    if (vn > 0.0f) {
      float tx = vel.x - ox * vn, tz = vel.z - oz * vn;
      vel.x = tx * (1.0f - u.friction * 0.35f);
      vel.z = tz * (1.0f - u.friction * 0.35f);
    }
  }
  return contact{pos, vel};
}
// bed_particle_radius (shader.rs:128): dx * bed_sample_radius_dx
__device__ __forceinline__ float bed_particle_radius(const cs_params& u) { return u.dx * CS_BED_SAMPLE_RADIUS_DX; }
// coffee_filter_floor_y (shader.rs:136-146) of the geometry's paper filter
__device__ __forceinline__ float coffee_filter_floor_y(const cs_params& u) {
  const cs_geometry& g = u.geo;
  const float fb = g.filter_center_y + g.filter_bot_y, ft = g.filter_center_y + g.filter_top_y;
  const float bco = cs_max(u.contact_offset, bed_particle_radius(u));
  const float fh = cs_max(ft - fb, 1e-6f);
  const float slope = g.filter_top_radius / fh;
  return fb + (bco + g.filter_thickness) / cs_max(slope, 1e-6f);
}
// resolve_scene_obstacles (shader.rs:744-823): the cone, then for bed grains the
// paper filter and its apex floor, then the cup.
__device__ contact scene_obstacles(const cs_params& u, cs_f3 pos, cs_f3 vel, bool is_bed) {
  const cs_geometry& g = u.geo;
  if (g.cone_enabled && pos.y <= g.cone_top_y && pos.y >= g.cone_bot_y) {
    float t = cs_clamp((pos.y - g.cone_bot_y) / (g.cone_top_y - g.cone_bot_y), 0.0f, 1.0f);
    float cone_r = cs_mix(g.cone_bot_radius, g.cone_top_radius, t) - u.contact_offset;
    contact c = radial_barrier(u, pos, vel, cone_r);
    pos = c.pos;
    vel = c.vel;
  }
  if (is_bed && g.filter_enabled) {
    const float ft = g.filter_center_y + g.filter_top_y, fb = g.filter_center_y + g.filter_bot_y;
    const float bco = cs_max(u.contact_offset, bed_particle_radius(u));
    const float floor_y = coffee_filter_floor_y(u);
    if (pos.y <= ft && pos.y >= fb) {
      const float t = cs_clamp((pos.y - fb) / (ft - fb), 0.0f, 1.0f);
      const float outer_r = cs_mix(0.0f, g.filter_top_radius, t);
      const float fr = cs_max(outer_r - g.filter_thickness, 0.0f) - bco;
      contact c = radial_barrier(u, pos, vel, cs_max(fr, 0.0f));
      pos = c.pos;
      vel = c.vel;
    }
    if (pos.y <= floor_y) {
      pos.y = floor_y;
      pos.x = 0.0f;
      pos.z = 0.0f;
      if (vel.y < 0.0f) {
        vel.y = 0.0f;
        vel.x *= 1.0f - u.friction * 0.55f;
        vel.z *= 1.0f - u.friction * 0.55f;
      }
      vel.x *= 1.0f - u.friction * 0.55f;
      vel.z *= 1.0f - u.friction * 0.55f;
    }
  }
  if (g.cup_enabled && pos.y <= g.cup_top_y) {
    float cup_r = g.cup_radius - u.wall_half - u.contact_offset;
    contact c = radial_barrier(u, pos, vel, cup_r);
    pos = c.pos;
    vel = c.vel;
    float floor_y = g.cup_bottom_y + u.wall_half + u.contact_offset;
    if (pos.y < floor_y) {
      pos.y = floor_y;
      if (vel.y < 0.0f) {
        vel.y = 0.0f;
        vel.x *= 1.0f - u.friction * 0.55f;
        vel.z *= 1.0f - u.friction * 0.55f;
      }
    }
  }
  return contact{pos, vel};
}
__device__ contact sdf_contact(const cs_params& u, const float* __restrict__ sdf, cs_f3 pos, cs_f3 vel,
                               bool is_bed = false) {
  float sv = sample_sdf(u, sdf, pos);
  if (sv < u.contact_offset) {
    cs_f3 n = sdf_gradient(u, sdf, pos);
    if (len3(n) > 1e-6f) {
      pos = pos + n * (u.contact_offset - sv);
      float vn = dot3(vel, n);
      if (vn < 0.0f) {
        vel = vel - (n * vn) * (1.0f + u.restitution);
        cs_f3 vt = vel - n * dot3(vel, n);
        float vtl = len3(vt);
        if (vtl > 1e-6f) {
          float fi = cs_min(u.friction * fabsf(vn), vtl);
          vel = vel - vt * (fi / vtl);
        }
      }
    }
  }
  return scene_obstacles(u, pos, vel, is_bed);
}

// ---- kind/fill word ------------------------------------------------------------
__device__ __forceinline__ bool kf_fluid(float kf) { return kf >= 0.0f; }
__device__ __forceinline__ bool kf_solid(float kf) { return kf == CS_KF_SOLID; }
__device__ __forceinline__ bool kf_bed(float kf) { return kf == CS_KF_BED; }
__device__ __forceinline__ float kf_fill(float kf) { return kf == CS_KF_BED ? 1.0f : kf; }
// pressure_face_weight (shader.rs:306-323) for a non-solid neighbour.
__device__ __forceinline__ float face_w(float self_kf, float self_fill, float nkf) {
  if (kf_bed(self_kf) || kf_bed(nkf)) return 1.0f;
  if (!kf_fluid(nkf)) return self_fill;
  return cs_min(self_fill, kf_fill(nkf));
}

__constant__ static int c_off6[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};

// volume_projection_target_divergence (shader.rs:597-619)
__device__ float volume_target_div(const cs_params& u, float rest, float cur) {
  float cv = u.dx * u.dx * u.dx;
  float eps = cs_max(cv * 1e-5f, u.p_vol * 1e-3f);
  if (rest <= eps) return 0.0f;
  float j = cur / rest;
  float comp = cs_clamp(1.0f - j, 0.0f, 0.75f);
  float expd = cs_clamp(j - 1.0f, 0.0f, 0.75f);
  float target = comp * u.j_alpha - expd * u.j_expand_alpha;
  float packed = cs_max(rest, cur) / cs_max(cv, 1e-8f);
  float over = cs_max(packed - u.max_rest_fraction, 0.0f);
  return cs_max(target, over * u.j_alpha);
}

// ---- P2G per-particle arithmetic (shader.rs:955-1052) ---------------------------
// Shared by every P2G layout so they compute the same fixed-point terms.
struct P2GParticle {
  int bx, by, bz;           // base node floor(g - 0.5)
  cs_f3 fx;                 // g - base
  float wx[3], wy[3], wz[3];
  cs_f3 ac0, ac1, ac2, mv;  // m * C columns, m * v
  float mass_p, rpv, cpv;   // mass, rest volume, current volume
};

// Loads particle `pid`; false when it does not scatter (bed phase or below the
// active-mass threshold, shader.rs:967).
__device__ __forceinline__ bool p2g_setup(const cs_params& u, const float4* __restrict__ part,
                                          const float4* __restrict__ aff, uint32_t pid, P2GParticle& q) {
  const float4 P = part[2 * pid], Vv = part[2 * pid + 1];
  const float4 A0 = aff[3 * pid], A1 = aff[3 * pid + 1], A2 = aff[3 * pid + 2];
  const float mass_p = Vv.w, J = P.w, phase = A0.w;
  if (!(phase < 0.5f) || mass_p <= u.nominal_mass * 0.10f) return false;
  const cs_f3 gp = (f3(P.x, P.y, P.z) - f3(u.origin[0], u.origin[1], u.origin[2])) * u.inv_dx;
  q.bx = cs_f2i(floorf(gp.x - 0.5f));
  q.by = cs_f2i(floorf(gp.y - 0.5f));
  q.bz = cs_f2i(floorf(gp.z - 0.5f));
  const cs_f3 fx = f3(gp.x - (float)q.bx, gp.y - (float)q.by, gp.z - (float)q.bz);
  q.fx = fx;
  q.wx[0] = 0.5f * (1.5f - fx.x) * (1.5f - fx.x);
  q.wx[1] = 0.75f - (fx.x - 1.0f) * (fx.x - 1.0f);
  q.wx[2] = 0.5f * (fx.x - 0.5f) * (fx.x - 0.5f);
  q.wy[0] = 0.5f * (1.5f - fx.y) * (1.5f - fx.y);
  q.wy[1] = 0.75f - (fx.y - 1.0f) * (fx.y - 1.0f);
  q.wy[2] = 0.5f * (fx.y - 0.5f) * (fx.y - 0.5f);
  q.wz[0] = 0.5f * (1.5f - fx.z) * (1.5f - fx.z);
  q.wz[1] = 0.75f - (fx.z - 1.0f) * (fx.z - 1.0f);
  q.wz[2] = 0.5f * (fx.z - 0.5f) * (fx.z - 0.5f);
  q.ac0 = f3(mass_p * A0.x, mass_p * A0.y, mass_p * A0.z);
  q.ac1 = f3(mass_p * A1.x, mass_p * A1.y, mass_p * A1.z);
  q.ac2 = f3(mass_p * A2.x, mass_p * A2.y, mass_p * A2.z);
  q.mv = f3(Vv.x, Vv.y, Vv.z) * mass_p;
  q.mass_p = mass_p;
  q.rpv = u.p_vol * mass_p / cs_max(u.nominal_mass, 1e-6f);
  q.cpv = q.rpv * cs_clamp(J, 0.40f, 2.00f);
  return true;
}

// The six fixed-point terms of node (base + (i, j, k)): mass, momentum x/y/z,
// rest volume, current volume (shader.rs:1013-1048).  Returns 1 when a term
// exceeds the overflow probe's limit.
__device__ __forceinline__ uint32_t p2g_node(const cs_params& u, const P2GParticle& q, int i, int j, int k,
                                             int32_t v[6]) {
  const float w = q.wx[i] * q.wy[j] * q.wz[k];
  const cs_f3 dpos = (f3((float)i, (float)j, (float)k) - q.fx) * u.dx;
  const cs_f3 mom = (q.mv + f3(dot3(q.ac0, dpos), dot3(q.ac1, dpos), dot3(q.ac2, dpos))) * w;
  const float fp = u.fp_scale;
  const float mfp = (w * q.mass_p) * fp, rfp = (w * q.rpv) * fp, cfp = (w * q.cpv) * fp;
  const float mxf = mom.x * fp, myf = mom.y * fp, mzf = mom.z * fp;
  const float lim = 1.0e9f;
  const uint32_t over = (fabsf(mfp) > lim || fabsf(mxf) > lim || fabsf(myf) > lim || fabsf(mzf) > lim ||
                         fabsf(rfp) > lim || fabsf(cfp) > lim) ? 1u : 0u;
  v[0] = cs_f2i(mfp);
  v[1] = cs_f2i(mxf);
  v[2] = cs_f2i(myf);
  v[3] = cs_f2i(mzf);
  v[4] = cs_f2i(rfp);
  v[5] = cs_f2i(cfp);
  return over;
}

// p2g (shader.rs:955-1052) for one particle: its 162 terms added to the grid
// lanes with global i32 atomics.
__device__ __forceinline__ void p2g_atomic_particle(const cs_params& u, const float4* __restrict__ part,
                                                    const float4* __restrict__ aff, int32_t* __restrict__ lanes,
                                                    uint32_t* __restrict__ metrics, uint32_t pid) {
  if (pid >= u.num_particles) return;
  P2GParticle q;
  if (!p2g_setup(u, part, aff, pid, q)) return;
  uint32_t nc = u.nc;
  uint32_t overflow = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        int cx = q.bx + i, cy = q.by + j, cz = q.bz + k;
        if (!on_grid(u, cx, cy, cz)) continue;
        int32_t v[6];
        overflow += p2g_node(u, q, i, j, k, v);
        uint32_t ci = cell_index(u, (uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
#pragma unroll
        for (int l = 0; l < 6; ++l) atomicAdd(&lanes[(size_t)l * nc + ci], v[l]);
      }
  if (overflow) atomicAdd(&metrics[CS_M_OVERFLOW], overflow);
}

}  // namespace
