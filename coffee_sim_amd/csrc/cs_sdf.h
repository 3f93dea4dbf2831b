// cs_sdf.h — the scene's signed distance field and solid-class mask, one
// voxel or cell at a time, shared by the host generator (cs_scene.cpp) and the
// device generator (cs_kernels.hip) so both give the same bits.
#pragma once
#include "cs_common.h"

// Interior distance of one obstacle (state.rs:225-268), CS_SDF_NO_CONSTRAINT
// outside its vertical extent.
CS_HD float cs_sdf_interior(const cs_obstacle& o, float px, float py, float pz) {
  if (o.kind == CS_OBSTACLE_TRUNCATED_CONE) {
    float local_y = py - o.center[1];
    if (local_y > o.top_y || local_y < o.bot_y) return CS_SDF_NO_CONSTRAINT;
    float height = o.top_y - o.bot_y;
    if (height < 1e-10f) return CS_SDF_NO_CONSTRAINT;
    float t = (local_y - o.bot_y) / height;
    float cone_r = o.bot_radius + (o.top_radius - o.bot_radius) * t;
    float dx = px - o.center[0], dz = pz - o.center[2];
    return cone_r - sqrtf(dx * dx + dz * dz);
  }
  if (o.kind == CS_OBSTACLE_CYLINDER) {
    float local_y = py - o.center[1];
    if (local_y > o.top_y) return CS_SDF_NO_CONSTRAINT;
    float dx = px - o.center[0], dz = pz - o.center[2];
    float radial = o.radius - sqrtf(dx * dx + dz * dz);
    float floor_d = local_y - o.bot_y;
    return fminf(radial, floor_d);
  }
  return CS_SDF_NO_CONSTRAINT;
}

// Voxel (ix, iy, iz) of the sdf_res^3 field (generate_sdf_data, state.rs:206-224):
// the obstacles in order, the first one seen replacing the "no constraint"
// value and later ones taking the max, each shifted by half the wall thickness.
CS_HD float cs_sdf_voxel(const cs_settings& s, uint32_t ix, uint32_t iy, uint32_t iz) {
  const float n = (float)s.sdf_res;
  const float bx = s.bounds_size[0], by = s.bounds_size[1], bz = s.bounds_size[2];
  const float half = s.wall_thickness * 0.5f;
  const float px = -bx * 0.5f + ((float)ix + 0.5f) * bx / n;
  const float py = -by * 0.5f + ((float)iy + 0.5f) * by / n;
  const float pz = -bz * 0.5f + ((float)iz + 0.5f) * bz / n;
  float v = CS_SDF_NO_CONSTRAINT;
  for (uint32_t oi = 0; oi < s.num_obstacles && oi < CS_MAX_OBSTACLES; ++oi) {
    const float sd = cs_sdf_interior(s.obstacles[oi], px, py, pz) - half;
    if (v >= CS_SDF_NO_CONSTRAINT - 1.0f) v = sd;
    else v = fmaxf(v, sd);
  }
  return v;
}

CS_HD float cs_sdf_texel(const float* d, uint32_t res, int x, int y, int z) {
  const int mx = (int)res - 1;
  x = x < 0 ? 0 : (x > mx ? mx : x);
  y = y < 0 ? 0 : (y > mx ? mx : y);
  z = z < 0 ? 0 : (z > mx ? mx : z);
  return d[((size_t)z * res + (size_t)y) * res + (size_t)x];
}

// sample_sdf_from_data (state.rs:287-320): Rust's fract (x - trunc x) and the
// lerp form c + (d - c) * f.
CS_HD float cs_sdf_sample(const cs_settings& s, const float* d, float px, float py, float pz) {
  const float res = (float)s.sdf_res;
  const float bx = s.bounds_size[0], by = s.bounds_size[1], bz = s.bounds_size[2];
  float ux = (px + bx * 0.5f) / bx * res - 0.5f;
  float uy = (py + by * 0.5f) / by * res - 0.5f;
  float uz = (pz + bz * 0.5f) / bz * res - 0.5f;
  int b0 = cs_f2i(floorf(ux)), b1 = cs_f2i(floorf(uy)), b2 = cs_f2i(floorf(uz));
  float fx = ux - truncf(ux), fy = uy - truncf(uy), fz = uz - truncf(uz);
  const uint32_t r = s.sdf_res;
  float c000 = cs_sdf_texel(d, r, b0, b1, b2), c100 = cs_sdf_texel(d, r, b0 + 1, b1, b2);
  float c010 = cs_sdf_texel(d, r, b0, b1 + 1, b2), c110 = cs_sdf_texel(d, r, b0 + 1, b1 + 1, b2);
  float c001 = cs_sdf_texel(d, r, b0, b1, b2 + 1), c101 = cs_sdf_texel(d, r, b0 + 1, b1, b2 + 1);
  float c011 = cs_sdf_texel(d, r, b0, b1 + 1, b2 + 1), c111 = cs_sdf_texel(d, r, b0 + 1, b1 + 1, b2 + 1);
  float c00 = c000 + (c100 - c000) * fx;
  float c10 = c010 + (c110 - c010) * fx;
  float c01 = c001 + (c101 - c001) * fx;
  float c11 = c011 + (c111 - c011) * fx;
  float c0 = c00 + (c10 - c00) * fy;
  float c1 = c01 + (c11 - c01) * fy;
  return c0 + (c1 - c0) * fz;
}

// Cell (ix, iy, iz) of the class mask (generate_sdf_class_data, state.rs:322-344):
// 1 where the field sampled at the cell centre is negative.
CS_HD uint8_t cs_sdf_class_cell(const cs_settings& s, const float* sdf, uint32_t ix, uint32_t iy, uint32_t iz) {
  const float dx = s.bounds_size[0] / (float)s.grid_dims[0];
  const float ox = -s.bounds_size[0] * 0.5f, oy = -s.bounds_size[1] * 0.5f, oz = -s.bounds_size[2] * 0.5f;
  const float cx = ox + ((float)ix + 0.5f) * dx;
  const float cy = oy + ((float)iy + 0.5f) * dx;
  const float cz = oz + ((float)iz + 0.5f) * dx;
  return cs_sdf_sample(s, sdf, cx, cy, cz) < 0.0f ? 1u : 0u;
}
