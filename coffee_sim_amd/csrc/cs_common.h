// cs_common.h — parameter block and binary32 helpers shared by the host code
// and the gfx950 kernels of the MLS-MPM substep.
//
// Numeric contract (identical on host and device, so every pass is
// reproducible bit for bit): binary32 arithmetic, no FMA contraction
// (-ffp-contract=off), correctly rounded division and sqrt, NaN-avoiding
// min/max whose ties return the first operand, float->i32 that truncates,
// saturates and maps NaN to 0, and WGSL's mix(a,b,t) = a*(1-t) + b*t.
#pragma once

#include <math.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/coffee_sim.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CS_HD __host__ __device__ __forceinline__
#else
#define CS_HD static inline
#endif

// Fixed-point encoding of the grid lanes (state.rs:8-18).
#define CS_FP_SCALE 262144.0f
#define CS_FP_VALUE_LIMIT 8192.0f
#define CS_METRICS_DIV_FP_SCALE 1024.0f
#define CS_SDF_NO_CONSTRAINT 999.0f

// Render sample radii in cells (DEFAULT_BREW, brew_config.rs:50-51; read by
// water_particle_radius / bed_particle_radius, shader.rs:127-128).
#define CS_WATER_SAMPLE_RADIUS_DX 0.18f
#define CS_BED_SAMPLE_RADIUS_DX 0.62f
// DEFAULT_BREW (brew_config.rs:59-60) and extraction_params.x (mod.rs:645)
#define CS_BED_ABSORPTION_RATE 1.6f
#define CS_TARGET_BED_RETENTION_ML 42.0f
#define CS_EXTRACTION_RATE 0.01f

// Cell kinds (shader.rs:174-178).
#define CS_CELL_AIR 0
#define CS_CELL_SURFACE 1
#define CS_CELL_INTERIOR 2
#define CS_CELL_BED 3
#define CS_CELL_SOLID 4

// Packed per-cell kind/fill word used by the pressure passes: fluid cells hold
// their liquid fill fraction (surface cells, in [0,1]), 1.0 (interior) or 2.0
// (bed, fill 1); air is -1 and solid is -2.
#define CS_KF_AIR -1.0f
#define CS_KF_SOLID -2.0f
#define CS_KF_BED 2.0f

// Metric slots (shader.rs:67-71) plus this build's residual bookkeeping.
#define CS_M_MAX_ABS_DIV 0
#define CS_M_FLUID_CELLS 1
#define CS_M_DIV_CLAMP 2
#define CS_M_P_CLAMP 3
#define CS_M_OVERFLOW 4
// Slots the reference leaves free: max|div| and max|div - target| over the fluid
// cells after project_pressure + boundary_project, and max|div - target| before it.
#define CS_M_POST_MAX_ABS_DIV 5
#define CS_M_POST_MAX_ABS_RES 6
#define CS_M_PRE_MAX_ABS_RES 7

CS_HD float cs_min(float a, float b) { return (a != a) ? b : ((b < a) ? b : a); }
CS_HD float cs_max(float a, float b) { return (a != a) ? b : ((b > a) ? b : a); }
CS_HD float cs_clamp(float x, float lo, float hi) { return cs_min(cs_max(x, lo), hi); }
CS_HD float cs_mix(float a, float b, float t) { return a * (1.0f - t) + b * t; }
CS_HD float cs_smoothstep(float lo, float hi, float x) {
  float t = cs_clamp((x - lo) / (hi - lo), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
CS_HD int32_t cs_f2i(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return 2147483647;
  if (x <= -2147483648.0f) return (-2147483647 - 1);
  return (int32_t)x;
}
CS_HD uint32_t cs_f2u(float x) {
  if (x != x || x <= 0.0f) return 0u;
  if (x >= 4294967296.0f) return 0xffffffffu;
  return (uint32_t)x;
}

// exp(x): Cody-Waite reduction x = k ln2 + r with ln2 split so k*ln2_hi is
// exact, then the degree-7 Taylor polynomial of e^r in Horner form
// (|r| <= ln2/2, truncation below 1e-8 relative), scaled by 2^k through the
// exponent field.
CS_HD float cs_expf(float x) {
  if (x != x) return x;
  if (x > 88.0f) x = 88.0f;
  if (x < -87.0f) x = -87.0f;
  const float log2e = 1.44269502f;
  const float ln2_hi = 0.693145751953125f;
  const float ln2_lo = 1.42860677e-06f;
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
  uint32_t bits = (uint32_t)(k + 127) << 23;
  float scale;
  __builtin_memcpy(&scale, &bits, 4);
  return p * scale;
}

// sin and cos of a float argument, each rounded once to float; the same
// double-precision steps on the host and the device, so both give the same
// bits.  Cody-Waite reduction x = k pi/2 + r with pi/2 = c1 + c2 + c3, where c1
// and c2 hold 22 significant bits so k c1 and k c2 are exact for |k| < 2^31
// (|x| < 3.3e9); then the Taylor series of sin r (through r^17) and cos r
// (through r^18) on |r| <= pi/4 in Horner form (truncation below 1e-18), and the
// quadrant k mod 4.  Used for the emitter's golden angle theta = sample * 2.399963.
CS_HD void cs_sincosf(float x, float* sn, float* cs) {
  const double xd = (double)x;
  const double k = rint(xd * 0x1.45f306dc9c883p-1);
  const double r = ((xd - k * 0x1.921fb8p+0) - k * -0x1.5dde98p-23) - k * 0x1.8469898cc5170p-48;
  const double r2 = r * r;
  double ps = 0x1.952c77030ad4ap-49;         // 1/17!
  ps = ps * r2 - 0x1.ae7f3e733b81fp-41;      // 1/15!
  ps = ps * r2 + 0x1.6124613a86d09p-33;      // 1/13!
  ps = ps * r2 - 0x1.ae64567f544e4p-26;      // 1/11!
  ps = ps * r2 + 0x1.71de3a556c734p-19;      // 1/9!
  ps = ps * r2 - 0x1.a01a01a01a01ap-13;      // 1/7!
  ps = ps * r2 + 0x1.1111111111111p-7;       // 1/5!
  ps = ps * r2 - 0x1.5555555555555p-3;       // 1/3!
  const double sr = r + (r * r2) * ps;
  double pc = -0x1.6827863b97d97p-53;        // -1/18!
  pc = pc * r2 + 0x1.ae7f3e733b81fp-45;      // 1/16!
  pc = pc * r2 - 0x1.93974a8c07c9dp-37;      // 1/14!
  pc = pc * r2 + 0x1.1eed8eff8d898p-29;      // 1/12!
  pc = pc * r2 - 0x1.27e4fb7789f5cp-22;      // 1/10!
  pc = pc * r2 + 0x1.a01a01a01a01ap-16;      // 1/8!
  pc = pc * r2 - 0x1.6c16c16c16c17p-10;      // 1/6!
  pc = pc * r2 + 0x1.5555555555555p-5;       // 1/4!
  pc = pc * r2 - 0.5;
  const double cr = 1.0 + r2 * pc;
  const long long q = (long long)k & 3;
  const double sd = q == 0 ? sr : (q == 1 ? cr : (q == 2 ? -sr : -cr));
  const double cd = q == 0 ? cr : (q == 1 ? -sr : (q == 2 ? -cr : sr));
  *sn = (float)sd;
  *cs = (float)cd;
}

// One substep's emission (InflowState::emit_particles, inflow.rs:165-219):
// `count` particles of the golden-angle disk, samples seq0.., written to
// particle slots first...  The frame (spout tangents, emit origin, velocity)
// is formed once on the host.
struct cs_emit_frame {
  float t1[3], t2[3], origin[3], vel[3];
  float jet_radius, dt, mass;
  uint32_t seq0, count, first;
};
// Particle i of the frame: pos.xyz, J = 1, vel.xyz, mass (the loop body of inflow.rs:195-214).
CS_HD void cs_emit_particle(const cs_emit_frame& f, uint32_t i, float* p8) {
  const float golden = 2.399963f, radial_irr = 0.7548777f;
  const uint32_t sample = f.seq0 + i;  // wrapping_add
  const float sf = (float)sample;
  float ru = (sf + 0.5f) * radial_irr;
  ru = ru - truncf(ru);  // Rust fract
  const float r = f.jet_radius * sqrtf(ru);
  const float theta = sf * golden;
  const float age = (((float)i + 0.5f) / (float)f.count) * f.dt;
  float st, ct;
  cs_sincosf(theta, &st, &ct);
  for (int c = 0; c < 3; ++c) {
    const float off = f.t1[c] * r * ct + f.t2[c] * r * st;
    p8[c] = f.origin[c] + off - f.vel[c] * age;
  }
  p8[3] = 1.0f;
  p8[4] = f.vel[0];
  p8[5] = f.vel[1];
  p8[6] = f.vel[2];
  p8[7] = f.mass;
}

struct cs_f3 {
  float x, y, z;
};
CS_HD cs_f3 f3(float x, float y, float z) { return cs_f3{x, y, z}; }
CS_HD cs_f3 operator+(cs_f3 a, cs_f3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
CS_HD cs_f3 operator-(cs_f3 a, cs_f3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
CS_HD cs_f3 operator*(cs_f3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
CS_HD cs_f3 operator/(cs_f3 a, float s) { return f3(a.x / s, a.y / s, a.z / s); }
CS_HD float dot3(cs_f3 a, cs_f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
CS_HD float len3(cs_f3 a) { return sqrtf(dot3(a, a)); }
CS_HD cs_f3 mix3(cs_f3 a, cs_f3 b, float t) {
  return f3(cs_mix(a.x, b.x, t), cs_mix(a.y, b.y, t), cs_mix(a.z, b.z, t));
}

// Per-substep parameter block: the reference's 256 B MpmUniforms
// (state.rs:31-56, values from write_uniforms mod.rs:569-670) plus the
// geometry and bed extensions.
// Side of the cubic bricks the pressure solve is tiled into.
#define CS_BRICK 8

struct cs_params {
  uint32_t gx, gy, gz, nc;
  uint32_t nbx, nby, nbz, nbricks;  // CS_BRICK^3 bricks covering the grid
  uint32_t num_particles, num_bed_cells, max_particles, use_sdf_cache;
  float dt, gravity, dx, inv_dx;
  float origin[3];
  float bmax[3];
  float bulk_K, viscosity, nominal_mass, p_vol;
  float fp_scale, inv_fp_scale, vel_cap;
  float sdf_res_f;
  uint32_t sdf_res;
  float friction, restitution, contact_offset;
  float water_nu, min_k;
  float div_clamp, pressure_clamp, metrics_div_fp_scale;
  float j_alpha, j_expand_alpha, max_rest_fraction;
  float wall_half;
  cs_geometry geo;
};

// The device-resident parameter block of a single-scene simulation (the
// reference writes its uniform buffer once per substep, mod.rs:569-670).  The
// host writes the part before `nactive` with one copy ahead of each substep's
// launches; the substep's kernels read their parameters here, not from their
// arguments, so one recorded launch sequence (a graph) serves every substep.
// The device writes the rest.
struct cs_dparams {
  cs_params u;
  cs_emit_frame ef;     // this substep's emission: ef.count particles from slot ef.first
  uint32_t epoch_base;  // the persistent solve's epochs are epoch_base + 1, 2, ...
  uint32_t slot;        // the substep's entry in the simulation's solve records
  uint32_t cap_bricks;  // active bricks the launches are sized for; a substep with more halts
  uint32_t reserved;
  // written on the device
  uint32_t nactive;  // the substep's active bricks (the brick scan)
  uint32_t halt;     // set (to slot + 1) by the brick scan when nactive > cap_bricks and then
                     // kept: every kernel that reads the block returns at once
};
#define CS_DPARAMS_HOST_BYTES offsetof(cs_dparams, nactive)

// What a launcher takes: the host copy of the parameters (launch shapes) and
// the device block the kernels read.
struct cs_pblock {
  cs_params u;
  cs_dparams* d;
};

// A bed of particles as the reaction impulses see it (shader.rs:458-515):
// the momentum the bed takes from the water goes to the grains stamped in
// the cells around, through bed_delta lanes 1-3.  n = 0: no grains (a field
// bed or none), no reaction.
struct cs_grains {
  const int32_t* lookup;  // per cell: the stamping bed particle, or -1
  const float* aff;       // 12 floats per bed particle (.w of col0 = phase)
  int32_t* delta;         // bed_delta [4][n]: water credit, impulse x, y, z (i32 fixed point)
  uint32_t n;
};
#define CS_BED_REACTION_ALPHA 0.04f        // shader.rs:76
#define CS_BED_REACTION_IMPULSE_CAP 0.012f // shader.rs:77

// Static per-bed-cell record derived once from cs_bed_record (field mode of
// shader.rs:361-666): everything of the bed formulas that does not depend on
// the particles.
struct cs_bed_cell {
  float k_record;       // bed_compacted_permeability of the cell's record
  float v_record[3];    // record velocity (G2P bed velocity sum)
  float pore_capacity;  // shader.rs:654-658
  float overfill_resp;  // shader.rs:660-662
  float porosity;       // clamp(bed_compacted_porosity, 0.08, 1)
  float solid_fraction; // 1 - porosity
  float solid_div;      // bed_velocity_divergence(cell)
  float bmv[3];         // bed_matrix_velocity_cell(cell)
};
