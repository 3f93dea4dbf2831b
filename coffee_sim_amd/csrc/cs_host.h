// cs_host.h — host-side pieces of the MLS-MPM substep that run on the CPU:
// unit constants, settings presets, SDF / class-mask generation, the inflow
// emitter, the per-substep parameter block and the static bed table.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "cs_common.h"

namespace cs {

// units.rs:10-34 evaluated in binary32.
float sim_units_per_meter();
float max_velocity();
float gravity_sim_units();
float ml_per_sim_unit_cubed();
float kozeny_carman_permeability(float grind_diameter_um, float porosity);

void spout_aim_at(cs_spout* sp, float x, float y, float z);
void spout_translate_origin_to(cs_spout* sp, float x, float y, float z);

int settings_preset(int preset, cs_settings* out);
int validate_settings(const cs_settings& s, std::string* why);

void generate_sdf(const cs_settings& s, float* out);
void generate_sdf_class(const cs_settings& s, const float* sdf, uint8_t* out);

// InflowState (inflow.rs:69-75) plus the fixed-count emitter's progress.
struct Inflow {
  float kettle_angle_deg = 0.0f;
  float flow_rate = 0.0f;
  float exit_speed = 0.0f;
  float accumulator = 0.0f;
  uint32_t emission_sequence = 0;
  uint32_t emitted_total = 0;
};

float pour_fraction_from_angle(float angle, const cs_spout& sp);
float effective_head_from_angle(float angle, const cs_spout& sp);
float exit_speed_cap_from_angle(float angle, const cs_spout& sp);
float exit_speed_from_head(float head, float max_exit_speed);
float flow_rate_from_speed(float v, float nozzle_radius, float cd, float vol_to_ml, float max_flow);

// InflowState::emit_particles (inflow.rs:122-241) up to the particle loop: the
// flow model and the capacity check decide how many particles this substep
// emits (the return value; `dropped` accumulates particles refused for
// capacity), and `f` receives the disk frame for them at slots current...
// The inflow state advances as if they were emitted.
uint32_t emit_plan(const cs_settings& s, Inflow& in, float dt, uint32_t current, cs_emit_frame* f,
                   uint32_t* dropped);
// emit_plan, then the particles on the host (8 floats each; the affine of new
// particles is all zero) appended to `out`.
uint32_t emit(const cs_settings& s, Inflow& in, float dt, uint32_t current, std::vector<float>& out,
              uint32_t* dropped);

// write_uniforms (mod.rs:569-670).
void make_params(const cs_settings& s, uint32_t num_particles, uint32_t num_bed_cells, float dt, cs_params* u);

// Static per-cell bed table (field mode).  `cell_index` receives, per cell, the
// index into `table` or -1.  Returns false when a lookup is out of range.
bool build_bed_table(const cs_settings& s, const int32_t* lookup, const cs_bed_record* records,
                     uint32_t num_records, std::vector<int32_t>& cell_index, std::vector<cs_bed_cell>& table);

}  // namespace cs
