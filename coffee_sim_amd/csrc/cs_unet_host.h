// cs_unet_host.h — what the simulation (cs_sim.cpp) uses of the learned warm start.
#pragma once

#include <stdint.h>

#include <string>

#include "../../include/coffee_sim.h"

// Records `msg` as the calling thread's cs_last_error() (cs_sim.cpp).
void cs_set_last_error(const char* msg);
int cs_unet_fail(const std::string& msg);

int cs_unet_side(const cs_unet* net);
cs_ctx* cs_unet_context(const cs_unet* net);
float cs_unet_rhs_scale(const cs_unet* net);
float cs_unet_pressure_scale(const cs_unet* net);
// Device buffer of four features per cell that cs_unet_predict_device reads.
float* cs_unet_feature_buffer(cs_unet* net);
// Runs the network on the feature buffer; writes the predicted pressure (in
// pressure units) for every cell to the device array `pressure_out`.
int cs_unet_predict_device(cs_unet* net, float* pressure_out);
// Appends a training sample and returns its device buffers and a zeroed device
// counter for the number of masked cells; cs_unet_finish_sample reads the count.
int cs_unet_new_sample(cs_unet* net, float** f4, float** target, uint8_t** mask, uint32_t** count);
int cs_unet_finish_sample(cs_unet* net);
