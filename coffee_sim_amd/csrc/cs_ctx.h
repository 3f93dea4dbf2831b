// cs_ctx.h — the device context shared by the simulation and the learned warm start.
#pragma once

#include <hip/hip_runtime_api.h>

#include "../../include/coffee_sim.h"

struct cs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
};
