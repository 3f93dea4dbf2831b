// cs_unet.h — launchers of the learned warm start's kernels (cs_unet_kernels.hip).
#pragma once

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#define CS_UNET_BIAS_ROWS 1024  // workgroups (partial rows) of the bias-gradient reduction

// 3x3x3 'same' convolution over an s^3 grid (s a multiple of 8), bf16 [voxel][channel]:
// w is [27][cout][cin]; bias may be null; with `mask` the input counts only where the
// mask (a ReLU output of the same shape as `in`) is positive.  cin % 16 == 0,
// cout in {16, 32, 48, 64, 96}.
hipError_t cs_unet_launch_conv(hipStream_t st, const uint16_t* in, int cin, int s, const uint16_t* w,
                               const float* bias, const uint16_t* mask, uint16_t* out, int cout, int relu);
// Weight and bias gradient of a convolution whose ReLU output is `outv`, given the
// gradient `gout` of that output: dw [27][cin][cout], db [cout] (fp32).
// partial: cs_unet_wgrad_floats(s, cin, cout) floats; bias_part: CS_UNET_BIAS_ROWS * cout.
#define CS_UNET_WGRAD_TARGET 512     // weight-gradient workgroups per launch
#define CS_UNET_WGRAD_PER_GROUP 16   // slots summed per thread in the first reduction pass
int cs_unet_wgrad_slots(int s, int cin, int cout);
// Per-slot partials plus the first reduction pass's group sums.
inline size_t cs_unet_wgrad_floats(int s, int cin, int cout) {
  const int slots = cs_unet_wgrad_slots(s, cin, cout);
  return (size_t)(slots + (slots + CS_UNET_WGRAD_PER_GROUP - 1) / CS_UNET_WGRAD_PER_GROUP) * 27 * cin * cout;
}
hipError_t cs_unet_launch_wgrad(hipStream_t st, const uint16_t* in, int cin, const uint16_t* gout,
                                const uint16_t* outv, int cout, int s, float* partial, float* bias_part, float* dw,
                                float* db);
hipError_t cs_unet_launch_pool(hipStream_t st, const uint16_t* in, int c, int s, uint16_t* out);
hipError_t cs_unet_launch_pool_back(hipStream_t st, const uint16_t* g_out, int c, int s, const uint16_t* add,
                                    int add_stride, int add_off, uint16_t* g_in);
hipError_t cs_unet_launch_up(hipStream_t st, const uint16_t* in, int c, int s, uint16_t* cat, int stride, int off);
hipError_t cs_unet_launch_up_back(hipStream_t st, const uint16_t* g_cat, int stride, int off, int c, int s,
                                  uint16_t* g_in);
hipError_t cs_unet_launch_put(hipStream_t st, const uint16_t* in, int c, size_t nv, uint16_t* cat, int stride,
                              int off);
hipError_t cs_unet_launch_head(hipStream_t st, const uint16_t* x, const float* wh, size_t nv, float* pred);
// block_sse: one float per 256 voxels; sse: 1 float.
hipError_t cs_unet_launch_loss(hipStream_t st, const float* pred, const float* target, const uint8_t* mask, size_t nv,
                               float inv_count, float* gpred, float* block_sse, float* sse);
// block_part: 17 floats per 256 voxels; dwh: 17 floats (16 weights, bias).
hipError_t cs_unet_launch_head_back(hipStream_t st, const uint16_t* x, const float* wh, const float* gpred, size_t nv,
                                    uint16_t* g_x, float* block_part, float* dwh);
hipError_t cs_unet_launch_prep_weights(hipStream_t st, const float* w, int cin, int cout, uint16_t* wf, uint16_t* wd);
hipError_t cs_unet_launch_adam(hipStream_t st, float* p, const float* g, float* m, float* v, size_t n, float lr_t,
                               float b1, float b2, float eps);
hipError_t cs_unet_launch_scale(hipStream_t st, const float* pred, size_t nv, float scale, float* out);
// Four float features per voxel -> the 16 bf16 input channels (12 zero).
hipError_t cs_unet_launch_pack_features(hipStream_t st, const float* f4, size_t nv, uint16_t* x);
