// cs_unet.cpp — the learned pressure warm start: parameters, forward and
// backward passes, Adam, and the device-resident training set (C ABI in
// include/coffee_sim.h).  The network has no reference counterpart; it stands in
// for the zero pressure every reference solve starts from (shader.rs:1187).
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <string>
#include <vector>

#include "cs_ctx.h"
#include "cs_unet.h"
#include "cs_unet_host.h"

namespace {

struct ConvSpec {
  int cin, cout, level;  // level: grid side n >> level
};
// e1a e1b | e2a e2b | ba bb | d2a d2b | d1a d1b
constexpr ConvSpec kConvs[10] = {{16, 16, 0}, {16, 16, 0}, {16, 32, 1}, {32, 32, 1}, {32, 64, 2},
                                 {64, 64, 2},  {96, 32, 1}, {32, 32, 1}, {48, 16, 0}, {16, 16, 0}};
enum { E1A, E1B, E2A, E2B, BA, BB, D2A, D2B, D1A, D1B };

#define U_HIP(call)                                                            \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) return cs_unet_fail(std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t count) {
  *p = nullptr;
  return hipMalloc(reinterpret_cast<void**>(p), (count ? count : 1) * sizeof(T));
}

struct Sample {
  float* f4 = nullptr;      // [nv][4]
  float* target = nullptr;  // [nv]
  uint8_t* mask = nullptr;  // [nv]
  float inv_count = 1.0f;
};

}  // namespace

struct cs_unet {
  cs_ctx* ctx = nullptr;
  int n = 0;
  size_t nv = 0;
  size_t np = 0;
  size_t woff[10] = {}, boff[10] = {}, head_off = 0;
  float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr;
  uint16_t* wf[10] = {};
  uint16_t* wd[10] = {};
  uint16_t *x = nullptr, *e1a = nullptr, *e1b = nullptr, *p1 = nullptr, *e2a = nullptr, *e2b = nullptr,
           *p2 = nullptr, *ba = nullptr, *bb = nullptr, *c2 = nullptr, *d2a = nullptr, *d2b = nullptr,
           *c1 = nullptr, *d1a = nullptr, *d1b = nullptr;
  uint16_t *g_d1b = nullptr, *g_d1a = nullptr, *g_c1 = nullptr, *g_d2b = nullptr, *g_d2a = nullptr,
           *g_c2 = nullptr, *g_bb = nullptr, *g_ba = nullptr, *g_p2 = nullptr, *g_e2b = nullptr, *g_e2a = nullptr,
           *g_p1 = nullptr, *g_e1b = nullptr, *g_e1a = nullptr;
  float *pred = nullptr, *gpred = nullptr, *wpart = nullptr, *bpart = nullptr, *block_sse = nullptr,
        *sse = nullptr, *head_part = nullptr;
  float *f4 = nullptr, *target = nullptr;
  uint8_t* mask = nullptr;
  uint32_t* count = nullptr;  // device counter for captures
  float* h_scalar = nullptr;  // pinned
  int step = 0;
  float rhs_scale = 1.0f, p_scale = 1.0f;
  std::vector<Sample> data;
};

namespace {

thread_local std::string g_unet_error;

int side(const cs_unet* u, int level) { return u->n >> level; }
size_t vox(const cs_unet* u, int level) {
  const size_t s = (size_t)side(u, level);
  return s * s * s;
}

void free_unet(cs_unet* u) {
  if (!u) return;
  void* ptrs[] = {u->P,     u->G,     u->M,     u->V,     u->x,     u->e1a,   u->e1b,   u->p1,      u->e2a,
                  u->e2b,   u->p2,    u->ba,    u->bb,    u->c2,    u->d2a,   u->d2b,   u->c1,      u->d1a,
                  u->d1b,   u->g_d1b, u->g_d1a, u->g_c1,  u->g_d2b, u->g_d2a, u->g_c2,  u->g_bb,    u->g_ba,
                  u->g_p2,  u->g_e2b, u->g_e2a, u->g_p1,  u->g_e1b, u->g_e1a, u->pred,  u->gpred,   u->wpart,
                  u->bpart, u->block_sse, u->sse, u->head_part, u->f4, u->target, u->mask, u->count};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (int k = 0; k < 10; ++k) {
    if (u->wf[k]) (void)hipFree(u->wf[k]);
    if (u->wd[k]) (void)hipFree(u->wd[k]);
  }
  for (auto& smp : u->data) {
    if (smp.f4) (void)hipFree(smp.f4);
    if (smp.target) (void)hipFree(smp.target);
    if (smp.mask) (void)hipFree(smp.mask);
  }
  if (u->h_scalar) (void)hipHostFree(u->h_scalar);
  delete u;
}

int prep_weights(cs_unet* u) {
  hipStream_t st = u->ctx->stream;
  for (int k = 0; k < 10; ++k)
    U_HIP(cs_unet_launch_prep_weights(st, u->P + u->woff[k], kConvs[k].cin, kConvs[k].cout, u->wf[k], u->wd[k]));
  return CS_OK;
}

int conv_fwd(cs_unet* u, int k, const uint16_t* in, uint16_t* out) {
  const ConvSpec& c = kConvs[k];
  U_HIP(cs_unet_launch_conv(u->ctx->stream, in, c.cin, side(u, c.level), u->wf[k], u->P + u->boff[k], nullptr, out,
                            c.cout, 1));
  return CS_OK;
}

// Weight/bias gradient of convolution k, and (unless `gin` is null) the
// gradient of its input, given the gradient `gout` of its ReLU output `out`.
int conv_back(cs_unet* u, int k, const uint16_t* in, const uint16_t* out, const uint16_t* gout, uint16_t* gin) {
  const ConvSpec& c = kConvs[k];
  hipStream_t st = u->ctx->stream;
  const int s = side(u, c.level);
  U_HIP(cs_unet_launch_wgrad(st, in, c.cin, gout, out, c.cout, s, u->wpart, u->bpart, u->G + u->woff[k],
                             u->G + u->boff[k]));
  if (gin) U_HIP(cs_unet_launch_conv(st, gout, c.cout, s, u->wd[k], nullptr, out, gin, c.cin, 0));
  return CS_OK;
}

int forward(cs_unet* u) {
  hipStream_t st = u->ctx->stream;
  int rc;
  if ((rc = conv_fwd(u, E1A, u->x, u->e1a))) return rc;
  if ((rc = conv_fwd(u, E1B, u->e1a, u->e1b))) return rc;
  U_HIP(cs_unet_launch_pool(st, u->e1b, 16, side(u, 0), u->p1));
  if ((rc = conv_fwd(u, E2A, u->p1, u->e2a))) return rc;
  if ((rc = conv_fwd(u, E2B, u->e2a, u->e2b))) return rc;
  U_HIP(cs_unet_launch_pool(st, u->e2b, 32, side(u, 1), u->p2));
  if ((rc = conv_fwd(u, BA, u->p2, u->ba))) return rc;
  if ((rc = conv_fwd(u, BB, u->ba, u->bb))) return rc;
  U_HIP(cs_unet_launch_up(st, u->bb, 64, side(u, 2), u->c2, 96, 0));
  U_HIP(cs_unet_launch_put(st, u->e2b, 32, vox(u, 1), u->c2, 96, 64));
  if ((rc = conv_fwd(u, D2A, u->c2, u->d2a))) return rc;
  if ((rc = conv_fwd(u, D2B, u->d2a, u->d2b))) return rc;
  U_HIP(cs_unet_launch_up(st, u->d2b, 32, side(u, 1), u->c1, 48, 0));
  U_HIP(cs_unet_launch_put(st, u->e1b, 16, vox(u, 0), u->c1, 48, 32));
  if ((rc = conv_fwd(u, D1A, u->c1, u->d1a))) return rc;
  if ((rc = conv_fwd(u, D1B, u->d1a, u->d1b))) return rc;
  U_HIP(cs_unet_launch_head(st, u->d1b, u->P + u->head_off, u->nv, u->pred));
  return CS_OK;
}

// Loss against (target, mask) and every parameter gradient into G.
int backward(cs_unet* u, const float* target, const uint8_t* mask, float inv_count) {
  hipStream_t st = u->ctx->stream;
  int rc;
  U_HIP(cs_unet_launch_loss(st, u->pred, target, mask, u->nv, inv_count, u->gpred, u->block_sse, u->sse));
  U_HIP(cs_unet_launch_head_back(st, u->d1b, u->P + u->head_off, u->gpred, u->nv, u->g_d1b, u->head_part,
                                 u->G + u->head_off));
  if ((rc = conv_back(u, D1B, u->d1a, u->d1b, u->g_d1b, u->g_d1a))) return rc;
  if ((rc = conv_back(u, D1A, u->c1, u->d1a, u->g_d1a, u->g_c1))) return rc;
  U_HIP(cs_unet_launch_up_back(st, u->g_c1, 48, 0, 32, side(u, 1), u->g_d2b));
  if ((rc = conv_back(u, D2B, u->d2a, u->d2b, u->g_d2b, u->g_d2a))) return rc;
  if ((rc = conv_back(u, D2A, u->c2, u->d2a, u->g_d2a, u->g_c2))) return rc;
  U_HIP(cs_unet_launch_up_back(st, u->g_c2, 96, 0, 64, side(u, 2), u->g_bb));
  if ((rc = conv_back(u, BB, u->ba, u->bb, u->g_bb, u->g_ba))) return rc;
  if ((rc = conv_back(u, BA, u->p2, u->ba, u->g_ba, u->g_p2))) return rc;
  U_HIP(cs_unet_launch_pool_back(st, u->g_p2, 32, side(u, 1), u->g_c2, 96, 64, u->g_e2b));
  if ((rc = conv_back(u, E2B, u->e2a, u->e2b, u->g_e2b, u->g_e2a))) return rc;
  if ((rc = conv_back(u, E2A, u->p1, u->e2a, u->g_e2a, u->g_p1))) return rc;
  U_HIP(cs_unet_launch_pool_back(st, u->g_p1, 16, side(u, 0), u->g_c1, 48, 32, u->g_e1b));
  if ((rc = conv_back(u, E1B, u->e1a, u->e1b, u->g_e1b, u->g_e1a))) return rc;
  if ((rc = conv_back(u, E1A, u->x, u->e1a, u->g_e1a, nullptr))) return rc;
  return CS_OK;
}

int adam(cs_unet* u, float lr) {
  ++u->step;
  const float b1 = 0.9f, b2 = 0.999f;
  const float lr_t = lr * sqrtf(1.0f - powf(b2, (float)u->step)) / (1.0f - powf(b1, (float)u->step));
  U_HIP(cs_unet_launch_adam(u->ctx->stream, u->P, u->G, u->M, u->V, u->np, lr_t, b1, b2, 1e-8f));
  return prep_weights(u);
}

int read_loss(cs_unet* u, float* loss, float inv_count) {
  U_HIP(hipMemcpyAsync(u->h_scalar, u->sse, sizeof(float), hipMemcpyDeviceToHost, u->ctx->stream));
  U_HIP(hipStreamSynchronize(u->ctx->stream));
  if (loss) *loss = u->h_scalar[0] * inv_count;
  return CS_OK;
}

// Uploads host features/target/mask into the working buffers; returns 1/count.
int upload(cs_unet* u, const float* f4, const float* target, const uint8_t* mask, float* inv_count) {
  hipStream_t st = u->ctx->stream;
  U_HIP(hipMemcpyAsync(u->f4, f4, u->nv * 4 * sizeof(float), hipMemcpyHostToDevice, st));
  if (target) U_HIP(hipMemcpyAsync(u->target, target, u->nv * sizeof(float), hipMemcpyHostToDevice, st));
  if (mask) {
    U_HIP(hipMemcpyAsync(u->mask, mask, u->nv, hipMemcpyHostToDevice, st));
    size_t cnt = 0;
    for (size_t i = 0; i < u->nv; ++i) cnt += mask[i] ? 1u : 0u;
    *inv_count = 1.0f / (float)std::max<size_t>(cnt, 1);
  }
  U_HIP(cs_unet_launch_pack_features(st, u->f4, u->nv, u->x));
  return CS_OK;
}

}  // namespace

int cs_unet_fail(const std::string& msg) {
  g_unet_error = msg;
  cs_set_last_error(msg.c_str());
  return CS_ERR_DEVICE;
}

// ---- internal API used by the simulation ----------------------------------------------------
int cs_unet_side(const cs_unet* u) { return u ? u->n : 0; }
cs_ctx* cs_unet_context(const cs_unet* u) { return u ? u->ctx : nullptr; }
float cs_unet_rhs_scale(const cs_unet* u) { return u->rhs_scale; }
float cs_unet_pressure_scale(const cs_unet* u) { return u->p_scale; }
float* cs_unet_feature_buffer(cs_unet* u) { return u->f4; }

int cs_unet_predict_device(cs_unet* u, float* pressure_out) {
  hipStream_t st = u->ctx->stream;
  U_HIP(cs_unet_launch_pack_features(st, u->f4, u->nv, u->x));
  int rc = forward(u);
  if (rc) return rc;
  U_HIP(cs_unet_launch_scale(st, u->pred, u->nv, u->p_scale, pressure_out));
  return CS_OK;
}

int cs_unet_new_sample(cs_unet* u, float** f4, float** target, uint8_t** mask, uint32_t** count) {
  Sample smp;
  hipError_t e = dalloc(&smp.f4, u->nv * 4);
  if (e == hipSuccess) e = dalloc(&smp.target, u->nv);
  if (e == hipSuccess) e = dalloc(&smp.mask, u->nv);
  if (e != hipSuccess) {
    if (smp.f4) (void)hipFree(smp.f4);
    if (smp.target) (void)hipFree(smp.target);
    if (smp.mask) (void)hipFree(smp.mask);
    g_unet_error = "training sample allocation failed";
    cs_set_last_error(g_unet_error.c_str());
    return CS_ERR_NOMEM;
  }
  U_HIP(hipMemsetAsync(u->count, 0, sizeof(uint32_t), u->ctx->stream));
  u->data.push_back(smp);
  *f4 = smp.f4;
  *target = smp.target;
  *mask = smp.mask;
  *count = u->count;
  return CS_OK;
}

int cs_unet_finish_sample(cs_unet* u) {
  uint32_t c = 0;
  U_HIP(hipMemcpyAsync(u->h_scalar, u->count, sizeof(uint32_t), hipMemcpyDeviceToHost, u->ctx->stream));
  U_HIP(hipStreamSynchronize(u->ctx->stream));
  memcpy(&c, u->h_scalar, sizeof(c));
  u->data.back().inv_count = 1.0f / (float)std::max<uint32_t>(c, 1u);
  return CS_OK;
}

// ---- C ABI ------------------------------------------------------------------------------------
extern "C" {

int cs_unet_create(cs_ctx* ctx, uint32_t n, uint64_t seed, cs_unet** out) {
  if (!ctx || !out) return CS_ERR_INVALID;
  *out = nullptr;
  if (n < 32 || n % 32 != 0 || n > 512) {
    cs_set_last_error("cs_unet_create: the grid side must be a multiple of 32 (32..512)");
    return CS_ERR_INVALID;
  }
  U_HIP(hipSetDevice(ctx->device));
  cs_unet* u = new cs_unet();
  u->ctx = ctx;
  u->n = (int)n;
  u->nv = (size_t)n * n * n;
  size_t o = 0;
  for (int k = 0; k < 10; ++k) {
    u->woff[k] = o;
    o += (size_t)27 * kConvs[k].cin * kConvs[k].cout;
    u->boff[k] = o;
    o += (size_t)kConvs[k].cout;
  }
  u->head_off = o;
  o += 17;
  u->np = o;
  const size_t v0 = vox(u, 0), v1 = vox(u, 1), v2 = vox(u, 2);
  size_t wmax = 0;
  for (int k = 0; k < 10; ++k)
    wmax = std::max(wmax, cs_unet_wgrad_floats(side(u, kConvs[k].level), kConvs[k].cin, kConvs[k].cout));
  hipError_t e = hipSuccess;
#define UA(p, cnt) \
  if (e == hipSuccess) e = dalloc(&(p), (cnt));
  UA(u->P, u->np) UA(u->G, u->np) UA(u->M, u->np) UA(u->V, u->np)
  for (int k = 0; k < 10; ++k) {
    UA(u->wf[k], (size_t)27 * kConvs[k].cin * kConvs[k].cout)
    UA(u->wd[k], (size_t)27 * kConvs[k].cin * kConvs[k].cout)
  }
  UA(u->x, v0 * 16) UA(u->e1a, v0 * 16) UA(u->e1b, v0 * 16) UA(u->p1, v1 * 16) UA(u->e2a, v1 * 32)
  UA(u->e2b, v1 * 32) UA(u->p2, v2 * 32) UA(u->ba, v2 * 64) UA(u->bb, v2 * 64) UA(u->c2, v1 * 96)
  UA(u->d2a, v1 * 32) UA(u->d2b, v1 * 32) UA(u->c1, v0 * 48) UA(u->d1a, v0 * 16) UA(u->d1b, v0 * 16)
  UA(u->g_d1b, v0 * 16) UA(u->g_d1a, v0 * 16) UA(u->g_c1, v0 * 48) UA(u->g_d2b, v1 * 32) UA(u->g_d2a, v1 * 32)
  UA(u->g_c2, v1 * 96) UA(u->g_bb, v2 * 64) UA(u->g_ba, v2 * 64) UA(u->g_p2, v2 * 32) UA(u->g_e2b, v1 * 32)
  UA(u->g_e2a, v1 * 32) UA(u->g_p1, v1 * 16) UA(u->g_e1b, v0 * 16) UA(u->g_e1a, v0 * 16)
  UA(u->pred, v0) UA(u->gpred, v0) UA(u->wpart, wmax) UA(u->bpart, (size_t)CS_UNET_BIAS_ROWS * 96)
  UA(u->block_sse, (v0 + 255) / 256) UA(u->sse, 1) UA(u->head_part, ((v0 + 255) / 256) * 17)
  UA(u->f4, v0 * 4) UA(u->target, v0) UA(u->mask, v0) UA(u->count, 1)
#undef UA
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&u->h_scalar), 16);
  if (e != hipSuccess) {
    free_unet(u);
    cs_set_last_error("cs_unet_create: device allocation failed");
    return CS_ERR_NOMEM;
  }
  // He-normal weights (std sqrt(2 / fan_in)), zero biases, head std sqrt(1/16)
  std::vector<float> p(u->np, 0.0f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  for (int k = 0; k < 10; ++k) {
    const float sd = sqrtf(2.0f / (27.0f * (float)kConvs[k].cin));
    for (size_t i = 0; i < (size_t)27 * kConvs[k].cin * kConvs[k].cout; ++i) p[u->woff[k] + i] = sd * nd(rng);
  }
  for (int c = 0; c < 16; ++c) p[u->head_off + c] = 0.25f * nd(rng);
  const int rc = cs_unet_set_params(u, p.data());
  if (rc) {
    free_unet(u);
    return rc;
  }
  *out = u;
  return CS_OK;
}

int cs_unet_destroy(cs_unet* u) {
  if (!u) return CS_ERR_INVALID;
  (void)hipSetDevice(u->ctx->device);
  (void)hipStreamSynchronize(u->ctx->stream);
  free_unet(u);
  return CS_OK;
}

int cs_unet_param_count(const cs_unet* u, uint32_t* count) {
  if (!u || !count) return CS_ERR_INVALID;
  *count = (uint32_t)u->np;
  return CS_OK;
}

int cs_unet_get_params(cs_unet* u, float* out) {
  if (!u || !out) return CS_ERR_INVALID;
  U_HIP(hipMemcpyAsync(out, u->P, u->np * sizeof(float), hipMemcpyDeviceToHost, u->ctx->stream));
  U_HIP(hipStreamSynchronize(u->ctx->stream));
  return CS_OK;
}

int cs_unet_set_params(cs_unet* u, const float* in) {
  if (!u || !in) return CS_ERR_INVALID;
  hipStream_t st = u->ctx->stream;
  U_HIP(hipMemcpyAsync(u->P, in, u->np * sizeof(float), hipMemcpyHostToDevice, st));
  U_HIP(hipMemsetAsync(u->M, 0, u->np * sizeof(float), st));
  U_HIP(hipMemsetAsync(u->V, 0, u->np * sizeof(float), st));
  u->step = 0;
  const int rc = prep_weights(u);
  if (rc) return rc;
  U_HIP(hipStreamSynchronize(st));
  return CS_OK;
}

int cs_unet_set_scales(cs_unet* u, float rhs_scale, float pressure_scale) {
  if (!u || !(rhs_scale > 0.0f) || !(pressure_scale > 0.0f)) return CS_ERR_INVALID;
  u->rhs_scale = rhs_scale;
  u->p_scale = pressure_scale;
  return CS_OK;
}

int cs_unet_get_scales(const cs_unet* u, float* rhs_scale, float* pressure_scale) {
  if (!u || !rhs_scale || !pressure_scale) return CS_ERR_INVALID;
  *rhs_scale = u->rhs_scale;
  *pressure_scale = u->p_scale;
  return CS_OK;
}

int cs_unet_forward(cs_unet* u, const float* f4, float* pred) {
  if (!u || !f4 || !pred) return CS_ERR_INVALID;
  float inv = 1.0f;
  int rc = upload(u, f4, nullptr, nullptr, &inv);
  if (!rc) rc = forward(u);
  if (rc) return rc;
  U_HIP(hipMemcpyAsync(pred, u->pred, u->nv * sizeof(float), hipMemcpyDeviceToHost, u->ctx->stream));
  U_HIP(hipStreamSynchronize(u->ctx->stream));
  return CS_OK;
}

int cs_unet_gradient(cs_unet* u, const float* f4, const float* target, const uint8_t* mask, float* loss,
                     float* grad) {
  if (!u || !f4 || !target || !mask) return CS_ERR_INVALID;
  float inv = 1.0f;
  int rc = upload(u, f4, target, mask, &inv);
  if (!rc) rc = forward(u);
  if (!rc) rc = backward(u, u->target, u->mask, inv);
  if (!rc) rc = read_loss(u, loss, inv);
  if (rc) return rc;
  if (grad) {
    U_HIP(hipMemcpyAsync(grad, u->G, u->np * sizeof(float), hipMemcpyDeviceToHost, u->ctx->stream));
    U_HIP(hipStreamSynchronize(u->ctx->stream));
  }
  return CS_OK;
}

int cs_unet_train_step(cs_unet* u, const float* f4, const float* target, const uint8_t* mask, float lr, float* loss) {
  if (!u || !f4 || !target || !mask) return CS_ERR_INVALID;
  float inv = 1.0f;
  int rc = upload(u, f4, target, mask, &inv);
  if (!rc) rc = forward(u);
  if (!rc) rc = backward(u, u->target, u->mask, inv);
  if (!rc) rc = adam(u, lr);
  if (!rc) rc = read_loss(u, loss, inv);
  return rc;
}

int cs_unet_benchmark(cs_unet* u, uint32_t reps, double* forward_us, double* gradient_us) {
  if (!u || reps == 0) return CS_ERR_INVALID;
  hipStream_t st = u->ctx->stream;
  hipEvent_t ev[3] = {};
  for (auto& e : ev) U_HIP(hipEventCreate(&e));
  int rc = CS_OK;
  const float inv = 1.0f / (float)u->nv;
  U_HIP(hipEventRecord(ev[0], st));
  for (uint32_t r = 0; r < reps && rc == CS_OK; ++r) rc = forward(u);
  U_HIP(hipEventRecord(ev[1], st));
  for (uint32_t r = 0; r < reps && rc == CS_OK; ++r) {
    rc = forward(u);
    if (!rc) rc = backward(u, u->target, u->mask, inv);
  }
  U_HIP(hipEventRecord(ev[2], st));
  U_HIP(hipEventSynchronize(ev[2]));
  float ms[2] = {0.0f, 0.0f};
  U_HIP(hipEventElapsedTime(&ms[0], ev[0], ev[1]));
  U_HIP(hipEventElapsedTime(&ms[1], ev[1], ev[2]));
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (rc) return rc;
  if (forward_us) *forward_us = 1000.0 * ms[0] / reps;
  if (gradient_us) *gradient_us = 1000.0 * ms[1] / reps;
  return CS_OK;
}

int cs_unet_dataset_size(const cs_unet* u, uint32_t* count) {
  if (!u || !count) return CS_ERR_INVALID;
  *count = (uint32_t)u->data.size();
  return CS_OK;
}

int cs_unet_dataset_clear(cs_unet* u) {
  if (!u) return CS_ERR_INVALID;
  (void)hipStreamSynchronize(u->ctx->stream);
  for (auto& smp : u->data) {
    (void)hipFree(smp.f4);
    (void)hipFree(smp.target);
    (void)hipFree(smp.mask);
  }
  u->data.clear();
  return CS_OK;
}

int cs_unet_read_sample(cs_unet* u, uint32_t i, float* f4, float* target, uint8_t* mask) {
  if (!u || i >= u->data.size()) return CS_ERR_INVALID;
  hipStream_t st = u->ctx->stream;
  const Sample& smp = u->data[i];
  if (f4) U_HIP(hipMemcpyAsync(f4, smp.f4, u->nv * 4 * sizeof(float), hipMemcpyDeviceToHost, st));
  if (target) U_HIP(hipMemcpyAsync(target, smp.target, u->nv * sizeof(float), hipMemcpyDeviceToHost, st));
  if (mask) U_HIP(hipMemcpyAsync(mask, smp.mask, u->nv, hipMemcpyDeviceToHost, st));
  U_HIP(hipStreamSynchronize(st));
  return CS_OK;
}

int cs_unet_train_epoch(cs_unet* u, float lr, uint64_t seed, float* mean_loss) {
  if (!u) return CS_ERR_INVALID;
  if (u->data.empty()) {
    cs_set_last_error("cs_unet_train_epoch: no captured samples");
    return CS_ERR_STATE;
  }
  std::vector<uint32_t> order(u->data.size());
  for (uint32_t i = 0; i < order.size(); ++i) order[i] = i;
  std::mt19937_64 rng(seed);
  std::shuffle(order.begin(), order.end(), rng);
  hipStream_t st = u->ctx->stream;
  double total = 0.0;
  for (uint32_t i : order) {
    const Sample& smp = u->data[i];
    U_HIP(cs_unet_launch_pack_features(st, smp.f4, u->nv, u->x));
    int rc = forward(u);
    if (!rc) rc = backward(u, smp.target, smp.mask, smp.inv_count);
    if (!rc) rc = adam(u, lr);
    float l = 0.0f;
    if (!rc) rc = read_loss(u, &l, smp.inv_count);
    if (rc) return rc;
    total += l;
  }
  if (mean_loss) *mean_loss = (float)(total / (double)order.size());
  return CS_OK;
}

}  // extern "C"
