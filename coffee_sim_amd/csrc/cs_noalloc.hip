// cs_noalloc.hip — cs_sim_pack_render_noalloc: the render packing of
// cs_sim_pack_render without any device allocation, for callers that cannot
// allocate (code under stream capture, tight frame loops, hosts that forbid it).
//
// The packed vertices are staged in `scratch`, a file-scope device array of a
// fixed size that each GPU the library runs on holds until the process ends.
// Its device address is looked up once per device (hipGetSymbolAddress, with
// that device current) and kept; host code passes that address, never the
// array's name.  Every call on a GPU shares that GPU's array: calls that
// overlap on one GPU (from two threads, or two streams in flight) race.
// Nothing in it needs to start at zero: every call writes each element it
// copies back.
#include <hip/hip_runtime.h>

#include <mutex>

#include "cs_host.h"
#include "cs_kernels.h"
#include "cs_sim_impl.h"

#define CS_NOALLOC_RENDER_MAX (1u << 18)  // particles (bed + water) one call packs: 8 MiB

static __device__ __align__(256) float scratch[(size_t)CS_NOALLOC_RENDER_MAX * 8];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch[kMaxDevices] = {};

// The scratch array's address on `device` (current), looked up on first use.
int scratch_on(int device, float** out) {
  if (device < 0 || device >= kMaxDevices) return cs_fail(CS_ERR_DEVICE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return cs_hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch[device] = static_cast<float*>(p);
  }
  *out = g_scratch[device];
  return CS_OK;
}

}  // namespace

extern "C" int cs_sim_pack_render_noalloc(cs_sim* m, float* render8) {
  if (!m || !render8) return cs_fail(CS_ERR_INVALID, "null argument");
  const uint64_t n = (uint64_t)m->num_bed + m->num_water;
  if (n > CS_NOALLOC_RENDER_MAX)
    return cs_fail(CS_ERR_RANGE, "cs_sim_pack_render_noalloc packs at most 262144 particles (bed + water)");
  hipError_t e = hipSetDevice(m->ctx->device);
  if (e != hipSuccess) return cs_hip_fail(e, "hipSetDevice");
  float* ptr = nullptr;
  int rc = scratch_on(m->ctx->device, &ptr);
  if (rc) return rc;
  if (n == 0) return CS_OK;
  hipStream_t st = m->ctx->stream;
  cs_params u;
  // the reference's order: bed particles first, then the water (mod.rs:239-249)
  if (m->num_bed) {
    cs::make_params(m->s, m->num_bed, m->num_bed_cells, 1.0e-4f, &u);  // dt does not enter the packing
    e = cs_launch_pack_render(st, u, m->bed_part, m->bed_aff, m->num_bed, m->bed_ext, ptr);
    if (e != hipSuccess) return cs_hip_fail(e, "pack_render (bed)");
  }
  cs::make_params(m->s, m->num_water, m->num_bed_cells, 1.0e-4f, &u);
  e = cs_launch_pack_render(st, u, m->part, m->aff, 0u, nullptr, ptr + (size_t)m->num_bed * 8);
  if (e != hipSuccess) return cs_hip_fail(e, "pack_render");
  e = hipMemcpyAsync(render8, ptr, (size_t)n * 32, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return cs_hip_fail(e, "cs_sim_pack_render_noalloc copy-back");
  return CS_OK;
}
