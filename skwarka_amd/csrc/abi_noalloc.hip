// skw_closest_hit_noalloc: closest hit for callers that cannot allocate device memory (code under
// HIP graph capture, serving loops, harnesses that forbid hipMalloc). The launch's working memory
// (the ray-claim counter and, for trees deeper than the LDS stacks, the global traversal stack)
// lives in a fixed file-scope device array instead of the scene's per-stream counters and the
// library's memory pool. Everything else is skw_closest_hit's: the same kernels, the same walks,
// bit-identical hit records.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.hpp"

using namespace skw;

namespace {

// 256 B for the counter, the rest for a global traversal stack. Every GPU the library runs on
// holds its own copy of this array until the process ends.
constexpr uint64_t kScratchBytes = 16ull << 20;
constexpr uint64_t kStackOffset = 256;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per device. In host code
// the array's name is only a stand-in for the symbol: the address comes from hipGetSymbolAddress.
skw_status scratch_ptr(void **out) {
  int dev = 0;
  SKW_HIP(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(SKW_UNSUPPORTED, "device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    SKW_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) return fail(SKW_HIP_ERROR, "hipGetSymbolAddress(scratch) returned NULL");
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return SKW_OK;
}

}  // namespace

extern "C" skw_status skw_closest_hit_noalloc(const skw_scene *scene, const skw_ray *d_rays, uint64_t n,
                                              skw_hit *d_hits, void *stream) {
  if (!scene) return fail(SKW_INVALID_ARG, "scene must not be NULL");
  if (!scene->on_device) return fail(SKW_UNSUPPORTED, "scene was created host-only");
  if (n == 0) return SKW_OK;
  if (!d_rays || !d_hits) return fail(SKW_INVALID_ARG, "ray/result pointers must not be NULL");
  const uint64_t need = scene->host.stack_need;
  if (need > 128 && (uint64_t)64 * need * sizeof(uint32_t) > kScratchBytes - kStackOffset)
    return fail(SKW_UNSUPPORTED, "the tree's traversal stack (" + std::to_string(need) +
                                     " entries per ray) exceeds the fixed working memory of " +
                                     std::to_string(kScratchBytes >> 20) + " MiB");
  int dev = 0;
  SKW_HIP(hipGetDevice(&dev), "hipGetDevice");
  if (dev != scene->device) return fail(SKW_INVALID_ARG, "the scene lives on another device than the current one");
  void *base = nullptr;
  skw_status s = scratch_ptr(&base);
  if (s != SKW_OK) return s;
  TraceStorage fixed;
  fixed.counter = static_cast<uint64_t *>(base);
  fixed.gstack = reinterpret_cast<uint32_t *>(static_cast<unsigned char *>(base) + kStackOffset);
  fixed.gstack_bytes = kScratchBytes - kStackOffset;
  return enqueue_trace(scene, d_rays, n, nullptr, d_hits, nullptr, reinterpret_cast<hipStream_t>(stream), false,
                       nullptr, nullptr, true, nullptr, &fixed);
}
