// Photon-map gather kernels on gfx950: one query per lane (all of a launch's queries start at
// once, so no query waits behind another), the pending far sides of the lookup in LDS (global
// memory for trees deeper than 64), heap slots per query contiguous in device memory. The lookup,
// heap and estimates are gather_device.hpp's.
#include <hip/hip_runtime.h>

#include "gather_device.hpp"
#include "kernels.hpp"

namespace skw {
namespace {

using namespace gath;

template <int MODE, int S, bool STATS>
__global__ __launch_bounds__(64) void k_gather_lds(GatherArgs a) {
  __shared__ uint32_t lds[S * kWave];
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nq) return;
  LdsStack<S> stk{lds + (threadIdx.x & (kWave - 1))};
  gather_one<MODE, STATS>(a, i, stk);
}

template <int MODE, bool STATS>
__global__ __launch_bounds__(64) void k_gather_global(GatherArgs a) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nq) return;
  GlobalStack stk{a.gstack + i, a.nq};
  gather_one<MODE, STATS>(a, i, stk);
}

template <int MODE, int S, bool STATS>
hipError_t launch_lds(const GatherArgs &a, uint32_t grid, hipStream_t stream) {
  hipLaunchKernelGGL((k_gather_lds<MODE, S, STATS>), dim3(grid), dim3(kWave), 0, stream, a);
  return hipGetLastError();
}

// The LDS stack holds stack_need entries per lane (4 bytes each), so it bounds occupancy together
// with the registers: nearest lookups (MODE 1) hold 56 VGPRs (36 waves per CU), so shallow trees
// get a short stack (32 entries: 8 KB per wave, 20 waves per CU; 24: 26; 20: 32; 18: 35; 16: 36).
// k-nearest lookups (MODE 0) hold 96 VGPRs (20 waves per CU) whatever the stack. (An instance for
// sets without an estimate holds 80 VGPRs; with a 24-entry stack it ran the final gather's 32-nearest
// lookups at 24 waves per CU no faster.)
template <int MODE, bool STATS>
hipError_t launch_mode(const GatherArgs &a, hipStream_t stream) {
  const uint32_t grid = (uint32_t)((a.nq + kWave - 1) / kWave);
  const uint32_t need = a.stack_need;
  if constexpr (MODE == 1) {
    if (need <= 16) return launch_lds<MODE, 16, STATS>(a, grid, stream);
    if (need <= 18) return launch_lds<MODE, 18, STATS>(a, grid, stream);
    if (need <= 20) return launch_lds<MODE, 20, STATS>(a, grid, stream);
    if (need <= 24) return launch_lds<MODE, 24, STATS>(a, grid, stream);
  }
  if (need <= 32) return launch_lds<MODE, 32, STATS>(a, grid, stream);
  if (need <= 64) return launch_lds<MODE, 64, STATS>(a, grid, stream);
  hipLaunchKernelGGL((k_gather_global<MODE, STATS>), dim3(grid), dim3(kWave), 0, stream, a);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_gather(const GatherArgs &args, int mode, hipStream_t stream) {
  if (args.nq == 0) return hipSuccess;
  if (args.stats) return mode == 1 ? launch_mode<1, true>(args, stream) : launch_mode<0, true>(args, stream);
  return mode == 1 ? launch_mode<1, false>(args, stream) : launch_mode<0, false>(args, stream);
}

namespace {
// out = (a + b) + c in float (a may be null: read as 0).
__global__ __launch_bounds__(256) void k_irr_sum(const float *a, const float *b, const float *c, float *out,
                                                 uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = a ? a[i] : 0.0f;
  out[i] = (x + b[i]) + c[i];
}
}  // namespace

namespace {
constexpr uint64_t kNboxMax = 2048;  // larger subtrees get the whole cube (they are rarely prunable)

// Node b's subtree holds the positions [b, e): e - 1 is its last node in pre-order, reached by
// following right children (or the left child where there is no right one) down to a leaf.
__global__ __launch_bounds__(256) void k_kd_nbox(const DevKdNode *__restrict__ nodes,
                                                 const DevPhoton *__restrict__ photons, const double *__restrict__ cdir,
                                                 uint64_t n, uint64_t *__restrict__ nbox,
                                                 DevNearRec *__restrict__ near) {
  const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  uint64_t cur = b;
  for (int d = 0; d < 64; ++d) {
    const uint32_t w = __float_as_uint(reinterpret_cast<const float4 *>(nodes)[cur].w);
    if ((w & 3u) == 3u) break;
    const uint64_t right = w >> 3;
    if (right < n)
      cur = right;
    else if (w & 4u)
      cur = cur + 1;
    else
      break;
  }
  const uint64_t e = cur + 1;
  double lo[3] = {1.0, 1.0, 1.0}, hi[3] = {-1.0, -1.0, -1.0};
  if (e - b > kNboxMax) {
    lo[0] = lo[1] = lo[2] = -1.0;
    hi[0] = hi[1] = hi[2] = 1.0;
  } else {
    for (uint64_t p = b; p < e; ++p) {
      const double *nv = cdir + 3 * (size_t)(photons[p].codes >> 16);
      for (int c = 0; c < 3; ++c) {
        lo[c] = fmin(lo[c], nv[c]);
        hi[c] = fmax(hi[c], nv[c]);
      }
    }
  }
  uint64_t box = 0;
  for (int c = 0; c < 3; ++c) {
    // outward to multiples of 1/127 (a margin keeps the rounding of x * 127 outward too)
    const int ql = (int)fmax(-127.0, floor(lo[c] * 127.0 - 1e-6));
    const int qh = (int)fmin(127.0, ceil(hi[c] * 127.0 + 1e-6));
    box |= (uint64_t)(uint8_t)(int8_t)ql << (8 * c);
    box |= (uint64_t)(uint8_t)(int8_t)qh << (8 * (c + 3));
  }
  nbox[b] = box;
  if (near) {
    DevNearRec r;
    r.node = reinterpret_cast<const float4 *>(nodes)[b];
    const double *nv = cdir + 3 * (size_t)(photons[b].codes >> 16);
    r.nv[0] = nv[0];
    r.nv[1] = nv[1];
    r.nv[2] = nv[2];
    r.nbox = box;
    r.pad[0] = 0;
    r.pad[1] = 0;
    near[b] = r;
  }
}
}  // namespace

hipError_t launch_kd_nbox(const DevKdNode *nodes, const DevPhoton *photons, const double *cdir, uint64_t n,
                          uint64_t *nbox, DevNearRec *near, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_kd_nbox, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, nodes, photons, cdir, n,
                     nbox, near);
  return hipGetLastError();
}

hipError_t launch_irr_sum(const float *a, const float *b, const float *c, float *out, uint64_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_irr_sum, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, a, b, c, out, n);
  return hipGetLastError();
}

}  // namespace skw
