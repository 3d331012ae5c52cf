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

template <int MODE, bool STATS>
hipError_t launch_mode(const GatherArgs &a, hipStream_t stream) {
  const uint32_t grid = (uint32_t)((a.nq + kWave - 1) / kWave);
  if (a.stack_need <= 32)
    hipLaunchKernelGGL((k_gather_lds<MODE, 32, STATS>), dim3(grid), dim3(kWave), 0, stream, a);
  else if (a.stack_need <= 64)
    hipLaunchKernelGGL((k_gather_lds<MODE, 64, STATS>), dim3(grid), dim3(kWave), 0, stream, a);
  else
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

hipError_t launch_irr_sum(const float *a, const float *b, const float *c, float *out, uint64_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_irr_sum, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, a, b, c, out, n);
  return hipGetLastError();
}

}  // namespace skw
