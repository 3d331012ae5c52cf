// Key-value radix sort (rocPRIM through hipCUB) used to order lookups spatially.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "kernels.hpp"

namespace skw {

hipError_t sort_pairs_u32(void *temp, size_t &temp_bytes, const uint32_t *keys_in, uint32_t *keys_out,
                          const uint32_t *vals_in, uint32_t *vals_out, uint64_t n, int begin_bit, int end_bit,
                          hipStream_t stream) {
  return hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys_in, keys_out, vals_in, vals_out, (int)n,
                                            begin_bit, end_bit, stream);
}

}  // namespace skw
