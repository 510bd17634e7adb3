// Device-wide sort and scans used to build join groups (rocPRIM, header
// only).  Kept in its own translation unit: these templates dominate
// compile time and rarely change.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "hb_internal.h"

namespace hb {

hipError_t sort_pairs_u64(void *temp, size_t &temp_bytes, const uint64_t *keys_in,
                          uint64_t *keys_out, const uint32_t *vals_in, uint32_t *vals_out,
                          uint32_t n, int end_bit, hipStream_t stream) {
  return rocprim::radix_sort_pairs(temp, temp_bytes, keys_in, keys_out, vals_in, vals_out,
                                   (size_t)n, 0, end_bit, stream);
}

hipError_t exscan_u32(void *temp, size_t &temp_bytes, const uint32_t *in, uint32_t *out,
                      uint32_t n, hipStream_t stream) {
  return rocprim::exclusive_scan(temp, temp_bytes, in, out, (uint32_t)0, (size_t)n,
                                 rocprim::plus<uint32_t>(), stream);
}

hipError_t exscan_u64(void *temp, size_t &temp_bytes, const uint64_t *in, uint64_t *out,
                      uint32_t n, hipStream_t stream) {
  return rocprim::exclusive_scan(temp, temp_bytes, in, out, (uint64_t)0, (size_t)n,
                                 rocprim::plus<uint64_t>(), stream);
}

}  // namespace hb
