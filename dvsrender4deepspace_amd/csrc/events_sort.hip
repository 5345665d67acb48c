// Device scan and stable sort used by event compaction (rocPRIM primitives).
#include <cstring>
#include <hip/hip_runtime.h>
#include <rocprim/rocprim.hpp>

#include "events_sort.h"

namespace dvsamd {

hipError_t inclusive_scan_int(void *temp, size_t &temp_bytes, const int *in, int *out, size_t n,
                              hipStream_t stream) {
  return rocprim::inclusive_scan(temp, temp_bytes, in, out, n, rocprim::plus<int>(), stream);
}

hipError_t sort_pairs_u64(void *temp, size_t &temp_bytes, const unsigned long long *keys_in,
                          unsigned long long *keys_out, const unsigned *values_in, unsigned *values_out,
                          size_t n, int begin_bit, int end_bit, hipStream_t stream) {
  return rocprim::radix_sort_pairs(temp, temp_bytes, keys_in, keys_out, values_in, values_out, n,
                                   (unsigned)begin_bit, (unsigned)end_bit, stream);
}

}  // namespace dvsamd
