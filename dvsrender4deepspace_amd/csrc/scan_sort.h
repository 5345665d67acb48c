// Hand-written device prefix sum and stable radix sort (scan_sort.hip).
#pragma once
#include <cstddef>
#include <hip/hip_runtime.h>

namespace dvsamd {
// Inclusive prefix sum of n ints (in == out allowed).  With temp == nullptr
// only sets temp_bytes, the scratch the call needs.
hipError_t inclusive_scan_int(void *temp, size_t &temp_bytes, const int *in, int *out, size_t n,
                              hipStream_t stream);
// Stable LSD radix sort of (key, value) pairs by key bits [begin_bit,
// end_bit); the inputs are left unchanged.  Same temp_bytes protocol.
hipError_t sort_pairs_u64(void *temp, size_t &temp_bytes, const unsigned long long *keys_in,
                          unsigned long long *keys_out, const unsigned *values_in, unsigned *values_out,
                          size_t n, int begin_bit, int end_bit, hipStream_t stream);
// Stable sort of (key, value) pairs whose keys all lie in [key_lo, key_lo +
// 2^key_bits): by the bucket (key - key_lo) >> (key_bits - bucket_bits)
// (bucket_bits bits, so ceil(bucket_bits / 8) passes), then inside each
// bucket by the whole key.  A bucket of more than 64 pairs is left in
// input order and sets *flag (the caller then sorts by all key bits).  Same
// temp_bytes protocol as sort_pairs_u64; the output may not alias the input.
hipError_t sort_pairs_u64_bucketed(void *temp, size_t &temp_bytes, const unsigned long long *keys_in,
                                   unsigned long long *keys_out, const unsigned *values_in, unsigned *values_out,
                                   size_t n, unsigned long long key_lo, int key_bits, int bucket_bits, int *flag,
                                   hipStream_t stream);
}  // namespace dvsamd
