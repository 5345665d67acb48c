// Wave64 prefix sums and a stable LSD radix sort, written for CDNA4.
//
// Prefix sum (reduce, then scan): tiles of 2,048 items (256 threads x 8
// consecutive items).  Pass 1 reduces each tile; pass 2, one block, scans
// the tile sums; pass 3 rescans each tile in registers, the 64 lanes of a
// wave with DPP row shifts (__shfl_up), the four waves of a block through
// LDS, and adds the tile's prefix.  A single tile needs pass 3 only.  The
// scan may run in place: pass 3 reads a thread's items before it writes them.
//
// Radix sort: 8-bit digits from begin_bit to end_bit, one pass per digit:
// per-tile digit histograms (LDS atomics), an inclusive scan of the
// histograms in digit-major order (the prefix sum above) that gives every
// (digit, tile) its output base, and a scatter that ranks equal digits
// stably: a tile is consumed in 8 rounds of 256 consecutive items; inside a
// wave, the lanes holding the same digit are found with eight ballots (one
// per digit bit), a lane's rank is the number of such lanes below it, and
// per-wave digit counts in LDS order the four waves of a round and carry
// the running count from round to round.  Equal keys keep their input
// order, so sorting by time after compaction in (segment, y, x) order gives
// the (segment, t, y, x) order of the readout.
//
// Bucketed sort, for keys known to lie in [lo, lo + 2^bits): the radix
// passes above over the top bucket_bits of (key - lo) only, then one pass in
// which every item ranks itself among its bucket's items by the whole key.
// A bucket of more than 64 items is reported to the caller, which then sorts
// on all bits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "scan_sort.h"

namespace dvsamd {
namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;
constexpr int kWaves = kThreads / 64;
constexpr int kDigits = 256;

__device__ inline int wave_inclusive(int v) {
  const int lane = __lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(v, off, 64);
    if (lane >= off) v += y;
  }
  return v;
}

// Exclusive prefix of v over the block (256 threads) and the block total.
__device__ inline int block_exclusive(int v, int *wave_sums, int &total) {
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  const int incl = wave_inclusive(v);
  if (lane == 63) wave_sums[wave] = incl;
  __syncthreads();
  int before = 0;
  total = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    const int s = wave_sums[w];
    before += w < wave ? s : 0;
    total += s;
  }
  __syncthreads();  // wave_sums is reused by the caller's next call
  return before + incl - v;
}

__global__ __launch_bounds__(kThreads) void tile_sums_kernel(const int *in, long long n, int *sums) {
  __shared__ int wave_sums[kWaves];
  const long long base = (long long)blockIdx.x * kTile + (long long)threadIdx.x * kItems;
  int s = 0;
#pragma unroll
  for (int j = 0; j < kItems; ++j) s += base + j < n ? in[base + j] : 0;
  int total;
  block_exclusive(s, wave_sums, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// One block of 256 threads: exclusive prefix of the tile sums, in place.
__global__ __launch_bounds__(kThreads) void scan_sums_kernel(int *sums, int count) {
  __shared__ int wave_sums[kWaves];
  int carry = 0;
  for (int at = 0; at < count; at += kThreads) {
    const int i = at + (int)threadIdx.x;
    const int v = i < count ? sums[i] : 0;
    int total;
    const int ex = block_exclusive(v, wave_sums, total);
    if (i < count) sums[i] = carry + ex;
    carry += total;
  }
}

__global__ __launch_bounds__(kThreads) void tile_scan_kernel(const int *in, int *out, long long n,
                                                             const int *prefix) {
  __shared__ int wave_sums[kWaves];
  const long long base = (long long)blockIdx.x * kTile + (long long)threadIdx.x * kItems;
  int v[kItems];
  int run = 0;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    v[j] = base + j < n ? in[base + j] : 0;
    run += v[j];
    v[j] = run;
  }
  int total;
  const int before = block_exclusive(run, wave_sums, total) + (prefix ? prefix[blockIdx.x] : 0);
#pragma unroll
  for (int j = 0; j < kItems; ++j)
    if (base + j < n) out[base + j] = before + v[j];
}

// The digit of a pass: bits [shift, shift + 8) clipped to the sort's end bit
// (digit_mask), so that bits at or above end_bit never take part.  A
// bucketed sort sorts by (key - sub) >> pre instead of the key itself.
struct KeyView {
  unsigned long long sub;
  int pre;
};
__device__ inline unsigned digit_of(unsigned long long key, int shift, unsigned digit_mask, KeyView kv) {
  return (unsigned)(((key - kv.sub) >> kv.pre) >> shift) & digit_mask;
}

// hist[d * tiles + tile] = number of keys of the tile with digit d.
__global__ __launch_bounds__(kThreads) void digit_hist_kernel(const unsigned long long *keys, long long n, int shift,
                                                              unsigned digit_mask, KeyView kv, int tiles, int *hist) {
  __shared__ int counts[kDigits];
  counts[threadIdx.x] = 0;
  __syncthreads();
  const long long base = (long long)blockIdx.x * kTile;
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const long long i = base + r * kThreads + threadIdx.x;
    if (i < n) atomicAdd(&counts[digit_of(keys[i], shift, digit_mask, kv)], 1);
  }
  __syncthreads();
  hist[(long long)threadIdx.x * tiles + blockIdx.x] = counts[threadIdx.x];
}

// Stable scatter of one digit pass; `incl` is the inclusive digit-major scan
// of `hist`.
__global__ __launch_bounds__(kThreads) void digit_scatter_kernel(const unsigned long long *keys_in,
                                                                 const unsigned *vals_in, long long n, int shift,
                                                                 unsigned digit_mask, KeyView kv, int tiles,
                                                                 const int *hist, const int *incl,
                                                                 unsigned long long *keys_out, unsigned *vals_out) {
  __shared__ int base[kDigits];
  __shared__ int run[kDigits];
  __shared__ int wcount[kWaves][kDigits];
  const int t = threadIdx.x, lane = __lane_id(), wave = t >> 6;
  {
    const long long h = (long long)t * tiles + blockIdx.x;
    base[t] = incl[h] - hist[h];
    run[t] = 0;
  }
  const unsigned long long below = (1ull << lane) - 1;
  const long long tile = (long long)blockIdx.x * kTile;
  for (int r = 0; r < kItems; ++r) {
#pragma unroll
    for (int w = 0; w < kWaves; ++w) wcount[w][t] = 0;
    __syncthreads();
    const long long i = tile + r * kThreads + t;
    const bool valid = i < n;
    const unsigned long long key = valid ? keys_in[i] : 0ull;
    const unsigned val = valid ? vals_in[i] : 0u;
    const unsigned d = digit_of(key, shift, digit_mask, kv);
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const unsigned long long m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const int rank = __popcll(peers & below);
    if (valid && rank == 0) wcount[wave][d] = __popcll(peers);
    __syncthreads();
    {  // thread t owns digit t: per-wave offsets of this round, then the carry
      int acc = run[t];
#pragma unroll
      for (int w = 0; w < kWaves; ++w) {
        const int c = wcount[w][t];
        wcount[w][t] = acc;
        acc += c;
      }
      run[t] = acc;
    }
    __syncthreads();
    if (valid) {
      const long long pos = (long long)base[d] + wcount[wave][d] + rank;
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads();
  }
}

__global__ void copy_pairs_kernel(const unsigned long long *keys_in, const unsigned *vals_in, long long n,
                                  unsigned long long *keys_out, unsigned *vals_out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    keys_out[i] = keys_in[i];
    vals_out[i] = vals_in[i];
  }
}

// After a stable sort by bucket b(key) = (key - sub) >> pre: every item
// finds its bucket's extent (the buckets are contiguous and ascending: a
// short walk, then a binary search, at most kBucketCap items either side),
// ranks itself among the bucket's items by (key, position) and lands at the
// bucket's start plus that rank, out of place.  Equal keys keep their order,
// so the output is the stable sort by the whole key.  An item of a bucket
// of more than kBucketCap items is copied in place and sets *flag (the
// caller then sorts on all key bits).
constexpr int kBucketCap = 1024;
constexpr int kBucketWalk = 8;  // linear steps before the binary search
__global__ __launch_bounds__(256) void bucket_rank_kernel(const unsigned long long *keys_in, const unsigned *vals_in,
                                                          long long n, KeyView kv, unsigned long long *keys_out,
                                                          unsigned *vals_out, int *flag) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long key = keys_in[i];
  auto bucket = [&](long long j) { return (keys_in[j] - kv.sub) >> kv.pre; };
  const unsigned long long b = bucket(i);
  // first index s of the bucket within [i - kBucketCap, i]
  long long s = i;
  for (int k = 0; k < kBucketWalk && s > 0 && i - s < kBucketCap && bucket(s - 1) == b; ++k) --s;
  if (s > 0 && i - s < kBucketCap && bucket(s - 1) == b) {
    long long lo = i - kBucketCap > 0 ? i - kBucketCap : 0, hi = s - 1;  // bucket(hi) == b
    while (lo < hi) {
      const long long m = lo + (hi - lo) / 2;
      if (bucket(m) == b) hi = m; else lo = m + 1;
    }
    s = lo;
  }
  // one past the last index e of the bucket within [i + 1, s + kBucketCap + 1]
  long long e = i + 1;
  for (int k = 0; k < kBucketWalk && e < n && e - s <= kBucketCap && bucket(e) == b; ++k) ++e;
  if (e < n && e - s <= kBucketCap && bucket(e) == b) {
    long long lo = e, hi = s + kBucketCap + 1 < n ? s + kBucketCap + 1 : n;  // bucket(lo) == b
    while (hi - lo > 1) {  // invariant: bucket(lo) == b, and hi == n or bucket(hi) != b or hi is past the cap
      const long long m = lo + (hi - lo) / 2;
      if (bucket(m) == b) lo = m; else hi = m;
    }
    e = lo + 1;
  }
  const bool over = (s > 0 && bucket(s - 1) == b) || e - s > kBucketCap || (e < n && bucket(e) == b);
  long long at = i;
  if (over) {
    atomicOr(flag, 1);
  } else {
    at = s;
    for (long long j = s; j < e; ++j) {
      const unsigned long long k = keys_in[j];
      at += (k < key || (k == key && j < i)) ? 1 : 0;
    }
  }
  keys_out[at] = key;
  vals_out[at] = vals_in[i];
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
inline long long tiles_of(long long n) { return (n + kTile - 1) / kTile; }

// Launch sequence of an inclusive scan with `sums` scratch (tiles ints).
hipError_t scan_launch(const int *in, int *out, long long n, int *sums, hipStream_t st) {
  const long long tiles = tiles_of(n);
  if (tiles == 1) {
    tile_scan_kernel<<<1, kThreads, 0, st>>>(in, out, n, nullptr);
    return hipGetLastError();
  }
  tile_sums_kernel<<<(unsigned)tiles, kThreads, 0, st>>>(in, n, sums);
  scan_sums_kernel<<<1, kThreads, 0, st>>>(sums, (int)tiles);
  tile_scan_kernel<<<(unsigned)tiles, kThreads, 0, st>>>(in, out, n, sums);
  return hipGetLastError();
}

}  // namespace

hipError_t inclusive_scan_int(void *temp, size_t &temp_bytes, const int *in, int *out, size_t n, hipStream_t st) {
  const long long tiles = tiles_of((long long)n);
  const size_t need = align256(sizeof(int) * (size_t)std::max<long long>(tiles, 1));
  if (!temp) {
    temp_bytes = need;
    return hipSuccess;
  }
  if (temp_bytes < need) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  if (tiles > 0x7FFFFFFFLL) return hipErrorInvalidValue;
  return scan_launch(in, out, (long long)n, static_cast<int *>(temp), st);
}

namespace {
hipError_t sort_passes(void *temp, size_t &temp_bytes, const unsigned long long *keys_in, unsigned long long *keys_out,
                       const unsigned *values_in, unsigned *values_out, size_t n, int begin_bit, int end_bit,
                       KeyView kv, hipStream_t st) {
  const long long tiles = std::max<long long>(tiles_of((long long)n), 1);
  const size_t hist_n = (size_t)kDigits * tiles;
  const size_t b_keys = align256(sizeof(unsigned long long) * std::max<size_t>(n, 1));
  const size_t b_vals = align256(sizeof(unsigned) * std::max<size_t>(n, 1));
  const size_t b_hist = align256(sizeof(int) * hist_n);
  const size_t b_sums = align256(sizeof(int) * (size_t)tiles_of((long long)hist_n));
  const size_t need = b_keys + b_vals + 2 * b_hist + b_sums;
  if (!temp) {
    temp_bytes = need;
    return hipSuccess;
  }
  if (temp_bytes < need || begin_bit < 0 || end_bit > 64 || begin_bit > end_bit) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  char *p = static_cast<char *>(temp);
  unsigned long long *tmp_keys = reinterpret_cast<unsigned long long *>(p);
  unsigned *tmp_vals = reinterpret_cast<unsigned *>(p + b_keys);
  int *hist = reinterpret_cast<int *>(p + b_keys + b_vals);
  int *incl = reinterpret_cast<int *>(p + b_keys + b_vals + b_hist);
  int *sums = reinterpret_cast<int *>(p + b_keys + b_vals + 2 * b_hist);
  const long long count = (long long)n;
  const int passes = (end_bit - begin_bit + 7) / 8;
  if (passes == 0) {
    copy_pairs_kernel<<<(unsigned)((count + 255) / 256), 256, 0, st>>>(keys_in, values_in, count, keys_out,
                                                                        values_out);
    return hipGetLastError();
  }
  // Ping-pong so that the last pass writes keys_out / values_out.
  const unsigned long long *src_k = keys_in;
  const unsigned *src_v = values_in;
  for (int pass = 0; pass < passes; ++pass) {
    const bool to_out = ((passes - 1 - pass) % 2) == 0;
    unsigned long long *dst_k = to_out ? keys_out : tmp_keys;
    unsigned *dst_v = to_out ? values_out : tmp_vals;
    const int shift = begin_bit + 8 * pass;
    const int width = std::min(8, end_bit - shift);
    const unsigned digit_mask = (1u << width) - 1u;
    digit_hist_kernel<<<(unsigned)tiles, kThreads, 0, st>>>(src_k, count, shift, digit_mask, kv, (int)tiles, hist);
    hipError_t e = scan_launch(hist, incl, (long long)hist_n, sums, st);
    if (e != hipSuccess) return e;
    digit_scatter_kernel<<<(unsigned)tiles, kThreads, 0, st>>>(src_k, src_v, count, shift, digit_mask, kv,
                                                               (int)tiles, hist, incl, dst_k, dst_v);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    src_k = dst_k;
    src_v = dst_v;
  }
  return hipSuccess;
}
}  // namespace

hipError_t sort_pairs_u64(void *temp, size_t &temp_bytes, const unsigned long long *keys_in,
                          unsigned long long *keys_out, const unsigned *values_in, unsigned *values_out, size_t n,
                          int begin_bit, int end_bit, hipStream_t st) {
  return sort_passes(temp, temp_bytes, keys_in, keys_out, values_in, values_out, n, begin_bit, end_bit,
                     KeyView{0, 0}, st);
}

hipError_t sort_pairs_u64_bucketed(void *temp, size_t &temp_bytes, const unsigned long long *keys_in,
                                   unsigned long long *keys_out, const unsigned *values_in, unsigned *values_out,
                                   size_t n, unsigned long long key_lo, int key_bits, int bucket_bits, int *flag,
                                   hipStream_t st) {
  if (key_bits < 0 || key_bits > 64 || bucket_bits <= 0 || bucket_bits > 64) return hipErrorInvalidValue;
  const int pre = key_bits > bucket_bits ? key_bits - bucket_bits : 0;
  const KeyView kv{key_lo, pre};
  const int bits = std::min(bucket_bits, key_bits);
  // [bucket-sorted keys][their values][the radix passes' scratch]
  const size_t b_keys = align256(sizeof(unsigned long long) * std::max<size_t>(n, 1));
  const size_t b_vals = align256(sizeof(unsigned) * std::max<size_t>(n, 1));
  size_t pass_bytes = 0;
  hipError_t e = sort_passes(nullptr, pass_bytes, keys_in, keys_out, values_in, values_out, n, 0, bits, kv, st);
  if (e != hipSuccess) return e;
  if (!temp) {
    temp_bytes = b_keys + b_vals + pass_bytes;
    return hipSuccess;
  }
  if (temp_bytes < b_keys + b_vals + pass_bytes || !flag) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  if (pre == 0)  // the bucket is the whole key
    return sort_passes(static_cast<char *>(temp) + b_keys + b_vals, pass_bytes, keys_in, keys_out, values_in,
                       values_out, n, 0, bits, kv, st);
  char *p = static_cast<char *>(temp);
  unsigned long long *bkeys = reinterpret_cast<unsigned long long *>(p);
  unsigned *bvals = reinterpret_cast<unsigned *>(p + b_keys);
  e = sort_passes(p + b_keys + b_vals, pass_bytes, keys_in, bkeys, values_in, bvals, n, 0, bits, kv, st);
  if (e != hipSuccess) return e;
  bucket_rank_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(bkeys, bvals, (long long)n, kv, keys_out,
                                                                   values_out, flag);
  return hipGetLastError();
}

}  // namespace dvsamd
