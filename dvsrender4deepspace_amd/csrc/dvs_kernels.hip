// gfx950 kernels of the DVS rendering path.
//
//  trace_kernel<MODE>  one lane per sub-ray; 256-lane blocks hold 256/ss^2
//                      pixel samples; per-lane traversal stack in LDS; ordered
//                      FP64 reduction of each sample's sub-rays in LDS.
//                      MODE_CELLS   pixel x circuit-cell radiance (sensor),
//                      MODE_FRAME   frame radiance + depth,
//                      MODE_DEPTH   event-time depth (primary rays only).
//  *_multi kernels     the engine step over a table of segments (sensors):
//                      trace_cells_multi   pixel x cell radiance per segment,
//                      cell_rates_multi    radiance -> photocurrent with the
//                                          cell's Philox shot-noise draw,
//                      circuit_multi       one lane per pixel: FP64 receptor
//                                          integration, comparator crossings,
//                                          Philox shot/reset noise,
//                      gather/pose/depth/permute for device event compaction.
#include <hip/hip_runtime.h>

#include "common.h"
#include "device_circuit.hpp"
#include "device_optics.hpp"
#include "kernels.h"

namespace dvsdev {

// optical_state_at at each cell midpoint (device_optics.cuh:13-26 as called
// by generate_pixels, src/cuda/event_sensor.cu:41).
__global__ void cell_pose_kernel(int cells, const double *t0, const double *t1, const int *iv_index,
                                 const dvs_optical_interval *ivs, Pose *poses) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cells) return;
  poses[c] = pose_at(ivs[iv_index[c]], (t0[c] + t1[c]) * .5);
}

__global__ void state_pose_kernel(int n, const dvs_optical_state *states, Pose *poses) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) poses[i] = pose_of(states[i]);
}

template <int MODE, bool WAVE>
__global__ __launch_bounds__(256) void trace_kernel(DevScene s, const Pose *poses, int cells, int npix,
                                                    long long items, const int *item_pixel, double *rad,
                                                    float *img, float *dep, int *status,
                                                    unsigned long long *shadow_count) {
  // LDS: each lane's traversal stack occupies words [k*256 + lane]; once the
  // lane's sub-ray is done, its first three stack words hold its radiance
  // (two halves of the double) and depth for the ordered per-sample sum.
  extern __shared__ double lds_words[];
  int *words = reinterpret_cast<int *>(lds_words);
  int *stack = words + threadIdx.x;
  const int R = s.samples * s.samples, slots = 256 / R;
  const int slot = threadIdx.x / R, sub = threadIdx.x - slot * R;
  const long long q = (long long)blockIdx.x * slots + slot;
  const bool active = slot < slots && q < items;
  SubSample r{0.0, -1.f, 0, 0};
  long long out_index = 0;
  int pixel = 0, pose_index = 0;
  if (active) {
    if (MODE == MODE_CELLS) {
      pixel = (int)(q / cells);
      pose_index = (int)(q - (long long)pixel * cells);
      out_index = (long long)pose_index * npix + pixel;
    } else if (MODE == MODE_FRAME) {
      pose_index = (int)(q / npix);
      pixel = (int)(q - (long long)pose_index * npix);
      out_index = q;
    } else {
      pixel = item_pixel[q];
      pose_index = (int)q;
      out_index = q;
    }
  }
  if (WAVE) {
    const Pose pose = poses[pose_index];
    r = sample_subray_wave<MODE != MODE_DEPTH>(s, pixel, pose, sub % s.samples, sub / s.samples, active);
  } else if (active) {
    const Pose pose = poses[pose_index];
    r = sample_subray<MODE != MODE_DEPTH>(s, pixel, pose, sub % s.samples, sub / s.samples, stack, 256);
  }
  stack[0] = __double2loint(r.radiance);
  stack[256] = __double2hiint(r.radiance);
  stack[512] = __float_as_int(r.depth);
  // shadow rays counted per block: one atomic on the shared counter per
  // block instead of per wave
  __shared__ int block_shadows[4];
  const unsigned long long shadows = __ballot(r.shadow);
  const unsigned long long invalid = __ballot(r.invalid);
  if ((threadIdx.x & 63) == 0) {
    block_shadows[threadIdx.x >> 6] = __popcll(shadows);
    if (invalid) atomicOr(status, 2);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int n = block_shadows[0] + block_shadows[1] + block_shadows[2] + block_shadows[3];
    if (n) atomicAdd(shadow_count, (unsigned long long)n);
  }
  if (!active || sub != 0) return;
  // Sequential sum in sub-ray order j = sy*ss + sx (sample_pixel,
  // device_optics.cuh:75-93; sample_warps, optical_sampler.cu:19-21).
  double total = 0, ranges = 0;
  int hits = 0;
  for (int j = 0; j < R; ++j) {
    const int lane = threadIdx.x + j;
    total += __hiloint2double(words[256 + lane], words[lane]);
    const float d = __int_as_float(words[512 + lane]);
    if (d >= 0) { ranges += d; ++hits; }
  }
  const double mean = total / R;
  const float depth = hits ? (float)(ranges / hits) : -1.f;
  if (MODE == MODE_CELLS) {
    rad[out_index] = mean;
  } else if (MODE == MODE_FRAME) {
    if (rad) rad[out_index] = mean;
    if (img) img[out_index] = (float)mean;
    if (dep) dep[out_index] = depth;
  } else {
    dep[out_index] = depth;
  }
}

// Settled receptor and mismatch thresholds (initialize_pixels,
// src/cuda/event_sensor.cu:18-28).
__global__ void init_kernel(int npix, const double *radiance, PhotoParams photo, CircuitParams p,
                            double threshold_std, double threshold_min, uint64_t seed,
                            dvs_pixel_state *states, int *status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix) return;
  const double r = radiance[i];
  if (!isfinite(r)) { atomicOr(status, 2); return; }
  dvs_pixel_state s;
  s.receptor_log = s.reference_log = log(fmax(1e-30, r * photo.electrons_per_radiance + photo.dark_electrons));
  s.receptor_inverse = exp(-s.receptor_log);
  s.refractory_until = -1;
  double zp = 0, zn = 0;
  if (threshold_std != 0) {
    NoiseStream rs(seed, (uint64_t)i, kStreamThreshold, 0);
    zp = rs.normal();
    zn = rs.normal();
  }
  s.threshold_pos = fmax(threshold_min, p.threshold_pos + threshold_std * zp);
  s.threshold_neg = fmax(threshold_min, p.threshold_neg + threshold_std * zn);
  s.reset_pending = 0;
  s.reset_block = 0;
  s.shot_block = 0;
  states[i] = s;
}


// ---------------------------------------------------------------------------
// Multi-segment engine step
// ---------------------------------------------------------------------------

// Radiance of every (pixel, circuit cell) of every segment: the optical half
// of generate_pixels (src/cuda/event_sensor.cu:39-42).  Items of one block
// are consecutive cells of one pixel (temporally coherent rays).
template <bool WAVE>
__global__ __launch_bounds__(256) void trace_cells_multi(const SegDev *segs, const long long *block_begin, int nseg,
                                                         int *status, unsigned long long *shadow_count) {
  extern __shared__ double lds_words[];
  int *words = reinterpret_cast<int *>(lds_words);
  int *stack = words + threadIdx.x;
  const int si = segment_of<long long>(block_begin, nseg, (long long)blockIdx.x);
  const SegDev &g = segs[si];
  const DevScene &s = g.scene;
  const int R = s.samples * s.samples, slots = 256 / R;
  const int slot = threadIdx.x / R, sub = threadIdx.x - slot * R;
  const long long q = ((long long)blockIdx.x - block_begin[si]) * slots + slot;
  const int cells = g.cells, npix = g.npix;
  const bool active = slot < slots && q < (long long)npix * cells;
  SubSample r{0.0, -1.f, 0, 0};
  int pixel = 0, cell = 0, px = 0;
  if (active) {
    pixel = (int)(q / cells);
    cell = (int)(q - (long long)pixel * cells);
    px = g.pixel_list ? g.pixel_list[pixel] : pixel;
  }
  // a block's lanes belong to one segment: the pose of cell 0 stands in for
  // the lanes without a sample
  if (WAVE)
    r = sample_subray_wave<true>(s, px, g.poses[cell], sub % s.samples, sub / s.samples, active);
  else if (active)
    r = sample_subray<true>(s, px, g.poses[cell], sub % s.samples, sub / s.samples, stack, 256);
  const LdsInts lw = (LdsInts)words;  // the lanes' radiances, low then high words
  lw[threadIdx.x] = __double2loint(r.radiance);
  lw[256 + threadIdx.x] = __double2hiint(r.radiance);
  // shadow rays counted per block: one atomic on the shared counter per
  // block instead of per wave
  __shared__ int block_shadows[4];
  const unsigned long long shadows = __ballot(r.shadow);
  const unsigned long long invalid = __ballot(r.invalid);
  if ((threadIdx.x & 63) == 0) {
    block_shadows[threadIdx.x >> 6] = __popcll(shadows);
    if (invalid) atomicOr(status, 2);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int n = block_shadows[0] + block_shadows[1] + block_shadows[2] + block_shadows[3];
    if (n) atomicAdd(shadow_count, (unsigned long long)n);
  }
  if (!active || sub != 0) return;
  double total = 0;
  for (int j = 0; j < R; ++j) total += __hiloint2double(lw[256 + threadIdx.x + j], lw[threadIdx.x + j]);
  g.rad[(size_t)cell * npix + pixel] = total / R;
}

// Radiance -> photocurrent in place for every (item, cell) of a chunk table:
// x over the items (pixels, or pixel-list entries) of all segments, y over
// the chunk's cells.  The shot-noise draw of a cell uses the pixel's cell
// serial (cells integrated so far, advanced by the circuit after each chunk)
// plus the cell's index in the chunk.
__global__ __launch_bounds__(256) void cell_rates_multi(const SegDev *segs, const int *item_begin, int nseg,
                                                        int items, int *status, int *rare) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  const int c = (int)blockIdx.y;
  bool marked = false;
  if (k < items) {
    const int si = segment_of<int>(item_begin, nseg, k);
    const SegDev &g = segs[si];
    const int item = k - item_begin[si];
    if (c < g.cells && item < g.npix) {
      const int px = g.pixel_list ? g.pixel_list[item] : item;
      double *cell = g.rad + (size_t)c * g.npix + item;
      const double light = *cell;
      if (!isfinite(light)) atomicOr(status, 2);
      const uint64_t serial = g.states[px].shot_block + (uint64_t)c;
      uint64_t shared[4] = {0, 0, 0, 0};
      if (g.photo.shot_noise) shot_block(g.seed, (uint64_t)px, serial / kShotCellsPerBlock, shared);
      const CellMaps &m = g.maps[c];
      const double rate = cell_rate_first(light, g.photo, m.t1 - m.t0, m.rate[0], m.rate[1], serial, shared);
      *cell = rate;
      marked = rate < 0;
    }
  }
  // one flag per wave: does any lane leave a draw to cell_rates_rare
  const unsigned long long any = __ballot(marked);
  if ((threadIdx.x & 63) == 0) rare[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 4 + (threadIdx.x >> 6)] = any != 0;
}

// Finishes the cells cell_rates_multi marked, in the waves that flagged any.
__global__ __launch_bounds__(256) void cell_rates_rare(const SegDev *segs, const int *item_begin, int nseg, int items,
                                                       const int *rare) {
  if (!rare[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 4 + (threadIdx.x >> 6)]) return;
  const int k = blockIdx.x * 256 + threadIdx.x;
  const int c = (int)blockIdx.y;
  if (k >= items) return;
  const int si = segment_of<int>(item_begin, nseg, k);
  const SegDev &g = segs[si];
  const int item = k - item_begin[si];
  if (c >= g.cells || item >= g.npix) return;
  const int px = g.pixel_list ? g.pixel_list[item] : item;
  double *cell = g.rad + (size_t)c * g.npix + item;
  const double v = *cell;
  if (v < 0) *cell = cell_rate_rest(v, g.t1[c] - g.t0[c], g.seed, (uint64_t)px, g.states[px].shot_block + (uint64_t)c);
}

// The circuit half of generate_pixels (event_sensor.cu:39-49): one lane per
// pixel integrates all cells of this step in order.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void circuit_multi(const SegDev *segs, const int *block_begin, int nseg,
                                                     int *status) {
  const int si = segment_of<int>(block_begin, nseg, (int)blockIdx.x);
  const SegDev &g = segs[si];
  const int i = ((int)blockIdx.x - block_begin[si]) * 256 + threadIdx.x;
  // lanes past the segment's pixels stay in the cell loop (they hold cells of
  // the wave's shared tiles) without a pixel of their own
  const bool own = i < g.npix;
  dvs_pixel_state st = own ? g.states[i] : dvs_pixel_state{};
  const int count = circuit_cells<false>(st, g.params, g.seed, (uint64_t)i, own, g.rad + (own ? i : 0),
                                         (size_t)g.npix, nullptr, 0, false, g.t0, g.t1, g.maps, 0, g.cells, g.slots + (size_t)(own ? i : 0) * g.max_events,
                                  own ? g.counts[i] : 0, g.max_events);
  if (!own) return;
  if (count < 0) { atomicOr(status, count == -1 ? 1 : 2); g.counts[i] = 0; return; }
  st.shot_block += (uint64_t)g.cells;
  g.states[i] = st;
  g.counts[i] = count;
}

__global__ void cell_maps_kernel(int cells, const double *t0, const double *t1, const int *cell_begin, int nseg,
                                 const double *charge, CellMaps *out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cells) return;
  out[c] = cell_maps(t0[c], t1[c], charge[segment_of<int>(cell_begin, nseg, c)]);
}

// Compaction: global pixel p's crossings go to [ends[p] - count, ends[p]);
// the global pixel order is segment-major, pixel-major within a segment.
// Compaction of every pixel's crossings to its place in the prefix sum
// `ends`, with the order-preserving key of its time; a key outside
// [key_lo, key_hi] (a time outside the step's span) sets *key_flag.
__global__ void gather_multi(const SegDev *segs, const int *pixel_begin, int nseg, int total_pixels, const int *ends,
                             unsigned long long key_lo, unsigned long long key_hi, double *time, int *seg,
                             int *pixel, int *polarity, unsigned long long *keys, unsigned *order, int *key_flag) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= total_pixels) return;
  const int si = segment_of<int>(pixel_begin, nseg, p);
  const SegDev &g = segs[si];
  const int i = p - pixel_begin[si];
  const int n = g.counts[i], base = ends[p] - n;
  for (int j = 0; j < n; ++j) {
    const CrossingEvent e = g.slots[(size_t)i * g.max_events + j];
    const int k = base + j;
    time[k] = e.time;
    seg[k] = si;
    pixel[k] = i;
    polarity[k] = e.polarity;
    // order-preserving key of an IEEE double
    const unsigned long long bits = (unsigned long long)__double_as_longlong(e.time);
    const unsigned long long key = (bits >> 63) ? ~bits : (bits | 0x8000000000000000ULL);
    keys[k] = key;
    order[k] = (unsigned)k;
    if (key < key_lo || key > key_hi) atomicOr(key_flag, 1);
  }
}

// Pose at an event: the first interval of the step whose end is not before
// the event (the forward scan of generate_pixels, event_sensor.cu:54).
__global__ void event_poses_multi(const SegDev *segs, int events, const double *time, const int *seg, Pose *poses) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= events) return;
  const SegDev &g = segs[seg[e]];
  const double t = time[e];
  int lo = 0, hi = g.iv_count - 1;
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (t > g.ivs[m].t1) lo = m + 1; else hi = m;
  }
  poses[e] = pose_at(g.ivs[lo], t);
}

// Event-time depth (sample_pixel at the event time, event_sensor.cu:55-58):
// primary rays only, mean hit range of the sub-rays, -1 if none hit.
template <bool WAVE>
__global__ __launch_bounds__(256) void event_depth_multi(const SegDev *segs, int events, int samples, const int *seg,
                                                         const int *pixel, const Pose *poses, float *depth,
                                                         int *status) {
  extern __shared__ double lds_words[];
  int *words = reinterpret_cast<int *>(lds_words);
  int *stack = words + threadIdx.x;
  const int R = samples * samples, slots = 256 / R;
  const int slot = threadIdx.x / R, sub = threadIdx.x - slot * R;
  const long long e = (long long)blockIdx.x * slots + slot;
  const bool active = slot < slots && e < events;
  SubSample r{0.0, -1.f, 0, 0};
  if (WAVE) {
    const long long k = active ? e : 0;  // lanes without an event stand in for event 0
    r = sample_subray_wave<false>(segs[seg[k]].scene, pixel[k], poses[k], sub % samples, sub / samples, active);
  } else if (active) {
    const SegDev &g = segs[seg[e]];
    r = sample_subray<false>(g.scene, pixel[e], poses[e], sub % samples, sub / samples, stack, 256);
  }
  stack[0] = __float_as_int(r.depth);
  if (__ballot(r.invalid) && (threadIdx.x & 63) == 0) atomicOr(status, 2);
  __syncthreads();
  if (!active || sub != 0) return;
  double ranges = 0;
  int hits = 0;
  for (int j = 0; j < R; ++j) {
    const float d = __int_as_float(words[threadIdx.x + j]);
    if (d >= 0) { ranges += d; ++hits; }
  }
  depth[e] = hits ? (float)(ranges / hits) : -1.f;
}

__global__ void segment_keys_kernel(int events, const unsigned *order, const int *seg, unsigned long long *keys) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < events) keys[k] = (unsigned long long)seg[order[k]];
}

__global__ void permute_multi(const SegDev *segs, int events, const unsigned *order, const double *time,
                              const int *seg, const int *pixel, const int *polarity, const float *depth,
                              dvs_event *out, int *out_seg) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= events) return;
  const unsigned j = order[k];
  const int width = segs[seg[j]].scene.width;
  dvs_event e;
  e.timestamp = time[j];
  e.x = (int16_t)(pixel[j] % width);
  e.y = (int16_t)(pixel[j] / width);
  e.polarity = (int8_t)polarity[j];
  e.depth = depth[j];
  out[k] = e;
  out_seg[k] = seg[j];
}

__global__ void fill_int_kernel(int *p, int value, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

// Motion-blur exposure e, pixel p: the cells [offsets[e], offsets[e+1]) of
// `frames` (cell-major, npix floats each) summed in cell order with their
// weights, as bin/cam_blur does (apps/main_cam_blur.cpp:76: float sum of
// frame x float(weight), each product and sum rounded to float) or, with
// `wide`, as bin/cam_blur_windows does (apps/main_cam_blur_windows.cpp:77,81:
// double sum of weight x frame, cast to float at the end).  The products and
// sums are explicitly rounded so that no multiply-add is contracted.
__global__ __launch_bounds__(256) void blur_accumulate_kernel(int npix, const int *offsets, const float *frames,
                                                            const double *weights, int wide, float *out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int e = blockIdx.y;
  if (p >= npix) return;
  const int begin = offsets[e], end = offsets[e + 1];
  if (wide) {
    double acc = 0.0;
    for (int c = begin; c < end; ++c)
      acc = __dadd_rn(acc, __dmul_rn(weights[c], (double)frames[(size_t)c * npix + p]));
    out[(size_t)e * npix + p] = (float)acc;
  } else {
    float acc = 0.0f;
    for (int c = begin; c < end; ++c)
      acc = __fadd_rn(acc, __fmul_rn(frames[(size_t)c * npix + p], (float)weights[c]));
    out[(size_t)e * npix + p] = acc;
  }
}

// Block k moves staged copy k: 16-byte words when source, destination and
// size allow, else 4-byte words, else bytes.
__global__ __launch_bounds__(256) void staged_scatter_kernel(const char *mirror, dvsamd::ScatterArgs a) {
  const int k = blockIdx.x;
  const char *src = mirror + a.src[k];
  char *dst = a.dst[k];
  const unsigned bytes = a.bytes[k];
  const uintptr_t align = (uintptr_t)src | (uintptr_t)dst | (uintptr_t)bytes;
  if (align % 16 == 0) {
    const uint4 *s = reinterpret_cast<const uint4 *>(src);
    uint4 *d = reinterpret_cast<uint4 *>(dst);
    for (unsigned i = threadIdx.x; i < bytes / 16; i += blockDim.x) d[i] = s[i];
  } else if (align % 4 == 0) {
    const unsigned *s = reinterpret_cast<const unsigned *>(src);
    unsigned *d = reinterpret_cast<unsigned *>(dst);
    for (unsigned i = threadIdx.x; i < bytes / 4; i += blockDim.x) d[i] = s[i];
  } else {
    for (unsigned i = threadIdx.x; i < bytes; i += blockDim.x) dst[i] = src[i];
  }
}

// ---- unit kernels -------------------------------------------------------
// The ray kernels' form of the reflectance (hapke_fast), whose constants
// the host computes as for a scene.
__global__ void hapke_kernel(int64_t n, const float *ci, const float *ce, const float *cg, const float *alb,
                             dvs_hapke h, HapkeConsts c, float *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = hapke_fast(ci[i], ce[i], cg[i], alb[i], h, c);
}
__global__ void invert_hapke_kernel(int64_t n, const float *inc, const float *emi, const float *phase,
                                    const float *target, dvs_hapke h, float *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = invert_hapke_if(inc[i], emi[i], phase[i], target[i], h);
}
__global__ void receptor_kernel(int64_t n, const double *u, const double *rate, const double *dt, double charge,
                                double *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = receptor_log_at(u[i], rate[i], dt[i], charge);
}
// advance_pixel of n independent pixels, one lane each; `normal` is the
// value of every reset-noise draw.
__global__ void advance_kernel(int64_t n, dvs_pixel_state *states, const dvs_pixel_params *params,
                               const double *rates, const double *t0, const double *t1, int capacity, double normal,
                               int *counts, dvs_crossing *events) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dvs_pixel_state st = states[i];
  PixelRegs r{st.receptor_log, st.reference_log, st.refractory_until, st.threshold_pos, st.threshold_neg,
              st.reset_pending};
  const dvs_pixel_params q = params[i];
  const CircuitParams p{q.threshold_pos, q.threshold_neg, q.charge_scale, q.refractory_time, q.reference_leak,
                        q.reset_noise_std};
  CrossingEvent *out = reinterpret_cast<CrossingEvent *>(events + i * capacity);
  const int count = advance_pixel(r, p, rates[i], t0[i], t1[i], out, capacity, [&] { return normal; });
  counts[i] = count;
  st.receptor_log = r.receptor_log;
  st.reference_log = r.reference_log;
  st.refractory_until = r.refractory_until;
  st.reset_pending = r.reset_pending;
  st.receptor_inverse = exp(-r.receptor_log);
  states[i] = st;
}
__global__ void poisson_kernel(int64_t n, const double *lambda, const double *u, double *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = poisson_inversion_screened(lambda[i], u[i]);
}
__global__ void noise_kernel(uint64_t seed, int64_t pixels, double lambda, int draws, double *poisson,
                             double *normal) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= pixels) return;
  NoiseStream shot(seed, (uint64_t)i, kStreamShot, 0), reset(seed, (uint64_t)i, kStreamReset, 0);
  for (int d = 0; d < draws; ++d) {
    if (poisson) poisson[i * draws + d] = shot.poisson(lambda);
    if (normal) normal[i * draws + d] = reset.normal();
  }
}
template <bool WAVE>
__global__ __launch_bounds__(256) void rays_kernel(DevScene s, int64_t n, const float *o, const float *d,
                                                   const int *self, int shadow, int *hit, int *tri, float *t,
                                                   float *u, float *v) {
  extern __shared__ double lds_words[];
  int *stack = reinterpret_cast<int *>(lds_words) + threadIdx.x;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (WAVE) {
    const bool active = i < n;
    const int64_t k = active ? i : 0;
    const f3 origin = mk(o[3 * k], o[3 * k + 1], o[3 * k + 2]);
    const f3 dir = mk(d[3 * k], d[3 * k + 1], d[3 * k + 2]);
    const f3 idir = inverse_direction(dir);
    if (shadow) {
      const int me = self[k];
      unsigned long long occluded = 0;
      for_each_geometry(s, __ballot(active), [&](const WaveGeom &g, unsigned long long group) {
        occluded |= wave_any(g, group, origin, dir, idir, me, kRayEps, 1e30f);
      });
      if (active) hit[i] = (occluded & lane_bit()) ? 1 : 0;
      return;
    }
    Hit h{0, -1, -1, 0.f, 0.f, 0.f};
    for_each_geometry(s, __ballot(active), [&](const WaveGeom &g, unsigned long long group) {
      wave_nearest(g, group, origin, dir, idir, h);
    });
    if (active) { hit[i] = h.hit; tri[i] = h.tri; t[i] = h.t; u[i] = h.u; v[i] = h.v; }
    return;
  }
  if (i >= n) return;
  const f3 origin = mk(o[3 * i], o[3 * i + 1], o[3 * i + 2]);
  const f3 dir = mk(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
  NoCount nc;
  if (shadow) {
    hit[i] = trace_any(s, origin, dir, self[i], kRayEps, 1e30f, stack, 256, nc) ? 1 : 0;
    return;
  }
  const Hit h = trace_nearest(s, origin, dir, stack, 256, nc);
  hit[i] = h.hit; tri[i] = h.tri; t[i] = h.t; u[i] = h.u; v[i] = h.v;
}

// Traversal work of every sub-ray of one frame (diagnostic): counts of
// node-pair visits, leaf visits and triangle tests for primary and shadow
// rays, summed per wave then atomically.
__global__ __launch_bounds__(256) void traversal_stats_kernel(DevScene s, const Pose *pose, int npix,
                                                              unsigned long long *totals) {
  extern __shared__ double lds_words[];
  int *stack = reinterpret_cast<int *>(lds_words) + threadIdx.x;
  const int R = s.samples * s.samples;
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  WorkCount cp, cs;
  int shadow = 0, primary = 0;
  if (g < (long long)npix * R) {
    const int pixel = (int)(g / R), sub = (int)(g - (long long)pixel * R);
    const SubSample r = sample_subray<true, WorkCount>(s, pixel, *pose, sub % s.samples, sub / s.samples, stack, 256,
                                                       &cp, &cs);
    shadow = r.shadow;
    primary = 1;
  }
  unsigned long long v[8] = {(unsigned long long)primary, cp.pairs, cp.leaves, cp.triangles,
                             (unsigned long long)shadow, cs.pairs, cs.leaves, cs.triangles};
  for (int k = 0; k < 8; ++k) {
    unsigned long long x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off);
    if ((threadIdx.x & 63) == 0 && x) atomicAdd(totals + k, x);
  }
}

}  // namespace dvsdev

// ---- launchers (host side of this translation unit) ----------------------
namespace dvsamd {
using namespace dvsdev;

static size_t trace_lds_bytes(const DevScene &s) {
  // per lane: max(stack depth, 3 words of the radiance/depth hand-off)
  return (size_t)(s.stack_depth < 3 ? 3 : s.stack_depth) * 256 * sizeof(int);
}

hipError_t launch_cell_poses(int cells, const double *t0, const double *t1, const int *iv_index,
                             const dvs_optical_interval *ivs, Pose *poses, hipStream_t st) {
  if (cells <= 0) return hipSuccess;
  cell_pose_kernel<<<(cells + 255) / 256, 256, 0, st>>>(cells, t0, t1, iv_index, ivs, poses);
  return hipGetLastError();
}
hipError_t launch_state_poses(int n, const dvs_optical_state *states, Pose *poses, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  state_pose_kernel<<<(n + 255) / 256, 256, 0, st>>>(n, states, poses);
  return hipGetLastError();
}
hipError_t launch_trace(int mode, const DevScene &s, const Pose *poses, int cells, int npix, long long items,
                        const int *item_pixel, double *rad, float *img, float *dep, int *status,
                        unsigned long long *shadow_count, hipStream_t st) {
  if (items <= 0) return hipSuccess;
  const int R = s.samples * s.samples, slots = 256 / R;
  const long long blocks = (items + slots - 1) / slots;
  if (blocks > kMaxBlocks256) return hipErrorInvalidConfiguration;
  // wave packets: only the three words of each lane's hand-off
  const size_t lds = s.packet ? 3 * 256 * sizeof(int) : trace_lds_bytes(s);
  const unsigned b = (unsigned)blocks;
  switch (mode * 2 + (s.packet ? 1 : 0)) {
    case MODE_CELLS * 2:
      trace_kernel<MODE_CELLS, false><<<b, 256, lds, st>>>(s, poses, cells, npix, items, item_pixel, rad, img, dep,
                                                          status, shadow_count);
      break;
    case MODE_CELLS * 2 + 1:
      trace_kernel<MODE_CELLS, true><<<b, 256, lds, st>>>(s, poses, cells, npix, items, item_pixel, rad, img, dep,
                                                         status, shadow_count);
      break;
    case MODE_FRAME * 2:
      trace_kernel<MODE_FRAME, false><<<b, 256, lds, st>>>(s, poses, cells, npix, items, item_pixel, rad, img, dep,
                                                          status, shadow_count);
      break;
    case MODE_FRAME * 2 + 1:
      trace_kernel<MODE_FRAME, true><<<b, 256, lds, st>>>(s, poses, cells, npix, items, item_pixel, rad, img, dep,
                                                         status, shadow_count);
      break;
    case MODE_DEPTH * 2:
      trace_kernel<MODE_DEPTH, false><<<b, 256, lds, st>>>(s, poses, cells, npix, items, item_pixel, rad, img, dep,
                                                          status, shadow_count);
      break;
    default:
      trace_kernel<MODE_DEPTH, true><<<b, 256, lds, st>>>(s, poses, cells, npix, items, item_pixel, rad, img, dep,
                                                         status, shadow_count);
  }
  return hipGetLastError();
}
hipError_t launch_init(int npix, const double *radiance, const PhotoParams &photo, const CircuitParams &p,
                       double threshold_std, double threshold_min, uint64_t seed, dvs_pixel_state *states,
                       int *status, hipStream_t st) {
  init_kernel<<<(npix + 255) / 256, 256, 0, st>>>(npix, radiance, photo, p, threshold_std, threshold_min, seed,
                                                  states, status);
  return hipGetLastError();
}
hipError_t launch_trace_cells_multi(const SegDev *segs, const long long *block_begin, int nseg, long long blocks,
                                    size_t lds, int *status, unsigned long long *shadow_count, hipStream_t st,
                                    bool wave) {
  if (blocks <= 0) return hipSuccess;
  if (blocks > kMaxBlocks256) return hipErrorInvalidConfiguration;
  if (wave)  // the two words of each lane's radiance hand-off
    trace_cells_multi<true><<<(unsigned)blocks, 256, 2 * 256 * sizeof(int), st>>>(segs, block_begin, nseg, status,
                                                                                  shadow_count);
  else
    trace_cells_multi<false><<<(unsigned)blocks, 256, lds, st>>>(segs, block_begin, nseg, status, shadow_count);
  return hipGetLastError();
}
hipError_t launch_cell_maps(int cells, const double *t0, const double *t1, const int *cell_begin, int nseg,
                            const double *charge, CellMaps *out, hipStream_t st) {
  if (cells <= 0) return hipSuccess;
  cell_maps_kernel<<<(unsigned)((cells + 255) / 256), 256, 0, st>>>(cells, t0, t1, cell_begin, nseg, charge, out);
  return hipGetLastError();
}
size_t cell_rates_flags(int items, int cells) {
  return items <= 0 || cells <= 0 ? 0 : (size_t)((items + 255) / 256) * 4 * (size_t)cells;
}
hipError_t launch_cell_rates_multi(const SegDev *segs, const int *item_begin, int nseg, int items, int cells,
                                   int *status, int *rare, size_t rare_capacity, hipStream_t st) {
  if (items <= 0 || cells <= 0) return hipSuccess;
  if (cells > 65535 || cell_rates_flags(items, cells) > rare_capacity) return hipErrorInvalidConfiguration;
  const dim3 grid((unsigned)((items + 255) / 256), (unsigned)cells);
  cell_rates_multi<<<grid, 256, 0, st>>>(segs, item_begin, nseg, items, status, rare);
  cell_rates_rare<<<grid, 256, 0, st>>>(segs, item_begin, nseg, items, rare);
  return hipGetLastError();
}
hipError_t launch_circuit_multi(const SegDev *segs, const int *block_begin, int nseg, int blocks, int *status,
                                hipStream_t st) {
  if (blocks <= 0) return hipSuccess;
  circuit_multi<<<blocks, 256, 0, st>>>(segs, block_begin, nseg, status);
  return hipGetLastError();
}
hipError_t launch_gather_multi(const SegDev *segs, const int *pixel_begin, int nseg, int total_pixels,
                               const int *ends, unsigned long long key_lo, unsigned long long key_hi, double *time,
                               int *seg, int *pixel, int *polarity, unsigned long long *keys, unsigned *order,
                               int *key_flag, hipStream_t st) {
  if (total_pixels <= 0) return hipSuccess;
  gather_multi<<<(total_pixels + 255) / 256, 256, 0, st>>>(segs, pixel_begin, nseg, total_pixels, ends, key_lo,
                                                           key_hi, time, seg, pixel, polarity, keys, order, key_flag);
  return hipGetLastError();
}
hipError_t launch_event_poses_multi(const SegDev *segs, int events, const double *time, const int *seg,
                                    Pose *poses, hipStream_t st) {
  if (events <= 0) return hipSuccess;
  event_poses_multi<<<(events + 255) / 256, 256, 0, st>>>(segs, events, time, seg, poses);
  return hipGetLastError();
}
hipError_t launch_event_depth_multi(const SegDev *segs, int events, int samples, size_t lds, const int *seg,
                                    const int *pixel, const Pose *poses, float *depth, int *status, hipStream_t st,
                                    bool wave) {
  if (events <= 0) return hipSuccess;
  const int slots = 256 / (samples * samples);
  const unsigned blocks = (unsigned)((events + slots - 1) / slots);
  if (wave)  // the one word of each lane's depth hand-off
    event_depth_multi<true><<<blocks, 256, 256 * sizeof(int), st>>>(segs, events, samples, seg, pixel, poses, depth,
                                                                    status);
  else
    event_depth_multi<false><<<blocks, 256, lds, st>>>(segs, events, samples, seg, pixel, poses, depth, status);
  return hipGetLastError();
}
hipError_t launch_segment_keys(int events, const unsigned *order, const int *seg, unsigned long long *keys,
                               hipStream_t st) {
  if (events <= 0) return hipSuccess;
  segment_keys_kernel<<<(events + 255) / 256, 256, 0, st>>>(events, order, seg, keys);
  return hipGetLastError();
}
hipError_t launch_permute_multi(const SegDev *segs, int events, const unsigned *order, const double *time,
                                const int *seg, const int *pixel, const int *polarity, const float *depth,
                                dvs_event *out, int *out_seg, hipStream_t st) {
  if (events <= 0) return hipSuccess;
  permute_multi<<<(events + 255) / 256, 256, 0, st>>>(segs, events, order, time, seg, pixel, polarity, depth, out,
                                                       out_seg);
  return hipGetLastError();
}
hipError_t launch_fill_int(int *p, int value, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  fill_int_kernel<<<(n + 255) / 256, 256, 0, st>>>(p, value, n);
  return hipGetLastError();
}
hipError_t launch_blur_accumulate(int npix, int exposures, const int *offsets, const float *frames,
                                  const double *weights, int wide, float *out, hipStream_t st) {
  if (npix <= 0 || exposures <= 0) return hipSuccess;
  if (exposures > 65535) return hipErrorInvalidConfiguration;
  blur_accumulate_kernel<<<dim3((npix + 255) / 256, exposures), 256, 0, st>>>(npix, offsets, frames, weights, wide,
                                                                             out);
  return hipGetLastError();
}
hipError_t launch_staged_scatter(const char *mirror, const ScatterArgs &a, hipStream_t st) {
  if (a.count <= 0) return hipSuccess;
  if (a.count > kScatterCopies) return hipErrorInvalidValue;
  staged_scatter_kernel<<<a.count, 256, 0, st>>>(mirror, a);
  return hipGetLastError();
}
hipError_t launch_hapke(int64_t n, const float *ci, const float *ce, const float *cg, const float *alb,
                        const dvs_hapke &h, float *out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hapke_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, ci, ce, cg, alb, h, hapke_constants(h), out);
  return hipGetLastError();
}
hipError_t launch_invert_hapke(int64_t n, const float *inc, const float *emi, const float *phase, const float *target,
                               const dvs_hapke &h, float *out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  invert_hapke_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, inc, emi, phase, target, h, out);
  return hipGetLastError();
}
hipError_t launch_receptor(int64_t n, const double *u, const double *rate, const double *dt, double charge,
                           double *out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  receptor_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, u, rate, dt, charge, out);
  return hipGetLastError();
}
hipError_t launch_advance(int64_t n, dvs_pixel_state *states, const dvs_pixel_params *params, const double *rates,
                          const double *t0, const double *t1, int capacity, double normal, int *counts,
                          dvs_crossing *events, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  advance_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, states, params, rates, t0, t1, capacity, normal,
                                                                counts, events);
  return hipGetLastError();
}
hipError_t launch_poisson(int64_t n, const double *lambda, const double *u, double *out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  poisson_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, lambda, u, out);
  return hipGetLastError();
}
hipError_t launch_noise(uint64_t seed, int64_t pixels, double lambda, int draws, double *poisson, double *normal,
                        hipStream_t st) {
  if (pixels <= 0) return hipSuccess;
  noise_kernel<<<(unsigned)((pixels + 255) / 256), 256, 0, st>>>(seed, pixels, lambda, draws, poisson, normal);
  return hipGetLastError();
}
hipError_t launch_traversal_stats(const DevScene &s, const Pose *pose, int npix, unsigned long long *totals,
                                  hipStream_t st) {
  const long long items = (long long)npix * s.samples * s.samples;
  const size_t lds = (size_t)s.stack_depth * 256 * sizeof(int);
  traversal_stats_kernel<<<(unsigned)((items + 255) / 256), 256, lds, st>>>(s, pose, npix, totals);
  return hipGetLastError();
}
hipError_t launch_rays(const DevScene &s, int64_t n, const float *o, const float *d, const int *self, int shadow,
                       int *hit, int *tri, float *t, float *u, float *v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (s.packet)
    rays_kernel<true><<<blocks, 256, 0, st>>>(s, n, o, d, self, shadow, hit, tri, t, u, v);
  else
    rays_kernel<false><<<blocks, 256, (size_t)s.stack_depth * 256 * sizeof(int), st>>>(s, n, o, d, self, shadow, hit,
                                                                                       tri, t, u, v);
  return hipGetLastError();
}

}  // namespace dvsamd
