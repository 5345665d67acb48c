// gfx950 kernels of the adaptive optics step (structure and equivalence
// argument in adaptive.h).
//
//  adapt_trace<LAYOUT>   one lane per sub-ray, ordered FP64 mean per pixel
//                        sample (warp_radiance, optical_sampler.cu:31-38)
//  adapt_decide          one lane per pending segment: the accept / exact /
//                        bisect rule of refine_pixels (optical_sampler.cu:
//                        50-79); resolved and bisected segments compacted by
//                        wave ballot + one atomic per wave
//  adapt_children        the two halves of every bisected segment
//  adapt_counts / term_bucket / place_knots
//                        knot counts, prefix offsets and depth-first knot order
//  adapt_exact_*         identify_exact_cells (optical_sampler.cu:114-131)
//  adapt_gate            advance_mean_probes (event_sensor.cu:138-150)
//  adapt_rates           photocurrent of every (cell, pixel) from knots or
//                        exact cells, with the cell's shot-noise draw
//  adapt_integrate       integrate_optical_knots and integrate_fallback_tile
//                        (event_sensor.cu:88-132): one lane per pixel, the
//                        serial receptor chain only
#include <hip/hip_runtime.h>

#include "adaptive.h"
#include "device_circuit.hpp"
#include "device_optics.hpp"

namespace dvsdev {

// native_pose (optical_sampler.cu:26-30): the pose at t in the first interval
// whose end is not before t.
__device__ inline Pose native_pose(const dvs_optical_interval *ivs, int count, double t) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int m = (lo + hi) / 2;
    if (t > ivs[m].t1) lo = m + 1; else hi = m;
  }
  return pose_at(ivs[lo], t);
}

// The five probe poses of every segment.  The end pose is kept (cur_end);
// the start probe of a segment reuses the previous window's end radiance
// when its pose equals the kept end pose bit for bit (prev_end, given only
// while the segment layout is unchanged): the trace is a function of the
// pose, the scene and the pixel, so the radiance is the one a trace gives.
__device__ __forceinline__ bool same_pose(const Pose &a, const Pose &b) {
  const unsigned *x = reinterpret_cast<const unsigned *>(&a), *y = reinterpret_cast<const unsigned *>(&b);
  bool same = true;
  for (int w = 0; w < 11; ++w) same &= x[w] == y[w];  // every field but the padding
  return same;
}
__global__ void adapt_probe_poses(int nseg, const SegDev *segs, const AdaptDev *ad, Pose *out, const Pose *prev_end,
                                  Pose *cur_end, int *reuse) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nseg * 5) return;
  const int s = k / 5, j = k - s * 5;
  const double start = ad[s].start, end = ad[s].end;
  const Pose pose = native_pose(segs[s].ivs, segs[s].iv_count, start + (end - start) * j * .25);
  out[k] = pose;
  if (j == 4) cur_end[s] = pose;
  if (j == 0) reuse[s] = prev_end != nullptr && same_pose(pose, prev_end[s]);
}

__global__ void adapt_cell_frac(int cells, const double *t0, const double *t1, const int *cell_begin, int nseg,
                                const AdaptDev *ad, double *frac) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cells) return;
  const AdaptDev &a = ad[segment_of<int>(cell_begin, nseg, c)];
  frac[c] = ((t0[c] + t1[c]) * .5 - a.start) / (a.end - a.start);
}

// LAYOUT_PROBE: segment s owns items [5 b, 5 e) of its pixels [b, e), pixel
// major, so that the rays of a wave are one pixel's probes at nearby times.
// Without reuse, local item l is probe l % 5 of pixel b + l / 5.  With reuse
// (the start pose equals the previous window's end pose), the first 4 (e - b)
// items are probes 1..4 of each pixel (one wave per pixel at 4x4 sub-rays),
// and the last e - b items only copy the previous end probe into probe 0, in
// blocks that trace nothing.  Probe j of pixel p goes to out[p * 5 + j].
// THREADS lanes per block (64, 128 or 256): THREADS / ss^2 pixel samples,
// each lane one sub-ray, with its traversal stack in LDS at stride THREADS
// (WAVE false) or in the wave-packet traversal (WAVE true, every lane of
// the block joins it).
template <int LAYOUT, int THREADS, bool WAVE>
__global__ __launch_bounds__(THREADS) void adapt_trace(const SegDev *segs, const int *pixel_begin, int nseg,
                                                   long long items, const int *item_pixel, const int *item_pose,
                                                   const Pose *poses, double *out, int *status,
                                                   unsigned long long *shadow_count, const int *reuse,
                                                   const double *prev_probes) {
  extern __shared__ double lds_words[];
  int *words = reinterpret_cast<int *>(lds_words);
  int *stack = words + threadIdx.x;
  const int samples = segs[0].scene.samples, R = samples * samples, slots = THREADS / R;
  const int slot = threadIdx.x / R, sub = threadIdx.x - slot * R;
  const long long q = (long long)blockIdx.x * slots + slot;
  const bool active = slot < slots && q < items;
  SubSample r{0.0, -1.f, 0, 0};
  long long at = q;  // output index
  bool reused = false;
  int gp = 0;
  Pose pose{};
  if (active) {
    if (LAYOUT == LAYOUT_PROBE) {
      const int si = segment_of<int>(pixel_begin, nseg, (int)(q / 5));
      const int b = pixel_begin[si], n = pixel_begin[si + 1] - b;
      const int l = (int)(q - 5LL * b);
      int j;
      if (!reuse[si]) {
        gp = b + l / 5;
        j = l % 5;
      } else if (l < 4 * n) {
        gp = b + l / 4;
        j = 1 + l % 4;
      } else {
        gp = b + (l - 4 * n);
        j = 0;
        reused = true;
      }
      pose = poses[si * 5 + j];
      at = (long long)gp * 5 + j;
    } else if (LAYOUT == LAYOUT_QUERY) {
      gp = item_pixel[q];
      pose = poses[q];
    } else {
      gp = item_pixel[q];
      pose = poses[item_pose[q]];
    }
  }
  const int si = segment_of<int>(pixel_begin, nseg, gp);
  if (WAVE)
    r = sample_subray_wave<true>(segs[si].scene, gp - pixel_begin[si], pose, sub % samples, sub / samples,
                                 active && !reused);
  else if (active && !reused)
    r = sample_subray<true>(segs[si].scene, gp - pixel_begin[si], pose, sub % samples, sub / samples, stack,
                            THREADS);
  const LdsInts lw = (LdsInts)words;  // the lanes' radiances, low then high words
  lw[threadIdx.x] = __double2loint(r.radiance);
  lw[THREADS + threadIdx.x] = __double2hiint(r.radiance);
  // shadow rays counted per block: one atomic on the shared counter per
  // block instead of per wave
  __shared__ int block_shadows[THREADS / 64];
  const unsigned long long shadows = __ballot(r.shadow);
  const unsigned long long invalid = __ballot(r.invalid);
  if ((threadIdx.x & 63) == 0) {
    block_shadows[threadIdx.x >> 6] = __popcll(shadows);
    if (invalid) atomicOr(status, 2);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int n = 0;
#pragma unroll
    for (int w = 0; w < THREADS / 64; ++w) n += block_shadows[w];
    if (n) atomicAdd(shadow_count, (unsigned long long)n);
  }
  if (!active || sub != 0) return;
  if (reused) {
    const double kept = prev_probes[at + 4];  // the previous window's end probe of this pixel
    if (!isfinite(kept)) atomicOr(status, 2);
    out[at] = kept;
    return;
  }
  double total = 0;
  for (int j = 0; j < R; ++j) total += __hiloint2double(lw[THREADS + threadIdx.x + j], lw[threadIdx.x + j]);
  if (!isfinite(total)) atomicOr(status, 2);
  out[at] = total / R;
}

// Two root segments per pixel: [0, 1/2] is popped first with one pending
// sibling on the depth-first stack, [1/2, 1] with none (refine_pixels, :46-49).
__global__ void adapt_init(int pixels, AdaptBufs b, AdaptEntry *entries, int *clear, int clear_words,
                           int *event_counts) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < clear_words) clear[p] = 0;
  if (p >= pixels) return;
  event_counts[p] = 0;
  const double *r = b.probes + (size_t)p * 5;
  entries[2 * p] = AdaptEntry{0.0, .5, r[0], r[1], r[2], -1.0, -1.0, p, 1};
  entries[2 * p + 1] = AdaptEntry{.5, 1.0, r[2], r[3], r[4], -1.0, -1.0, p, 0};
  b.bound[p] = 3;  // knot {0, r0} + two pending segments
  b.flags[p] = 0;
  b.bisects[p] = 0;
  b.exact_terms[p] = 0;
}

__global__ __launch_bounds__(256) void adapt_decide(int bound, const int *prev_level, const AdaptEntry *entries,
                                                    const SegDev *segs, const AdaptDev *ad, const int *pixel_begin,
                                                    int nseg, AdaptBufs b, AdaptTerm *terms,
                                                    unsigned long long *keys, AdaptEntry *bisected, int *query_pixel,
                                                    Pose *query_pose, AdaptEntry *direct, int *head, int *level_ctr) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = prev_level ? min(bound, 2 * (prev_level[0] + prev_level[1])) : bound;
  int decision = -1;  // -1 dropped, ADAPT_ACCEPT, ADAPT_EXACT, 2 bisect
  AdaptEntry e{};
  int si = 0;
  if (k < n) {
    e = entries[k];
    si = segment_of<int>(pixel_begin, nseg, e.pixel);
    const AdaptDev &A = ad[si];
    if (!(__hip_atomic_load(&b.flags[e.pixel], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 1)) {
      if (__hip_atomic_load(&b.bound[e.pixel], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > A.capacity) {
        atomicOr(&b.flags[e.pixel], 1);  // the knot count can only exceed the capacity
      } else {
        const SegDev &g = segs[si];
        const dvs_pixel_state &st = g.states[e.pixel - pixel_begin[si]];
        const PhotoParams ph = g.photo;
        const double charge = g.params.charge_scale;
        const double actual = e.mid * ph.electrons_per_radiance + ph.dark_electrons;
        const double estimate = (e.left + e.right) * .5 * ph.electrons_per_radiance + ph.dark_electrons;
        const double residual = fabs(log(fmax(1e-30, actual) / fmax(1e-30, estimate)));
        const double leaf = (e.b - e.a) * (A.end - A.start);
        // both pieces start from the pixel's voltage: exp(u) once for the two
        double va = st.receptor_log, vb = va;
        if (!(leaf <= 0)) {  // receptor_log_at's guard
          const double x = exp(st.receptor_log);
          va = receptor_log_from(st.receptor_log, x, actual, leaf, charge);
          vb = receptor_log_from(st.receptor_log, x, estimate, leaf, charge);
        }
        const double verr = fabs(va - vb);
        const double budget = 1e-3 * fmin(st.threshold_pos, st.threshold_neg);
        if (verr <= budget && (residual <= A.tolerance || fmax(actual, estimate) * leaf < 1.0))
          decision = ADAPT_ACCEPT;
        else if (leaf <= ph.integration_step || e.zeros + 2 > 16)
          decision = ADAPT_EXACT;
        else
          decision = 2;
      }
    }
  }
  const bool resolved = decision == ADAPT_ACCEPT || decision == ADAPT_EXACT;
  const bool ahead = e.lspec >= 0;  // the halves' midpoints are already traced
  const unsigned long long tmask = __ballot(resolved);
  const unsigned long long emask = __ballot(decision == ADAPT_EXACT);
  const unsigned long long bmask = __ballot(decision == 2 && !ahead);
  const unsigned long long dmask = __ballot(decision == 2 && ahead);
  // One returning atomic per counter and block: atomics on one address are
  // served one after another, so per-wave ones would cost each wave a round
  // trip behind all the others.  The block's four waves take consecutive
  // ranges of its share (no result depends on the compaction order).
  __shared__ int wave_count[4][3], block_base[3];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    wave_count[wave][0] = __popcll(tmask);
    wave_count[wave][1] = __popcll(bmask);
    wave_count[wave][2] = __popcll(dmask);
    if (emask) atomicAdd(&head[1], __popcll(emask));
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int total = wave_count[0][threadIdx.x] + wave_count[1][threadIdx.x] + wave_count[2][threadIdx.x] +
                      wave_count[3][threadIdx.x];
    int *ctr = threadIdx.x == 0 ? &head[0] : &level_ctr[threadIdx.x - 1];
    block_base[threadIdx.x] = total ? atomicAdd(ctr, total) : 0;
  }
  __syncthreads();
  int tbase = block_base[0], bbase = block_base[1], dbase = block_base[2];
  for (int w = 0; w < wave; ++w) {
    tbase += wave_count[w][0];
    bbase += wave_count[w][1];
    dbase += wave_count[w][2];
  }
  const unsigned long long below = (1ull << lane) - 1ull;
  if (resolved) {
    const int idx = tbase + __popcll(tmask & below);
    terms[idx] = AdaptTerm{e.a, e.b, e.mid, e.right, e.pixel, decision};
    // start fractions are dyadic with at most 21 bits below the point
    keys[idx] = ((unsigned long long)e.pixel << 24) | (unsigned long long)(e.a * 16777216.0);
    // the bound grows by one from 3: it passes the capacity exactly once
    if (decision == ADAPT_ACCEPT) {
      if (atomicAdd(&b.bound[e.pixel], 1) == ad[si].capacity) atomicAdd(&head[2], 1);
    } else {
      atomicAdd(&b.exact_terms[e.pixel], 1);
    }
  } else if (decision == 2) {
    if (atomicAdd(&b.bound[e.pixel], 1) == ad[si].capacity) atomicAdd(&head[2], 1);
    atomicAdd(&b.bisects[e.pixel], 1);
    const double mid = (e.a + e.b) * .5;
    if (ahead) {
      const int idx = dbase + __popcll(dmask & below);
      direct[2 * idx] = AdaptEntry{e.a, mid, e.left, e.lspec, e.mid, -1.0, -1.0, e.pixel, e.zeros + 1};
      direct[2 * idx + 1] = AdaptEntry{mid, e.b, e.mid, e.rspec, e.right, -1.0, -1.0, e.pixel, e.zeros};
    } else {
      const int idx = bbase + __popcll(bmask & below);
      bisected[idx] = e;
      const AdaptDev &A = ad[si];
      const SegDev &g = segs[si];
      // the halves' midpoints, then the midpoints of their halves, with the
      // operations a half's own bisection would use
      const double l_mid = (e.a + mid) * .5, r_mid = (mid + e.b) * .5;
      const double at[kBisectQueries] = {l_mid, r_mid, (e.a + l_mid) * .5, (l_mid + mid) * .5, (mid + r_mid) * .5,
                                         (r_mid + e.b) * .5};
#pragma unroll
      for (int j = 0; j < kBisectQueries; ++j) {
        query_pixel[kBisectQueries * idx + j] = e.pixel;
        query_pose[kBisectQueries * idx + j] = native_pose(g.ivs, g.iv_count, A.start + (A.end - A.start) * at[j]);
      }
    }
  }
}

// Halves of bisected segment k: the left half is popped before the right
// one, with the right half pending on the stack (refine_pixels, :76-78).
__global__ void adapt_children(int n, const AdaptEntry *bisected, const double *query_rad, AdaptEntry *next) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const AdaptEntry e = bisected[k];
  const double mid = (e.a + e.b) * .5;
  const double *q = query_rad + (size_t)kBisectQueries * k;
  next[2 * k] = AdaptEntry{e.a, mid, e.left, q[0], e.mid, q[2], q[3], e.pixel, e.zeros + 1};
  next[2 * k + 1] = AdaptEntry{mid, e.b, e.mid, q[1], e.right, q[4], q[5], e.pixel, e.zeros};
}

__global__ void adapt_counts(int pixels, const AdaptDev *ad, const int *pixel_begin, int nseg, AdaptBufs b,
                             int *fill, int clear_exact, int *stats, int stat_words) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < stat_words) stats[p] = 0;
  if (p >= pixels) return;
  fill[p] = 0;
  if (clear_exact) {
    b.ecount[p] = 0;
    b.eoff[p] = 0;
  }
  const AdaptDev &A = ad[segment_of<int>(pixel_begin, nseg, p)];
  int f = b.flags[p];
  if (b.bound[p] > A.capacity) f |= 1;
  b.flags[p] = f;
  b.kcount[p] = (f & 1) ? 0 : b.bound[p];
}

// Knots in time order without a sort.  A knot pixel's resolved segments
// partition its window, and segment t's knots start at
//   base + 1 + (knots of the pixel's segments that start before t),
// base being the pixel's first knot (the window start).  Each segment first
// drops a word {start fraction, knots} into a bucket inside its pixel's knot
// range (a pixel has fewer segments than knots), then counts the bucket.
// Start fractions are dyadic with at most 21 bits below the point, so the
// 24-bit key of the segment key word is exact.
__global__ void adapt_term_bucket(int n, const unsigned long long *keys, const AdaptTerm *terms, AdaptBufs b,
                                  int *fill, unsigned *bucket) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int p = (int)(keys[t] >> 24);
  if (b.flags[p] & 1) return;
  const int base = b.koff[p] - b.kcount[p];
  const unsigned knots = terms[t].kind == ADAPT_ACCEPT ? 2u : 1u;
  bucket[base + atomicAdd(&fill[p], 1)] = ((unsigned)(keys[t] & 0xFFFFFFu) << 2) | knots;
}

__global__ void adapt_place_knots(int n, const unsigned long long *keys, const AdaptTerm *terms, AdaptBufs b,
                                  const int *fill, const unsigned *bucket) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int p = (int)(keys[t] >> 24);
  if (b.flags[p] & 1) return;
  const int base = b.koff[p] - b.kcount[p], m = fill[p];
  const unsigned mine = (unsigned)(keys[t] & 0xFFFFFFu);
  int before = 0;
  for (int i = 0; i < m; ++i) {
    const unsigned w = bucket[base + i];
    before += (w >> 2) < mine ? (int)(w & 3u) : 0;
  }
  const int pos = base + 1 + before;
  const AdaptTerm T = terms[t];
  if (T.kind == ADAPT_ACCEPT) {
    b.knots[pos] = make_double2((T.a + T.b) * .5, T.mid);
    b.knots[pos + 1] = make_double2(T.b, T.right);
  } else {
    b.knots[pos] = make_double2(-T.b, T.right);
  }
  if (mine == 0) b.knots[base] = make_double2(0.0, b.probes[(size_t)p * 5]);
}

// Knot bracket of a cell midpoint fraction (integrate_optical_knots, :99).
__device__ __forceinline__ int knot_bracket(const double2 *v, int size, double fraction, int a) {
  while (a + 2 < size && fraction > fabs(v[a + 1].x)) ++a;
  return a;
}

__global__ void adapt_exact_count(int pixels, const SegDev *segs, const AdaptDev *ad, const int *pixel_begin,
                                  int nseg, AdaptBufs b) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= pixels) return;
  b.ecount[p] = 0;
  if (b.flags[p] & 1) return;
  const int si = segment_of<int>(pixel_begin, nseg, p);
  const AdaptDev &A = ad[si];
  const int cells = segs[si].cells, size = b.kcount[p];
  const double2 *v = b.knots + (b.koff[p] - size);
  bool any = false;  // only pixels with an exact segment have exact cells
  for (int k = 1; k < size && !any; ++k) any = v[k].x < 0;
  if (!any) return;
  const double *frac = b.cell_frac + A.cell_offset;
  int a = 0, used = 0;
  for (int c = 0; c < cells; ++c) {
    a = knot_bracket(v, size, frac[c], a);
    if (v[a + 1].x < 0) {
      if (used == A.capacity) {  // identify_exact_cells overflow: whole-pixel tracing
        b.flags[p] |= 1;
        return;
      }
      ++used;
    }
  }
  b.ecount[p] = used;
}

__global__ void adapt_exact_list(int pixels, const SegDev *segs, const AdaptDev *ad, const int *pixel_begin,
                                 int nseg, AdaptBufs b, int *item_pixel, int *item_pose) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= pixels || b.ecount[p] == 0) return;
  const int si = segment_of<int>(pixel_begin, nseg, p);
  const AdaptDev &A = ad[si];
  const int cells = segs[si].cells, size = b.kcount[p];
  const double2 *v = b.knots + (b.koff[p] - size);
  const double *frac = b.cell_frac + A.cell_offset;
  int a = 0, at = b.eoff[p] - b.ecount[p];
  for (int c = 0; c < cells; ++c) {
    a = knot_bracket(v, size, frac[c], a);
    if (v[a + 1].x < 0) {
      b.ecell[at] = c;
      item_pixel[at] = p;
      item_pose[at] = A.cell_offset + c;
      ++at;
    }
  }
}

__global__ void adapt_fallback_flags(int pixels, const int *flags, int *out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < pixels) out[p] = flags[p] & 1;
}

__global__ void adapt_fallback_list(int pixels, const int *pixel_begin, int nseg, const int *scan, AdaptBufs b,
                                    int *list) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= pixels) return;
  if (!(b.flags[p] & 1)) {
    b.fb_index[p] = -1;
    return;
  }
  const int si = segment_of<int>(pixel_begin, nseg, p);
  const int begin = pixel_begin[si] ? scan[pixel_begin[si] - 1] : 0;
  list[scan[p] - 1] = p - pixel_begin[si];
  b.fb_index[p] = scan[p] - 1 - begin;
}

__global__ void adapt_segment_ends(int nseg, const int *pixel_begin, const int *scan, int *out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  out[s] = pixel_begin[s + 1] > 0 ? scan[pixel_begin[s + 1] - 1] : 0;
}

// One block per segment: OpticalRefinementStats over its pixels.
// Per-segment refinement statistics: kStatBlocks blocks per segment each
// reduce a strided share of its pixels and add it to out (cleared first).
constexpr int kStatBlocks = 64;
__global__ __launch_bounds__(256) void adapt_segment_stats(const int *pixel_begin, AdaptBufs b, AdaptSegStats *out) {
  __shared__ unsigned long long q[256];
  __shared__ int fb[256], mk[256], ex[256], nc[256];
  const int s = blockIdx.y;
  unsigned long long queries = 0;
  int fallback = 0, max_knots = 0, exact = 0, cells = 0;
  for (int p = pixel_begin[s] + (int)blockIdx.x * 256 + threadIdx.x; p < pixel_begin[s + 1]; p += kStatBlocks * 256) {
    queries += 5 + 2ull * (unsigned)b.bisects[p];
    fallback += b.flags[p] & 1;
    max_knots = max(max_knots, b.kcount[p]);
    exact += b.exact_terms[p];
    cells += b.ecount[p];
  }
  q[threadIdx.x] = queries;
  fb[threadIdx.x] = fallback;
  mk[threadIdx.x] = max_knots;
  ex[threadIdx.x] = exact;
  nc[threadIdx.x] = cells;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      q[threadIdx.x] += q[threadIdx.x + w];
      fb[threadIdx.x] += fb[threadIdx.x + w];
      mk[threadIdx.x] = max(mk[threadIdx.x], mk[threadIdx.x + w]);
      ex[threadIdx.x] += ex[threadIdx.x + w];
      nc[threadIdx.x] += nc[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && (q[0] | fb[0] | mk[0] | ex[0] | nc[0])) {
    AdaptSegStats &o = out[s];
    atomicAdd(&o.traced_pixel_samples, q[0]);
    atomicAdd(&o.fallback_pixels, fb[0]);
    atomicMax(&o.maximum_knots, mk[0]);
    atomicAdd(&o.exact_segments, ex[0]);
    atomicAdd(&o.native_cell_queries, nc[0]);
  }
}

// Receptor log voltage of each probe pixel under native and under cached
// (knot-interpolated) radiance (advance_mean_probes, event_sensor.cu:138-150).
// Two lanes per probe pixel (list entry i of gate g): lane 0 integrates the
// receptor under the native radiance, lane 1 under the interpolated one, on
// the same instructions with the inputs selected per lane, so the wave does
// not diverge.  The error goes to errors[g * 1024 + slots[g * 1024 + i]].
__global__ void adapt_gate(const SegDev *gate_segs, const int *gate_seg, const int *slots, const AdaptDev *ad,
                           AdaptBufs b, double *errors) {
  const int g = blockIdx.x / 8, t = (blockIdx.x & 7) * 256 + threadIdx.x;
  const int i = t >> 1, chain = t & 1;
  const SegDev &G = gate_segs[g];
  if (i >= G.npix) return;  // both lanes of the pair
  const AdaptDev &A = ad[gate_seg[g]];
  const int local = G.pixel_list[i], gp = G.pixel_offset + local;
  const double charge = G.params.charge_scale;
  const PhotoParams ph = G.photo;
  double u = G.states[local].receptor_log;
  const bool knots = !(b.flags[gp] & 1);
  const int size = b.kcount[gp];
  const double2 *v = b.knots + (b.koff[gp] - size);
  const double *frac = b.cell_frac + A.cell_offset;
  const int cells = G.cells;
  double *out = errors + g * 1024 + slots[g * 1024 + i];
  if (cells <= 0) {
    if (chain == 0) *out = 0.0;
    return;
  }
  // The receptor chain is serial over the cells: the inputs of the next
  // group of cells are loaded while this group computes, and the knot pair
  // around the cell stays in registers (cell fractions ascend, so the
  // bracket only walks forward).
  constexpr int kGroup = 4;
  auto at = [&](int c) { return min(c, cells - 1); };
  double light_n[kGroup], f_n[kGroup], dt_n[kGroup];
#pragma unroll
  for (int j = 0; j < kGroup; ++j) {
    light_n[j] = G.rad[(size_t)at(j) * G.npix + i];
    f_n[j] = frac[at(j)];
    dt_n[j] = G.t1[at(j)] - G.t0[at(j)];
  }
  int a = 0;
  double2 va = make_double2(0.0, 0.0), vb = va;
  if (knots && size >= 2) {
    va = v[0];
    vb = v[1];
  }
  for (int c0 = 0; c0 < cells; c0 += kGroup) {
    double light[kGroup], f[kGroup], dt[kGroup];
#pragma unroll
    for (int j = 0; j < kGroup; ++j) {
      light[j] = light_n[j];
      f[j] = f_n[j];
      dt[j] = dt_n[j];
    }
#pragma unroll
    for (int j = 0; j < kGroup; ++j) {
      const int c = at(c0 + kGroup + j);
      light_n[j] = G.rad[(size_t)c * G.npix + i];
      f_n[j] = frac[c];
      dt_n[j] = G.t1[c] - G.t0[c];
    }
#pragma unroll
    for (int j = 0; j < kGroup; ++j) {
      if (c0 + j < cells) {
        double cached = light[j];
        if (knots) {
          while (a + 2 < size && f[j] > fabs(vb.x)) {  // knot_bracket
            ++a;
            va = vb;
            vb = v[a + 1];
          }
          const double w = (f[j] - fabs(va.x)) / (fabs(vb.x) - fabs(va.x));
          if (vb.x >= 0) cached = va.y + (vb.y - va.y) * w;
        }
        const double r = chain ? fmax(0.0, cached) : light[j];
        u = receptor_log_at(u, r * ph.electrons_per_radiance + ph.dark_electrons, dt[j], charge);
      }
    }
  }
  const double other = __shfl_xor(u, 1);
  if (chain == 0) *out = fabs(u - other);
}

// Knot bracket by bisection: the index the walk above stops at, i.e. the
// first a with fraction <= |v[a+1].x|, at most size - 2 (|x| is ascending).
__device__ __forceinline__ int knot_bracket_search(const double2 *v, int size, double fraction) {
  int lo = 1, hi = size - 1;  // first j in [1, size-1] with fraction <= |v[j].x|, else size-1
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (fraction > fabs(v[m].x)) lo = m + 1; else hi = m;
  }
  return lo - 1;
}

// Photocurrent of every (cell, pixel) of cells [c0, c1) for pixels that
// interpolate knots (integrate_optical_knots, event_sensor.cu:97-105): knot
// or exact-cell radiance, then the cell's shot-noise draw.  A thread takes
// one pixel and `run` consecutive cells (x over pixels, y over cell runs; a
// multiple of kCodeCells, kRateCells unless tuned): one bisection finds the first cell's knot bracket and exact-cell
// position, then both walk forward as in the reference, so each knot is read
// once per run while the rate stores stay coalesced across pixels.
constexpr int kRateCells = 32;  // the default run; launches take any multiple of kCodeCells
static_assert(kRateCells % kCodeCells == 0, "rate runs start at whole code words");
// Returns whether any cell of the run was left to adapt_rates_rare.
__device__ __forceinline__ bool adapt_rates_run(const SegDev *segs, const AdaptDev *ad, const int *pixel_begin,
                                                int nseg, int pixels, const AdaptBufs &b, int c0, int c1, int run,
                                                int *status) {
  const int gp = blockIdx.x * 256 + threadIdx.x;
  if (gp >= pixels) return false;
  const int si = segment_of<int>(pixel_begin, nseg, gp);
  const SegDev &g = segs[si];
  const AdaptDev &A = ad[si];
  const int cb = c0 + (int)blockIdx.y * run, ce = min(min(c1, g.cells), cb + run);
  if (cb >= ce || !A.accepted || (b.flags[gp] & 1)) return false;
  const int i = gp - pixel_begin[si];
  const int size = b.kcount[gp];
  const double2 *v = b.knots + (b.koff[gp] - size);
  const double *frac = b.cell_frac + A.cell_offset;
  int a = knot_bracket_search(v, size, frac[cb]);
  const int ec = b.ecount[gp], eo = b.eoff[gp] - ec;
  int e = 0, hi = ec;  // first exact cell >= cb
  while (e < hi) {
    const int m = (e + hi) >> 1;
    if (b.ecell[eo + m] < cb) e = m + 1; else hi = m;
  }
  // cell serials from the segment's count, not the pixel state: the circuit
  // of the previous chunk may be updating the state while this runs
  const uint64_t serial0 = A.cell_base;
  // the segment's fields in registers (the stores below could alias them)
  const PhotoParams ph = g.photo;
  const uint64_t seed = g.seed;
  const CellMaps *maps = g.maps;
  double *out = b.rates + gp;
  uint64_t *out_codes = b.codes + gp;
  uint64_t shared[4] = {0, 0, 0, 0}, shared_index = ~0ull;
  uint64_t word = 0;  // codes of the current group of kCodeCells cells
  bool marked = false;
  // the knot pair around the cell stays in registers (knot_bracket's walk),
  // and the next cell's fraction and map words are loaded a cell ahead
  double2 va = v[a], vb = v[a + 1];
  // (entries are 16-byte aligned: {t0, t1} and {rate[0], rate[1]})
  double f_next = frac[cb];
  double2 t_next = reinterpret_cast<const double2 *>(maps + cb)[0];
  double2 r_next = reinterpret_cast<const double2 *>(maps + cb)[1];
  for (int c = cb; c < ce; ++c) {
    const double f = f_next;
    const double2 t = t_next, r = r_next;
    const int cn = min(c + 1, ce - 1);
    f_next = frac[cn];
    t_next = reinterpret_cast<const double2 *>(maps + cn)[0];
    r_next = reinterpret_cast<const double2 *>(maps + cn)[1];
    while (a + 2 < size && f > fabs(vb.x)) {
      ++a;
      va = vb;
      vb = v[a + 1];
    }
    const double w = (f - fabs(va.x)) / (fabs(vb.x) - fabs(va.x));
    double light = va.y + (vb.y - va.y) * w;
    if (e < ec && b.ecell[eo + e] == c) light = b.erad[eo + e++];
    light = fmax(0.0, light);
    if (!isfinite(light)) atomicOr(status, 2);
    const uint64_t serial = serial0 + (uint64_t)c;
    if (ph.shot_noise && serial / kShotCellsPerBlock != shared_index) {
      shared_index = serial / kShotCellsPerBlock;
      shot_block(seed, (uint64_t)i, shared_index, shared);
    }
    double escape = 0;
    const int code = cell_code_first(light, ph, t.y - t.x, r.x, r.y, serial, shared, escape);
    if (code == kCodeEscape) {
      out[(size_t)(c - c0) * pixels] = escape;
      marked |= escape < 0;
    }
    // cb - c0 is a multiple of kCodeCells: runs start at whole groups
    const int j = (c - c0) % kCodeCells;
    word |= (uint64_t)code << (8 * j);
    if (j == kCodeCells - 1 || c == ce - 1) {
      out_codes[(size_t)((c - c0) / kCodeCells) * pixels] = word;
      word = 0;
    }
  }
  return marked;
}
__device__ __forceinline__ size_t rate_wave() {
  return ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 4 + (threadIdx.x >> 6);
}
__global__ __launch_bounds__(256) void adapt_rates(const SegDev *segs, const AdaptDev *ad, const int *pixel_begin,
                                                   int nseg, int pixels, AdaptBufs b, int c0, int c1, int run,
                                                   int *status, int *rare) {
  const bool marked = adapt_rates_run(segs, ad, pixel_begin, nseg, pixels, b, c0, c1, run, status);
  // one flag per wave: does any lane leave a draw to adapt_rates_rare
  const unsigned long long any = __ballot(marked);
  if ((threadIdx.x & 63) == 0) rare[rate_wave()] = any != 0;
}
// The draws adapt_rates left to the cells' own streams (lambda >= 10), in
// the waves that flagged any.
__global__ __launch_bounds__(256) void adapt_rates_rare(const SegDev *segs, const AdaptDev *ad,
                                                        const int *pixel_begin, int nseg, int pixels, AdaptBufs b,
                                                        int c0, int c1, int run, const int *rare) {
  if (!rare[rate_wave()]) return;
  const int gp = blockIdx.x * 256 + threadIdx.x;
  if (gp >= pixels) return;
  const int si = segment_of<int>(pixel_begin, nseg, gp);
  const SegDev &g = segs[si];
  const AdaptDev &A = ad[si];
  const int cb = c0 + (int)blockIdx.y * run, ce = min(min(c1, g.cells), cb + run);
  if (cb >= ce || !A.accepted || (b.flags[gp] & 1)) return;
  const int i = gp - pixel_begin[si];
  double *out = b.rates + gp;
  const uint64_t *codes = b.codes + gp;
  for (int c = cb; c < ce; ++c) {
    // only escaped cells have a double; the first phase marked its draws -lambda
    const uint64_t w = codes[(size_t)((c - c0) / kCodeCells) * pixels];
    if (((w >> (8 * ((c - c0) % kCodeCells))) & 0xff) != (uint64_t)kCodeEscape) continue;
    double &r = out[(size_t)(c - c0) * pixels];
    if (r < 0) r = cell_rate_rest(r, g.t1[c] - g.t0[c], g.seed, (uint64_t)i, A.cell_base + (uint64_t)c);
  }
}

// Circuit of cells [c0, c1) of every pixel of an accepted window: the
// photocurrent of a knot pixel from adapt_rates, of a fallback pixel from
// its native trace; comparator and reset noise as in the native circuit.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void adapt_integrate(const SegDev *segs, const AdaptDev *ad, const int *block_begin,
                                                       int nseg, int pixels, AdaptBufs b, int c0, int c1,
                                                       int *status) {
  const int si = segment_of<int>(block_begin, nseg, (int)blockIdx.x);
  const SegDev &g = segs[si];
  const AdaptDev &A = ad[si];
  const int i = ((int)blockIdx.x - block_begin[si]) * 256 + threadIdx.x;
  if (!A.accepted) return;  // the whole block: blocks do not straddle segments
  // lanes past the segment's pixels stay in the cell loop (they hold cells of
  // the wave's shared tiles) without a pixel of their own
  const bool own = i < g.npix;
  const int gp = g.pixel_offset + (own ? i : 0);
  dvs_pixel_state st = own ? g.states[i] : dvs_pixel_state{};
  const bool fallback = b.flags[gp] & 1;
  const double *rates = fallback ? A.fb_rad + b.fb_index[gp] : b.rates + gp;
  const size_t stride = fallback ? (size_t)A.fb_count : (size_t)pixels;
  const int c_end = min(c1, g.cells);
  // knot pixels read photocurrent codes (their escapes from b.rates); a
  // fallback pixel's photocurrents are all doubles in A.fb_rad
  const int count = circuit_cells<true>(st, g.params, g.seed, (uint64_t)i, own, rates, stride, b.codes + gp,
                                        (size_t)pixels, !fallback, g.t0, g.t1, g.maps, c0, c_end,
                                        g.slots + (size_t)(own ? i : 0) * g.max_events, own ? g.counts[i] : 0,
                                        g.max_events);
  if (!own) return;
  if (count < 0) { atomicOr(status, count == -1 ? 1 : 2); g.counts[i] = 0; return; }
  st.shot_block += (uint64_t)max(0, c_end - c0);
  g.states[i] = st;
  g.counts[i] = count;
}

// verify_knots (optical_sampler.cu:86-103), in two parts around a native
// trace of the probes.  Pose j of `times` sits at window fraction
// (j + .5) / times.
__global__ void adapt_verify_poses(const SegDev *seg, double start, double end, int times, Pose *poses) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= times) return;
  const double fraction = (j + .5) / times;
  poses[j] = native_pose(seg->ivs, seg->iv_count, start + (end - start) * fraction);
}
// Log error of the knot interpolation against the traced radiance of each
// probe; item q is probe time q / positions.  -1 where the pixel has no
// knots (native fallback) or the time falls in an exact segment.
__global__ void adapt_verify_errors(int items, int positions, int times, const int *item_pixel, const double *light,
                                    AdaptBufs b, PhotoParams photo, double *errors) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= items) return;
  const int gp = item_pixel[q];
  const double fraction = (q / positions + .5) / times;
  if (b.flags[gp] & 1) { errors[q] = -1; return; }
  const int size = b.kcount[gp];
  const double2 *v = b.knots + (b.koff[gp] - size);
  int a = 0;
  while (a + 2 < size && fraction > fabs(v[a + 1].x)) ++a;
  if (v[a + 1].x < 0) { errors[q] = -1; return; }
  const double weight = (fraction - fabs(v[a].x)) / (fabs(v[a + 1].x) - fabs(v[a].x));
  const double estimate = v[a].y + (v[a + 1].y - v[a].y) * weight;
  const double actual = light[q] * photo.electrons_per_radiance + photo.dark_electrons;
  const double predicted = estimate * photo.electrons_per_radiance + photo.dark_electrons;
  errors[q] = fabs(log(fmax(1e-30, actual) / fmax(1e-30, predicted)));
}

}  // namespace dvsdev

namespace dvsamd {
using namespace dvsdev;

hipError_t launch_adapt_verify_poses(const SegDev *seg, double start, double end, int times, Pose *poses,
                                     hipStream_t st) {
  if (times <= 0) return hipSuccess;
  adapt_verify_poses<<<1, 64, 0, st>>>(seg, start, end, times, poses);
  return hipGetLastError();
}
hipError_t launch_adapt_verify_errors(int items, int positions, int times, const int *item_pixel,
                                      const double *light, AdaptBufs b, const PhotoParams &photo, double *errors,
                                      hipStream_t st) {
  if (items <= 0) return hipSuccess;
  adapt_verify_errors<<<(items + 255) / 256, 256, 0, st>>>(items, positions, times, item_pixel, light, b, photo,
                                                            errors);
  return hipGetLastError();
}

namespace {
inline unsigned grid(long long n, int block = 256) { return (unsigned)((n + block - 1) / block); }
}  // namespace

hipError_t launch_adapt_probe_poses(int nseg, const SegDev *segs, const AdaptDev *ad, Pose *out,
                                    const Pose *prev_end, Pose *cur_end, int *reuse, hipStream_t st) {
  if (nseg <= 0) return hipSuccess;
  adapt_probe_poses<<<grid(nseg * 5), 256, 0, st>>>(nseg, segs, ad, out, prev_end, cur_end, reuse);
  return hipGetLastError();
}
hipError_t launch_adapt_cell_frac(int cells, const double *t0, const double *t1, const int *cell_begin, int nseg,
                                  const AdaptDev *ad, double *frac, hipStream_t st) {
  if (cells <= 0) return hipSuccess;
  adapt_cell_frac<<<grid(cells), 256, 0, st>>>(cells, t0, t1, cell_begin, nseg, ad, frac);
  return hipGetLastError();
}
namespace {
template <int THREADS, bool WAVE>
void adapt_trace_as(int layout, long long blocks, size_t lds, const SegDev *segs, const int *pixel_begin, int nseg,
                    long long items, const int *item_pixel, const int *item_pose, const Pose *poses, double *out,
                    int *status, unsigned long long *shadow_count, hipStream_t st, const int *reuse,
                    const double *prev_probes) {
  const dim3 g((unsigned)blocks);
  switch (layout) {
    case LAYOUT_PROBE:
      adapt_trace<LAYOUT_PROBE, THREADS, WAVE><<<g, THREADS, lds, st>>>(segs, pixel_begin, nseg, items, item_pixel,
                                                                          item_pose, poses, out, status,
                                                                          shadow_count, reuse, prev_probes);
      break;
    case LAYOUT_QUERY:
      adapt_trace<LAYOUT_QUERY, THREADS, WAVE><<<g, THREADS, lds, st>>>(segs, pixel_begin, nseg, items, item_pixel,
                                                                          item_pose, poses, out, status,
                                                                          shadow_count, reuse, prev_probes);
      break;
    default:
      adapt_trace<LAYOUT_CELL, THREADS, WAVE><<<g, THREADS, lds, st>>>(segs, pixel_begin, nseg, items, item_pixel,
                                                                         item_pose, poses, out, status,
                                                                         shadow_count, reuse, prev_probes);
  }
}
template <int THREADS>
void adapt_trace_form(bool wave, int layout, long long blocks, size_t lds, const SegDev *segs,
                      const int *pixel_begin, int nseg, long long items, const int *item_pixel,
                      const int *item_pose, const Pose *poses, double *out, int *status,
                      unsigned long long *shadow_count, hipStream_t st, const int *reuse,
                      const double *prev_probes) {
  if (wave)
    adapt_trace_as<THREADS, true>(layout, blocks, lds, segs, pixel_begin, nseg, items, item_pixel, item_pose, poses,
                                  out, status, shadow_count, st, reuse, prev_probes);
  else
    adapt_trace_as<THREADS, false>(layout, blocks, lds, segs, pixel_begin, nseg, items, item_pixel, item_pose,
                                   poses, out, status, shadow_count, st, reuse, prev_probes);
}
}  // namespace

hipError_t launch_adapt_trace(int layout, const SegDev *segs, const int *pixel_begin, int nseg, int samples,
                              long long items, const int *item_pixel, const int *item_pose, const Pose *poses,
                              size_t lds, double *out, int *status, unsigned long long *shadow_count,
                              hipStream_t st, const int *reuse, const double *prev_probes, int threads, bool wave) {
  if (layout == LAYOUT_PROBE && (items % 5 != 0 || !reuse)) return hipErrorInvalidValue;
  if (threads != 64 && threads != 128 && threads != 256) return hipErrorInvalidValue;
  if (samples * samples > threads) threads = 256;  // a block holds at least one pixel sample
  if (items <= 0) return hipSuccess;
  const int slots = threads / (samples * samples);
  const long long blocks = (items + slots - 1) / slots;
  if (blocks > kMaxBlocks256) return hipErrorInvalidConfiguration;
  // per-lane stacks: lds holds a stack of lds / 1024 levels for 256 lanes,
  // the same levels for `threads` lanes; wave packets: the two words of each
  // lane's radiance hand-off
  const size_t bytes = wave ? 2 * sizeof(int) * (size_t)threads : lds / 256 * (size_t)threads;
  if (threads == 64)
    adapt_trace_form<64>(wave, layout, blocks, bytes, segs, pixel_begin, nseg, items, item_pixel, item_pose, poses,
                         out, status, shadow_count, st, reuse, prev_probes);
  else if (threads == 128)
    adapt_trace_form<128>(wave, layout, blocks, bytes, segs, pixel_begin, nseg, items, item_pixel, item_pose, poses,
                          out, status, shadow_count, st, reuse, prev_probes);
  else
    adapt_trace_form<256>(wave, layout, blocks, bytes, segs, pixel_begin, nseg, items, item_pixel, item_pose, poses,
                          out, status, shadow_count, st, reuse, prev_probes);
  return hipGetLastError();
}
hipError_t launch_adapt_init(int pixels, AdaptBufs b, AdaptEntry *entries, int *clear, int clear_words,
                             int *event_counts, hipStream_t st) {
  const int n = max(pixels, clear_words);
  if (n <= 0) return hipSuccess;
  adapt_init<<<grid(n), 256, 0, st>>>(pixels, b, entries, clear, clear_words, event_counts);
  return hipGetLastError();
}
hipError_t launch_adapt_decide(int n, const int *prev_level, const AdaptEntry *entries, const SegDev *segs,
                               const AdaptDev *ad, const int *pixel_begin, int nseg, AdaptBufs b, AdaptTerm *terms,
                               unsigned long long *keys, AdaptEntry *bisected, int *query_pixel, Pose *query_pose,
                               AdaptEntry *direct, int *head, int *level_ctr, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  adapt_decide<<<grid(n), 256, 0, st>>>(n, prev_level, entries, segs, ad, pixel_begin, nseg, b, terms, keys,
                                         bisected, query_pixel, query_pose, direct, head, level_ctr);
  return hipGetLastError();
}
hipError_t launch_adapt_children(int n, const AdaptEntry *bisected, const double *query_rad, AdaptEntry *next,
                                 hipStream_t st) {
  if (n <= 0) return hipSuccess;
  adapt_children<<<grid(n), 256, 0, st>>>(n, bisected, query_rad, next);
  return hipGetLastError();
}
hipError_t launch_adapt_counts(int pixels, const AdaptDev *ad, const int *pixel_begin, int nseg, AdaptBufs b,
                               int *fill, bool clear_exact, AdaptSegStats *stats, hipStream_t st) {
  static_assert(sizeof(AdaptSegStats) % sizeof(int) == 0, "statistics are cleared word by word");
  const int stat_words = nseg * (int)(sizeof(AdaptSegStats) / sizeof(int));
  const int n = max(pixels, stat_words);
  if (n <= 0) return hipSuccess;
  adapt_counts<<<grid(n), 256, 0, st>>>(pixels, ad, pixel_begin, nseg, b, fill, clear_exact ? 1 : 0,
                                         reinterpret_cast<int *>(stats), stat_words);
  return hipGetLastError();
}
hipError_t launch_adapt_knots(int n, const unsigned long long *keys, const AdaptTerm *terms, AdaptBufs b, int *fill,
                              unsigned *bucket, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  adapt_term_bucket<<<grid(n), 256, 0, st>>>(n, keys, terms, b, fill, bucket);
  adapt_place_knots<<<grid(n), 256, 0, st>>>(n, keys, terms, b, fill, bucket);
  return hipGetLastError();
}
hipError_t launch_adapt_exact_count(int pixels, const SegDev *segs, const AdaptDev *ad, const int *pixel_begin,
                                    int nseg, AdaptBufs b, hipStream_t st) {
  if (pixels <= 0) return hipSuccess;
  adapt_exact_count<<<grid(pixels), 256, 0, st>>>(pixels, segs, ad, pixel_begin, nseg, b);
  return hipGetLastError();
}
hipError_t launch_adapt_exact_list(int pixels, const SegDev *segs, const AdaptDev *ad, const int *pixel_begin,
                                   int nseg, AdaptBufs b, int *item_pixel, int *item_pose, hipStream_t st) {
  if (pixels <= 0) return hipSuccess;
  adapt_exact_list<<<grid(pixels), 256, 0, st>>>(pixels, segs, ad, pixel_begin, nseg, b, item_pixel, item_pose);
  return hipGetLastError();
}
hipError_t launch_adapt_fallback_flags(int pixels, const int *flags, int *out, hipStream_t st) {
  if (pixels <= 0) return hipSuccess;
  adapt_fallback_flags<<<grid(pixels), 256, 0, st>>>(pixels, flags, out);
  return hipGetLastError();
}
hipError_t launch_adapt_fallback_list(int pixels, const AdaptDev *, const int *pixel_begin, int nseg,
                                      const int *scan, AdaptBufs b, int *list, hipStream_t st) {
  if (pixels <= 0) return hipSuccess;
  adapt_fallback_list<<<grid(pixels), 256, 0, st>>>(pixels, pixel_begin, nseg, scan, b, list);
  return hipGetLastError();
}
hipError_t launch_adapt_segment_ends(int nseg, const int *pixel_begin, const int *scan, int *out, hipStream_t st) {
  if (nseg <= 0) return hipSuccess;
  adapt_segment_ends<<<grid(nseg), 256, 0, st>>>(nseg, pixel_begin, scan, out);
  return hipGetLastError();
}
hipError_t launch_adapt_segment_stats(int nseg, const int *pixel_begin, AdaptBufs b, AdaptSegStats *out,
                                      hipStream_t st) {
  if (nseg <= 0) return hipSuccess;
  adapt_segment_stats<<<dim3(kStatBlocks, nseg), 256, 0, st>>>(pixel_begin, b, out);
  return hipGetLastError();
}
hipError_t launch_adapt_gate(const SegDev *gate_segs, const int *gate_seg, const int *slots, int ngate,
                             const AdaptDev *ad, AdaptBufs b, double *errors, hipStream_t st) {
  if (ngate <= 0) return hipSuccess;
  adapt_gate<<<ngate * 8, 256, 0, st>>>(gate_segs, gate_seg, slots, ad, b, errors);
  return hipGetLastError();
}
size_t adapt_rates_flags(int pixels, int cells, int run) {
  if (run < kCodeCells) run = kCodeCells;
  return pixels <= 0 || cells <= 0 ? 0 : (size_t)grid(pixels) * 4 * (size_t)((cells + run - 1) / run);
}
hipError_t launch_adapt_rates(const SegDev *segs, const AdaptDev *ad, const int *pixel_begin, int nseg, int pixels,
                              AdaptBufs b, int c0, int c1, int *status, int *rare, size_t rare_capacity,
                              hipStream_t st, int run) {
  if (run < kCodeCells || run % kCodeCells != 0) return hipErrorInvalidValue;
  if (pixels <= 0 || c1 <= c0) return hipSuccess;
  if (adapt_rates_flags(pixels, c1 - c0, run) > rare_capacity) return hipErrorInvalidConfiguration;
  const dim3 g(grid(pixels), (unsigned)((c1 - c0 + run - 1) / run));
  if (g.y > 65535) return hipErrorInvalidConfiguration;
  adapt_rates<<<g, 256, 0, st>>>(segs, ad, pixel_begin, nseg, pixels, b, c0, c1, run, status, rare);
  adapt_rates_rare<<<g, 256, 0, st>>>(segs, ad, pixel_begin, nseg, pixels, b, c0, c1, run, rare);
  return hipGetLastError();
}
hipError_t launch_adapt_integrate(const SegDev *segs, const AdaptDev *ad, const int *block_begin, int nseg,
                                  int blocks, int pixels, AdaptBufs b, int c0, int c1, int *status, hipStream_t st) {
  if (blocks <= 0 || c1 <= c0) return hipSuccess;
  adapt_integrate<<<blocks, 256, 0, st>>>(segs, ad, block_begin, nseg, pixels, b, c0, c1, status);
  return hipGetLastError();
}

}  // namespace dvsamd
