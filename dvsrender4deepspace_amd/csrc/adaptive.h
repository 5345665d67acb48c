// Adaptive optical refinement on gfx950: the OpticalSampler of the reference
// (include/dvs/optical_sampler.hpp, src/cuda/optical_sampler.cu) and the
// knot-interpolating circuit of EventSensor::process_adaptive
// (src/cuda/event_sensor.cu:88-150,306-374), restructured as wavefronts.
//
// The reference refines each pixel depth first (one warp per pixel, a
// 16-entry stack).  Here every unresolved segment of every pixel is one lane
// of a level-synchronous wavefront: a level decides all pending segments,
// compacts the resolved ones and the bisected ones by wave ballot, traces the
// bisected segments' new probes as one ray batch (with the midpoints of the
// halves' halves, so that the next level's bisections need no trace) and
// forms the next level.  The result equals the depth-first one:
//  * a segment's decision depends only on its own values and on the number
//    of left turns on its path (the depth-first stack height when it is
//    popped), which each entry carries;
//  * a pixel overflows its knot capacity in the depth-first walk exactly when
//    its final knot count exceeds the capacity.  The count's running lower
//    bound (1 + pending segments + knots written) only grows, so a pixel is
//    dropped as soon as the bound passes the capacity;
//  * the depth-first knot order is the order of the resolved segments' start
//    fractions.  A pixel's resolved segments partition its window, so a
//    segment's first knot follows the knots of the pixel's segments that
//    start before it: each segment is placed by counting those in a
//    per-pixel bucket, without a sort.
#pragma once
#include "kernels.h"

namespace dvsdev {

enum { ADAPT_ACCEPT = 0, ADAPT_EXACT = 1 };
enum { LAYOUT_PROBE = 0, LAYOUT_QUERY = 1, LAYOUT_CELL = 2 };

// Window of one segment in an adaptive step.
struct AdaptDev {
  double start, end, tolerance;
  unsigned long long cell_base;  // circuit cells the segment integrated before this window
  int capacity;          // knot and exact-cell capacity per pixel
  int accepted;          // 0 once the convergence gate rejected the window
  int cell_offset;       // first cell of the segment in the step's cell tables
  int fb_begin;          // first entry of the segment in the fallback pixel list
  int fb_count;          // fallback pixels of the segment
  int pad;
  const double *fb_rad;  // [cell - chunk start][fb_count] photocurrent of the fallback pixels (native trace)
};

// An unresolved time segment [a, b] of a pixel (fractions of the window) with
// its probe radiances and the number of left turns on its path.  lspec and
// rspec are the radiances at the midpoints of its two halves when they were
// traced ahead, with the queries of the bisection that made the segment
// (negative otherwise): if the segment bisects, its halves need no trace.
struct AdaptEntry {
  double a, b, left, mid, right;
  double lspec, rspec;
  int pixel, zeros;
};
// Queries per traced bisection: the midpoints of the two halves and,
// traced ahead, the midpoints of the halves' halves.
constexpr int kBisectQueries = 6;
// A resolved segment: accepted (two knots) or exact (one negative knot).
struct AdaptTerm {
  double a, b, mid, right;
  int pixel, kind;
};

// Per-step arrays over the step's global pixels.
struct AdaptBufs {
  const double *probes;  // [P][5] initial probe radiance
  int *bound;            // running lower bound of each pixel's final knot count
  int *flags;            // bit 0: whole-pixel native tracing (fallback)
  int *bisects;          // bisections per pixel (two probes each)
  int *exact_terms;      // exact segments per pixel
  int *kcount;           // knot count (0 for fallback pixels)
  int *koff;             // inclusive prefix of kcount
  double2 *knots;        // {fraction, radiance}; negative fraction: an exact segment ends here
  int *ecount;           // exact cells per pixel
  int *eoff;             // inclusive prefix of ecount
  int *ecell;            // exact cells (segment-local indices), ascending per pixel
  double *erad;          // their radiance
  int *fb_index;         // index of the pixel in its segment's fallback list, -1 otherwise
  const double *cell_frac;  // window fraction of every cell midpoint of the step
  double *rates;         // [cell - chunk start][pixel] photocurrent of knot pixels (escaped cells)
  uint64_t *codes;       // [(cell - chunk start) / 8][pixel] photocurrent codes of knot pixels
};

// Refinement counters of one window, cleared once per window.  Four head
// words, then the traced and untraced bisections of each level (level L >= 1
// at kRefineHeadWords + 2 (L - 1)).
//   head[0] resolved segments so far
//   head[1] exact segments so far
//   head[2] pixels whose knot-count bound has passed their capacity
//   head[3] unused
constexpr int kMaxRefineLevels = 64;
constexpr int kRefineHeadWords = 4;
constexpr int kRefineCounterWords = kRefineHeadWords + 2 * kMaxRefineLevels;

// OpticalRefinementStats of one segment (include/dvs/optical_sampler.hpp:16).
struct AdaptSegStats {
  unsigned long long traced_pixel_samples;
  int fallback_pixels, maximum_knots, exact_segments, native_cell_queries;
};

}  // namespace dvsdev

namespace dvsamd {
using dvsdev::AdaptBufs;
using dvsdev::kMaxRefineLevels;
using dvsdev::kRefineCounterWords;
using dvsdev::kRefineHeadWords;
using dvsdev::AdaptDev;
using dvsdev::AdaptEntry;
using dvsdev::AdaptSegStats;
using dvsdev::AdaptTerm;

hipError_t launch_adapt_probe_poses(int nseg, const SegDev *segs, const AdaptDev *ad, Pose *out,
                                    const Pose *prev_end, Pose *cur_end, int *reuse, hipStream_t st);
hipError_t launch_adapt_cell_frac(int cells, const double *t0, const double *t1, const int *cell_begin, int nseg,
                                  const AdaptDev *ad, double *frac, hipStream_t st);
// Radiance of `items` pixel samples: LAYOUT_PROBE covers the five probes of
// every pixel (items = 5 P; probe j of a pixel of segment s at pose
// poses[s * 5 + j], written to out[pixel * 5 + j]; where reuse[s] is set,
// probe 0 is the previous window's end probe, prev_probes[pixel * 5 + 4]);
// LAYOUT_QUERY item q is item_pixel[q] at poses[q]; LAYOUT_CELL item q is
// item_pixel[q] at poses[item_pose[q]].  `lds` is the block's LDS for 256
// lanes; blocks of `threads` lanes (64, 128 or 256) take a proportional share.
// `wave`: the wave-packet traversal (every segment's scene must allow it),
// which needs no LDS stack.
hipError_t launch_adapt_trace(int layout, const SegDev *segs, const int *pixel_begin, int nseg, int samples,
                              long long items, const int *item_pixel, const int *item_pose, const Pose *poses,
                              size_t lds, double *out, int *status, unsigned long long *shadow_count,
                              hipStream_t st, const int *reuse = nullptr, const double *prev_probes = nullptr,
                              int threads = 256, bool wave = false);
// Also clears clear[0, clear_words) and event_counts[0, pixels).
hipError_t launch_adapt_init(int pixels, AdaptBufs b, AdaptEntry *entries, int *clear, int clear_words,
                             int *event_counts, hipStream_t st);
// One level: resolved segments to terms; a bisected segment whose halves'
// midpoints were traced ahead writes its halves to `direct` (level_ctr[1]
// pairs), any other to `bisected` with kBisectQueries queries each
// (level_ctr[0]).  The level has n entries, or, with prev_level (the
// previous level's two counters), 2 (prev_level[0] + prev_level[1]) entries
// of which n is an upper bound.  head: the four head counter words.
hipError_t launch_adapt_decide(int n, const int *prev_level, const AdaptEntry *entries, const SegDev *segs,
                               const AdaptDev *ad, const int *pixel_begin, int nseg, AdaptBufs b, AdaptTerm *terms,
                               unsigned long long *keys, AdaptEntry *bisected, int *query_pixel, Pose *query_pose,
                               AdaptEntry *direct, int *head, int *level_ctr, hipStream_t st);
// The halves of the traced bisections, with their halves' midpoints.
hipError_t launch_adapt_children(int n, const AdaptEntry *bisected, const double *query_rad, AdaptEntry *next,
                                 hipStream_t st);
// Knot counts and fallback flags; also clears fill (P words), the exact
// counts when clear_exact, and the nseg statistics records.
hipError_t launch_adapt_counts(int pixels, const AdaptDev *ad, const int *pixel_begin, int nseg, AdaptBufs b,
                               int *fill, bool clear_exact, AdaptSegStats *stats, hipStream_t st);
// Knots of every knot pixel in time order from the resolved segments
// (terms, keys as written by adapt_decide); fill holds P zeroed counters and
// bucket at least as many words as knots.
hipError_t launch_adapt_knots(int n, const unsigned long long *keys, const AdaptTerm *terms, AdaptBufs b, int *fill,
                              unsigned *bucket, hipStream_t st);
hipError_t launch_adapt_exact_count(int pixels, const SegDev *segs, const AdaptDev *ad, const int *pixel_begin,
                                    int nseg, AdaptBufs b, hipStream_t st);
hipError_t launch_adapt_exact_list(int pixels, const SegDev *segs, const AdaptDev *ad, const int *pixel_begin,
                                   int nseg, AdaptBufs b, int *item_pixel, int *item_pose, hipStream_t st);
hipError_t launch_adapt_fallback_flags(int pixels, const int *flags, int *out, hipStream_t st);
hipError_t launch_adapt_fallback_list(int pixels, const AdaptDev *ad, const int *pixel_begin, int nseg,
                                      const int *scan, AdaptBufs b, int *list, hipStream_t st);
// out[s] = scan[pixel_begin[s + 1] - 1]: inclusive prefix at each segment's end.
hipError_t launch_adapt_segment_ends(int nseg, const int *pixel_begin, const int *scan, int *out, hipStream_t st);
// Adds each segment's statistics into out (cleared by adapt_counts).
hipError_t launch_adapt_segment_stats(int nseg, const int *pixel_begin, AdaptBufs b, AdaptSegStats *out,
                                      hipStream_t st);
// Convergence gate: gate g is segment gate_seg[g] with its probe pixels as
// gate_segs[g].pixel_list and their native cell radiance in gate_segs[g].rad;
// the error of list entry i goes to errors[g * 1024 + slots[g * 1024 + i]].
hipError_t launch_adapt_gate(const SegDev *gate_segs, const int *gate_seg, const int *slots, int ngate,
                             const AdaptDev *ad, AdaptBufs b, double *errors, hipStream_t st);
// Photocurrent of cells [c0, c1) of the knot pixels in runs of `run` cells
// per thread (a multiple of 8); `rare` holds adapt_rates_flags(pixels, c1 -
// c0, run) per-wave flags of the second phase.
size_t adapt_rates_flags(int pixels, int cells, int run);
hipError_t launch_adapt_rates(const SegDev *segs, const AdaptDev *ad, const int *pixel_begin, int nseg, int pixels,
                              AdaptBufs b, int c0, int c1, int *status, int *rare, size_t rare_capacity,
                              hipStream_t st, int run);
hipError_t launch_adapt_integrate(const SegDev *segs, const AdaptDev *ad, const int *block_begin, int nseg,
                                  int blocks, int pixels, AdaptBufs b, int c0, int c1, int *status, hipStream_t st);
// verify_knots (optical_sampler.cu:86-103): the probe poses of a window, and
// the knot-interpolation log errors of traced probes.
hipError_t launch_adapt_verify_poses(const SegDev *seg, double start, double end, int times, Pose *poses,
                                     hipStream_t st);
hipError_t launch_adapt_verify_errors(int items, int positions, int times, const int *item_pixel,
                                      const double *light, AdaptBufs b, const PhotoParams &photo, double *errors,
                                      hipStream_t st);

}  // namespace dvsamd
