// Launchers of the gfx950 kernels (dvs_kernels.hip), called by the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include "device_circuit.hpp"
#include "device_optics.hpp"

namespace dvsdev {

// A dispatch counts its grid in work-items as a 32-bit value: launches of
// 256-thread blocks stay below this many blocks.
constexpr long long kMaxBlocks256 = 0xffffffffLL / 256;
enum { MODE_CELLS = 0, MODE_FRAME = 1, MODE_DEPTH = 2 };

// One segment (one sensor on one scene view) of an engine step, resident in
// device memory; every multi-segment kernel finds its segment from blockIdx.
struct SegDev {
  DevScene scene;
  const dvs_optical_interval *ivs;  // this step's intervals of the segment
  const double *t0, *t1;            // this step's circuit cells of the segment
  const CellMaps *maps;             // their whole-cell receptor maps
  const Pose *poses;                // cell midpoint poses
  double *rad;                      // [cells][npix] radiance of this step
  dvs_pixel_state *states;
  CrossingEvent *slots;             // npix * max_events crossings
  int *counts;                      // npix crossing counts (segment's slice)
  CircuitParams params;
  PhotoParams photo;
  uint64_t seed;
  int iv_count, cells, npix, max_events;
  int pixel_offset;                 // first global pixel index of the segment
  int pad;
  const int *pixel_list;            // when set, item pixel k traces sensor pixel pixel_list[k]
};

// Last segment whose first block (or pixel) is <= b (begin[0] == 0).
template <class T>
__device__ __forceinline__ int segment_of(const T *begin, int nseg, T b) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (begin[m] <= b) lo = m; else hi = m - 1;
  }
  return lo;
}
}

namespace dvsamd {
using dvsdev::CellMaps;
using dvsdev::CircuitParams;
using dvsdev::CrossingEvent;
using dvsdev::DevScene;
using dvsdev::PhotoParams;
using dvsdev::Pose;
using dvsdev::SegDev;

// Multi-segment launches.  block_begin / pixel_begin hold nseg+1 prefix
// offsets (in blocks) of each segment's work; the grids are their totals.
// `wave`: the wave-packet traversal (every segment's scene must allow it);
// `lds` (per-lane stacks) is then not used.
hipError_t launch_trace_cells_multi(const SegDev *segs, const long long *block_begin, int nseg, long long blocks,
                                    size_t lds, int *status, unsigned long long *shadow_count, hipStream_t st,
                                    bool wave = false);
// cell_maps of every cell of a step; cell_begin holds nseg+1 prefix offsets
// of the segments' cells and charge their charge scales.
hipError_t launch_cell_maps(int cells, const double *t0, const double *t1, const int *cell_begin, int nseg,
                            const double *charge, CellMaps *out, hipStream_t st);
// Radiance -> photocurrent in place over `items` table items (prefix
// item_begin) and up to `cells` cells (at most 65535).  `rare` holds
// cell_rates_flags(items, cells) per-wave flags of the second phase.
size_t cell_rates_flags(int items, int cells);
hipError_t launch_cell_rates_multi(const SegDev *segs, const int *item_begin, int nseg, int items, int cells,
                                   int *status, int *rare, size_t rare_capacity, hipStream_t st);
hipError_t launch_circuit_multi(const SegDev *segs, const int *block_begin, int nseg, int blocks, int *status,
                                hipStream_t st);
hipError_t launch_gather_multi(const SegDev *segs, const int *pixel_begin, int nseg, int total_pixels,
                               const int *ends, unsigned long long key_lo, unsigned long long key_hi, double *time,
                               int *seg, int *pixel, int *polarity, unsigned long long *keys, unsigned *order,
                               int *key_flag, hipStream_t st);
hipError_t launch_event_poses_multi(const SegDev *segs, int events, const double *time, const int *seg,
                                    Pose *poses, hipStream_t st);
hipError_t launch_event_depth_multi(const SegDev *segs, int events, int samples, size_t lds, const int *seg,
                                    const int *pixel, const Pose *poses, float *depth, int *status, hipStream_t st,
                                    bool wave = false);
hipError_t launch_segment_keys(int events, const unsigned *order, const int *seg, unsigned long long *keys,
                               hipStream_t st);
hipError_t launch_permute_multi(const SegDev *segs, int events, const unsigned *order, const double *time,
                                const int *seg, const int *pixel, const int *polarity, const float *depth,
                                dvs_event *out, int *out_seg, hipStream_t st);
hipError_t launch_fill_int(int *p, int value, int n, hipStream_t st);
hipError_t launch_blur_accumulate(int npix, int exposures, const int *offsets, const float *frames,
                                  const double *weights, int wide, float *out, hipStream_t st);

hipError_t launch_cell_poses(int cells, const double *t0, const double *t1, const int *iv_index,
                             const dvs_optical_interval *ivs, Pose *poses, hipStream_t st);
hipError_t launch_state_poses(int n, const dvs_optical_state *states, Pose *poses, hipStream_t st);
hipError_t launch_trace(int mode, const DevScene &s, const Pose *poses, int cells, int npix, long long items,
                        const int *item_pixel, double *rad, float *img, float *dep, int *status,
                        unsigned long long *shadow_count, hipStream_t st);
hipError_t launch_init(int npix, const double *radiance, const PhotoParams &photo, const CircuitParams &p,
                       double threshold_std, double threshold_min, uint64_t seed, dvs_pixel_state *states,
                       int *status, hipStream_t st);
hipError_t launch_hapke(int64_t n, const float *ci, const float *ce, const float *cg, const float *alb,
                        const dvs_hapke &h, float *out, hipStream_t st);
hipError_t launch_poisson(int64_t n, const double *lambda, const double *u, double *out, hipStream_t st);
hipError_t launch_invert_hapke(int64_t n, const float *inc, const float *emi, const float *phase, const float *target,
                               const dvs_hapke &h, float *out, hipStream_t st);
hipError_t launch_advance(int64_t n, dvs_pixel_state *states, const dvs_pixel_params *params, const double *rates,
                          const double *t0, const double *t1, int capacity, double normal, int *counts,
                          dvs_crossing *events, hipStream_t st);
hipError_t launch_receptor(int64_t n, const double *u, const double *rate, const double *dt, double charge,
                           double *out, hipStream_t st);
hipError_t launch_noise(uint64_t seed, int64_t pixels, double lambda, int draws, double *poisson, double *normal,
                        hipStream_t st);
hipError_t launch_traversal_stats(const DevScene &s, const Pose *pose, int npix, unsigned long long *totals,
                                  hipStream_t st);
hipError_t launch_rays(const DevScene &s, int64_t n, const float *o, const float *d, const int *self, int shadow,
                       int *hit, int *tri, float *t, float *u, float *v, hipStream_t st);
}  // namespace dvsamd
