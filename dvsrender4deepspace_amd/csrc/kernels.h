// Launchers of the gfx950 kernels (dvs_kernels.hip), called by the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include "device_circuit.hpp"
#include "device_optics.hpp"

namespace dvsdev {
enum { MODE_CELLS = 0, MODE_FRAME = 1, MODE_DEPTH = 2 };
}

namespace dvsamd {
using dvsdev::CircuitParams;
using dvsdev::CrossingEvent;
using dvsdev::DevScene;
using dvsdev::PhotoParams;
using dvsdev::Pose;

hipError_t launch_cell_poses(int cells, const double *t0, const double *t1, const int *iv_index,
                             const dvs_optical_interval *ivs, Pose *poses, hipStream_t st);
hipError_t launch_state_poses(int n, const dvs_optical_state *states, Pose *poses, hipStream_t st);
hipError_t launch_event_poses(int events, const double *times, const dvs_optical_interval *ivs, int iv_count,
                              Pose *poses, hipStream_t st);
hipError_t launch_trace(int mode, const DevScene &s, const Pose *poses, int cells, int npix, long long items,
                        const int *item_pixel, double *rad, float *img, float *dep, int *status,
                        unsigned long long *shadow_count, hipStream_t st);
hipError_t launch_init(int npix, const double *radiance, const PhotoParams &photo, const CircuitParams &p,
                       double threshold_std, double threshold_min, uint64_t seed, dvs_pixel_state *states,
                       int *status, hipStream_t st);
hipError_t launch_circuit(int npix, int c_begin, int c_count, const double *t0, const double *t1,
                          const double *rad, dvs_pixel_state *states, const CircuitParams &p,
                          const PhotoParams &photo, uint64_t seed, int max_events, CrossingEvent *slots,
                          int *counts, int *status, hipStream_t st);
hipError_t launch_gather(int npix, const int *ends, const int *counts, int max_events, const CrossingEvent *slots,
                         double *time, int *pixel, int *polarity, unsigned long long *keys, unsigned *order,
                         hipStream_t st);
hipError_t launch_permute(int events, int width, const unsigned *order, const double *time, const int *pixel,
                          const int *polarity, const float *depth, dvs_event *out, hipStream_t st);
hipError_t launch_hapke(int64_t n, const float *ci, const float *ce, const float *cg, const float *alb,
                        const dvs_hapke &h, float *out, hipStream_t st);
hipError_t launch_receptor(int64_t n, const double *u, const double *rate, const double *dt, double charge,
                           double *out, hipStream_t st);
hipError_t launch_noise(uint64_t seed, int64_t pixels, double lambda, int draws, double *poisson, double *normal,
                        hipStream_t st);
hipError_t launch_traversal_stats(const DevScene &s, const Pose *pose, int npix, unsigned long long *totals,
                                  hipStream_t st);
hipError_t launch_rays(const DevScene &s, int64_t n, const float *o, const float *d, const int *self, int shadow,
                       int *hit, int *tri, float *t, float *u, float *v, hipStream_t st);
}  // namespace dvsamd
