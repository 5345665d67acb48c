// gfx950 device pixel circuit (FP64) and counter-based noise.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dvs_hip.h"

namespace dvsdev {

// PixelParameters (include/dvs/pixel_model.hpp:7-10).
struct CircuitParams {
  double threshold_pos, threshold_neg, charge_scale, refractory_time;
  double reference_leak, reset_noise_std;
};
// PhotoParameters (include/dvs/optical_sensor.hpp:21-25).
struct PhotoParams {
  double electrons_per_radiance, dark_electrons, integration_step;
  int shot_noise;
  int pad;
};
struct CrossingEvent { double time; int polarity; int pad; };

// Exact constant-rate receptor solution (receptor_log_at, pixel_model.hpp:18-26).
// receptor_log_from takes x = exp(u) from the caller, which evaluates it once
// for all pieces that start at u.
__host__ __device__ __forceinline__ double receptor_log_from(double u, double x, double rate, double dt, double charge) {
  if (dt <= 0) return u;
  rate = fmax(0.0, rate);
  if (rate == 0) return u - log1p(x * dt / charge);
  const double a = exp(-rate * dt / charge), b = -expm1(-rate * dt / charge);
  return -log(a / x + b / rate);
}
__host__ __device__ __forceinline__ double receptor_log_at(double u, double rate, double dt, double charge) {
  if (dt <= 0) return u;
  return receptor_log_from(u, exp(u), rate, dt, charge);
}

// ---------------------------------------------------------------------------
// Noise.  Philox4x64-10 (Salmon, Moraes, Dror, Shaw, SC'11) keyed by
// (seed, tag ^ stream), counter (block, pixel, word, 0).  Each draw starts a
// fresh block.  Shot noise is keyed per circuit cell: the draw of a pixel's
// n-th cell since initialisation uses word = n and blocks 0, 1, ... of its
// own, so it can be drawn by any thread at any time.  The inversion branch
// (lambda < 10, one uniform) instead takes a 32-bit uniform: half n mod 2 of
// word (n / 2) mod 4 of block n / 8 of the pixel's shared shot stream (word
// 0), so eight consecutive cells share one Philox block.  (The inversion
// resolves no probability near 2^-32 below lambda = 10.)  Reset noise (word
// 0) advances a per-pixel block counter, as
// its draws occur at crossings.  Batching, launch shape and other requests
// cannot change any of these streams.
// ---------------------------------------------------------------------------
enum { kStreamShot = 0, kStreamThreshold = 1, kStreamReset = 2, kStreamShotShared = 3 };
constexpr uint64_t kStreamTag = 0x4456535F4E4F4953ULL;

__device__ __forceinline__ void philox4x64(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3, uint64_t k0,
                                           uint64_t k1, uint64_t out[4]) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B97F4A7C15ULL; k1 += 0xBB67AE8584CAA73BULL; }
    const uint64_t lo0 = 0xD2E7470EE14C6C93ULL * c0, hi0 = __umul64hi(0xD2E7470EE14C6C93ULL, c0);
    const uint64_t lo1 = 0xCA5A826395121157ULL * c2, hi1 = __umul64hi(0xCA5A826395121157ULL, c2);
    const uint64_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
__device__ __forceinline__ double uniform_of(uint64_t x) {  // open interval (0, 1), 53 bits
  return ((double)(x >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
// The shared shot-noise block of cells 8 block .. 8 block + 7
// (block = cell serial / kShotCellsPerBlock).
constexpr int kShotCellsPerBlock = 8;
__device__ __forceinline__ void shot_block(uint64_t seed, uint64_t pixel, uint64_t block, uint64_t out[4]) {
  philox4x64(block, pixel, 0, 0, seed, kStreamTag ^ kStreamShotShared, out);
}
// The inversion uniform of cell `serial` in its shared block w.
__device__ __forceinline__ double shot_uniform(const uint64_t w[4], uint64_t serial) {
  const int j = (int)((serial >> 1) & 3);  // selected without a dynamic index so the block stays in registers
  const uint64_t x = j == 0 ? w[0] : j == 1 ? w[1] : j == 2 ? w[2] : w[3];
  const uint32_t h = (serial & 1) ? (uint32_t)(x >> 32) : (uint32_t)x;
  return ((double)h + 0.5) * (1.0 / 4294967296.0);
}
// Sequential-search inversion of Poisson(lambda), 0 < lambda < 10, from one
// uniform u (Devroye 1986, ch. X.3).
__device__ __forceinline__ double poisson_inversion(double lambda, double u) {
  double p = exp(-lambda), F = p;
  int k = 0;
  while (u > F && k < 1000) { ++k; p *= lambda / k; F += p; }
  return k;
}
// The count poisson_inversion(lambda, u) returns, for 0 < lambda < 10,
// mostly without FP64 arithmetic: the cumulative sums run in float and are
// compared with u outside a band of relative half-width kScreenBand around
// each sum.  The float sums stay within 1.1e-5 of the double ones (expf
// within 2 ulp; lambda's float rounding moves exp(-lambda) by at most
// 2^-24 lambda < 6e-7; each further term adds at most 2.5 ulp: the
// reciprocal, two products and lambda's rounding; kScreenTerms terms), so
// a decision outside the band is the double search's.  Inside the band, or
// past kScreenTerms terms, the double search itself decides.
constexpr float kScreenBand = 4e-5f;
constexpr int kScreenTerms = 30;
__device__ __forceinline__ double poisson_inversion_screened(double lambda, double u) {
  const float lf = (float)lambda, uf = (float)u;
  float p = expf(-lf), F = p;
  for (int k = 0; k < kScreenTerms; ++k) {
    if (uf < F * (1.0f - kScreenBand)) return k;  // u <= F_k for certain
    if (uf <= F * (1.0f + kScreenBand)) break;    // too close to call in float
    p *= lf * __builtin_amdgcn_rcpf((float)(k + 1));
    F += p;
  }
  return poisson_inversion(lambda, u);
}
struct NoiseStream {
  uint64_t seed, pixel, stream, block, word;
  uint64_t buf[4];
  int used;
  __device__ NoiseStream(uint64_t s, uint64_t p, uint64_t st, uint64_t b, uint64_t w = 0)
      : seed(s), pixel(p), stream(st), block(b), word(w), used(4) {}
  __device__ double uniform() {  // open interval (0, 1), 53 bits
    if (used == 4) { philox4x64(block++, pixel, word, 0, seed, kStreamTag ^ stream, buf); used = 0; }
    // select without a dynamic index so the block stays in registers
    const uint64_t x = used == 0 ? buf[0] : used == 1 ? buf[1] : used == 2 ? buf[2] : buf[3];
    ++used;
    return uniform_of(x);
  }
  // Box-Muller, one normal per fresh block.
  __device__ double normal() {
    used = 4;
    const double u1 = uniform(), u2 = uniform();
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
  // Poisson(lambda): sequential-search inversion below 10 (Devroye 1986,
  // ch. X.3), PTRS transformed rejection (Hoermann 1993) up to 1e6, and the
  // reference's rounded normal approximation above (photon_noise.cuh:4-9).
  // A NaN lambda gives NaN (the rejection loop would never accept it).
  __device__ double poisson(double lambda) {
    if (lambda != lambda) return lambda;
    if (lambda <= 0) return 0;
    if (lambda > 1e6) return fmax(0.0, floor(lambda + sqrt(lambda) * normal() + .5));
    used = 4;
    if (lambda < 10) return poisson_inversion(lambda, uniform());
    const double slam = sqrt(lambda), loglam = log(lambda);
    const double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2);
    for (;;) {
      const double U = uniform() - 0.5, V = uniform();
      const double us = 0.5 - fabs(U);
      const double k = floor((2 * a / us + b) * U + lambda + 0.43);
      if (us >= 0.07 && V <= vr) return k;
      if (k < 0 || (us < 0.013 && V > us)) continue;
      if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -lambda + k * loglam - lgamma(k + 1)) return k;
    }
  }
};

// The shot-noise draws outside the inversion branch, from the cell's own
// stream.  They are rare and need far more registers than the rest of the
// photocurrent (lgamma, Box-Muller), so they run in kernels of their own.
__device__ inline double shot_stream_draw(uint64_t seed, uint64_t pixel, uint64_t cell_serial, double lambda) {
  NoiseStream shot(seed, pixel, kStreamShot, 0, cell_serial);
  return shot.poisson(lambda);
}

// Photocurrent of one circuit cell (generate_pixels, event_sensor.cu:42-43):
// radiance -> electrons/s, with shot noise rate = Poisson(rate * dt) / dt
// (poisson_electrons, photon_noise.cuh:4-9) keyed by the cell serial, in two
// phases.  cell_rate_first decides every cell whose draw is the inversion
// branch (or needs no draw) and returns -lambda (< 0) for a cell whose draw
// needs the cell's own stream (lambda >= 10); cell_rate_rest finishes it.
// `shared` is shot_block(seed, pixel, cell_serial / kShotCellsPerBlock)
// (unused without noise).  cell_rate_first follows the cell maps below.
__device__ inline double cell_rate_rest(double marked, double dt, uint64_t seed, uint64_t pixel, uint64_t cell_serial) {
  return shot_stream_draw(seed, pixel, cell_serial, -marked) / dt;
}

// ---------------------------------------------------------------------------
// advance_pixel (include/dvs/pixel_model.hpp:35-102) in FP64: refractory hold,
// leakage-extremum split, analytic crossing root without leakage and a
// 48-step bisection with it, reset at the crossing voltage.
// ---------------------------------------------------------------------------
struct PixelRegs {
  double receptor_log, reference_log, refractory_until, threshold_pos, threshold_neg;
  int reset_pending;
};
template <class Normal>
__host__ __device__ inline int advance_pixel(PixelRegs &s, const CircuitParams &p, double rate, double t0, double t1,
                                    CrossingEvent *out, int capacity, Normal &&normal) {
  if (t1 <= t0) return 0;
  int count = 0;
  double cursor = t0, u = s.receptor_log;
  // Each pass advances one piece: the rest of a refractory hold, or the
  // free-running stretch up to the cell end (or the leakage extremum).  The
  // receptor solution is evaluated at one place per pass so that lanes of a
  // wave in different phases share it; the operations on each lane are those
  // of the reference's sequence (a hold ending inside a cell continues with
  // the free-running stretch in the next pass).
  while (cursor < t1) {
    const bool hold = cursor < s.refractory_until;
    if (!hold && s.reset_pending) {
      if (p.reset_noise_std > 0) s.reference_log += p.reset_noise_std * normal();
      s.reset_pending = 0;
    }
    const double x = exp(u);
    double end = 0, c0 = 0, boundary = 0, dt;
    if (hold) {
      end = fmin(t1, s.refractory_until);
      dt = end - cursor;
    } else {
      const double remaining = t1 - cursor;
      c0 = u - s.reference_log;
      boundary = remaining;
      if (p.reference_leak > 0) {
        const double target = rate + p.charge_scale * p.reference_leak;
        if (x > target && target > rate) {
          double extremum;
          if (rate > 0) extremum = -p.charge_scale / rate * log((1.0 / target - 1.0 / rate) / (1.0 / x - 1.0 / rate));
          else extremum = p.charge_scale * (1.0 / target - 1.0 / x);
          if (extremum > 1e-15 && extremum < boundary) boundary = extremum;
        }
      }
      dt = boundary;
    }
    // dt > 0 here: a hold ends after the cursor and a free piece is the rest
    // of the cell or an extremum above 1e-15
    const double un = receptor_log_from(u, x, rate, dt, p.charge_scale);
    if (hold) {
      u = un;
      cursor = end;
      s.reference_log = u;
      if ((cursor < t1 || cursor == s.refractory_until) && cursor >= s.refractory_until && s.reset_pending) {
        if (p.reset_noise_std > 0) s.reference_log += p.reset_noise_std * normal();
        s.reset_pending = 0;
      }
      continue;
    }
    const double ub = un;
    const double cb = ub - s.reference_log + p.reference_leak * boundary;
    int polarity = -1;
    double level = 0;
    if (c0 >= s.threshold_pos) { polarity = 1; level = s.threshold_pos; }
    else if (c0 <= -s.threshold_neg) { polarity = 0; level = -s.threshold_neg; }
    else if (cb >= s.threshold_pos) { polarity = 1; level = s.threshold_pos; }
    else if (cb <= -s.threshold_neg) { polarity = 0; level = -s.threshold_neg; }
    if (polarity < 0) {
      u = ub;
      s.reference_log -= p.reference_leak * boundary;
      cursor += boundary;
      continue;
    }
    if (count == capacity) return -1;
    double lo = 0, hi = boundary;
    if (c0 >= s.threshold_pos || c0 <= -s.threshold_neg) {
      hi = 0;
    } else if (p.reference_leak == 0) {
      const double target_inverse = exp(-(s.reference_log + level));
      const double initial_inverse = exp(-u);
      if (rate > 0) {
        const double stationary_inverse = 1.0 / rate;
        hi = -p.charge_scale / rate * log1p((target_inverse - initial_inverse) / (initial_inverse - stationary_inverse));
      } else {
        hi = p.charge_scale * (target_inverse - initial_inverse);
      }
      hi = fmin(boundary, fmax(0.0, hi));
    } else {
      for (int j = 0; j < 48; ++j) {
        const double mid = (lo + hi) * .5;
        const double c = receptor_log_from(u, x, rate, mid, p.charge_scale) - s.reference_log + p.reference_leak * mid;
        if (polarity ? c >= level : c <= level) hi = mid; else lo = mid;
      }
    }
    u = receptor_log_from(u, x, rate, hi, p.charge_scale);
    cursor += hi;
    out[count].time = cursor;
    out[count].polarity = polarity;
    ++count;
    s.reference_log = u;
    s.refractory_until = cursor + p.refractory_time;
    s.reset_pending = 1;
  }
  s.receptor_log = u;
  return count;
}

// ---------------------------------------------------------------------------
// The same circuit without reference leakage, integrated in the inverse
// receptor variable y = exp(-u).  Over a piece of length dt at photocurrent r
// the constant-rate solution above is affine in y:
//   y' = exp(-r dt / C) y + (-expm1(-r dt / C)) / r     (r > 0)
//   y' = y + dt / C                                      (r = 0)
// so a cell costs one such map instead of exp(u), log and log1p, and the
// comparator tests u - ref >= theta_pos and u - ref <= -theta_neg become
// y <= exp(-(ref + theta_pos)) and y >= exp(-(ref - theta_neg)).  The analytic
// crossing root of the reference is already written in these inverse
// voltages.  y moves from cell to cell by whole-cell maps and is kept in the
// pixel state, so the receptor trajectory does not depend on how steps,
// chunks or crossings split the time line; pieces inside a cell (hold ends,
// crossings) start from partial maps.  The results are the reference's up to
// rounding (the reference's own u-space sequence rounds differently).
// ---------------------------------------------------------------------------
struct InverseFactors {
  double a, q;
};
__device__ __forceinline__ InverseFactors inverse_factors(double rate, double dt, double charge) {
  const double e = -rate * dt / charge;
  return InverseFactors{exp(e), -expm1(e) / rate};
}
__device__ __forceinline__ double inverse_map(double y, double rate, double dt, double charge) {
  if (dt <= 0) return y;
  if (rate == 0) return y + dt / charge;
  const InverseFactors f = inverse_factors(rate, dt, charge);
  return f.a * y + f.q;
}

// One circuit cell as the inverse circuit reads it: its window and its
// whole-cell maps y' = a y + q for the photocurrents shot noise makes common
// there: map[0] = {1, dt / C} for zero (1 y = y exactly, so this is
// inverse_map's y + dt / C) and map[k] for k = 1, 2 electrons (rate[k - 1] =
// k / dt, the value the photocurrent kernels write).  They depend on the cell
// alone, so one small kernel evaluates them for every pixel, with the
// operations of inverse_map; a pixel whose photocurrent is none of these
// evaluates its own.  A cell without length gets a NaN map[0], which no
// range test passes.
constexpr int kMappedElectrons = 2;
struct CellMaps {
  double t0, t1;
  double rate[kMappedElectrons];
  double map[kMappedElectrons + 1][2];
};
constexpr int kCellWords = 4 + 2 * (kMappedElectrons + 1);  // doubles per CellMaps
static_assert(sizeof(CellMaps) == kCellWords * sizeof(double), "CellMaps is read as words");
static_assert(kCellWords % 2 == 0, "CellMaps entries stay 16-byte aligned");
__device__ __forceinline__ CellMaps cell_maps(double t0, double t1, double charge) {
  CellMaps m;
  const double dt = t1 - t0;
  m.t0 = t0;
  m.t1 = t1;
  m.map[0][0] = dt > 0 ? 1.0 : __builtin_nan("");
  m.map[0][1] = dt / charge;
#pragma unroll
  for (int k = 0; k < kMappedElectrons; ++k) {
    m.rate[k] = (double)(k + 1) / dt;
    const InverseFactors f = inverse_factors(m.rate[k], dt, charge);
    m.map[k + 1][0] = f.a;
    m.map[k + 1][1] = f.q;
  }
  return m;
}
// First phase of a cell's photocurrent (see cell_rate_rest) from the words
// of its CellMaps entry: dt = t1 - t0, and the counts one and two divided
// by dt (the entry's rate[0] and rate[1], the same quotients).
__device__ __forceinline__ double cell_rate_first(double light, const PhotoParams &ph, double dt, double one_per_dt,
                                                  double two_per_dt, uint64_t cell_serial, const uint64_t shared[4]) {
  const double rate = light * ph.electrons_per_radiance + ph.dark_electrons;
  if (!ph.shot_noise) return rate;
  const double lambda = rate * dt;
  if (lambda != lambda) return lambda;  // NaN photocurrent: the circuit reports it
  if (lambda <= 0) return 0.0 / dt;
  if (lambda < 10) {
    const double k = poisson_inversion_screened(lambda, shot_uniform(shared, cell_serial));
    if (k == 0 && dt > 0) return 0.0;  // 0 / dt
    if (k == 1) return one_per_dt;
    if (k == 2) return two_per_dt;
    return k / dt;
  }
  return -lambda;
}
// Photocurrent codes.  With shot noise a cell's photocurrent is k / dt for
// the drawn electron count k, so the adaptive photocurrent passes a byte per
// cell to the circuit instead of a double: k itself below kCodeEscape, whose
// photocurrent is 0 for k = 0 and k / dt otherwise (the quotients
// cell_rate_first returns, dt = t1 - t0 of the cell's window), and
// kCodeEscape for every other value, which travels as a double beside the
// codes (no shot noise, a draw left to the second phase, a NaN, dt <= 0).
// Eight consecutive cells of a pixel share a 64-bit word, cell j in byte j.
constexpr int kCodeEscape = 255;
constexpr int kCodeCells = 8;
__device__ __forceinline__ int cell_code_first(double light, const PhotoParams &ph, double dt, double one_per_dt,
                                               double two_per_dt, uint64_t cell_serial, const uint64_t shared[4],
                                               double &escape) {
  const double rate = light * ph.electrons_per_radiance + ph.dark_electrons;
  if (!ph.shot_noise) { escape = rate; return kCodeEscape; }
  const double lambda = rate * dt;
  if (lambda != lambda) { escape = lambda; return kCodeEscape; }
  if (lambda <= 0) {
    if (dt > 0) return 0;
    escape = 0.0 / dt;
    return kCodeEscape;
  }
  if (lambda < 10) {
    const double k = poisson_inversion_screened(lambda, shot_uniform(shared, cell_serial));
    if (dt > 0 && k < kCodeEscape) return (int)k;
    escape = k == 0 && dt > 0 ? 0.0 : k == 1 ? one_per_dt : k == 2 ? two_per_dt : k / dt;
    return kCodeEscape;
  }
  escape = -lambda;
  return kCodeEscape;
}
// The photocurrent of code k < kCodeEscape in a cell of length dt > 0.
__device__ __forceinline__ double code_rate(int k, double dt) { return k == 0 ? 0.0 : (double)k / dt; }

// inverse_map over the whole cell, dt = t1 - t0 > 0, rate >= 0.
__device__ __forceinline__ double inverse_map_cell(double y, double rate, double dt, double charge,
                                                   const CellMaps &m) {
  if (rate == 0) return y + m.map[0][1];
#pragma unroll
  for (int k = 0; k < kMappedElectrons; ++k)
    if (rate == m.rate[k]) return m.map[k + 1][0] * y + m.map[k + 1][1];
  const InverseFactors f = inverse_factors(rate, dt, charge);
  return f.a * y + f.q;
}

struct InverseRegs {
  double y, reference_log, refractory_until, threshold_pos, threshold_neg;
  double y_pos, y_neg;  // exp(-(reference_log + threshold_pos)), exp(-(reference_log - threshold_neg))
  int reset_pending;
  __device__ void set_reference(double ref) {
    reference_log = ref;
    y_pos = exp(-(ref + threshold_pos));
    y_neg = exp(-(ref - threshold_neg));
  }
};

// One cell [t0, t1].  A free-running cell whose start and end stay inside
// (y_pos, y_neg), or a cell inside a refractory hold, only moves y; other
// cells run the reference's pieces.  Within a hold the reference tracks the
// reference voltage to the receptor; that value is only read when the hold
// ends, so whole held cells leave it to the caller.
typedef __attribute__((address_space(1))) CrossingEvent *GlobalEvents;
// The reference's pieces of a cell with a crossing, a hold end or a pending
// reset; y0 and yb are y at t0 and t1.
__device__ __forceinline__ int cell_pieces(InverseRegs &s, const CircuitParams &p, double rate, double t0, double t1,
                                           double y0, double yb, GlobalEvents out, int capacity, NoiseStream &reset) {
  int count = 0;
  double cursor = t0, y = y0;
  while (cursor < t1) {
    const bool hold = cursor < s.refractory_until;
    if (!hold && s.reset_pending) {
      if (p.reset_noise_std > 0) s.set_reference(s.reference_log + p.reset_noise_std * reset.normal());
      s.reset_pending = 0;
    }
    if (hold) {
      const double end = fmin(t1, s.refractory_until);
      y = end == t1 ? yb : inverse_map(y, rate, end - cursor, p.charge_scale);
      cursor = end;
      double ref = -log(y);
      if ((cursor < t1 || cursor == s.refractory_until) && cursor >= s.refractory_until && s.reset_pending) {
        if (p.reset_noise_std > 0) ref += p.reset_noise_std * reset.normal();
        s.reset_pending = 0;
      }
      s.set_reference(ref);
      continue;
    }
    // free piece [cursor, t1]: comparator at both ends, as the reference
    int polarity = -1;
    double target = 0;
    bool at_start = true;
    if (y <= s.y_pos) { polarity = 1; target = s.y_pos; }
    else if (y >= s.y_neg) { polarity = 0; target = s.y_neg; }
    else {
      at_start = false;
      if (yb <= s.y_pos) { polarity = 1; target = s.y_pos; }
      else if (yb >= s.y_neg) { polarity = 0; target = s.y_neg; }
    }
    if (polarity < 0) break;
    if (count == capacity) return -1;
    double hi = 0;
    if (!at_start) {
      if (rate > 0) hi = -p.charge_scale / rate * log1p((target - y) / (y - 1.0 / rate));
      else hi = p.charge_scale * (target - y);
      hi = fmin(t1 - cursor, fmax(0.0, hi));
    }
    y = inverse_map(y, rate, hi, p.charge_scale);
    cursor += hi;
    out[count].time = cursor;
    out[count].polarity = polarity;
    ++count;
    s.set_reference(-log(y));
    s.refractory_until = cursor + p.refractory_time;
    s.reset_pending = 1;
  }
  return count;
}

__device__ __forceinline__ int advance_cell_inverse(InverseRegs &s, const CircuitParams &p, double rate,
                                                   const CellMaps &m, GlobalEvents out, int capacity,
                                                   NoiseStream &reset) {
  const double t0 = m.t0, t1 = m.t1;
  if (t1 <= t0) return 0;
  rate = fmax(0.0, rate);
  const double y0 = s.y, yb = inverse_map_cell(y0, rate, t1 - t0, p.charge_scale, m);
  s.y = yb;
  if (t0 >= s.refractory_until && !s.reset_pending) {
    if (y0 > s.y_pos && y0 < s.y_neg && yb > s.y_pos && yb < s.y_neg) return 0;
  } else if (s.refractory_until > t1) {
    return 0;
  }
  return cell_pieces(s, p, rate, t0, t1, y0, yb, out, capacity, reset);
}

// Global-memory views of the circuit's arrays.  Through generic pointers the
// compiler emits flat loads, whose waits also drain every prefetched load.
typedef const __attribute__((address_space(1))) double *GlobalDoubles;

// The value of x held by lane `lane` (wave-uniform) of the wave.
__device__ __forceinline__ double from_lane(double x, int lane) {
  const long long b = __double_as_longlong(x);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane), lo = __builtin_amdgcn_readlane((int)b, lane);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Photocurrent of cells [c_begin, c_end) at rates[(c - c_begin) * stride],
// loaded kCellPrefetch cells ahead of the circuit: a free-running cell is far
// shorter than an HBM round trip, so one wave needs that many loads in flight.
// The CellMaps of those cells are the same for every lane: the wave loads
// their kCellPrefetch * kCellWords doubles as two coalesced words per lane,
// one block ahead, and stages them in LDS for the block.  Every global load
// is issued unconditionally (from a clamped, valid address), so the waits
// before a block cover only the loads of the block before.
constexpr int kCellPrefetch = 8;
constexpr int kBlockWords = kCellPrefetch * kCellWords;
static_assert(kBlockWords <= 128, "a block's maps fit two words per lane");

// The circuit of one pixel over cells [c_begin, c_end): photocurrent of cell c
// at rates[(c - c_begin) * stride], windows and maps at maps[c] (cell_maps),
// crossings appended at out[count...] up to `capacity` in all.  Without
// leakage the inverse form above runs and the state carries y; with it the
// reference's sequence runs on u (t0, t1 hold the windows).  Returns the new
// count, -1 past capacity, -2 on a non-finite photocurrent.  Every lane of a
// 256-thread block calls it with the same cells (inactive lanes, past the
// segment's pixels, with active = false and a valid rates pointer): the lanes
// share the maps' loads.
//
// A block of kCellPrefetch cells first runs in bulk: without branches, every
// cell is its tabulated map and the comparator's range test.  A lane whose
// block has no crossing, no hold end, no pending reset and tabulated
// photocurrents only (nearly all of them) keeps that result, which is what
// the cell-by-cell path gives; the others run the block again cell by cell.
//
// CODED: the photocurrents arrive as codes (cell_code_first), word g of the
// pixel's cells c_begin + 8 g .. c_begin + 8 g + 7 at codes[g * code_stride], and
// escaped cells at rates[(c - c_begin) * stride] as above; coded = false
// takes every cell as escaped (the pixel has no codes).  The decoded
// photocurrent is the double the uncoded path would read.
typedef const __attribute__((address_space(1))) uint64_t *GlobalWords;
template <bool CODED>
__device__ inline int circuit_cells(dvs_pixel_state &st, const CircuitParams &p, uint64_t seed, uint64_t pixel,
                                    bool active, const double *rates, size_t stride, const uint64_t *codes,
                                    size_t code_stride, bool coded, const double *t0, const double *t1, const CellMaps *maps,
                                    int c_begin, int c_end, CrossingEvent *out_events, int count, int capacity) {
  const GlobalDoubles R = (GlobalDoubles)rates, T0 = (GlobalDoubles)t0, T1 = (GlobalDoubles)t1;
  const GlobalWords C = (GlobalWords)codes;
  const GlobalEvents out = (GlobalEvents)out_events;
  NoiseStream reset(seed, pixel, kStreamReset, st.reset_block);
  if (c_end <= c_begin) return count;
  // code of cell c (CODED): its byte of the word of its group of eight
  auto code_of = [&](int c) -> int {
    const uint64_t w = coded ? C[(size_t)((c - c_begin) / kCodeCells) * code_stride] : ~0ull;
    return (int)((w >> (8 * ((c - c_begin) % kCodeCells))) & 0xff);
  };
  if (p.reference_leak > 0) {
    if (!active) return 0;
    auto normal = [&]() { return reset.normal(); };
    PixelRegs s{st.receptor_log, st.reference_log, st.refractory_until, st.threshold_pos, st.threshold_neg,
                st.reset_pending};
    double next = 0;
    if (!CODED) next = R[0];
    for (int c = c_begin; c < c_end; ++c) {
      double rate = next;
      if (CODED) {
        const int k = code_of(c);
        rate = k == kCodeEscape ? R[(size_t)(c - c_begin) * stride] : code_rate(k, T1[c] - T0[c]);
      } else if (c + 1 < c_end) {
        next = R[(size_t)(c + 1 - c_begin) * stride];
      }
      if (!isfinite(rate)) return -2;
      const int emitted = advance_pixel(s, p, rate, T0[c], T1[c], out_events + count, capacity - count, normal);
      if (emitted < 0) return -1;
      count += emitted;
    }
    st.receptor_log = s.receptor_log;
    st.receptor_inverse = exp(-s.receptor_log);
    st.reference_log = s.reference_log;
    st.refractory_until = s.refractory_until;
    st.reset_pending = s.reset_pending;
    st.reset_block = (uint32_t)reset.block;
    return count;
  }
  __shared__ __attribute__((aligned(16))) double block_words[4][2 * 64];  // one block's CellMaps per wave
  double *L = block_words[threadIdx.x >> 6];
  const CircuitParams pr = p;  // in registers: stores to the crossings may alias p
  InverseRegs s;
  s.y = st.receptor_inverse;
  s.refractory_until = st.refractory_until;
  s.threshold_pos = st.threshold_pos;
  s.threshold_neg = st.threshold_neg;
  s.reset_pending = st.reset_pending;
  s.set_reference(st.reference_log);
  int code = 0;
  const int lane = (int)(threadIdx.x & 63);
  const GlobalDoubles W = (GlobalDoubles)maps;  // CellMaps as words
  const long long last_word = (long long)c_end * kCellWords - 1;
  // Clamped, branch-free addresses: every lane (inactive ones hold a valid
  // rates pointer) loads in every block, so the compiler can count the loads
  // in flight exactly.
  auto word = [&](int cb, int k) { return W[min((long long)cb * kCellWords + k, last_word)]; };
  auto rate_of = [&](int c) { return R[(size_t)(min(c, c_end - 1) - c_begin) * stride]; };
  // the code word of the block at cb (CODED; all escapes without codes)
  auto codes_of = [&](int cb) -> uint64_t {
    const uint64_t w = C[(size_t)((min(cb, c_end - 1) - c_begin) / kCodeCells) * code_stride];
    return coded ? w : ~0ull;
  };
  static_assert(kCellPrefetch == kCodeCells, "a block is one code word");
  // One block of cells from cb: integrates it with the photocurrents `use`
  // (CODED: the code word `use_codes`) and map words u0/u1 (loaded one block
  // earlier) while loading the next block's into `fill` (`fill_codes`) and
  // f0/f1.  The loop below alternates two register sets, so no value moves
  // between them: a move would wait for the load.
  auto block = [&](const double (&use)[kCellPrefetch], uint64_t use_codes, double u0, double u1,
                   double (&fill)[kCellPrefetch], uint64_t &fill_codes, double &f0, double &f1,
                   int cb) __attribute__((always_inline)) {
    if (CODED) {
      fill_codes = codes_of(cb + kCellPrefetch);
    } else {
#pragma unroll
      for (int j = 0; j < kCellPrefetch; ++j) fill[j] = rate_of(cb + kCellPrefetch + j);
    }
    f0 = word(cb + kCellPrefetch, lane);
    f1 = word(cb + kCellPrefetch, 64 + lane);
    // stage this block's maps; LDS operations of a wave complete in order
    __builtin_amdgcn_wave_barrier();
    L[lane] = u0;
    L[64 + lane] = u1;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int n = min(kCellPrefetch, c_end - cb);
    // bulk pass, in phases so that each phase's LDS reads are in flight
    // together: the tabulated photocurrents, the chosen maps, the y chain
    const double y0 = s.y;
    double y = y0;
    bool ok = active, in_range = y0 > s.y_pos && y0 < s.y_neg;
    double2 r[kCellPrefetch], aq[kCellPrefetch];
#pragma unroll
    for (int j = 0; j < kCellPrefetch; ++j) r[j] = *reinterpret_cast<const double2 *>(L + j * kCellWords + 2);
    int k[kCellPrefetch];
    if (CODED) {
      // every code of the block's cells tabulated (0, 1 or 2) in one test:
      // no byte with a bit above the lowest two, nor both of them (3); k =
      // code & 3 then picks the map (an untabulated code's index reads
      // words of the next cell in the staged block: the bulk result is then
      // not used)
      static_assert(kMappedElectrons == 2, "codes 0..2 are tabulated");
      const uint64_t live = n >= kCellPrefetch ? ~0ull : (1ull << (8 * n)) - 1;
      const uint64_t w = use_codes & live;
      ok = ok && (w & 0xFCFCFCFCFCFCFCFCull) == 0 && (w & (w >> 1) & 0x0101010101010101ull) == 0;
#pragma unroll
      for (int j = 0; j < kCellPrefetch; ++j) k[j] = (int)((use_codes >> (8 * j)) & 3);
    } else {
#pragma unroll
      for (int j = 0; j < kCellPrefetch; ++j) {
        const double rate = use[j];
        const bool one = rate == r[j].x, two = rate == r[j].y;
        k[j] = one ? 1 : two ? 2 : 0;
        ok = ok && (j >= n || one || two || rate == 0);
      }
    }
#pragma unroll
    for (int j = 0; j < kCellPrefetch; ++j) aq[j] = *reinterpret_cast<const double2 *>(L + j * kCellWords + 4 + 2 * k[j]);
    if (n == kCellPrefetch) {  // a whole block: no cell past the range
#pragma unroll
      for (int j = 0; j < kCellPrefetch; ++j) {
        y = aq[j].x * y + aq[j].y;
        in_range = in_range && y > s.y_pos && y < s.y_neg;
      }
    } else {
#pragma unroll
      for (int j = 0; j < kCellPrefetch; ++j) {
        if (j < n) {
          y = aq[j].x * y + aq[j].y;
          in_range = in_range && y > s.y_pos && y < s.y_neg;
        }
      }
    }
    const double first_t0 = L[0], last_t1 = L[(n - 1) * kCellWords + 1];
    const bool free = first_t0 >= s.refractory_until && !s.reset_pending;
    ok = ok && y == y && (free ? in_range : s.refractory_until > last_t1);
    if (ok) {
      s.y = y;
    } else if (active) {
      // cell by cell from the block's start
#pragma clang loop unroll(full)
      for (int j = 0; j < kCellPrefetch; ++j) {
        if (j < n && active) {
          CellMaps m;
          const double *w = L + j * kCellWords;
          m.t0 = w[0];
          m.t1 = w[1];
#pragma unroll
          for (int e = 0; e < kMappedElectrons; ++e) m.rate[e] = w[2 + e];
#pragma unroll
          for (int e = 0; e <= kMappedElectrons; ++e) {
            m.map[e][0] = w[2 + kMappedElectrons + 2 * e];
            m.map[e][1] = w[3 + kMappedElectrons + 2 * e];
          }
          double rate;
          if (CODED) {
            const int code = (int)((use_codes >> (8 * j)) & 0xff);
            rate = code == kCodeEscape ? R[(size_t)(cb + j - c_begin) * stride] : code_rate(code, m.t1 - m.t0);
          } else {
            rate = use[j];
          }
          if (!isfinite(rate)) {
            code = -2;
            active = false;
          } else {
            const int emitted = advance_cell_inverse(s, pr, rate, m, out + count, capacity - count, reset);
            if (emitted < 0) {
              code = -1;
              active = false;
            } else {
              count += emitted;
            }
          }
        }
      }
    }
  };
  double ra[kCellPrefetch], rb[kCellPrefetch];
  uint64_t ca = 0, cbw = 0;
  if (CODED) {
    ca = codes_of(c_begin);
  } else {
#pragma unroll
    for (int j = 0; j < kCellPrefetch; ++j) ra[j] = rate_of(c_begin + j);
  }
  double a0 = word(c_begin, lane), a1 = word(c_begin, 64 + lane), b0 = 0, b1 = 0;
  for (int cb = c_begin; cb < c_end; cb += 2 * kCellPrefetch) {
    block(ra, ca, a0, a1, rb, cbw, b0, b1, cb);
    if (cb + kCellPrefetch < c_end) block(rb, cbw, b0, b1, ra, ca, a0, a1, cb + kCellPrefetch);
  }
  if (code < 0) return code;
  st.receptor_inverse = s.y;
  st.receptor_log = -log(s.y);
  // inside a hold the reference voltage follows the receptor
  st.reference_log = s.refractory_until > T1[c_end - 1] ? st.receptor_log : s.reference_log;
  st.refractory_until = s.refractory_until;
  st.reset_pending = s.reset_pending;
  st.reset_block = (uint32_t)reset.block;
  return count;
}

}  // namespace dvsdev
