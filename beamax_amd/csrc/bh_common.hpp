// Shared device/host helpers for the beamax_amd HIP library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string>

#include "../../include/beamax_amd.h"

// ---------------------------------------------------------------------------
// error reporting: every entry point returns an int status; the message of the
// last failure on the calling thread is kept for bh_last_error().
// ---------------------------------------------------------------------------
namespace bh {
void set_error(const char* fmt, ...);
void clear_error();
}  // namespace bh

#define BH_FAIL(code, ...)            \
  do {                                \
    bh::set_error(__VA_ARGS__);       \
    return (code);                    \
  } while (0)

#define BH_HIP(call)                                                            \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess)                                                       \
      BH_FAIL(BH_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
              __FILE__, __LINE__);                                              \
  } while (0)

#define BH_LAUNCH_CHECK()                                                        \
  do {                                                                           \
    hipError_t e_ = hipGetLastError();                                           \
    if (e_ != hipSuccess)                                                        \
      BH_FAIL(BH_ERR_HIP, "kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), \
              __FILE__, __LINE__);                                               \
  } while (0)

// ---------------------------------------------------------------------------
// complex FP64 arithmetic on double2 (interleaved re, im)
// ---------------------------------------------------------------------------
struct cd {
  double re, im;
};

__host__ __device__ inline cd mkc(double r, double i = 0.0) { return cd{r, i}; }
__host__ __device__ inline cd operator+(cd a, cd b) { return cd{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline cd operator-(cd a, cd b) { return cd{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline cd operator-(cd a) { return cd{-a.re, -a.im}; }
__host__ __device__ inline cd operator*(cd a, cd b) {
  return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ inline cd operator*(double s, cd a) { return cd{s * a.re, s * a.im}; }
__host__ __device__ inline cd operator*(cd a, double s) { return cd{s * a.re, s * a.im}; }
// Smith's algorithm for complex division (robust scaling).
__host__ __device__ inline cd operator/(cd a, cd b) {
  if (fabs(b.re) >= fabs(b.im)) {
    double r = b.im / b.re, den = b.re + b.im * r;
    return cd{(a.re + a.im * r) / den, (a.im - a.re * r) / den};
  } else {
    double r = b.re / b.im, den = b.re * r + b.im;
    return cd{(a.re * r + a.im) / den, (a.im * r - a.re) / den};
  }
}
__host__ __device__ inline cd operator/(cd a, double s) { return cd{a.re / s, a.im / s}; }
__host__ __device__ inline double cabs_(cd a) { return hypot(a.re, a.im); }
// principal square root (branch cut on the negative real axis, sign of im kept)
__host__ __device__ inline cd csqrt_(cd z) {
  double r = hypot(z.re, z.im);
  if (r == 0.0) return cd{0.0, z.im};
  if (z.re >= 0.0) {
    double t = sqrt(0.5 * (r + z.re));
    return cd{t, z.im / (2.0 * t)};
  } else {
    double t = sqrt(0.5 * (r - z.re));
    return cd{fabs(z.im) / (2.0 * t), copysign(t, z.im)};
  }
}

// Python-style floor modulus for integers (result in [0, m)).
__host__ __device__ inline long long pmod(long long a, long long m) {
  long long r = a % m;
  return r < 0 ? r + m : r;
}
// floor division for integers
__host__ __device__ inline long long fdiv(long long a, long long b) {
  long long q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
  return q;
}

namespace bh {
constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 6.28318530717958647692;
constexpr double kLog2e = 1.44269504088896340736;
}  // namespace bh
