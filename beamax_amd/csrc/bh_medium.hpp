// Grid-valued sound speed: value, gradient and Hessian of the tensor-product
// interpolant of beamax/utils/interp.py:101-270 (make_c_function_from_grid):
// index coordinate u = (x - origin) / spacing, i0 = floor(u), t = u - i0,
// weights over the offsets (0, 1) (n-linear) or (-1, 0, 1, 2) (cubic
// B-spline on prefiltered coefficients), sample indices mapped by the
// boundary rule.  Derivatives are those of the same expression in x (floor
// has zero derivative): d/dx = (1/spacing) d/dt of the weights.
#pragma once

#include "bh_internal.hpp"

namespace {

enum { kSpeedLinear = 0, kSpeedBspline3 = 1 };
enum { kBndNearest = 0, kBndWrap = 1, kBndReflect = 2, kBndMirror = 3 };

struct SpeedDev {
  int ndim, method, boundary;
  long long N[3];
  double spacing[3], origin[3];
  const double* values;   // grid values (linear) or spline coefficients (bspline3), C order
  double lam;             // absorption coefficient of the amplitude equation
};

__device__ inline long long map_index(long long i, long long n, int boundary) {
  if (n == 1) return 0;
  switch (boundary) {
    case kBndWrap: {
      long long r = i % n;
      return r < 0 ? r + n : r;
    }
    case kBndReflect: {
      const long long period = 2 * n;
      long long r = i % period;
      if (r < 0) r += period;
      return r < n ? r : period - 1 - r;
    }
    case kBndMirror: {
      const long long period = 2 * n - 2;
      long long r = i % period;
      if (r < 0) r += period;
      return r < n ? r : period - r;
    }
    default:
      return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
  }
}

// weights w[k], first and second t-derivatives over the K offsets
template <int K>
__device__ inline void speed_weights(double t, double* w, double* w1, double* w2) {
  if constexpr (K == 2) {
    w[0] = 1.0 - t;
    w[1] = t;
    w1[0] = -1.0;
    w1[1] = 1.0;
    w2[0] = 0.0;
    w2[1] = 0.0;
  } else {
    const double s = 1.0 - t, t2 = t * t, t3 = t2 * t;
    w[0] = s * s * s / 6.0;
    w[1] = (3.0 * t3 - 6.0 * t2 + 4.0) / 6.0;
    w[2] = (-3.0 * t3 + 3.0 * t2 + 3.0 * t + 1.0) / 6.0;
    w[3] = t3 / 6.0;
    w1[0] = -0.5 * s * s;
    w1[1] = (9.0 * t2 - 12.0 * t) / 6.0;
    w1[2] = (-9.0 * t2 + 6.0 * t + 3.0) / 6.0;
    w1[3] = 0.5 * t2;
    w2[0] = s;
    w2[1] = 3.0 * t - 2.0;
    w2[2] = -3.0 * t + 1.0;
    w2[3] = t;
  }
}

// c(x), grad c (D) and Hessian (D x D, row-major)
template <int D, int K>
__device__ inline void speed_eval_k(const SpeedDev& S, const double* x, double& c, double* g,
                                    double* H) {
  constexpr int off0 = (K == 2) ? 0 : -1;
  long long i0[D];
  double w[D][K], w1[D][K], w2[D][K];
  for (int a = 0; a < D; ++a) {
    const double u = (x[a] - S.origin[a]) / S.spacing[a];
    const double f = floor(u);
    i0[a] = (long long)f;
    speed_weights<K>(u - f, w[a], w1[a], w2[a]);
  }
  c = 0.0;
  for (int a = 0; a < D; ++a) {
    g[a] = 0.0;
    for (int b = 0; b < D; ++b) H[a * D + b] = 0.0;
  }
  int kk[D];
  for (int a = 0; a < D; ++a) kk[a] = 0;
  constexpr int ncorner = (D == 1) ? K : (D == 2 ? K * K : K * K * K);
  for (int corner = 0; corner < ncorner; ++corner) {
    int rem = corner;
    long long flat = 0;
    for (int a = D - 1; a >= 0; --a) {
      kk[a] = rem % K;
      rem /= K;
    }
    for (int a = 0; a < D; ++a)
      flat = flat * S.N[a] + map_index(i0[a] + off0 + kk[a], S.N[a], S.boundary);
    const double v = S.values[flat];
    double wall = 1.0;
    for (int a = 0; a < D; ++a) wall *= w[a][kk[a]];
    c += wall * v;
    for (int a = 0; a < D; ++a) {
      double pa = w1[a][kk[a]];
      for (int b = 0; b < D; ++b)
        if (b != a) pa *= w[b][kk[b]];
      g[a] += pa * v;
      for (int b = a; b < D; ++b) {
        double pab;
        if (b == a) {
          pab = w2[a][kk[a]];
          for (int e = 0; e < D; ++e)
            if (e != a) pab *= w[e][kk[e]];
        } else {
          pab = w1[a][kk[a]] * w1[b][kk[b]];
          for (int e = 0; e < D; ++e)
            if (e != a && e != b) pab *= w[e][kk[e]];
        }
        H[a * D + b] += pab * v;
      }
    }
  }
  for (int a = 0; a < D; ++a) {
    g[a] /= S.spacing[a];
    for (int b = a; b < D; ++b) {
      H[a * D + b] /= S.spacing[a] * S.spacing[b];
      H[b * D + a] = H[a * D + b];
    }
  }
}

template <int D>
__device__ inline void speed_eval(const SpeedDev& S, const double* x, double& c, double* g,
                                  double* H) {
  if (S.method == kSpeedBspline3) speed_eval_k<D, 4>(S, x, c, g, H);
  else speed_eval_k<D, 2>(S, x, c, g, H);
}

}  // namespace
