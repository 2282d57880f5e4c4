// Point runs ("segments") and the closed-form envelope intervals shared by
// the culled summation kernels and the transposed forward sum.  A run is k
// points p0 + u v, u = 0..n-1, packed into 64-point groups (padding slots have
// col = -1).  Along a run a beam's envelope exponent delta^T Q delta is a
// quadratic in u, so the points where it stays below a bound form one index
// interval, found here in closed form.
#pragma once

#include "bh_beamcore.hpp"

namespace {

struct SegDev {
  long long nseg, ngroups, S;
  const double* p0;
  const double* v;
  const int32_t* n;
  const int32_t* g0;
  const int32_t* grp_seg;
  const double* pts;
  const int32_t* col;
};

// log2-scaled envelope matrix (full, symmetric) from a record
template <int D>
__device__ inline void env_matrix(const RecS<D>& R, double (&Q)[D][D]) {
  int q = 0;
  for (int a = 0; a < D; ++a) Q[a][a] = (double)R.qi[q++];
  for (int a = 0; a < D; ++a)
    for (int b = a + 1; b < D; ++b) {
      Q[a][b] = 0.5 * (double)R.qi[q];
      Q[b][a] = Q[a][b];
      ++q;
    }
}
template <int D>
__device__ inline void env_matrix(const RecD<D>& R, double (&Q)[D][D]) {
  const double s = 0.5 * R.w * bh::kLog2e;
  int q = 0;
  for (int a = 0; a < D; ++a) Q[a][a] = s * R.mi[q++];
  for (int a = 0; a < D; ++a)
    for (int b = a + 1; b < D; ++b) {
      Q[a][b] = 0.5 * s * R.mi[q];
      Q[b][a] = Q[a][b];
      ++q;
    }
}
template <int D>
__device__ inline bool rec_zero(const RecS<D>& R) { return !(R.amp > 0.f); }
template <int D>
__device__ inline bool rec_zero(const RecD<D>& R) { return !(R.amp > 0.0); }

// index interval [lo, hi] of segment k where exponent <= E; empty -> lo > hi
template <int D>
__device__ inline void seg_interval(const SegDev& G, long long k, const double* x,
                                    const double (&Q)[D][D], double E, int& lo, int& hi) {
  const int n = G.n[k];
  double d0[D], v[D], Qv[D];
  for (int a = 0; a < D; ++a) {
    d0[a] = G.p0[k * D + a] - x[a];
    v[a] = G.v[k * D + a];
  }
  for (int a = 0; a < D; ++a) {
    double s = 0.0;
    for (int b = 0; b < D; ++b) s += Q[a][b] * v[b];
    Qv[a] = s;
  }
  double A = 0.0, Bq = 0.0, C = 0.0;
  for (int a = 0; a < D; ++a) {
    A += v[a] * Qv[a];
    Bq += d0[a] * Qv[a];
    double s = 0.0;
    for (int b = 0; b < D; ++b) s += Q[a][b] * d0[b];
    C += d0[a] * s;
  }
  // exponent(u) = A u^2 + 2 Bq u + C on u in [0, n-1]
  if (!(A > 1e-300)) {
    if (C <= E || !(C == C)) { lo = 0; hi = n - 1; } else { lo = 1; hi = 0; }
    return;
  }
  const double disc = Bq * Bq - A * (C - E);
  if (!(disc >= 0.0)) {
    if (disc == disc) { lo = 1; hi = 0; } else { lo = 0; hi = n - 1; }   // NaN: keep all
    return;
  }
  const double r = sqrt(disc);
  const double ulo = (-Bq - r) / A, uhi = (-Bq + r) / A;
  const double flo = floor(ulo) - 1.0, fhi = ceil(uhi) + 1.0;
  lo = flo < 0.0 ? 0 : (flo > (double)(n - 1) ? n : (int)flo);
  hi = fhi > (double)(n - 1) ? n - 1 : (fhi < 0.0 ? -1 : (int)fhi);
}

int check_segments(const bh_segments* g, int ndim) {
  if (!g) BH_FAIL(BH_ERR_ARG, "segments: null");
  if (g->nseg <= 0 || g->ngroups <= 0 || g->S <= 0)
    BH_FAIL(BH_ERR_ARG, "segments: nseg=%lld ngroups=%lld S=%lld", (long long)g->nseg,
            (long long)g->ngroups, (long long)g->S);
  if (!g->seg_p0 || !g->seg_v || !g->seg_n || !g->seg_g0 || !g->grp_seg || !g->pts || !g->col)
    BH_FAIL(BH_ERR_ARG, "segments: null array");
  if (ndim < 1 || ndim > 3) BH_FAIL(BH_ERR_ARG, "segments: ndim=%d", ndim);
  return BH_OK;
}

SegDev to_segdev(const bh_segments* g) {
  SegDev d;
  d.nseg = g->nseg;
  d.ngroups = g->ngroups;
  d.S = g->S;
  d.p0 = g->seg_p0;
  d.v = g->seg_v;
  d.n = g->seg_n;
  d.g0 = g->seg_g0;
  d.grp_seg = g->grp_seg;
  d.pts = g->pts;
  d.col = g->col;
  return d;
}

}  // namespace
