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
  // sqrt(disc) and 1/A from the hardware estimates with one Newton step
  // each (relative error below 1e-13 for operands inside [1e-200, 1e200];
  // outside it the exact forms); the ends are widened by one point plus
  // 1e-12 of the magnitudes that form them, so the interval always holds
  // every point whose exponent is <= E
  double r, iA;
  if (disc > 1e-200 && disc < 1e200 && A > 1e-200 && A < 1e200) {
    const double y = __builtin_amdgcn_rsq(disc);
    const double r0 = disc * y;
    r = fma(0.5 * y, fma(-r0, r0, disc), r0);
    const double q = __builtin_amdgcn_rcp(A);
    iA = q * fma(-A, q, 2.0);
  } else {
    r = sqrt(disc);
    iA = 1.0 / A;
  }
  const double ulo = (-Bq - r) * iA, uhi = (-Bq + r) * iA;
  const double slack = 1e-12 * (fabs(Bq) + r) * iA;
  const double flo = floor(ulo - slack) - 1.0, fhi = ceil(uhi + slack) + 1.0;
  // flo and fhi are whole numbers: clamp to [0, n] and [-1, n - 1]
  lo = (int)fmin(fmax(flo, 0.0), (double)n);
  hi = (int)fmax(fmin(fhi, (double)(n - 1)), -1.0);
}

// LDS writes of a wave visible to all of its lanes (no workgroup barrier)
__device__ inline void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One (record, 64-point group) pair of the FP32-split sums in closed form
// along the group: lane l evaluates the point pref + l v, where
//   cycles(l)   = A + l (B + l C)                          (FP64, exact algebra)
// l is an integer, so A, B and C are stored reduced modulo one (|cycles| <
// 8200 over two groups) and A carries kPhaseBias = 1.5 2^20: every value of
// the polynomial then lies in [2^20, 2^21), where the FP64 spacing is 2^-32,
// and the low 32 bits of the result are the fractional cycle in units of
// 2^-32.  Their top 23 bits under the exponent of 1.0 (one v_alignbit_b32)
// give 1 + frac in FP32 for v_cos_f32, which reduces revolutions itself; the
// extra 2^-24 in the bias turns that truncation into rounding.
//   exponent(l) = Ce (l - l0)^2 + emin   (log2 units; = delta^T Qi delta)
// The envelope is expanded about its minimum along the line, so the FP32
// evaluation adds two non-negative terms (no cancellation).  It is stored as
// s = sqrt(Ce) and -l0 s, so that u = l s - l0 s is one FMA and the exponent
// another; negated (the argument of exp2), with log2(2|A|) folded in:
//   term(l) = cos(2 pi cycles(l)) * 2^(nlog - u(l)^2).
struct PairQ {
  double A, B, C;
  float s, nl0s, nlog, pad;
};

constexpr double kPhaseBias = 1572864.0 + 0x1p-24;

template <int D>
__device__ inline int sym_off(int a, int b) {   // packed index of (a < b)
  return D + (a * (2 * D - a - 1)) / 2 + (b - a - 1);
}

template <int D>
__device__ inline PairQ pair_quad(const RecS<D>& R, const double* pref, const double* v) {
  double d[D], Qd[D], Qv[D], Ed[D], Ev[D];
  for (int a = 0; a < D; ++a) d[a] = pref[a] - R.x[a];
  for (int a = 0; a < D; ++a) {
    double qd = 0.0, qv = 0.0, ed = 0.0, ev = 0.0;
    for (int b = 0; b < D; ++b) {
      const double qr = a == b ? R.qr[a] : 0.5 * R.qr[a < b ? sym_off<D>(a, b) : sym_off<D>(b, a)];
      const double qi =
          a == b ? (double)R.qi[a] : 0.5 * (double)R.qi[a < b ? sym_off<D>(a, b) : sym_off<D>(b, a)];
      qd += qr * d[b];
      qv += qr * v[b];
      ed += qi * d[b];
      ev += qi * v[b];
    }
    Qd[a] = qd;
    Qv[a] = qv;
    Ed[a] = ed;
    Ev[a] = ev;
  }
  PairQ c;
  double A = R.ph0, B = 0.0, C = 0.0, Ae = 0.0, Be = 0.0, Ce = 0.0;
  for (int a = 0; a < D; ++a) {
    A += d[a] * (R.k[a] + Qd[a]);
    B += v[a] * (R.k[a] + 2.0 * Qd[a]);
    C += v[a] * Qv[a];
    Ae += d[a] * Ed[a];
    Be += 2.0 * v[a] * Ed[a];
    Ce += v[a] * Ev[a];
  }
  c.A = __builtin_amdgcn_fract(A) + kPhaseBias;
  c.B = B - rint(B);
  c.C = C - rint(C);
  double emin = Ae, l0 = 0.0;
  if (Ce > 1e-300) {
    // l0 = -Be / (2 Ce): hardware reciprocal + one Newton step (l0 only
    // places the expansion point; emin is then consistent with it)
    const double x2 = 2.0 * Ce;
    double r = __builtin_amdgcn_rcp(x2);
    r = r * fma(-x2, r, 2.0);
    l0 = -Be * r;
    emin = fma(Be, l0, fma(Ce * l0, l0, Ae));   // exponent at l0
  }
  const float sc = __builtin_amdgcn_sqrtf((float)(Ce > 0.0 ? Ce : 0.0));
  c.s = sc;
  c.nl0s = (float)(-l0 * (double)sc);
  c.nlog = (float)(-emin) + __builtin_amdgcn_logf(R.amp);
  c.pad = 0.f;
  return c;
}

// the same pair at lanes l (pa) and l + 64 (pb): two adjacent groups of a run
__device__ inline void eval_quad2(const PairQ& c, double ld, float lf, float& pa, float& pb) {
  const double cyc0 = fma(ld, fma(ld, c.C, c.B), c.A);
  const double ld1 = ld + 64.0;
  const double cyc1 = fma(ld1, fma(ld1, c.C, c.B), c.A);
  const float fr0 = __uint_as_float(
      __builtin_amdgcn_alignbit(0x7Fu, (unsigned)__double_as_longlong(cyc0), 9u));
  const float fr1 = __uint_as_float(
      __builtin_amdgcn_alignbit(0x7Fu, (unsigned)__double_as_longlong(cyc1), 9u));
  const float u0 = fmaf(lf, c.s, c.nl0s);
  const float u1 = fmaf(64.f, c.s, u0);
  const float a0 = fmaf(-u0, u0, c.nlog);
  const float a1 = fmaf(-u1, u1, c.nlog);
  pa = fmaf(__builtin_amdgcn_cosf(fr0), __builtin_amdgcn_exp2f(a0), pa);
  pb = fmaf(__builtin_amdgcn_cosf(fr1), __builtin_amdgcn_exp2f(a1), pb);
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
