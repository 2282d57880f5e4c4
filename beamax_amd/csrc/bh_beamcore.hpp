// Beam state, evaluation records and the per-term evaluators shared by the
// summation kernels (included by several translation units; everything is
// internal to each).
#pragma once

#include "bh_internal.hpp"

namespace {

template <int D>
struct Sym {
  static constexpr int N = D * (D + 1) / 2;
};


// ---------------------------------------------------------------------------
// closed-form propagation (gb_solvers.py:194-422)
// ---------------------------------------------------------------------------
template <int D>
struct State {
  double x[D];
  double p[D];
  cd M[D][D];
  cd A;
};

template <int D>
__device__ inline void hom_diag_state(const double* x0, const double* p0, const cd* alpha, cd a0,
                                      double c0, double t, State<D>& S) {
  double np2 = 0.0;
  for (int a = 0; a < D; ++a) np2 += p0[a] * p0[a];
  const double np = sqrt(np2);
  for (int a = 0; a < D; ++a) {
    S.x[a] = x0[a] + (c0 * (p0[a] / np)) * t;
    S.p[a] = p0[a];
  }
  const double ct = c0 * t;
  const double np3 = np * np * np;
  cd Ai[D], Aiu[D], vAi[D];
  cd scal = mkc(1.0);
  for (int a = 0; a < D; ++a) {
    const cd A = mkc(1.0) + (ct * alpha[a]) / np;
    Ai[a] = mkc(1.0) / A;
    const double u = (-ct * p0[a]) / np3;
    Aiu[a] = Ai[a] * u;
    vAi[a] = (p0[a] * alpha[a]) * Ai[a];
    scal = scal + vAi[a] * u;
  }
  for (int a = 0; a < D; ++a)
    for (int b = 0; b < D; ++b) {
      cd y = -(Aiu[a] * vAi[b]) / scal;
      if (a == b) y = Ai[a] + y;
      S.M[a][b] = alpha[a] * y;
    }
  if constexpr (D == 3) {
    const double n = np;
    const cd a1 = alpha[0], a2 = alpha[1], a3 = alpha[2];
    const double tsc = c0 * t / (n * n * n * n);
    const cd t1 = (c0 * t) * (a2 * a3) + (a2 + a3) * n;
    const cd t2 = (c0 * t) * (a1 * a3) + (a1 + a3) * n;
    const cd t3 = (c0 * t) * (a1 * a2) + (a1 + a2) * n;
    const cd den = mkc(1.0) + tsc * ((p0[0] * p0[0]) * t1 + (p0[1] * p0[1]) * t2 +
                                     (p0[2] * p0[2]) * t3);
    S.A = a0 / csqrt_(den);
  } else if constexpr (D == 2) {
    cd ap = mkc(0.0);
    for (int a = 0; a < D; ++a) ap = ap + (p0[a] * p0[a]) * alpha[D - 1 - a];
    ap = ap / np3;
    S.A = a0 / csqrt_(mkc(1.0) + ct * ap);
  } else {
    S.A = a0;
  }
}

// ---------------------------------------------------------------------------
// evaluation records
// ---------------------------------------------------------------------------
template <int D>
struct RecS {  // FP32-split
  double x[D];
  double k[D];
  double qr[Sym<D>::N];
  double ph0;
  float qi[Sym<D>::N];
  float amp;   // 2|A|
};

template <int D>
struct RecD {  // FP64
  double x[D];
  double p[D];
  double mr[Sym<D>::N];
  double mi[Sym<D>::N];
  double w;
  double amp;
  double ang;
};

__device__ inline double safe_angle(cd A) {
  const double re = (A.re == 0.0 && A.im == 0.0) ? 1e-12 : A.re;
  return atan2(A.im, re);
}

template <int D>
__device__ inline void make_rec(const State<D>& S, double w, RecS<D>& R) {
  for (int a = 0; a < D; ++a) {
    R.x[a] = S.x[a];
    R.k[a] = w * S.p[a] * (1.0 / bh::kTwoPi);
  }
  const double sr = w * (1.0 / (2.0 * bh::kTwoPi));
  const double si = 0.5 * w * bh::kLog2e;
  int q = 0;
  for (int a = 0; a < D; ++a) {
    R.qr[q] = sr * S.M[a][a].re;
    R.qi[q] = (float)(si * S.M[a][a].im);
    ++q;
  }
  for (int a = 0; a < D; ++a)
    for (int b = a + 1; b < D; ++b) {
      R.qr[q] = sr * (S.M[a][b].re + S.M[b][a].re);
      R.qi[q] = (float)(si * (S.M[a][b].im + S.M[b][a].im));
      ++q;
    }
  // 2|A| in FP32: |A|^2 clamped to [1e-100, 1e100] (the FP32 result is 0 or
  // infinite outside) and its root from the hardware estimate with one
  // Newton step
  {
    const double a2 = fmin(fmax(S.A.re * S.A.re + S.A.im * S.A.im, 1e-100), 1e100);
    const double y = __builtin_amdgcn_rsq(a2);
    const double r0 = a2 * y;
    R.amp = (float)(2.0 * fma(0.5 * y, fma(-r0, r0, a2), r0));
  }
  // arg A in FP32 (a few 1e-8 cycles, below the FP32 rounding of the reduced
  // phase that the cosine takes); (0, 0) -> 0 as safe_angle
  const float are = (float)S.A.re, aim = (float)S.A.im;
  R.ph0 = (double)(atan2f(aim, (are == 0.f && aim == 0.f) ? 1e-12f : are) *
                   (float)(1.0 / bh::kTwoPi));
}

template <int D>
__device__ inline void make_rec(const State<D>& S, double w, RecD<D>& R) {
  for (int a = 0; a < D; ++a) {
    R.x[a] = S.x[a];
    R.p[a] = S.p[a];
  }
  int q = 0;
  for (int a = 0; a < D; ++a) {
    R.mr[q] = S.M[a][a].re;
    R.mi[q] = S.M[a][a].im;
    ++q;
  }
  for (int a = 0; a < D; ++a)
    for (int b = a + 1; b < D; ++b) {
      R.mr[q] = S.M[a][b].re + S.M[b][a].re;
      R.mi[q] = S.M[a][b].im + S.M[b][a].im;
      ++q;
    }
  R.w = w;
  R.amp = 2.0 * cabs_(S.A);
  R.ang = safe_angle(S.A);
}

template <int D>
__device__ inline void zero_rec(RecS<D>& R) {
  for (int a = 0; a < D; ++a) { R.x[a] = 0.0; R.k[a] = 0.0; }
  for (int q = 0; q < Sym<D>::N; ++q) { R.qr[q] = 0.0; R.qi[q] = 0.f; }
  R.ph0 = 0.0;
  R.amp = 0.f;
}
template <int D>
__device__ inline void zero_rec(RecD<D>& R) {
  for (int a = 0; a < D; ++a) { R.x[a] = 0.0; R.p[a] = 0.0; }
  for (int q = 0; q < Sym<D>::N; ++q) { R.mr[q] = 0.0; R.mi[q] = 0.0; }
  R.w = 0.0;
  R.amp = 0.0;
  R.ang = 0.0;
}

// delta = s - x with the optional minimum-image wrap (gb/core.py:16-36, 290-294)
template <int D, bool PER>
__device__ inline void make_delta(const double* s, const double* x, const double* dom,
                                  const int* per, double* dl) {
  for (int a = 0; a < D; ++a) {
    double dv = s[a] - x[a];
    if (PER) {
      if (per[a]) dv = dv - dom[a] * rint(dv / dom[a]);
    }
    dl[a] = dv;
  }
}

// FP32-split phase of one term: fr = fract(cycles) with
// cycles = sum_a dl_a (k_a + sum_b Q_ab dl_b) + ph0 formed in FP64 (Q
// symmetric), and the envelope exponent arg (log2 units) in FP32.
template <int D>
__device__ inline void split_phase(const RecS<D>& R, const double* dl, float& fr, float& arg) {
  double cyc = R.ph0;
  if constexpr (D == 1) {
    cyc = fma(dl[0], fma(R.qr[0], dl[0], R.k[0]), cyc);
    const float f0 = (float)dl[0];
    arg = -f0 * (R.qi[0] * f0);
  } else if constexpr (D == 2) {
    const double w0 = fma(R.qr[0], dl[0], fma(R.qr[2], dl[1], R.k[0]));
    const double w1 = fma(R.qr[1], dl[1], R.k[1]);
    cyc = fma(dl[0], w0, cyc);
    cyc = fma(dl[1], w1, cyc);
    const float f0 = (float)dl[0], f1 = (float)dl[1];
    const float e0 = fmaf(R.qi[0], f0, R.qi[2] * f1);
    arg = fmaf(-f0, e0, -f1 * (R.qi[1] * f1));
  } else {
    // qr layout: 00 11 22 01 02 12
    const double w0 = fma(R.qr[0], dl[0], fma(R.qr[3], dl[1], fma(R.qr[4], dl[2], R.k[0])));
    const double w1 = fma(R.qr[1], dl[1], fma(R.qr[5], dl[2], R.k[1]));
    const double w2 = fma(R.qr[2], dl[2], R.k[2]);
    cyc = fma(dl[0], w0, cyc);
    cyc = fma(dl[1], w1, cyc);
    cyc = fma(dl[2], w2, cyc);
    const float f0 = (float)dl[0], f1 = (float)dl[1], f2 = (float)dl[2];
    const float e0 = fmaf(R.qi[0], f0, fmaf(R.qi[3], f1, R.qi[4] * f2));
    const float e1 = fmaf(R.qi[1], f1, R.qi[5] * f2);
    arg = fmaf(-f0, e0, fmaf(-f1, e1, -f2 * (R.qi[2] * f2)));
  }
  // fract(cyc) in [0, 1): the hardware cos/sin take revolutions
  fr = (float)__builtin_amdgcn_fract(cyc);
}

// part += 2|A| cos(2 pi cyc) 2^-ex
template <int D>
__device__ inline void eval_split(const RecS<D>& R, const double* dl, float& part) {
  float fr, arg;
  split_phase<D>(R, dl, fr, arg);
  part = fmaf(R.amp * __builtin_amdgcn_cosf(fr), __builtin_amdgcn_exp2f(arg), part);
}

// (pr, pi) += r 2|A| 2^-ex (cos, -sin)(2 pi cyc): r times the conjugate of the
// complex term whose real part eval_split sums
template <int D>
__device__ inline void eval_split_conj(const RecS<D>& R, const double* dl, float r, float& pr,
                                       float& pi) {
  float fr, arg;
  split_phase<D>(R, dl, fr, arg);
  const float m = (r * R.amp) * __builtin_amdgcn_exp2f(arg);
  pr = fmaf(m, __builtin_amdgcn_cosf(fr), pr);
  pi = fmaf(-m, __builtin_amdgcn_sinf(fr), pi);
}

template <int D>
__device__ inline double eval_f64(const RecD<D>& R, const double* dl) {
  double lin = 0.0;
  for (int a = 0; a < D; ++a) lin += R.p[a] * dl[a];
  double q = 0.0, chi = 0.0;
  int idx = 0;
  for (int a = 0; a < D; ++a) {
    q += R.mr[idx] * dl[a] * dl[a];
    chi += R.mi[idx] * dl[a] * dl[a];
    ++idx;
  }
  for (int a = 0; a < D; ++a)
    for (int b = a + 1; b < D; ++b) {
      q += R.mr[idx] * dl[a] * dl[b];
      chi += R.mi[idx] * dl[a] * dl[b];
      ++idx;
    }
  const double ph = R.w * (lin + 0.5 * q) + R.ang;
  return R.amp * cos(ph) * exp(-R.w * (0.5 * chi));
}

template <int D>
__device__ inline void eval_f64_conj(const RecD<D>& R, const double* dl, double r, double& pr,
                                     double& pi) {
  double lin = 0.0;
  for (int a = 0; a < D; ++a) lin += R.p[a] * dl[a];
  double q = 0.0, chi = 0.0;
  int idx = 0;
  for (int a = 0; a < D; ++a) {
    q += R.mr[idx] * dl[a] * dl[a];
    chi += R.mi[idx] * dl[a] * dl[a];
    ++idx;
  }
  for (int a = 0; a < D; ++a)
    for (int b = a + 1; b < D; ++b) {
      q += R.mr[idx] * dl[a] * dl[b];
      chi += R.mi[idx] * dl[a] * dl[b];
      ++idx;
    }
  const double ph = R.w * (lin + 0.5 * q) + R.ang;
  const double m = r * R.amp * exp(-R.w * (0.5 * chi));
  double sn, cs;
  sincos(ph, &sn, &cs);
  pr += m * cs;
  pi -= m * sn;
}

template <int D, bool F64>
struct RecOf {
  using type = RecS<D>;
};
template <int D>
struct RecOf<D, true> {
  using type = RecD<D>;
};


}  // namespace

namespace {

// Closed form of solve_hom_diag (gb_solvers.py:306-422) in the expanded
// Sherman-Morrison form used by the culled sums: with s = c0 t / |p|,
//   w_a = alpha_a / (1 + s alpha_a),  q_a = w_a p_a,
//   kappa = (s / |p|^2) / (1 - (s / |p|^2) sum_j p_j^2 w_j),
//   M = diag(w) + kappa q q^T
// (identical to diag(alpha) (A + u v^T)^-1 with A = I + s diag(alpha),
// u = -s p / |p|^2, v = p o alpha), and the reference's amplitude formulas.
// 1/x from the hardware reciprocal and two Newton steps (about 1 ulp; the
// per-record closed forms below use it in place of IEEE division)
__device__ inline double frcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = r * fma(-x, r, 2.0);
  return r * fma(-x, r, 2.0);
}

// Lighter forms for the FP32-split records, whose phase terms need about
// 1e-10 relative and whose amplitude is rounded to FP32: one Newton step on
// the hardware estimates (measured on gfx950 over operands in [1e-30, 1e30]:
// estimates within 5.3e-8, after the step 2.2e-15 for 1/x and 4.2e-15 for the
// root; the root falls back to the IEEE form outside [1e-200, 1e200]).
__device__ inline double frcp1(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  return r * fma(-x, r, 2.0);
}
__device__ inline double fsqrt1(double x) {
  if (!(x > 1e-200 && x < 1e200)) return sqrt(x);
  const double y = __builtin_amdgcn_rsq(x);
  const double r0 = x * y;
  return fma(0.5 * y, fma(-r0, r0, x), r0);
}
template <bool LIGHT>
__device__ inline double rcp_(double x) { return LIGHT ? frcp1(x) : frcp(x); }
template <bool LIGHT>
__device__ inline double sqrt_(double x) { return LIGHT ? fsqrt1(x) : sqrt(x); }

// a / sqrt(z) (principal root) with the reciprocals of frcp: a conj(s) / |z|
template <bool LIGHT = false>
__device__ inline cd fdiv_csqrt(cd a, cd z) {
  const double r = sqrt_<LIGHT>(z.re * z.re + z.im * z.im);
  if (!(r > 0.0)) return a / csqrt_(z);
  // principal root: t = sqrt((r + |Re z|) / 2) and u = Im z / (2 t) give
  // (t, u) for Re z >= 0 and (|u|, t sign(Im z)) otherwise (no branch)
  const double t = sqrt_<LIGHT>(0.5 * (r + fabs(z.re)));
  const double u = z.im * rcp_<LIGHT>(2.0 * t);
  const bool right = z.re >= 0.0;
  const cd s = mkc(right ? t : fabs(u), right ? u : copysign(t, z.im));
  const double ir = rcp_<LIGHT>(r);
  return mkc((a.re * s.re + a.im * s.im) * ir, (a.im * s.re - a.re * s.im) * ir);
}

template <int D, bool LIGHT = false>
__device__ inline void hom_diag_state_fast(const double* x0, const double* p0, const cd* alpha,
                                           cd a0, double c0, double t, State<D>& S) {
  double np2 = 0.0;
  for (int a = 0; a < D; ++a) np2 += p0[a] * p0[a];
  const double np = sqrt_<LIGHT>(np2);
  const double inp = rcp_<LIGHT>(np);
  const double s = (c0 * t) * inp;
  const double sn = s * (inp * inp);
  cd w[D], q[D];
  cd den = mkc(1.0);
  for (int a = 0; a < D; ++a) {
    S.x[a] = x0[a] + (c0 * (p0[a] * inp)) * t;
    S.p[a] = p0[a];
    // w = alpha / (1 + s alpha)
    const double br = 1.0 + s * alpha[a].re, bi = s * alpha[a].im;
    const double r = rcp_<LIGHT>(br * br + bi * bi);
    w[a] = mkc((alpha[a].re * br + alpha[a].im * bi) * r, (alpha[a].im * br - alpha[a].re * bi) * r);
    q[a] = w[a] * p0[a];
    den = den - (sn * (p0[a] * p0[a])) * w[a];
  }
  const double dr = rcp_<LIGHT>(den.re * den.re + den.im * den.im);
  const cd kap = mkc(sn * den.re * dr, -sn * den.im * dr);
  // M = diag(w) + kap q q^T, with kap q formed once per row
  cd kq[D];
  for (int a = 0; a < D; ++a) kq[a] = kap * q[a];
  for (int a = 0; a < D; ++a)
    for (int b = a; b < D; ++b) {
      cd m = kq[a] * q[b];
      if (a == b) m = m + w[a];
      S.M[a][b] = m;
      S.M[b][a] = m;
    }
  if constexpr (D == 3) {
    const double i2 = inp * inp;
    const double tsc = c0 * t * (i2 * i2);
    const cd a1 = alpha[0], a2 = alpha[1], a3 = alpha[2];
    const cd t1 = (c0 * t) * (a2 * a3) + (a2 + a3) * np;
    const cd t2 = (c0 * t) * (a1 * a3) + (a1 + a3) * np;
    const cd t3 = (c0 * t) * (a1 * a2) + (a1 + a2) * np;
    const cd dd = mkc(1.0) + tsc * ((p0[0] * p0[0]) * t1 + (p0[1] * p0[1]) * t2 + (p0[2] * p0[2]) * t3);
    S.A = fdiv_csqrt<LIGHT>(a0, dd);
  } else if constexpr (D == 2) {
    const cd ap = ((p0[0] * p0[0]) * alpha[1] + (p0[1] * p0[1]) * alpha[0]) * (inp * inp * inp);
    S.A = fdiv_csqrt<LIGHT>(a0, mkc(1.0) + (c0 * t) * ap);
  } else {
    S.A = a0;
  }
}

}  // namespace

namespace {

// ---------------------------------------------------------------------------
// The FP32-split record of a 2-D closed-form beam with alpha = i b (b > 0, the
// form bh_forward_params gives every forward beam) in real arithmetic.  With
// s = c0 t / |p| and h = p / |p|:
//   w_a = i b_a / (1 + i s b_a) = (s b_a^2 + i b_a) / (1 + s^2 b_a^2),
//   den = 1 - s sum_a h_a^2 w_a,   M = diag(w) + (s / den) v v^T, v_a = w_a h_a
// (hom_diag_state_fast's Sherman-Morrison form with q = w o p and
// kappa q q^T = (s / den) v v^T), and c0 t ap = i s bh with
// bh = h_0^2 b_1 + h_1^2 b_0, so A = a0 (1 + i gamma)^(-1/2), gamma = s bh:
//   |A| = |a0| (1 + gamma^2)^(-1/4),  arg A = arg a0 - atan(gamma) / 2
// (the principal root: Re(1 + i gamma) = 1 > 0).  The time-independent part
// is FastEnt2 (fast_consts2); fast_rec2 adds the time.
// ---------------------------------------------------------------------------
struct FastEnt2 {
  double x0[2], h[2], b[2];
  double iP;     // 1 / |p|; <= 0: not of this form (the generic closed form applies)
  double bh;
  double ks;     // w |p| / (2 pi): k = ks h
  double om;     // w
  double amp0;   // 2 |a0|
  double arg0;   // arg(a0) / (2 pi)
};
static_assert(sizeof(FastEnt2) == 96, "FastEnt2 is read as six 16-byte words");

__device__ inline FastEnt2 fast_consts2(const double* x0, const double* p0, const cd* alpha, cd a0,
                                        double w) {
  FastEnt2 F;
  F.iP = -1.0;
  if (!(alpha[0].re == 0.0 && alpha[1].re == 0.0 && alpha[0].im > 0.0 && alpha[1].im > 0.0))
    return F;
  // 1 / |p| and 2 |a0| from the hardware estimates with one Newton step
  // (about 1e-15 relative; the IEEE forms outside [1e-200, 1e200]); arg a0
  // in FP32 like the general record's arg A (a few 1e-8 cycles)
  const double P2 = p0[0] * p0[0] + p0[1] * p0[1];
  if (!(P2 > 0.0)) return F;
  double iP;
  if (P2 > 1e-200 && P2 < 1e200) {
    const double y = __builtin_amdgcn_rsq(P2);
    iP = y * fma(-0.5 * P2 * y, y, 1.5);
  } else {
    iP = 1.0 / sqrt(P2);
  }
  F.iP = iP;
  for (int a = 0; a < 2; ++a) {
    F.x0[a] = x0[a];
    F.h[a] = p0[a] * iP;
    F.b[a] = alpha[a].im;
  }
  F.bh = F.h[0] * F.h[0] * F.b[1] + F.h[1] * F.h[1] * F.b[0];
  F.ks = w * (P2 * iP) * (1.0 / bh::kTwoPi);
  F.om = w;
  F.amp0 = 2.0 * fsqrt1(a0.re * a0.re + a0.im * a0.im);
  const float are = (float)a0.re, aim = (float)a0.im;
  F.arg0 = (double)(atan2f(aim, (are == 0.f && aim == 0.f) ? 1e-12f : are) *
                    (float)(1.0 / bh::kTwoPi));
  return F;
}

// the record of fast_consts2's beam at time t (c0 = the mode's signed speed)
__device__ inline void fast_rec2(const FastEnt2& F, double c0, double t, RecS<2>& R) {
  const double ct = c0 * t;
  const double s = ct * F.iP;
  double wr[2], wi[2], vr[2], vi[2];
  double sr = 0.0, si = 0.0;
  for (int a = 0; a < 2; ++a) {
    R.x[a] = fma(ct, F.h[a], F.x0[a]);
    R.k[a] = F.ks * F.h[a];
    const double sb = s * F.b[a];
    const double r = frcp1(fma(sb, sb, 1.0));
    wi[a] = F.b[a] * r;
    wr[a] = sb * wi[a];
    vr[a] = wr[a] * F.h[a];
    vi[a] = wi[a] * F.h[a];
    sr = fma(vr[a], F.h[a], sr);
    si = fma(vi[a], F.h[a], si);
  }
  // g = s / den, den = 1 - s (sr + i si)
  const double dr = fma(-s, sr, 1.0), di = -s * si;
  const double id2 = frcp1(fma(dr, dr, di * di));
  const double gr = s * dr * id2, gi = -s * di * id2;
  // g v_a, then M_aa = w_a + (g v_a) v_a and M_01 = (g v_0) v_1
  double Mr[3], Mi[3];
  const double g0r = gr * vr[0] - gi * vi[0], g0i = gr * vi[0] + gi * vr[0];
  const double g1r = gr * vr[1] - gi * vi[1], g1i = gr * vi[1] + gi * vr[1];
  Mr[0] = wr[0] + (g0r * vr[0] - g0i * vi[0]);
  Mi[0] = wi[0] + (g0r * vi[0] + g0i * vr[0]);
  Mr[1] = wr[1] + (g1r * vr[1] - g1i * vi[1]);
  Mi[1] = wi[1] + (g1r * vi[1] + g1i * vr[1]);
  Mr[2] = 2.0 * (g0r * vr[1] - g0i * vi[1]);
  Mi[2] = 2.0 * (g0r * vi[1] + g0i * vr[1]);
  const double cr = F.om * (1.0 / (2.0 * bh::kTwoPi));
  const double cm = 0.5 * F.om * bh::kLog2e;
  for (int q = 0; q < 3; ++q) {
    R.qr[q] = cr * Mr[q];
    R.qi[q] = (float)(cm * Mi[q]);
  }
  const float gm = (float)(s * F.bh);
  const float z = fmaf(gm, gm, 1.f);
  R.amp = (float)F.amp0 * __builtin_amdgcn_rsqf(__builtin_amdgcn_sqrtf(z));
  R.ph0 = F.arg0 - (double)(atanf(gm) * (float)(0.5 / bh::kTwoPi));
}

}  // namespace
