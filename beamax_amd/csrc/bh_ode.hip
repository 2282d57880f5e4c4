// Gaussian-beam ODEs in a heterogeneous medium (gb_solvers.py:787-1113):
//   H = mode c(x) |p|,  x' = H_p,  p' = -H_x,
//   M' = -(H_xx + H_xp M + M H_xp^T + M H_pp M),
//   A' = -A (c^2 tr M - H_x.H_p - H_p^T M H_p + lam H) / (2 H),
// with c the grid interpolant of bh_medium.hpp, integrated per beam (one
// thread each) by the Dormand-Prince 5(4) pair (Dormand and Prince 1980)
// with an adaptive step: error = RMS over the state entries (complex entries
// by modulus) of |y5 - y4| / (atol + rtol max(|y_n|, |y_{n+1}|)), accepted
// when <= 1, next step h * clip(0.9 err^(-1/5), 0.2, 10).  Steps are
// shortened to land on every output time, where the state is written.  The
// reference integrates with Tsit5 (another 5(4) pair) under the same
// tolerances, so both agree to the tolerance, not bit for bit.
#include "bh_medium.hpp"

namespace {

template <int D>
struct Ode {
  static constexpr int NS = 2 * D + 2 * D * D + 2;   // real state entries
};

// state layout: x[D], p[D], M[D][D] (re, im interleaved), A (re, im)
template <int D>
__device__ inline void ode_rhs(const SpeedDev& S, double mode, const double* y, double* f) {
  const double* x = y;
  const double* p = y + D;
  const double* M = y + 2 * D;          // complex, interleaved
  const double* A = y + 2 * D + 2 * D * D;
  double c, g[D], H[D * D];
  speed_eval<D>(S, x, c, g, H);
  double np2 = 0.0;
  for (int a = 0; a < D; ++a) np2 += p[a] * p[a];
  const double np = sqrt(np2);
  const double G = mode * c * np;
  double Gx[D], Gp[D];
  for (int a = 0; a < D; ++a) {
    Gx[a] = mode * g[a] * np;
    Gp[a] = mode * c * p[a] / np;
    f[a] = Gp[a];
    f[D + a] = -Gx[a];
  }
  // Hessian blocks (real): Gxx = mode H |p|, Gxp_ij = mode g_i p_j / |p|,
  // Gpp = mode c (I / |p| - p p^T / |p|^3)
  double Gxx[D][D], Gxp[D][D], Gpp[D][D];
  const double np3 = np2 * np;
  for (int a = 0; a < D; ++a)
    for (int b = 0; b < D; ++b) {
      Gxx[a][b] = mode * H[a * D + b] * np;
      Gxp[a][b] = mode * g[a] * p[b] / np;
      Gpp[a][b] = mode * c * ((a == b ? 1.0 / np : 0.0) - p[a] * p[b] / np3);
    }
  // MG = M Gpp (complex), then dM = -(Gxx + Gxp M + M Gxp^T + MG M)
  double MGr[D][D], MGi[D][D];
  for (int a = 0; a < D; ++a)
    for (int b = 0; b < D; ++b) {
      double sr = 0.0, si = 0.0;
      for (int e = 0; e < D; ++e) {
        sr += M[2 * (a * D + e)] * Gpp[e][b];
        si += M[2 * (a * D + e) + 1] * Gpp[e][b];
      }
      MGr[a][b] = sr;
      MGi[a][b] = si;
    }
  double trr = 0.0, tri = 0.0;
  for (int a = 0; a < D; ++a) {
    trr += M[2 * (a * D + a)];
    tri += M[2 * (a * D + a) + 1];
    for (int b = 0; b < D; ++b) {
      double rr = Gxx[a][b], ri = 0.0;
      for (int e = 0; e < D; ++e) {
        // Gxp M
        rr += Gxp[a][e] * M[2 * (e * D + b)];
        ri += Gxp[a][e] * M[2 * (e * D + b) + 1];
        // M Gxp^T
        rr += M[2 * (a * D + e)] * Gxp[b][e];
        ri += M[2 * (a * D + e) + 1] * Gxp[b][e];
        // (M Gpp) M
        const double mr = M[2 * (e * D + b)], mi = M[2 * (e * D + b) + 1];
        rr += MGr[a][e] * mr - MGi[a][e] * mi;
        ri += MGr[a][e] * mi + MGi[a][e] * mr;
      }
      f[2 * D + 2 * (a * D + b)] = -rr;
      f[2 * D + 2 * (a * D + b) + 1] = -ri;
    }
  }
  // s = c^2 tr M - Gx.Gp - Gp^T M Gp + lam G (complex); A' = -A s / (2 G)
  double gg = 0.0;
  for (int a = 0; a < D; ++a) gg += Gx[a] * Gp[a];
  double qr = 0.0, qi = 0.0;
  for (int a = 0; a < D; ++a)
    for (int b = 0; b < D; ++b) {
      qr += Gp[a] * M[2 * (a * D + b)] * Gp[b];
      qi += Gp[a] * M[2 * (a * D + b) + 1] * Gp[b];
    }
  const double sr = c * c * trr - gg - qr + S.lam * G;
  const double si = c * c * tri - qi;
  const double inv = -1.0 / (2.0 * G);
  f[2 * D + 2 * D * D] = inv * (A[0] * sr - A[1] * si);
  f[2 * D + 2 * D * D + 1] = inv * (A[0] * si + A[1] * sr);
}

// error norm: complex entries (M, A) by modulus, real entries (x, p) as is
template <int D>
__device__ inline double err_norm(const double* y0, const double* y1, const double* e,
                                  double rtol, double atol) {
  double s = 0.0;
  int n = 0;
  for (int a = 0; a < 2 * D; ++a) {
    const double sc = atol + rtol * fmax(fabs(y0[a]), fabs(y1[a]));
    const double r = e[a] / sc;
    s += r * r;
    ++n;
  }
  for (int q = 2 * D; q < Ode<D>::NS; q += 2) {
    const double m0 = sqrt(y0[q] * y0[q] + y0[q + 1] * y0[q + 1]);
    const double m1 = sqrt(y1[q] * y1[q] + y1[q + 1] * y1[q + 1]);
    const double sc = atol + rtol * fmax(m0, m1);
    const double r2 = (e[q] * e[q] + e[q + 1] * e[q + 1]) / (sc * sc);
    s += r2;
    ++n;
  }
  return sqrt(s / n);
}

// Advance y (with k1 = f(y) on entry and exit) over s in [s, sout], s < sout,
// for dy/ds = dir * f(y): dir = +1 integrates forward in time, -1 backward
// (s = t_start - t).  h is the proposed step, carried between calls.
// Returns 0, 1 (step limit) or 2 (non-finite error estimate).
template <int D>
__device__ inline int dopri_advance(const SpeedDev& S, double mode, double dir, double* y,
                                    double* k1, double& s, double sout, double& h,
                                    long long& steps, long long max_steps, double rtol,
                                    double atol) {
  constexpr int NS = Ode<D>::NS;
  double k2[NS], k3[NS], k4[NS], k5[NS], k6[NS], k7[NS], yt[NS], y5[NS], er[NS];
  auto rhs = [&](const double* yin, double* f) {
    ode_rhs<D>(S, mode, yin, f);
    for (int q = 0; q < NS; ++q) f[q] *= dir;
  };
  while (s < sout) {
    if (++steps > max_steps) return 1;
    const bool last = s + h >= sout;
    const double hs = last ? sout - s : h;
    // Dormand-Prince 5(4) stages
    for (int q = 0; q < NS; ++q) yt[q] = y[q] + hs * (1.0 / 5.0) * k1[q];
    rhs(yt, k2);
    for (int q = 0; q < NS; ++q) yt[q] = y[q] + hs * ((3.0 / 40.0) * k1[q] + (9.0 / 40.0) * k2[q]);
    rhs(yt, k3);
    for (int q = 0; q < NS; ++q)
      yt[q] = y[q] + hs * ((44.0 / 45.0) * k1[q] - (56.0 / 15.0) * k2[q] + (32.0 / 9.0) * k3[q]);
    rhs(yt, k4);
    for (int q = 0; q < NS; ++q)
      yt[q] = y[q] + hs * ((19372.0 / 6561.0) * k1[q] - (25360.0 / 2187.0) * k2[q] +
                           (64448.0 / 6561.0) * k3[q] - (212.0 / 729.0) * k4[q]);
    rhs(yt, k5);
    for (int q = 0; q < NS; ++q)
      yt[q] = y[q] + hs * ((9017.0 / 3168.0) * k1[q] - (355.0 / 33.0) * k2[q] +
                           (46732.0 / 5247.0) * k3[q] + (49.0 / 176.0) * k4[q] -
                           (5103.0 / 18656.0) * k5[q]);
    rhs(yt, k6);
    for (int q = 0; q < NS; ++q)
      y5[q] = y[q] + hs * ((35.0 / 384.0) * k1[q] + (500.0 / 1113.0) * k3[q] +
                           (125.0 / 192.0) * k4[q] - (2187.0 / 6784.0) * k5[q] +
                           (11.0 / 84.0) * k6[q]);
    rhs(y5, k7);
    // y5 - y4 with the embedded 4th-order weights
    for (int q = 0; q < NS; ++q)
      er[q] = hs * ((35.0 / 384.0 - 5179.0 / 57600.0) * k1[q] +
                    (500.0 / 1113.0 - 7571.0 / 16695.0) * k3[q] +
                    (125.0 / 192.0 - 393.0 / 640.0) * k4[q] +
                    (-2187.0 / 6784.0 + 92097.0 / 339200.0) * k5[q] +
                    (11.0 / 84.0 - 187.0 / 2100.0) * k6[q] - (1.0 / 40.0) * k7[q]);
    const double en = err_norm<D>(y, y5, er, rtol, atol);
    if (!(en == en)) return 2;
    double fac = en > 0.0 ? 0.9 * pow(en, -0.2) : 10.0;
    fac = fmin(10.0, fmax(0.2, fac));
    if (en <= 1.0) {
      s = last ? sout : s + hs;
      for (int q = 0; q < NS; ++q) {
        y[q] = y5[q];
        k1[q] = k7[q];   // first same as last
      }
      // a step shortened to land on an output keeps the earlier proposal
      h = (last && hs < h) ? fmax(hs * fac, h) : hs * fac;
    } else {
      h = hs * fac;
    }
  }
  return 0;
}

struct OdeArgs {
  bh_fwd_beams b;
  long long count, total;   // total = 2 count when mirrored
  int mirror;               // 1: mirrored copies (mode -1) appended; 0: all +1; -1: all -1
  SpeedDev S;
  const double* ts;
  long long nt;
  double rtol, atol;
  long long max_steps;
  double* traj;             // [nt][total][NS]
  int* status;              // per beam: 0 ok, 1 step limit, 2 non-finite state
};

template <int D>
__global__ __launch_bounds__(64) void k_ode_forward(OdeArgs A) {
  constexpr int NS = Ode<D>::NS;
  const long long bi = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (bi >= A.total) return;
  const bool copy = A.mirror > 0 && bi >= A.count;   // mirrored copy of beam bi - count
  const bool neg = A.mirror < 0 || copy;
  const long long j = copy ? bi - A.count : bi;
  const double mode = neg ? -1.0 : 1.0;
  double y[NS];
  for (int a = 0; a < D; ++a) {
    y[a] = A.b.x0[j * D + a];
    y[D + a] = A.b.p0[j * D + a];
  }
  for (int a = 0; a < D; ++a)
    for (int e = 0; e < D; ++e) {
      const bool dg = a == e;
      y[2 * D + 2 * (a * D + e)] = dg ? A.b.alpha[2 * (j * D + a)] : 0.0;
      y[2 * D + 2 * (a * D + e) + 1] = dg ? A.b.alpha[2 * (j * D + a) + 1] : 0.0;
    }
  y[2 * D + 2 * D * D] = A.b.a0[2 * j];
  y[2 * D + 2 * D * D + 1] = A.b.a0[2 * j + 1];
  auto emit = [&](long long k) {
    double* o = A.traj + (k * A.total + bi) * NS;
    for (int q = 0; q < NS; ++q) o[q] = y[q];
  };
  emit(0);
  double t = A.ts[0];
  double h = A.nt > 1 ? A.ts[1] - A.ts[0] : 0.0;
  long long steps = 0;
  int st = 0;
  double k1[NS];
  ode_rhs<D>(A.S, mode, y, k1);
  for (long long out = 1; out < A.nt; ++out) {
    st = dopri_advance<D>(A.S, mode, 1.0, y, k1, t, A.ts[out], h, steps, A.max_steps, A.rtol,
                          A.atol);
    if (st) break;
    emit(out);
  }
  // rows not reached (step limit / non-finite) are left unwritten: the
  // caller raises on a non-zero status
  A.status[bi] = st;
}

struct OdeStaticArgs {
  bh_static_beams b;
  long long count;
  SpeedDev S;
  const double* t_start;
  const double* t_end;   // null: every beam ends at t = 0
  const double* mode;
  double rtol, atol, dt0;
  long long max_steps;
  int* status;
};

// Beams (x, p, full complex M, A at t_start) taken to t_end (0 for the
// terminal-value beams of the adjoint) in place: solve_ODE_batch_t with
// ts_i = (t_start, t_end) (gb_solvers.py:1017-1113), whose right-hand side has
// no absorption term (lam = 0).  Beams with zero amplitude, t_end = t_start
// or mode 0 (no group velocity; the closed form leaves them in place) are not
// integrated.
template <int D>
__global__ __launch_bounds__(64) void k_ode_static(OdeStaticArgs A) {
  constexpr int NS = Ode<D>::NS;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= A.count) return;
  const double t0 = A.t_start[i];
  const double t1 = A.t_end ? A.t_end[i] : 0.0;
  const double mode = A.mode[i];
  const double ar = A.b.A[2 * i], ai = A.b.A[2 * i + 1];
  if (t1 == t0 || mode == 0.0 || (ar == 0.0 && ai == 0.0)) {
    A.status[i] = 0;
    return;
  }
  double y[NS];
  for (int a = 0; a < D; ++a) {
    y[a] = A.b.x[i * D + a];
    y[D + a] = A.b.p[i * D + a];
  }
  for (int q = 0; q < 2 * D * D; ++q) y[2 * D + q] = A.b.M[2 * i * D * D + q];
  y[2 * D + 2 * D * D] = ar;
  y[2 * D + 2 * D * D + 1] = ai;
  const double dir = t1 < t0 ? -1.0 : 1.0;
  const double span = fabs(t1 - t0);
  double h = A.dt0 > 0.0 ? A.dt0 : span;
  double sv = 0.0;
  long long steps = 0;
  double k1[NS];
  ode_rhs<D>(A.S, mode, y, k1);
  for (int q = 0; q < NS; ++q) k1[q] *= dir;
  const int st = dopri_advance<D>(A.S, mode, dir, y, k1, sv, span, h, steps, A.max_steps, A.rtol,
                                  A.atol);
  if (st == 0) {
    for (int a = 0; a < D; ++a) {
      A.b.x[i * D + a] = y[a];
      A.b.p[i * D + a] = y[D + a];
    }
    for (int q = 0; q < 2 * D * D; ++q) A.b.M[2 * i * D * D + q] = y[2 * D + q];
    A.b.A[2 * i] = y[2 * D + 2 * D * D];
    A.b.A[2 * i + 1] = y[2 * D + 2 * D * D + 1];
  }
  A.status[i] = st;
}

SpeedDev speed_dev(const bh_speed* sp, int ndim) {
  SpeedDev S{};
  S.ndim = sp->ndim;
  S.method = sp->method;
  S.boundary = sp->boundary;
  for (int i = 0; i < 3; ++i) {
    S.N[i] = i < ndim ? sp->N[i] : 1;
    S.spacing[i] = i < ndim ? sp->spacing[i] : 1.0;
    S.origin[i] = i < ndim ? sp->origin[i] : 0.0;
  }
  S.values = sp->values;
  S.lam = sp->lam;
  return S;
}

}  // namespace

extern "C" int bh_ode_forward(const bh_fwd_beams* b, int ndim, int mirror, const bh_speed* sp,
                              const double* ts, int64_t nt, double rtol, double atol,
                              int64_t max_steps, double* traj, int* status, bh_stream s) {
  bh::clear_error();
  if (!b || !sp || !ts || !traj || !status) BH_FAIL(BH_ERR_ARG, "bh_ode_forward: null pointer");
  if (ndim < 1 || ndim > 3 || sp->ndim != ndim)
    BH_FAIL(BH_ERR_ARG, "bh_ode_forward: ndim=%d, speed ndim=%d", ndim, sp->ndim);
  if (nt <= 0) BH_FAIL(BH_ERR_ARG, "bh_ode_forward: nt=%lld", (long long)nt);
  if (!(rtol > 0.0) || !(atol > 0.0) || max_steps <= 0)
    BH_FAIL(BH_ERR_ARG, "bh_ode_forward: rtol, atol and max_steps must be positive");
  if (!sp->values) BH_FAIL(BH_ERR_ARG, "bh_ode_forward: null speed table");
  if (b->count == 0) return BH_OK;
  OdeArgs a{};
  a.b = *b;
  a.count = b->count;
  a.mirror = mirror > 0 ? 1 : (mirror < 0 ? -1 : 0);
  a.total = b->count * (mirror > 0 ? 2 : 1);
  a.S = speed_dev(sp, ndim);
  a.ts = ts;
  a.nt = nt;
  a.rtol = rtol;
  a.atol = atol;
  a.max_steps = max_steps;
  a.traj = traj;
  a.status = status;
  hipStream_t st = (hipStream_t)s;
  const unsigned grid = bh::grid1d(a.total, 64);
  if (ndim == 1) hipLaunchKernelGGL(k_ode_forward<1>, dim3(grid), dim3(64), 0, st, a);
  else if (ndim == 2) hipLaunchKernelGGL(k_ode_forward<2>, dim3(grid), dim3(64), 0, st, a);
  else hipLaunchKernelGGL(k_ode_forward<3>, dim3(grid), dim3(64), 0, st, a);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

extern "C" int bh_ode_static(const bh_static_beams* b, int ndim, const bh_speed* sp,
                             const double* t_start, const double* t_end, const double* mode,
                             double lam, double rtol, double atol, double dt0,
                             int64_t max_steps, int* status, bh_stream s) {
  bh::clear_error();
  if (!b || !sp || !t_start || !mode || !status)
    BH_FAIL(BH_ERR_ARG, "bh_ode_static: null pointer");
  if (ndim < 1 || ndim > 3 || sp->ndim != ndim)
    BH_FAIL(BH_ERR_ARG, "bh_ode_static: ndim=%d, speed ndim=%d", ndim, sp->ndim);
  if (!(rtol > 0.0) || !(atol > 0.0) || max_steps <= 0)
    BH_FAIL(BH_ERR_ARG, "bh_ode_static: rtol, atol and max_steps must be positive");
  if (!sp->values) BH_FAIL(BH_ERR_ARG, "bh_ode_static: null speed table");
  if (b->count < 0) BH_FAIL(BH_ERR_ARG, "bh_ode_static: count=%lld", (long long)b->count);
  if (b->count == 0) return BH_OK;
  if (!b->x || !b->p || !b->M || !b->A) BH_FAIL(BH_ERR_ARG, "bh_ode_static: null beam table");
  OdeStaticArgs a{};
  a.b = *b;
  a.count = b->count;
  a.S = speed_dev(sp, ndim);
  a.S.lam = lam;
  a.t_start = t_start;
  a.t_end = t_end;
  a.mode = mode;
  a.rtol = rtol;
  a.atol = atol;
  a.dt0 = dt0;
  a.max_steps = max_steps;
  a.status = status;
  hipStream_t st = (hipStream_t)s;
  const unsigned grid = bh::grid1d(a.count, 64);
  if (ndim == 1) hipLaunchKernelGGL(k_ode_static<1>, dim3(grid), dim3(64), 0, st, a);
  else if (ndim == 2) hipLaunchKernelGGL(k_ode_static<2>, dim3(grid), dim3(64), 0, st, a);
  else hipLaunchKernelGGL(k_ode_static<3>, dim3(grid), dim3(64), 0, st, a);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

namespace {
template <int D>
__global__ void k_speed_eval(SpeedDev S, const double* x, long long n, double* c, double* g,
                             double* H) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double cv, gv[D], Hv[D * D];
  speed_eval<D>(S, x + i * D, cv, gv, Hv);
  c[i] = cv;
  for (int a = 0; a < D; ++a) g[i * D + a] = gv[a];
  for (int q = 0; q < D * D; ++q) H[i * D * D + q] = Hv[q];
}
}  // namespace

extern "C" int bh_speed_eval(const bh_speed* sp, const double* x, int64_t n, double* c,
                             double* grad, double* hess, bh_stream s) {
  bh::clear_error();
  if (!sp || !x || !c || !grad || !hess || !sp->values)
    BH_FAIL(BH_ERR_ARG, "bh_speed_eval: null pointer");
  if (sp->ndim < 1 || sp->ndim > 3) BH_FAIL(BH_ERR_ARG, "bh_speed_eval: ndim=%d", sp->ndim);
  if (sp->method != kSpeedLinear && sp->method != kSpeedBspline3)
    BH_FAIL(BH_ERR_ARG, "bh_speed_eval: method=%d", sp->method);
  if (sp->boundary < kBndNearest || sp->boundary > kBndMirror)
    BH_FAIL(BH_ERR_ARG, "bh_speed_eval: boundary=%d", sp->boundary);
  if (n <= 0) return BH_OK;
  SpeedDev S{};
  S.ndim = sp->ndim;
  S.method = sp->method;
  S.boundary = sp->boundary;
  for (int i = 0; i < 3; ++i) {
    S.N[i] = i < sp->ndim ? sp->N[i] : 1;
    S.spacing[i] = i < sp->ndim ? sp->spacing[i] : 1.0;
    S.origin[i] = i < sp->ndim ? sp->origin[i] : 0.0;
  }
  S.values = sp->values;
  S.lam = sp->lam;
  const unsigned grid = bh::grid1d(n, 128);
  hipStream_t st = (hipStream_t)s;
  if (sp->ndim == 1) hipLaunchKernelGGL(k_speed_eval<1>, dim3(grid), dim3(128), 0, st, S, x, (long long)n, c, grad, hess);
  else if (sp->ndim == 2) hipLaunchKernelGGL(k_speed_eval<2>, dim3(grid), dim3(128), 0, st, S, x, (long long)n, c, grad, hess);
  else hipLaunchKernelGGL(k_speed_eval<3>, dim3(grid), dim3(128), 0, st, S, x, (long long)n, c, grad, hess);
  BH_LAUNCH_CHECK();
  return BH_OK;
}
