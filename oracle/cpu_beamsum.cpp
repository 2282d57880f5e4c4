// FP64 C++17/OpenMP restatement of the MSGB real beam sums -- oracle / CPU
// baseline, test infrastructure only.
//
// forward: beams in batches of 128 (MSGBSolver batch_size); per batch the
// closed-form solve_hom_diag trajectories (beamax/gb/gb_solvers.py:194-422)
// are materialised, then sum_b 2|A| cos(w(p.d + d^T ReM d/2) + arg A)
// exp(-w d^T ImM d/2) is accumulated into the (Nt, S) FP64 array
// (beamax/gb/core.py:194-329).  Threads own disjoint sensor columns, beams are
// added in order, so the result does not depend on the thread count.
// static: the same field at one time level summed over points
// (beamax/gb/core.py:332-425 at t = 0).
// forward with a cut: the forward sum skipping the terms below 2^-E of their
// beam's 2|A| (the GPU's certified envelope cut), over straight runs of
// sensors: the CPU baseline of the culled GPU sums.
#include <omp.h>

#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

using cplx = std::complex<double>;

namespace {

struct Rec {
  double x[3], p[3], w, amp, ang;
  double mr[3][3], mi[3][3];
};

double safe_angle(cplx a) {
  double re = (a.real() == 0.0 && a.imag() == 0.0) ? 1e-12 : a.real();
  return std::atan2(a.imag(), re);
}

void traj(int d, const double* x0, const double* p0, const cplx* al, cplx a0, double c0, double t,
          double w, Rec& r) {
  double np2 = 0;
  for (int a = 0; a < d; ++a) np2 += p0[a] * p0[a];
  double np = std::sqrt(np2), np3 = np * np * np;
  for (int a = 0; a < d; ++a) {
    r.x[a] = x0[a] + c0 * (p0[a] / np) * t;
    r.p[a] = p0[a];
  }
  cplx Ai[3], Aiu[3], vAi[3], scal = 1.0;
  for (int a = 0; a < d; ++a) {
    cplx A = 1.0 + c0 * t * al[a] / np;
    Ai[a] = 1.0 / A;
    double u = -c0 * t * p0[a] / np3;
    Aiu[a] = Ai[a] * u;
    vAi[a] = p0[a] * al[a] * Ai[a];
    scal += vAi[a] * u;
  }
  for (int a = 0; a < d; ++a)
    for (int b = 0; b < d; ++b) {
      cplx y = -(Aiu[a] * vAi[b]) / scal;
      if (a == b) y += Ai[a];
      cplx m = al[a] * y;
      r.mr[a][b] = m.real();
      r.mi[a][b] = m.imag();
    }
  cplx A;
  if (d == 3) {
    double tsc = c0 * t / (np * np * np * np);
    cplx t1 = c0 * al[1] * al[2] * t + (al[1] + al[2]) * np;
    cplx t2 = c0 * al[0] * al[2] * t + (al[0] + al[2]) * np;
    cplx t3 = c0 * al[0] * al[1] * t + (al[0] + al[1]) * np;
    A = a0 / std::sqrt(1.0 + tsc * (p0[0] * p0[0] * t1 + p0[1] * p0[1] * t2 + p0[2] * p0[2] * t3));
  } else if (d == 2) {
    cplx ap = (p0[0] * p0[0] * al[1] + p0[1] * p0[1] * al[0]) / np3;
    A = a0 / std::sqrt(1.0 + c0 * t * ap);
  } else {
    A = a0;
  }
  r.w = w;
  r.amp = 2.0 * std::abs(A);
  r.ang = safe_angle(A);
}

inline double field(const Rec& r, int d, const double* s) {
  double dl[3];
  for (int a = 0; a < d; ++a) dl[a] = s[a] - r.x[a];
  double lin = 0, q = 0, chi = 0;
  for (int a = 0; a < d; ++a) {
    lin += dl[a] * r.p[a];
    for (int b = 0; b < d; ++b) {
      q += dl[a] * r.mr[a][b] * dl[b];
      chi += dl[a] * r.mi[a][b] * dl[b];
    }
  }
  double ph = r.w * (lin + 0.5 * q) + r.ang;
  return r.amp * std::cos(ph) * std::exp(-r.w * 0.5 * chi);
}

}  // namespace

extern "C" int oracle_forward_sum(int64_t B, int d, const double* x0, const double* p0,
                                  const double* alpha, const double* a0, const double* omega,
                                  const double* mode, double c, int64_t nt, const double* ts,
                                  int64_t S, const double* sensors, double* out, int nthreads,
                                  int64_t beam_begin, int64_t beam_end) {
  if (d < 1 || d > 3) return -1;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  if (beam_end > B || beam_end < 0) beam_end = B;
  const int64_t batch = 128;
  std::vector<Rec> recs(batch * nt);
  for (int64_t i = 0; i < nt * S; ++i) out[i] = 0.0;
  for (int64_t b0 = beam_begin; b0 < beam_end; b0 += batch) {
    const int64_t nb = std::min(batch, beam_end - b0);
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < nb * nt; ++q) {
      int64_t i = b0 + q / nt, t = q % nt;
      cplx al[3];
      for (int a = 0; a < d; ++a) al[a] = cplx(alpha[2 * (i * d + a)], alpha[2 * (i * d + a) + 1]);
      traj(d, x0 + i * d, p0 + i * d, al, cplx(a0[2 * i], a0[2 * i + 1]), c * mode[i], ts[t],
           omega[i], recs[q]);
    }
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < S; ++s) {
      const double* sp = sensors + s * d;
      for (int64_t bi = 0; bi < nb; ++bi)
        for (int64_t t = 0; t < nt; ++t) out[t * S + s] += field(recs[bi * nt + t], d, sp);
    }
  }
  return 0;
}

extern "C" int oracle_static_sum(int64_t B, int d, const double* x, const double* p,
                                 const double* M, const double* A, const double* omega, int64_t P,
                                 const double* pts, double* out, int nthreads) {
  if (d < 1 || d > 3) return -1;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  std::vector<Rec> recs(B);
  for (int64_t i = 0; i < B; ++i) {
    Rec& r = recs[i];
    for (int a = 0; a < d; ++a) {
      r.x[a] = x[i * d + a];
      r.p[a] = p[i * d + a];
      for (int b = 0; b < d; ++b) {
        r.mr[a][b] = M[2 * ((i * d + a) * d + b)];
        r.mi[a][b] = M[2 * ((i * d + a) * d + b) + 1];
      }
    }
    cplx Ai(A[2 * i], A[2 * i + 1]);
    r.w = omega[i];
    r.amp = 2.0 * std::abs(Ai);
    r.ang = safe_angle(Ai);
  }
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < P; ++q) {
    double acc = 0.0;
    for (int64_t i = 0; i < B; ++i) acc += field(recs[i], d, pts + q * d);
    out[q] = acc;
  }
  return 0;
}

// The forward sum with the certified envelope cut of the GPU's culled sums:
// a term is skipped when its envelope exponent delta^T (w log2(e) ImM / 2)
// delta exceeds E (log2 units), i.e. when it is below 2^-E of its beam's
// 2|A|.  The sensors are given as straight runs (point u of run r is
// p0_r + u v_r, output column cols[off_r + u]); along a run the exponent is
// a quadratic in u, so each (beam, time, run) evaluates exactly the index
// interval where it is <= E.  Threads own disjoint time rows and add beams in
// order, so the result does not depend on the thread count.  *terms counts
// the evaluated terms.
extern "C" int oracle_forward_sum_runs(int64_t B, int d, const double* x0, const double* p0,
                                       const double* alpha, const double* a0,
                                       const double* omega, const double* mode, double c,
                                       int64_t nt, const double* ts, int64_t nruns,
                                       const double* run_p0, const double* run_v,
                                       const int64_t* run_off, const int64_t* cols, int64_t S,
                                       double E, double* out, int nthreads, int64_t beam_begin,
                                       int64_t beam_end, int64_t* terms) {
  if (d < 1 || d > 3) return -1;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  if (beam_end > B || beam_end < 0) beam_end = B;
  const double log2e = 1.4426950408889634;
  for (int64_t i = 0; i < nt * S; ++i) out[i] = 0.0;
  int64_t count = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : count)
  for (int64_t t = 0; t < nt; ++t) {
    double* row = out + t * S;
    for (int64_t i = beam_begin; i < beam_end; ++i) {
      cplx al[3];
      for (int a = 0; a < d; ++a) al[a] = cplx(alpha[2 * (i * d + a)], alpha[2 * (i * d + a) + 1]);
      Rec r;
      traj(d, x0 + i * d, p0 + i * d, al, cplx(a0[2 * i], a0[2 * i + 1]), c * mode[i], ts[t],
           omega[i], r);
      if (!(r.amp > 0.0)) continue;
      double Q[3][3];
      for (int a = 0; a < d; ++a)
        for (int b = 0; b < d; ++b) Q[a][b] = 0.5 * r.w * log2e * r.mi[a][b];
      for (int64_t k = 0; k < nruns; ++k) {
        const int64_t n = run_off[k + 1] - run_off[k];
        double dl[3], v[3];
        for (int a = 0; a < d; ++a) {
          dl[a] = run_p0[k * d + a] - r.x[a];
          v[a] = run_v[k * d + a];
        }
        // exponent(u) = qa u^2 + 2 qb u + qc
        double qa = 0, qb = 0, qc = 0;
        for (int a = 0; a < d; ++a)
          for (int b = 0; b < d; ++b) {
            qa += v[a] * Q[a][b] * v[b];
            qb += dl[a] * Q[a][b] * v[b];
            qc += dl[a] * Q[a][b] * dl[b];
          }
        int64_t lo = 0, hi = n - 1;
        if (qa > 0.0) {
          const double disc = qb * qb - qa * (qc - E);
          if (disc < 0.0) continue;
          const double sq = std::sqrt(disc);
          lo = std::max<int64_t>(0, (int64_t)std::ceil((-qb - sq) / qa));
          hi = std::min<int64_t>(n - 1, (int64_t)std::floor((-qb + sq) / qa));
        } else if (qc > E) {
          continue;
        }
        for (int64_t u = lo; u <= hi; ++u) {
          double s[3];
          for (int a = 0; a < d; ++a) s[a] = run_p0[k * d + a] + (double)u * v[a];
          row[cols[run_off[k] + u]] += field(r, d, s);
        }
        if (hi >= lo) count += hi - lo + 1;
      }
    }
  }
  if (terms) *terms = count;
  return 0;
}

// The static (one time level) sum with the same cut over point runs:
// threads own disjoint runs, beams are added in order.
extern "C" int oracle_static_sum_runs(int64_t B, int d, const double* x, const double* p,
                                      const double* M, const double* A, const double* omega,
                                      int64_t nruns, const double* run_p0, const double* run_v,
                                      const int64_t* run_off, const int64_t* cols, int64_t P,
                                      double E, double* out, int nthreads, int64_t* terms) {
  if (d < 1 || d > 3) return -1;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  const double log2e = 1.4426950408889634;
  std::vector<Rec> recs(B);
  for (int64_t i = 0; i < B; ++i) {
    Rec& r = recs[i];
    for (int a = 0; a < d; ++a) {
      r.x[a] = x[i * d + a];
      r.p[a] = p[i * d + a];
      for (int b = 0; b < d; ++b) {
        r.mr[a][b] = M[2 * ((i * d + a) * d + b)];
        r.mi[a][b] = M[2 * ((i * d + a) * d + b) + 1];
      }
    }
    cplx Ai(A[2 * i], A[2 * i + 1]);
    r.w = omega[i];
    r.amp = 2.0 * std::abs(Ai);
    r.ang = safe_angle(Ai);
  }
  for (int64_t i = 0; i < P; ++i) out[i] = 0.0;
  int64_t count = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : count)
  for (int64_t k = 0; k < nruns; ++k) {
    const int64_t n = run_off[k + 1] - run_off[k];
    for (int64_t i = 0; i < B; ++i) {
      const Rec& r = recs[i];
      if (!(r.amp > 0.0)) continue;
      double dl[3], v[3];
      for (int a = 0; a < d; ++a) {
        dl[a] = run_p0[k * d + a] - r.x[a];
        v[a] = run_v[k * d + a];
      }
      double qa = 0, qb = 0, qc = 0;
      for (int a = 0; a < d; ++a)
        for (int b = 0; b < d; ++b) {
          const double q = 0.5 * r.w * log2e * r.mi[a][b];
          qa += v[a] * q * v[b];
          qb += dl[a] * q * v[b];
          qc += dl[a] * q * dl[b];
        }
      int64_t lo = 0, hi = n - 1;
      if (qa > 0.0) {
        const double disc = qb * qb - qa * (qc - E);
        if (disc < 0.0) continue;
        const double sq = std::sqrt(disc);
        lo = std::max<int64_t>(0, (int64_t)std::ceil((-qb - sq) / qa));
        hi = std::min<int64_t>(n - 1, (int64_t)std::floor((-qb + sq) / qa));
      } else if (qc > E) {
        continue;
      }
      for (int64_t u = lo; u <= hi; ++u) {
        double s[3];
        for (int a = 0; a < d; ++a) s[a] = run_p0[k * d + a] + (double)u * v[a];
        out[cols[run_off[k] + u]] += field(r, d, s);
      }
      if (hi >= lo) count += hi - lo + 1;
    }
  }
  if (terms) *terms = count;
  return 0;
}
