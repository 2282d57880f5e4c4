// Beam parameters, closed-form propagation and the real beam summations.
//
// Forward sum (the hot loop): a workgroup owns 4 time levels (one per wave) x
// 64*SPT sensors.  Beams stream through LDS in chunks of 64: each lane builds
// the per-(beam, time) record of one beam for its wave's time level from the
// closed-form solve_hom_diag solution (FP64), then every lane evaluates that
// chunk at its SPT sensors.  Records broadcast from LDS, sensors stay in
// registers, rows of the (Nt, S) output are written coalesced.
//
// FP32-split evaluation: the Gaussian phase in cycles,
//   cyc = delta.(k + Qr delta) + arg(A)/2pi,  k = w p / 2pi,  Qr = w ReM / 4pi,
// is formed in FP64 (delta = s - x(t) in FP64), reduced to [-1/2, 1/2] in FP64
// and only then rounded to FP32 for the hardware cos (which takes
// revolutions).  The envelope exponent delta^T Qi delta (Qi = w log2e ImM / 2)
// and everything after are FP32; partial sums over 64 beams are FP32 and are
// accumulated across chunks in FP64.
#include "bh_beamcore.hpp"
#include "bh_medium.hpp"

namespace {

// Per-device status word of this translation unit's kernels (bit
// BH_STATUS_FWD_INDEX / BH_STATUS_ADJ_INDEX: a beam index outside [-1, total)
// reached a parameter kernel).  Read and cleared by bh_device_status.
__device__ unsigned int g_status = 0u;

__device__ inline void flag_status(unsigned bit) { atomicOr(&g_status, bit); }

constexpr double kQuietNaN = __builtin_nan("");

using bh::TilingDev;

// ---------------------------------------------------------------------------
// common decoders
// ---------------------------------------------------------------------------
struct Decoded {
  int lvl;
  long long box;
  long long k[BH_MAX_DIM];
};

__device__ inline Decoded decode_flat(const TilingDev& T, long long flat) {
  Decoded r;
  int l = 0;
  while (l + 1 < T.num_levels && flat >= T.coeffs_cumsum[l + 1]) ++l;
  r.lvl = l;
  long long P = 1;
  for (int a = 0; a < T.ndim; ++a) P *= T.support[l][a];
  const long long local = flat - T.coeffs_cumsum[l];
  r.box = T.boxes_cumsum[l] + local / P;
  long long jf = local % P;
  for (int a = T.ndim - 1; a >= 0; --a) {
    r.k[a] = jf % T.support[l][a];
    jf /= T.support[l][a];
  }
  return r;
}

// transforms.py:193-226
__device__ inline cd frame_phase(const TilingDev& T, const int32_t* centre, int lvl,
                                 const long long* k) {
  double cyc = 0.0;
  for (int a = 0; a < T.ndim; ++a) {
    const long long sup = (long long)T.box_length[lvl] * T.redundancy * T.aspect[a];
    const long long half = sup / 2;
    const long long par = fdiv((long long)centre[a] + T.N[a] / 2, half);
    const long long roll = (1 - 2 * (par & 1)) * (half / 2);
    cyc += (double)(((long long)centre[a] - roll) * k[a]) / (double)sup;
  }
  cyc = cyc - floor(cyc);
  double sv, cv;
  sincospi(2.0 * cyc, &sv, &cv);
  return mkc(cv, sv);
}

// forward_solver_utils.py:255-373 (compute_params, sign +1) times the coefficient
__global__ __launch_bounds__(128) void k_forward_params(TilingDev T, const int32_t* __restrict__ centres, double L0,
                                 double L1, double L2, const long long* __restrict__ idx,
                                 const double2* __restrict__ vals, long long k,
                                 bh_fwd_beams out) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= k) return;
  const int d = T.ndim;
  const double L[3] = {L0, L1, L2};
  const long long flat = idx[i];
  if (flat < 0 || flat >= T.coeffs_cumsum[T.num_levels]) {
    // -1 is the reference's padding index (zero coefficient): an inert beam
    // of zero amplitude.  Any other index outside the coefficient vector is
    // a caller error: the beam's amplitude is NaN, so every output it
    // touches is NaN, and the status word records it.
    const double amp = flat == -1 ? 0.0 : kQuietNaN;
    if (flat != -1) flag_status(BH_STATUS_FWD_INDEX);
    for (int a = 0; a < d; ++a) {
      out.p0[i * d + a] = a == 0 ? bh::kTwoPi : 0.0;
      out.x0[i * d + a] = 0.0;
      out.alpha[2 * (i * d + a)] = 0.0;
      out.alpha[2 * (i * d + a) + 1] = 1.0;
    }
    out.a0[2 * i] = amp;
    out.a0[2 * i + 1] = amp;
    out.omega[i] = 1.0;
    return;
  }
  const Decoded q = decode_flat(T, flat);
  const int32_t* c = centres + q.box * d;
  double ch[3], n2 = 0.0;
  for (int a = 0; a < d; ++a) { ch[a] = (double)c[a] / L[a]; n2 += ch[a] * ch[a]; }
  const double norm = sqrt(n2);
  cd amp = mkc(1.0);
  double a0r = 1.0;
  for (int a = 0; a < d; ++a) {
    out.p0[i * d + a] = 2.0 * bh::kPi * ch[a] / norm;
    const double bl = (double)T.box_length[q.lvl] / L[a] * (double)T.aspect[a];
    const double Ll = bl * T.redundancy;
    const double sig = bl / 2.0;
    const double ps = bh::kPi * sig;
    out.alpha[2 * (i * d + a)] = 0.0;
    out.alpha[2 * (i * d + a) + 1] = 2.0 * ps * ps / norm;
    a0r *= sqrt((bh::kPi * L[a]) / (Ll * (double)T.N[a])) * sig;
    const double sup = (double)((long long)T.box_length[q.lvl] * T.aspect[a] * T.redundancy);
    out.x0[i * d + a] = (double)q.k[a] * L[a] / sup;
  }
  amp = a0r * frame_phase(T, c, q.lvl, q.k);
  const double2 v = vals[i];
  amp = amp * mkc(v.x, v.y);
  out.a0[2 * i] = amp.re;
  out.a0[2 * i + 1] = amp.im;
  out.omega[i] = norm;
}

// ---------------------------------------------------------------------------
// forward summation kernel
// ---------------------------------------------------------------------------
struct FwdSumArgs {
  bh_fwd_beams b;
  long long count;
  long long total;  // count * (mirror ? 2 : 1)
  double c;
  const double* ts;
  long long nt;
  const double* sensors;
  long long S;
  double dom[3];
  int per[3];
  void* out;
};

constexpr int kChunkB = 64;

template <int D, int SPT, bool F64, bool PER>
__global__ __launch_bounds__(256) void k_forward_sum(FwdSumArgs A) {
  using Rec = typename RecOf<D, F64>::type;
  __shared__ Rec rec[4][kChunkB];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const long long t = (long long)blockIdx.y * 4 + wv;
  const bool tval = t < A.nt;
  const double tv = tval ? A.ts[t] : 0.0;
  const long long s0 = (long long)blockIdx.x * (64 * SPT) + lane;
  double sp[SPT][D];
  for (int i = 0; i < SPT; ++i) {
    const long long s = s0 + 64 * i;
    for (int a = 0; a < D; ++a) sp[i][a] = (s < A.S) ? A.sensors[s * D + a] : 0.0;
  }
  double acc[SPT];
  for (int i = 0; i < SPT; ++i) acc[i] = 0.0;
  for (long long cb = 0; cb < A.total; cb += kChunkB) {
    __syncthreads();
    {
      const long long bi = cb + lane;
      Rec& R = rec[wv][lane];
      if (tval && bi < A.total) {
        const bool neg = bi >= A.count;
        const long long j = neg ? bi - A.count : bi;
        double x0[D], p0[D];
        cd al[D];
        for (int a = 0; a < D; ++a) {
          x0[a] = A.b.x0[j * D + a];
          p0[a] = A.b.p0[j * D + a];
          al[a] = mkc(A.b.alpha[2 * (j * D + a)], A.b.alpha[2 * (j * D + a) + 1]);
        }
        const cd a0 = mkc(A.b.a0[2 * j], A.b.a0[2 * j + 1]);
        State<D> St;
        hom_diag_state<D>(x0, p0, al, a0, neg ? -A.c : A.c, tv, St);
        if (PER) {
          for (int a = 0; a < D; ++a)
            if (A.per[a]) St.x[a] = St.x[a] - A.dom[a] * floor(St.x[a] / A.dom[a]);
        }
        make_rec<D>(St, A.b.omega[j], R);
      } else {
        zero_rec<D>(R);
      }
    }
    __syncthreads();
    if (tval) {
      const int nb = (A.total - cb) < kChunkB ? (int)(A.total - cb) : kChunkB;
      if constexpr (F64) {
        for (int bb = 0; bb < nb; ++bb) {
          const Rec R = rec[wv][bb];
          for (int i = 0; i < SPT; ++i) {
            double dl[D];
            make_delta<D, PER>(sp[i], R.x, A.dom, A.per, dl);
            acc[i] += eval_f64<D>(R, dl);
          }
        }
      } else {
        float part[SPT];
        for (int i = 0; i < SPT; ++i) part[i] = 0.f;
        for (int bb = 0; bb < nb; ++bb) {
          const Rec R = rec[wv][bb];
          for (int i = 0; i < SPT; ++i) {
            double dl[D];
            make_delta<D, PER>(sp[i], R.x, A.dom, A.per, dl);
            eval_split<D>(R, dl, part[i]);
          }
        }
        for (int i = 0; i < SPT; ++i) acc[i] += (double)part[i];
      }
    }
  }
  if (tval) {
    for (int i = 0; i < SPT; ++i) {
      const long long s = s0 + 64 * i;
      if (s < A.S) {
        if constexpr (F64) ((double*)A.out)[t * A.S + s] = acc[i];
        else ((float*)A.out)[t * A.S + s] = (float)acc[i];
      }
    }
  }
}

// materialised trajectories (SolverFn): one thread per (beam, time)
template <int D>
__global__ __launch_bounds__(128) void k_hom_diag(long long nb, const double* x0, const double* p0, const double* alpha,
                           const double* a0, const double* mode, double c, const double* ts,
                           long long nt, double* xt, double* pt, double* Mt, double* At) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= nb * nt) return;
  const long long b = i / nt, t = i - b * nt;
  double xx[D], pp[D];
  cd al[D];
  for (int a = 0; a < D; ++a) {
    xx[a] = x0[b * D + a];
    pp[a] = p0[b * D + a];
    al[a] = mkc(alpha[2 * (b * D + a)], alpha[2 * (b * D + a) + 1]);
  }
  State<D> S;
  hom_diag_state<D>(xx, pp, al, mkc(a0[2 * b], a0[2 * b + 1]), c * mode[b], ts[t], S);
  for (int a = 0; a < D; ++a) {
    xt[i * D + a] = S.x[a];
    pt[i * D + a] = S.p[a];
    for (int e = 0; e < D; ++e) {
      Mt[2 * ((i * D + a) * D + e)] = S.M[a][e].re;
      Mt[2 * ((i * D + a) * D + e) + 1] = S.M[a][e].im;
    }
  }
  At[2 * i] = S.A.re;
  At[2 * i + 1] = S.A.im;
}

// ---------------------------------------------------------------------------
// adjoint beams (tr_solver_utils.py:205-435, adjoint_solver_utils.py, and
// solve_hom_TR at t = 0, gb_solvers.py:484-594)
// ---------------------------------------------------------------------------
template <int D>
__device__ inline void cinv_det(const cd (&Z)[D][D], cd (&Zi)[D][D], cd& det) {
  if constexpr (D == 1) {
    det = Z[0][0];
    Zi[0][0] = mkc(1.0) / Z[0][0];
  } else if constexpr (D == 2) {
    det = Z[0][0] * Z[1][1] - Z[0][1] * Z[1][0];
    Zi[0][0] = Z[1][1] / det;
    Zi[0][1] = -Z[0][1] / det;
    Zi[1][0] = -Z[1][0] / det;
    Zi[1][1] = Z[0][0] / det;
  } else {
    const cd c00 = Z[1][1] * Z[2][2] - Z[1][2] * Z[2][1];
    const cd c01 = Z[1][2] * Z[2][0] - Z[1][0] * Z[2][2];
    const cd c02 = Z[1][0] * Z[2][1] - Z[1][1] * Z[2][0];
    det = Z[0][0] * c00 + Z[0][1] * c01 + Z[0][2] * c02;
    Zi[0][0] = c00 / det;
    Zi[1][0] = c01 / det;
    Zi[2][0] = c02 / det;
    Zi[0][1] = (Z[0][2] * Z[2][1] - Z[0][1] * Z[2][2]) / det;
    Zi[1][1] = (Z[0][0] * Z[2][2] - Z[0][2] * Z[2][0]) / det;
    Zi[2][1] = (Z[0][1] * Z[2][0] - Z[0][0] * Z[2][1]) / det;
    Zi[0][2] = (Z[0][1] * Z[1][2] - Z[0][2] * Z[1][1]) / det;
    Zi[1][2] = (Z[0][2] * Z[1][0] - Z[0][0] * Z[1][2]) / det;
    Zi[2][2] = (Z[0][0] * Z[1][1] - Z[0][1] * Z[1][0]) / det;
  }
}

// The closed form of solve_hom_TR (gb_solvers.py:484-594) over a time
// offset dt from the reference state (xT, pT, MT, aT): x = xT + c0 p^ dt,
// Z = I + c0 dt P_perp MT / |p| (P_perp = I - p p^T / |p|^2), M = MT Z^-1,
// A = aT / sqrt(det Z) with the principal root.
template <int D>
__device__ inline void hom_tr_step(const double* xT, const double* pT, const cd (&MT)[D][D], cd aT,
                                   double c0, double dt, double* x, cd (&M)[D][D], cd& A) {
  double np2 = 0.0;
  for (int a = 0; a < D; ++a) np2 += pT[a] * pT[a];
  const double np = sqrt(np2);
  for (int a = 0; a < D; ++a) x[a] = xT[a] + (c0 * (pT[a] / np)) * dt;
  cd Z[D][D], Zi[D][D], det;
  for (int a = 0; a < D; ++a)
    for (int b = 0; b < D; ++b) {
      cd pm = mkc(0.0);
      for (int e = 0; e < D; ++e) {
        const double pp = ((a == e) ? 1.0 : 0.0) - pT[a] * pT[e] / np2;
        pm = pm + pp * MT[e][b];
      }
      Z[a][b] = ((c0 * dt) * pm) / np;
      if (a == b) Z[a][b] = mkc(1.0) + Z[a][b];
    }
  cinv_det<D>(Z, Zi, det);
  for (int a = 0; a < D; ++a)
    for (int b = 0; b < D; ++b) {
      cd s = mkc(0.0);
      for (int e = 0; e < D; ++e) s = s + MT[a][e] * Zi[e][b];
      M[a][b] = s;
    }
  A = aT / csqrt_(det);
}

// solve_hom_TR as a SolverFn: thread = (beam, time), states at ts[beam][j]
// measured from ts[beam][0]
template <int D>
__global__ __launch_bounds__(128) void k_hom_tr(long long nb, const double* xT, const double* pT,
                                                const double* mT, const double* aT,
                                                const double* mode, double c, const double* ts,
                                                long long nt, double* xt, double* pt, double* Mt,
                                                double* At) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= nb * nt) return;
  const long long k = i / nt, j = i - k * nt;
  double x0[D], p0[D];
  cd MT[D][D];
  for (int a = 0; a < D; ++a) {
    x0[a] = xT[k * D + a];
    p0[a] = pT[k * D + a];
    for (int b = 0; b < D; ++b)
      MT[a][b] = mkc(mT[2 * ((k * D + a) * D + b)], mT[2 * ((k * D + a) * D + b) + 1]);
  }
  double x[D];
  cd M[D][D], A;
  hom_tr_step<D>(x0, p0, MT, mkc(aT[2 * k], aT[2 * k + 1]), c * mode[k],
                 ts[k * nt + j] - ts[k * nt], x, M, A);
  for (int a = 0; a < D; ++a) {
    xt[i * D + a] = x[a];
    pt[i * D + a] = p0[a];
    for (int b = 0; b < D; ++b) {
      Mt[2 * ((i * D + a) * D + b)] = M[a][b].re;
      Mt[2 * ((i * D + a) * D + b) + 1] = M[a][b].im;
    }
  }
  At[2 * i] = A.re;
  At[2 * i + 1] = A.im;
}

struct AdjArgs {
  TilingDev T;
  const int32_t* centres;
  double L[3];
  const long long* idx;
  const double2* vals;
  long long k;
  int normal_axis;
  double normal_value;
  double inward;
  double c;
  double guard;
  int swap;
  int time_reversal;   // 1: compute_TR_parameters only (no B^-1, no exclusion)
  bh_static_beams out;
  // heterogeneous medium: c = c(x_T) from the speed model, and the beams are
  // left at their terminal time (t_start) for the ODE integrator
  int have_speed;
  SpeedDev S;
  double* t_start;   // per beam (terminal output only)
  double* mode;      // per beam signum (terminal output only)
};

template <int D>
__global__ __launch_bounds__(128) void k_adjoint_params(AdjArgs A) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= A.k) return;
  const TilingDev& T = A.T;
  const long long flat = A.idx[i];
  if (flat < 0 || flat >= T.coeffs_cumsum[T.num_levels]) {
    // -1: an inert padding beam; any other index outside the coefficient
    // vector: NaN amplitude and the status word (as in k_forward_params)
    const double amp = flat == -1 ? 0.0 : kQuietNaN;
    if (flat != -1) flag_status(BH_STATUS_ADJ_INDEX);
    for (int a = 0; a < D; ++a) {
      A.out.x[i * D + a] = 0.0;
      A.out.p[i * D + a] = 1.0;
      for (int b = 0; b < D; ++b) {
        A.out.M[2 * ((i * D + a) * D + b)] = 0.0;
        A.out.M[2 * ((i * D + a) * D + b) + 1] = a == b ? 1.0 : 0.0;
      }
    }
    A.out.A[2 * i] = amp;
    A.out.A[2 * i + 1] = amp;
    A.out.omega[i] = 1.0;
    if (A.t_start) {
      A.t_start[i] = 0.0;
      A.mode[i] = 1.0;
    }
    return;
  }
  const Decoded q = decode_flat(T, flat);
  const int32_t* cen = A.centres + q.box * D;
  double ch[D], bl[D], Ll[D], sig[D];
  for (int a = 0; a < D; ++a) {
    ch[a] = (double)cen[a] / A.L[a];
    bl[a] = (double)T.box_length[q.lvl] / A.L[a] * (double)T.aspect[a];
    Ll[a] = bl[a] * T.redundancy;
    sig[a] = bl[a] / 2.0;
  }
  const double tau = ch[0];
  const double sgn_tau = (tau > 0.0) ? 1.0 : ((tau < 0.0) ? -1.0 : 0.0);
  const double signum = -sgn_tau;
  const double om = fmax(fabs(tau), 1e-6);
  const double t_start = (double)q.k[0] / Ll[0];
  int perm[D];
  for (int a = 0; a < D; ++a) perm[a] = a;
  if (A.swap && D >= 2) { perm[0] = 1; perm[1] = 0; }
  // detector position (tangential coordinates from the packet index)
  double xT[D];
  {
    int tj = 0;
    for (int a = 0; a < D; ++a) {
      if (D == 1 || a == A.normal_axis) {
        xT[a] = A.normal_value;
      } else {
        xT[a] = (double)q.k[tj + 1] / Ll[tj + 1];
        ++tj;
      }
    }
  }
  // local speed at x_T (tr_solver_utils.py:336-343, adjoint_solver_utils.py:185)
  double c = A.c;
  double gk[D];   // grad c at x_T in the kernel frame (zero in a homogeneous medium)
  for (int a = 0; a < D; ++a) gk[a] = 0.0;
  if (A.have_speed) {
    double xp[D], g[D], H[D * D];
    for (int a = 0; a < D; ++a) xp[a] = xT[perm[a]];   // physical frame
    speed_eval<D>(A.S, xp, c, g, H);
    for (int a = 0; a < D; ++a) gk[a] = g[perm[a]];    // perm is an involution
  }
  // unit momentum on the detector plane
  double pu[D];
  double ptan[D > 1 ? D - 1 : 1];
  double pn;
  if (D == 1) {
    pn = 1.0 / c;   // |xi_tau| / c with xi_tau = +-1 (tr_solver_utils.py:335-341)
    pu[0] = A.inward * pn;
  } else {
    const double ta = fmax(fabs(tau), 1e-6);
    double tn2 = 0.0;
    for (int j = 0; j < D - 1; ++j) {
      ptan[j] = c * ch[j + 1] / ta;
      tn2 += ptan[j] * ptan[j];
    }
    pn = sqrt(fmax(1.0 - tn2, 0.0));
    int tj = 0;
    for (int a = 0; a < D; ++a) {
      if (a == A.normal_axis) {
        pu[a] = A.inward * pn;
      } else {
        pu[a] = ptan[tj];
        ++tj;
      }
    }
  }
  double pT[D];
  for (int a = 0; a < D; ++a) pT[a] = (D == 1) ? bh::kTwoPi * pu[a] : bh::kTwoPi * pu[a] / c;
  cd alpha[D];
  for (int a = 0; a < D; ++a) {
    const double ps = bh::kPi * sig[a];
    alpha[a] = mkc(0.0, 2.0 * ps * ps / om);
  }
  // mT_inverse (tr_solver_utils.py:116-170) with xdot = H_p, pdot = -H_x at
  // x_T, axis 0 as the normal; pdot = 0 in a homogeneous medium
  double np2 = 0.0;
  for (int a = 0; a < D; ++a) np2 += pT[a] * pT[a];
  double np = sqrt(np2);
  double xd[D], pd[D];
  for (int a = 0; a < D; ++a) {
    xd[a] = signum * c * pT[a] / np;
    pd[a] = -(signum * gk[a] * np);
  }
  cd MT[D][D];
  for (int a = 0; a < D; ++a)
    for (int b = 0; b < D; ++b) MT[a][b] = mkc(0.0);
  {
    double pdx = pd[0] * xd[0];
    for (int j = 1; j < D; ++j) pdx -= pd[j] * xd[j];
    cd num = alpha[0] + mkc(pdx);
    for (int j = 1; j < D; ++j) num = num + (alpha[j] * xd[j]) * xd[j];
    MT[0][0] = num / (xd[0] * xd[0]);
    for (int j = 1; j < D; ++j) {
      MT[0][j] = (mkc(pd[j]) - alpha[j] * xd[j]) / xd[0];
      MT[j][0] = MT[0][j];
      MT[j][j] = alpha[j];
    }
  }
  double amp_r = 1.0;
  for (int a = 0; a < D; ++a)
    amp_r *= sqrt((bh::kPi * A.L[a]) / (Ll[a] * (double)T.N[a])) * sig[a];
  cd aT = amp_r * frame_phase(T, cen, q.lvl, q.k);
  bool bad = fabs(pT[A.normal_axis]) == 0.0;
  // principal B^-1 (adjoint_solver_utils.py:17-72), negated for the advanced propagator
  const double Om = bh::kTwoPi * tau;
  double qsq = 0.0;
  for (int j = 1; j < D; ++j) {
    const double Q = bh::kTwoPi * ch[j];
    qsq += Q * Q;
  }
  const double rad = Om * Om - c * c * qsq;
  const double gam = sqrt(fmax(rad, 0.0));
  const bool valid = (rad > 0.0) && (fabs(Om) > 0.0) && (gam > A.guard * fabs(Om));
  const double sgo = (Om > 0.0) ? 1.0 : ((Om < 0.0) ? -1.0 : 0.0);
  const cd binv = A.time_reversal ? mkc(1.0) : (valid ? mkc(0.0, sgo / (2.0 * c * gam)) : mkc(0.0));
  if (bad) aT = mkc(0.0);
  if (bad || (!valid && !A.time_reversal)) {
    for (int a = 0; a < D; ++a) {
      pT[a] = 1.0;
      for (int b = 0; b < D; ++b) MT[a][b] = (a == b) ? mkc(0.0, 1.0) : mkc(0.0);
    }
  }
  aT = aT * binv;
  const double2 v = A.vals[i];
  aT = aT * mkc(v.x, v.y);
  if (A.t_start) {
    // terminal state in the physical frame; the medium ODE takes it to t = 0
    for (int a = 0; a < D; ++a) {
      A.out.x[i * D + a] = xT[perm[a]];
      A.out.p[i * D + a] = pT[perm[a]];
      for (int b = 0; b < D; ++b) {
        A.out.M[2 * ((i * D + a) * D + b)] = MT[perm[a]][perm[b]].re;
        A.out.M[2 * ((i * D + a) * D + b) + 1] = MT[perm[a]][perm[b]].im;
      }
    }
    A.out.A[2 * i] = aT.re;
    A.out.A[2 * i + 1] = aT.im;
    A.out.omega[i] = om;
    A.t_start[i] = t_start;
    A.mode[i] = signum;
    return;
  }
  // propagate t_start -> 0
  double x[D];
  cd M0[D][D], A0;
  hom_tr_step<D>(xT, pT, MT, aT, c * signum, 0.0 - t_start, x, M0, A0);
  for (int a = 0; a < D; ++a) {
    A.out.x[i * D + a] = x[perm[a]];
    A.out.p[i * D + a] = pT[perm[a]];
    for (int b = 0; b < D; ++b) {
      A.out.M[2 * ((i * D + a) * D + b)] = M0[perm[a]][perm[b]].re;
      A.out.M[2 * ((i * D + a) * D + b) + 1] = M0[perm[a]][perm[b]].im;
    }
  }
  A.out.A[2 * i] = A0.re;
  A.out.A[2 * i + 1] = A0.im;
  A.out.omega[i] = om;
}

// ---------------------------------------------------------------------------
// static (single time level) summation over points
// ---------------------------------------------------------------------------
template <int D, bool F64>
__global__ void k_static_records(bh_static_beams b, void* recs) {
  using Rec = typename RecOf<D, F64>::type;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= b.count) return;
  State<D> S;
  for (int a = 0; a < D; ++a) {
    S.x[a] = b.x[i * D + a];
    S.p[a] = b.p[i * D + a];
    for (int e = 0; e < D; ++e)
      S.M[a][e] = mkc(b.M[2 * ((i * D + a) * D + e)], b.M[2 * ((i * D + a) * D + e) + 1]);
  }
  S.A = mkc(b.A[2 * i], b.A[2 * i + 1]);
  make_rec<D>(S, b.omega[i], ((Rec*)recs)[i]);
}

template <int D, int PPT, bool F64>
__global__ __launch_bounds__(256) void k_static_sum(const void* __restrict__ recs_, long long count,
                                                    const double* __restrict__ pts, long long P,
                                                    double scale, void* out) {
  using Rec = typename RecOf<D, F64>::type;
  const Rec* recs = (const Rec*)recs_;
  __shared__ Rec rec[kChunkB];
  const long long p0 = (long long)blockIdx.x * (256 * PPT) + threadIdx.x;
  double xp[PPT][D];
  for (int i = 0; i < PPT; ++i) {
    const long long p = p0 + 256 * i;
    for (int a = 0; a < D; ++a) xp[i][a] = (p < P) ? pts[p * D + a] : 0.0;
  }
  double acc[PPT];
  for (int i = 0; i < PPT; ++i) acc[i] = 0.0;
  for (long long cb = 0; cb < count; cb += kChunkB) {
    __syncthreads();
    if (threadIdx.x < kChunkB) {
      if (cb + threadIdx.x < count) rec[threadIdx.x] = recs[cb + threadIdx.x];
      else zero_rec<D>(rec[threadIdx.x]);
    }
    __syncthreads();
    const int nb = (count - cb) < kChunkB ? (int)(count - cb) : kChunkB;
    if constexpr (F64) {
      for (int bb = 0; bb < nb; ++bb) {
        const Rec R = rec[bb];
        for (int i = 0; i < PPT; ++i) {
          double dl[D];
          for (int a = 0; a < D; ++a) dl[a] = xp[i][a] - R.x[a];
          acc[i] += eval_f64<D>(R, dl);
        }
      }
    } else {
      float part[PPT];
      for (int i = 0; i < PPT; ++i) part[i] = 0.f;
      for (int bb = 0; bb < nb; ++bb) {
        const Rec R = rec[bb];
        for (int i = 0; i < PPT; ++i) {
          double dl[D];
          for (int a = 0; a < D; ++a) dl[a] = xp[i][a] - R.x[a];
          eval_split<D>(R, dl, part[i]);
        }
      }
      for (int i = 0; i < PPT; ++i) acc[i] += (double)part[i];
    }
  }
  for (int i = 0; i < PPT; ++i) {
    const long long p = p0 + 256 * i;
    if (p < P) {
      if constexpr (F64) ((double*)out)[p] = scale * acc[i];
      else ((float*)out)[p] = (float)(scale * acc[i]);
    }
  }
}

template <int D, int SPT, bool F64>
void launch_fwd(const FwdSumArgs& a, bool per, hipStream_t s) {
  dim3 grid((unsigned)((a.S + 64 * SPT - 1) / (64 * SPT)), (unsigned)((a.nt + 3) / 4));
  if (per) hipLaunchKernelGGL((k_forward_sum<D, SPT, F64, true>), grid, dim3(256), 0, s, a);
  else hipLaunchKernelGGL((k_forward_sum<D, SPT, F64, false>), grid, dim3(256), 0, s, a);
}

template <int D, bool F64>
void launch_fwd_spt(const FwdSumArgs& a, bool per, hipStream_t s) {
  // sensors per lane: enough blocks for 256 CUs first, then register reuse
  const long long tb = (a.nt + 3) / 4;
  if (a.S >= 256 && tb * ((a.S + 255) / 256) >= 1024) launch_fwd<D, 4, F64>(a, per, s);
  else if (a.S >= 128 && tb * ((a.S + 127) / 128) >= 512) launch_fwd<D, 2, F64>(a, per, s);
  else launch_fwd<D, 1, F64>(a, per, s);
}

}  // namespace

namespace {
// zero fill in 4-byte words (a kernel rather than hipMemsetAsync, so the
// call stays correct inside a captured graph)
__global__ void k_zero4(void* out, long long n4) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n4) ((unsigned*)out)[i] = 0u;
}
}  // namespace

extern "C" int bh_forward_params(const bh_tiling* t, const double* grid_size,
                                 const int64_t* idx, const double* vals, int64_t k,
                                 bh_fwd_beams out, bh_stream s) {
  bh::clear_error();
  if (int rc = bh::check_tiling(t)) return rc;
  if (k < 0) BH_FAIL(BH_ERR_ARG, "bh_forward_params: k=%lld", (long long)k);
  if (k == 0) return BH_OK;
  if (!grid_size || !idx || !vals || !out.x0 || !out.p0 || !out.omega || !out.alpha || !out.a0)
    BH_FAIL(BH_ERR_ARG, "bh_forward_params: null pointer");
  const bh::TilingDev T = bh::to_dev(t);
  const double L1 = t->ndim > 1 ? grid_size[1] : 1.0, L2 = t->ndim > 2 ? grid_size[2] : 1.0;
  hipLaunchKernelGGL(k_forward_params, dim3(bh::grid1d(k, 128)), dim3(128), 0, (hipStream_t)s,
                     T, t->centres, grid_size[0], L1, L2, (const long long*)idx,
                     (const double2*)vals, (long long)k, out);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

extern "C" int bh_forward_sum(const bh_fwd_beams* b, int ndim, int mirror, double c,
                              const double* ts, int64_t nt, const double* sensors, int64_t S,
                              const double* domain_size, const int32_t* periodic, int precision,
                              void* out, bh_stream s) {
  bh::clear_error();
  if (!b || !ts || !sensors || !out) BH_FAIL(BH_ERR_ARG, "bh_forward_sum: null pointer");
  if (ndim < 1 || ndim > 3) BH_FAIL(BH_ERR_ARG, "bh_forward_sum: ndim=%d", ndim);
  if (nt <= 0 || S <= 0) BH_FAIL(BH_ERR_ARG, "bh_forward_sum: nt=%lld S=%lld", (long long)nt,
                                 (long long)S);
  if ((nt + 3) / 4 > 65535) BH_FAIL(BH_ERR_UNSUPPORTED, "bh_forward_sum: nt=%lld too large",
                                    (long long)nt);
  if (precision != BH_PRECISION_FP64 && precision != BH_PRECISION_FP32_SPLIT)
    BH_FAIL(BH_ERR_ARG, "bh_forward_sum: precision=%d", precision);
  FwdSumArgs a;
  a.b = *b;
  a.count = b->count;
  a.total = b->count * (mirror ? 2 : 1);
  a.c = c;
  a.ts = ts;
  a.nt = nt;
  a.sensors = sensors;
  a.S = S;
  bool per = false;
  for (int i = 0; i < 3; ++i) {
    a.dom[i] = (domain_size && i < ndim) ? domain_size[i] : 1.0;
    a.per[i] = (periodic && i < ndim) ? (periodic[i] != 0) : 0;
    per = per || a.per[i];
  }
  a.out = out;
  hipStream_t st = (hipStream_t)s;
  if (a.total == 0) {
    const long long n4 = (long long)nt * S * (precision == BH_PRECISION_FP64 ? 2 : 1);
    hipLaunchKernelGGL(k_zero4, dim3(bh::grid1d(n4, 256)), dim3(256), 0, st, out, n4);
    BH_LAUNCH_CHECK();
    return BH_OK;
  }
  if (!b->x0 || !b->p0 || !b->omega || !b->alpha || !b->a0)
    BH_FAIL(BH_ERR_ARG, "bh_forward_sum: null beam array");
  const bool f64 = precision == BH_PRECISION_FP64;
  if (ndim == 1) { if (f64) launch_fwd_spt<1, true>(a, per, st); else launch_fwd_spt<1, false>(a, per, st); }
  else if (ndim == 2) { if (f64) launch_fwd_spt<2, true>(a, per, st); else launch_fwd_spt<2, false>(a, per, st); }
  else { if (f64) launch_fwd_spt<3, true>(a, per, st); else launch_fwd_spt<3, false>(a, per, st); }
  BH_LAUNCH_CHECK();
  return BH_OK;
}

extern "C" int bh_solve_hom_TR(int64_t b, int ndim, const double* xT, const double* pT,
                               const double* mT, const double* aT, const double* mode, double c,
                               const double* ts, int64_t nt, double* xt, double* pt, double* Mt,
                               double* At, bh_stream s) {
  bh::clear_error();
  if (ndim < 1 || ndim > 3) BH_FAIL(BH_ERR_ARG, "bh_solve_hom_TR: ndim=%d", ndim);
  if (b < 0 || nt < 0) BH_FAIL(BH_ERR_ARG, "bh_solve_hom_TR: b=%lld nt=%lld", (long long)b,
                               (long long)nt);
  if (b == 0 || nt == 0) return BH_OK;
  if (!xT || !pT || !mT || !aT || !mode || !ts || !xt || !pt || !Mt || !At)
    BH_FAIL(BH_ERR_ARG, "bh_solve_hom_TR: null pointer");
  const unsigned g = bh::grid1d(b * nt, 128);
  hipStream_t st = (hipStream_t)s;
  if (ndim == 1) hipLaunchKernelGGL(k_hom_tr<1>, dim3(g), dim3(128), 0, st, (long long)b, xT, pT, mT, aT, mode, c, ts, (long long)nt, xt, pt, Mt, At);
  else if (ndim == 2) hipLaunchKernelGGL(k_hom_tr<2>, dim3(g), dim3(128), 0, st, (long long)b, xT, pT, mT, aT, mode, c, ts, (long long)nt, xt, pt, Mt, At);
  else hipLaunchKernelGGL(k_hom_tr<3>, dim3(g), dim3(128), 0, st, (long long)b, xT, pT, mT, aT, mode, c, ts, (long long)nt, xt, pt, Mt, At);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

extern "C" int bh_solve_hom_diag(int64_t b, int ndim, const double* x0, const double* p0,
                                 const double* alpha, const double* a0, const double* mode,
                                 double c, const double* ts, int64_t nt, double* xt, double* pt,
                                 double* Mt, double* At, bh_stream s) {
  bh::clear_error();
  if (ndim < 1 || ndim > 3) BH_FAIL(BH_ERR_ARG, "bh_solve_hom_diag: ndim=%d", ndim);
  if (b <= 0 || nt <= 0) return BH_OK;
  if (!x0 || !p0 || !alpha || !a0 || !mode || !ts || !xt || !pt || !Mt || !At)
    BH_FAIL(BH_ERR_ARG, "bh_solve_hom_diag: null pointer");
  const unsigned g = bh::grid1d(b * nt, 128);
  hipStream_t st = (hipStream_t)s;
  if (ndim == 1) hipLaunchKernelGGL(k_hom_diag<1>, dim3(g), dim3(128), 0, st, (long long)b, x0, p0, alpha, a0, mode, c, ts, (long long)nt, xt, pt, Mt, At);
  else if (ndim == 2) hipLaunchKernelGGL(k_hom_diag<2>, dim3(g), dim3(128), 0, st, (long long)b, x0, p0, alpha, a0, mode, c, ts, (long long)nt, xt, pt, Mt, At);
  else hipLaunchKernelGGL(k_hom_diag<3>, dim3(g), dim3(128), 0, st, (long long)b, x0, p0, alpha, a0, mode, c, ts, (long long)nt, xt, pt, Mt, At);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

static int boundary_params(const bh_tiling* t, const double* data_grid_size, const int64_t* idx,
                           const double* vals, int64_t k, int normal_axis, double normal_value,
                           double inward_sign, double c, double guard, int swap_axes, int tr,
                           bh_static_beams out, bh_stream s, const bh_speed* sp = nullptr,
                           double* t_start = nullptr, double* mode = nullptr);

extern "C" int bh_adjoint_params(const bh_tiling* t, const double* data_grid_size,
                                 const int64_t* idx, const double* vals, int64_t k,
                                 int normal_axis, double normal_value, double inward_sign,
                                 double c, double guard, int swap_axes, bh_static_beams out,
                                 bh_stream s) {
  return boundary_params(t, data_grid_size, idx, vals, k, normal_axis, normal_value, inward_sign,
                         c, guard, swap_axes, 0, out, s);
}

extern "C" int bh_tr_params(const bh_tiling* t, const double* data_grid_size, const int64_t* idx,
                            const double* vals, int64_t k, int normal_axis, double normal_value,
                            double inward_sign, double c, int swap_axes, bh_static_beams out,
                            bh_stream s) {
  return boundary_params(t, data_grid_size, idx, vals, k, normal_axis, normal_value, inward_sign,
                         c, 0.0, swap_axes, 1, out, s);
}

extern "C" int bh_boundary_terminal(const bh_tiling* t, const double* data_grid_size,
                                    const int64_t* idx, const double* vals, int64_t k,
                                    int normal_axis, double normal_value, double inward_sign,
                                    const bh_speed* sp, double guard, int time_reversal,
                                    int swap_axes, bh_static_beams out, double* t_start,
                                    double* mode, bh_stream s) {
  if (!sp || !t_start || !mode) {
    bh::clear_error();
    BH_FAIL(BH_ERR_ARG, "bh_boundary_terminal: null speed model or output");
  }
  return boundary_params(t, data_grid_size, idx, vals, k, normal_axis, normal_value, inward_sign,
                         1.0, guard, swap_axes, time_reversal ? 1 : 0, out, s, sp, t_start, mode);
}

static int boundary_params(const bh_tiling* t, const double* data_grid_size, const int64_t* idx,
                           const double* vals, int64_t k, int normal_axis, double normal_value,
                           double inward_sign, double c, double guard, int swap_axes, int tr,
                           bh_static_beams out, bh_stream s, const bh_speed* sp, double* t_start,
                           double* mode) {
  bh::clear_error();
  if (int rc = bh::check_tiling(t)) return rc;
  if (k < 0) BH_FAIL(BH_ERR_ARG, "bh_adjoint_params: k=%lld", (long long)k);
  if (k == 0) return BH_OK;
  if (!data_grid_size || !idx || !vals || !out.x || !out.p || !out.M || !out.A || !out.omega)
    BH_FAIL(BH_ERR_ARG, "bh_adjoint_params: null pointer");
  if (normal_axis < 0 || normal_axis >= t->ndim)
    BH_FAIL(BH_ERR_ARG, "bh_adjoint_params: normal_axis=%d", normal_axis);
  if (!(c > 0.0)) BH_FAIL(BH_ERR_ARG, "bh_adjoint_params: c must be positive");
  AdjArgs a;
  a.T = bh::to_dev(t);
  a.centres = t->centres;
  for (int i = 0; i < 3; ++i) a.L[i] = i < t->ndim ? data_grid_size[i] : 1.0;
  a.idx = (const long long*)idx;
  a.vals = (const double2*)vals;
  a.k = k;
  a.normal_axis = normal_axis;
  a.normal_value = normal_value;
  a.inward = inward_sign;
  a.c = c;
  a.guard = guard;
  a.swap = swap_axes;
  a.time_reversal = tr;
  a.out = out;
  a.have_speed = sp ? 1 : 0;
  a.S = SpeedDev{};
  if (sp) {
    if (sp->ndim != t->ndim || !sp->values)
      BH_FAIL(BH_ERR_ARG, "bh_boundary_terminal: speed model ndim=%d for %d-D data", sp->ndim,
              t->ndim);
    a.S.ndim = sp->ndim;
    a.S.method = sp->method;
    a.S.boundary = sp->boundary;
    for (int i = 0; i < 3; ++i) {
      a.S.N[i] = i < sp->ndim ? sp->N[i] : 1;
      a.S.spacing[i] = i < sp->ndim ? sp->spacing[i] : 1.0;
      a.S.origin[i] = i < sp->ndim ? sp->origin[i] : 0.0;
    }
    a.S.values = sp->values;
    a.S.lam = sp->lam;
  }
  a.t_start = t_start;
  a.mode = mode;
  const unsigned g = bh::grid1d(k, 128);
  hipStream_t st = (hipStream_t)s;
  if (t->ndim == 1) hipLaunchKernelGGL(k_adjoint_params<1>, dim3(g), dim3(128), 0, st, a);
  else if (t->ndim == 2) hipLaunchKernelGGL(k_adjoint_params<2>, dim3(g), dim3(128), 0, st, a);
  else hipLaunchKernelGGL(k_adjoint_params<3>, dim3(g), dim3(128), 0, st, a);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

extern "C" size_t bh_static_sum_workspace_size(int64_t count, int ndim) {
  if (count <= 0 || ndim < 1 || ndim > 3) return 0;
  size_t r = 0;
  if (ndim == 1) r = sizeof(RecD<1>);
  else if (ndim == 2) r = sizeof(RecD<2>);
  else r = sizeof(RecD<3>);
  return (size_t)count * r;
}

extern "C" int bh_static_sum(const bh_static_beams* b, int ndim, const double* points,
                             int64_t P, double scale, int precision, void* out, void* ws,
                             size_t ws_bytes, bh_stream s) {
  bh::clear_error();
  if (!b || !points || !out) BH_FAIL(BH_ERR_ARG, "bh_static_sum: null pointer");
  if (ndim < 1 || ndim > 3) BH_FAIL(BH_ERR_ARG, "bh_static_sum: ndim=%d", ndim);
  if (precision != BH_PRECISION_FP64 && precision != BH_PRECISION_FP32_SPLIT)
    BH_FAIL(BH_ERR_ARG, "bh_static_sum: precision=%d", precision);
  if (P <= 0) return BH_OK;
  hipStream_t st = (hipStream_t)s;
  const bool f64 = precision == BH_PRECISION_FP64;
  if (b->count == 0) {
    const long long n4 = (long long)P * (f64 ? 2 : 1);
    hipLaunchKernelGGL(k_zero4, dim3(bh::grid1d(n4, 256)), dim3(256), 0, st, out, n4);
    BH_LAUNCH_CHECK();
    return BH_OK;
  }
  const size_t need = bh_static_sum_workspace_size(b->count, ndim);
  if (!ws || ws_bytes < need) BH_FAIL(BH_ERR_WORKSPACE, "bh_static_sum: workspace %zu < %zu",
                                      ws_bytes, need);
  const unsigned gr = bh::grid1d(b->count, 128);
  const long long cnt = b->count;
#define BH_STATIC(D, F)                                                                    \
  do {                                                                                     \
    hipLaunchKernelGGL((k_static_records<D, F>), dim3(gr), dim3(128), 0, st, *b, ws);       \
    const int ppt = (P >= 256LL * 4 * 1024) ? 4 : ((P >= 256LL * 2 * 512) ? 2 : 1);         \
    if (ppt == 4)                                                                          \
      hipLaunchKernelGGL((k_static_sum<D, 4, F>), dim3(bh::grid1d(P, 1024)), dim3(256), 0, \
                         st, ws, cnt, points, (long long)P, scale, out);                   \
    else if (ppt == 2)                                                                     \
      hipLaunchKernelGGL((k_static_sum<D, 2, F>), dim3(bh::grid1d(P, 512)), dim3(256), 0,  \
                         st, ws, cnt, points, (long long)P, scale, out);                   \
    else                                                                                   \
      hipLaunchKernelGGL((k_static_sum<D, 1, F>), dim3(bh::grid1d(P, 256)), dim3(256), 0,  \
                         st, ws, cnt, points, (long long)P, scale, out);                   \
  } while (0)
  if (ndim == 1) { if (f64) BH_STATIC(1, true); else BH_STATIC(1, false); }
  else if (ndim == 2) { if (f64) BH_STATIC(2, true); else BH_STATIC(2, false); }
  else { if (f64) BH_STATIC(3, true); else BH_STATIC(3, false); }
#undef BH_STATIC
  BH_LAUNCH_CHECK();
  return BH_OK;
}

namespace bh {
int beams_status(int clear, unsigned* bits, hipStream_t s) {
  unsigned v = 0u;
  if (hipMemcpyFromSymbolAsync(&v, HIP_SYMBOL(g_status), sizeof(v), 0, hipMemcpyDeviceToHost,
                               s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    BH_FAIL(BH_ERR_HIP, "bh_device_status: reading the status word failed");
  if (clear && v) {
    const unsigned zero = 0u;
    if (hipMemcpyToSymbolAsync(HIP_SYMBOL(g_status), &zero, sizeof(zero), 0,
                               hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      BH_FAIL(BH_ERR_HIP, "bh_device_status: clearing the status word failed");
  }
  *bits |= v;
  return BH_OK;
}
}  // namespace bh

extern "C" int bh_device_status(int clear, uint32_t* status, bh_stream s) {
  bh::clear_error();
  if (!status) BH_FAIL(BH_ERR_ARG, "bh_device_status: null status");
  unsigned bits = 0u;
  if (int rc = bh::beams_status(clear, &bits, (hipStream_t)s)) return rc;
  if (int rc = bh::transpose_status(clear, &bits, (hipStream_t)s)) return rc;
  *status = bits;
  return BH_OK;
}
