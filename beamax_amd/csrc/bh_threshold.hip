// Coefficient thresholding strategies other than top_n
// (forward_solver_utils.py:14-178): the selected indices in ascending index
// order and their values, with the count written to device memory.
//
//   hard           |c| > val
//   percentile     |c| > P, P = the val-th percentile of |c| (linear
//                  interpolation between order statistics, jnp.percentile)
//   hard_reassign  |c| >= val max|c| (and max|c| > 0, |c| > 0); values scaled
//                  by sqrt(sum |c|^2 / sum of the retained |c|^2)
//   bao_energy     |c |centre|^d| > val (integer box centre of the coefficient)
//   perc_max_abs   |c| > val max|c|
//
// |c| is sqrt(re*re + im*im) with each operation correctly rounded (the
// top_n key).  Reductions (max, energies) and the ordered compaction use
// fixed orders, so results are deterministic.  The reference pads the
// percentile and perc_max_abs outputs to n entries with index -1 and value 0
// (zero-amplitude beams); here only the selected entries are returned.
#include "bh_internal.hpp"

namespace {

constexpr int kTB = 256;
constexpr int kTI = 16;
constexpr int kTChunk = kTB * kTI;

enum { kHard = 0, kPercentile = 1, kReassign = 2, kBao = 3, kPercMaxAbs = 4 };

__device__ inline double mag(double2 v) {
  return __dsqrt_rn(__dadd_rn(__dmul_rn(v.x, v.x), __dmul_rn(v.y, v.y)));
}

struct ThrState {
  double max_abs;    // max |c|
  double energy;     // sum |c|^2
  double thr;        // comparison threshold
  double ratio;      // hard_reassign value scale
  unsigned long long klo, khi;   // percentile order statistics (keys)
};

// key[i] = bits of the compared magnitude; per-block max |c| and sum |c|^2
__global__ __launch_bounds__(kTB) void k_thr_keys(const double2* __restrict__ c, long long n,
                                                  int strategy, bh::TilingDev T,
                                                  const int32_t* __restrict__ centres,
                                                  unsigned long long* __restrict__ key,
                                                  double* __restrict__ part) {
  __shared__ double smax[kTB], ssum[kTB];
  double mx = 0.0, sm = 0.0;
  const long long base = (long long)blockIdx.x * kTChunk;
  for (int it = 0; it < kTI; ++it) {
    const long long i = base + (long long)it * kTB + threadIdx.x;
    if (i >= n) break;
    const double2 v = c[i];
    const double a = mag(v);
    mx = fmax(mx, a);
    sm += a * a;
    double kv = a;
    if (strategy == kBao) {
      // |centre|^d of the coefficient's box (integer Fourier-index centre)
      int lvl = 0;
      while (lvl + 1 < T.num_levels && i >= T.coeffs_cumsum[lvl + 1]) ++lvl;
      long long P = 1;
      for (int x = 0; x < T.ndim; ++x) P *= T.support[lvl][x];
      const long long box = T.boxes_cumsum[lvl] + (i - T.coeffs_cumsum[lvl]) / P;
      double n2 = 0.0;
      for (int x = 0; x < T.ndim; ++x) {
        const double q = (double)centres[box * T.ndim + x];
        n2 += q * q;
      }
      const double nx = sqrt(n2);
      double w = nx;
      for (int x = 1; x < T.ndim; ++x) w *= nx;
      kv = mag(make_double2(v.x * w, v.y * w));
    }
    key[i] = (unsigned long long)__double_as_longlong(kv);
  }
  smax[threadIdx.x] = mx;
  ssum[threadIdx.x] = sm;
  __syncthreads();
  for (int w = kTB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
      ssum[threadIdx.x] += ssum[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = smax[0];
    part[2 * blockIdx.x + 1] = ssum[0];
  }
}

// one thread: totals in block order, then the comparison threshold
__global__ void k_thr_state(const double* __restrict__ part, long long nblk, int strategy,
                            double val, long long n, ThrState* st) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double mx = 0.0, sm = 0.0;
  for (long long b = 0; b < nblk; ++b) {
    mx = fmax(mx, part[2 * b]);
    sm += part[2 * b + 1];
  }
  st->max_abs = mx;
  st->energy = sm;
  st->ratio = 1.0;
  double thr = val;
  if (strategy == kPercMaxAbs || strategy == kReassign) thr = mx * val;
  if (strategy == kPercentile) {
    // jnp.percentile, method 'linear': q = val / 100, pos = q (n - 1)
    const double pos = (val / 100.0) * (double)(n - 1);
    const double lo = floor(pos);
    const double hw = __dadd_rn(pos, -lo);
    const double a_lo = __longlong_as_double((long long)st->klo);
    const double a_hi = __longlong_as_double((long long)st->khi);
    // separately rounded products and sum (no fused multiply-add)
    thr = __dadd_rn(__dmul_rn(a_lo, __dadd_rn(1.0, -hw)), __dmul_rn(a_hi, hw));
  }
  st->thr = thr;
}

__device__ inline bool thr_flag(int strategy, double a, const ThrState& s) {
  if (strategy == kReassign) return s.max_abs > 0.0 && a >= s.thr && a > 0.0;
  return a > s.thr;
}

// per-block selected counts (and, for hard_reassign, retained energy)
__global__ __launch_bounds__(kTB) void k_thr_count(const unsigned long long* __restrict__ key,
                                                   long long n, int strategy,
                                                   const ThrState* __restrict__ st,
                                                   long long* __restrict__ cnt,
                                                   double* __restrict__ part) {
  __shared__ int scount;
  __shared__ double se[kTB];
  if (threadIdx.x == 0) scount = 0;
  __syncthreads();
  const ThrState s = *st;
  int c = 0;
  double e = 0.0;
  const long long base = (long long)blockIdx.x * kTChunk;
  for (int it = 0; it < kTI; ++it) {
    const long long i = base + (long long)it * kTB + threadIdx.x;
    if (i >= n) break;
    const double a = __longlong_as_double((long long)key[i]);
    if (thr_flag(strategy, a, s)) {
      ++c;
      e += a * a;
    }
  }
  atomicAdd(&scount, c);
  se[threadIdx.x] = e;
  __syncthreads();
  for (int w = kTB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) se[threadIdx.x] += se[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    cnt[blockIdx.x] = scount;
    part[blockIdx.x] = se[0];
  }
}

// one thread: exclusive offsets, total count, hard_reassign ratio
__global__ void k_thr_scan(long long* cnt, const double* __restrict__ part, long long nblk,
                           int strategy, ThrState* st, int64_t* count_out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  long long tot = 0;
  double kept = 0.0;
  for (long long b = 0; b < nblk; ++b) {
    const long long c = cnt[b];
    cnt[b] = tot;
    tot += c;
    kept += part[b];
  }
  if (strategy == kReassign) {
    // ratio = sqrt(total / retained); no retained energy -> nothing survives
    st->ratio = kept > 0.0 ? sqrt(st->energy / kept) : 0.0;
    if (!(kept > 0.0)) tot = 0;
  }
  *count_out = tot;
}

// ordered compaction (ascending index)
__global__ __launch_bounds__(kTB) void k_thr_compact(const double2* __restrict__ c,
                                                     const unsigned long long* __restrict__ key,
                                                     long long n, int strategy,
                                                     const ThrState* __restrict__ st,
                                                     const long long* __restrict__ off,
                                                     int64_t* __restrict__ idx_out,
                                                     double2* __restrict__ val_out) {
  __shared__ int wsum[kTB / 64];
  const ThrState s = *st;
  if (strategy == kReassign && !(s.ratio > 0.0)) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  long long pos = off[blockIdx.x];
  const long long base = (long long)blockIdx.x * kTChunk;
  for (int it = 0; it < kTI; ++it) {
    const long long i = base + (long long)it * kTB + threadIdx.x;
    bool f = false;
    if (i < n) f = thr_flag(strategy, __longlong_as_double((long long)key[i]), s);
    const unsigned long long b = __ballot(f);
    const int before = __popcll(b & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wv] = __popcll(b);
    __syncthreads();
    int wbase = 0, rowtot = 0;
    for (int q = 0; q < kTB / 64; ++q) {
      if (q < wv) wbase += wsum[q];
      rowtot += wsum[q];
    }
    if (f) {
      const long long o = pos + wbase + before;
      idx_out[o] = i;
      double2 v = c[i];
      if (strategy == kReassign) v = make_double2(v.x * s.ratio, v.y * s.ratio);
      val_out[o] = v;
    }
    pos += rowtot;
    __syncthreads();
  }
}

struct ThrWs {
  unsigned long long* key;
  double* part;
  long long* cnt;
  ThrState* st;
  void* kth;
  size_t bytes;
};

ThrWs carve_thr(void* ws, long long n) {
  const long long nblk = (n + kTChunk - 1) / kTChunk;
  char* p = (char*)ws;
  ThrWs w;
  auto take = [&](size_t bytes) {
    char* r = p;
    p += (bytes + 255) / 256 * 256;
    return (void*)r;
  };
  w.key = (unsigned long long*)take(n * 8);
  w.part = (double*)take(2 * nblk * 8);
  w.cnt = (long long*)take(nblk * 8);
  w.st = (ThrState*)take(sizeof(ThrState));
  w.kth = take(bh::kth_key_ws_bytes());
  w.bytes = (size_t)(p - (char*)ws);
  return w;
}

}  // namespace

extern "C" size_t bh_threshold_workspace_size(int64_t n) {
  if (n <= 0) return 0;
  return carve_thr(nullptr, n).bytes;
}

extern "C" int bh_threshold(const double* c, int64_t n, int strategy, double val,
                            const bh_tiling* t, int64_t* idx_out, double* val_out,
                            int64_t* count_out, void* ws, size_t ws_bytes, bh_stream s) {
  bh::clear_error();
  if (n <= 0) BH_FAIL(BH_ERR_ARG, "bh_threshold: empty input (n=%lld)", (long long)n);
  if (strategy < kHard || strategy > kPercMaxAbs)
    BH_FAIL(BH_ERR_ARG, "Invalid thresholding strategy: %d", strategy);
  if (!c || !idx_out || !val_out || !count_out || !ws)
    BH_FAIL(BH_ERR_ARG, "bh_threshold: null pointer");
  if (strategy == kPercentile && !(val >= 0.0 && val <= 100.0))
    BH_FAIL(BH_ERR_ARG, "percentile threshold must lie in [0, 100].");
  bh::TilingDev T{};
  const int32_t* centres = nullptr;
  if (strategy == kBao) {
    if (!t) BH_FAIL(BH_ERR_ARG, "bao_energy thresholding requires wpt.");
    if (int rc = bh::check_tiling(t)) return rc;
    if (t->coeffs_cumsum[t->num_levels] != n)
      BH_FAIL(BH_ERR_ARG, "bh_threshold: n=%lld does not match the tiling (%lld coefficients)",
              (long long)n, (long long)t->coeffs_cumsum[t->num_levels]);
    T = bh::to_dev(t);
    centres = t->centres;
  }
  const size_t need = bh_threshold_workspace_size(n);
  if (ws_bytes < need) BH_FAIL(BH_ERR_WORKSPACE, "bh_threshold: workspace %zu < %zu", ws_bytes,
                               need);
  hipStream_t st = (hipStream_t)s;
  ThrWs w = carve_thr(ws, n);
  const long long nblk = (n + kTChunk - 1) / kTChunk;
  hipLaunchKernelGGL(k_thr_keys, dim3((unsigned)nblk), dim3(kTB), 0, st, (const double2*)c,
                     (long long)n, strategy, T, centres, w.key, w.part);
  BH_LAUNCH_CHECK();
  if (strategy == kPercentile) {
    // order statistics a[lo], a[hi] (ascending) = the (n - lo)-th and
    // (n - hi)-th largest keys
    const double pos = (val / 100.0) * (double)(n - 1);
    long long lo = (long long)floor(pos);
    if (lo < 0) lo = 0;
    if (lo > n - 1) lo = n - 1;
    long long hi = (long long)ceil(pos);
    if (hi > n - 1) hi = n - 1;
    bh::kth_largest_key(w.key, n, n - lo, &w.st->klo, w.kth, st);
    bh::kth_largest_key(w.key, n, n - hi, &w.st->khi, w.kth, st);
    BH_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_thr_state, dim3(1), dim3(64), 0, st, w.part, nblk, strategy, val,
                     (long long)n, w.st);
  hipLaunchKernelGGL(k_thr_count, dim3((unsigned)nblk), dim3(kTB), 0, st, w.key, (long long)n,
                     strategy, w.st, w.cnt, w.part);
  hipLaunchKernelGGL(k_thr_scan, dim3(1), dim3(64), 0, st, w.cnt, w.part, nblk, strategy, w.st,
                     count_out);
  hipLaunchKernelGGL(k_thr_compact, dim3((unsigned)nblk), dim3(kTB), 0, st, (const double2*)c,
                     w.key, (long long)n, strategy, w.st, w.cnt, idx_out, (double2*)val_out);
  BH_LAUNCH_CHECK();
  return BH_OK;
}
