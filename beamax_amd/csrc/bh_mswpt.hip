// MSWPT (multiscale wave-packet transform) kernels, FP64.
//
// Layout: the grid spectrum lives unshifted (plain fftn order) in the
// workspace; fftshift / ifftshift are folded into index maps.  Coefficients
// are written in the reference's flat layout (levels concatenated, then box,
// then the C-order support index) and transformed in place by batched IFFTs
// over the box axes.
#include "bh_internal.hpp"

namespace bh {

int check_tiling(const bh_tiling* t) {
  if (t == nullptr) BH_FAIL(BH_ERR_ARG, "tiling is null");
  if (t->ndim < 1 || t->ndim > BH_MAX_DIM) BH_FAIL(BH_ERR_ARG, "tiling: ndim=%d", t->ndim);
  if (t->num_levels < 1 || t->num_levels > BH_MAX_LEVELS)
    BH_FAIL(BH_ERR_ARG, "tiling: num_levels=%d", t->num_levels);
  if (t->redundancy != 1 && t->redundancy != 2)
    BH_FAIL(BH_ERR_ARG, "tiling: redundancy=%d", t->redundancy);
  if (t->windowing != BH_WINDOW_RECTANGULAR && t->windowing != BH_WINDOW_RECTANGULAR_MIRROR)
    BH_FAIL(BH_ERR_ARG, "tiling: windowing=%d", t->windowing);
  if (t->centres == nullptr) BH_FAIL(BH_ERR_ARG, "tiling: centres is null");
  for (int a = 0; a < t->ndim; ++a)
    if (t->N[a] < 2 || (t->N[a] & 1)) BH_FAIL(BH_ERR_ARG, "tiling: N[%d]=%lld", a, (long long)t->N[a]);
  for (int l = 0; l < t->num_levels; ++l) {
    int64_t p = 1;
    for (int a = 0; a < t->ndim; ++a) {
      const int64_t S = t->support[l][a];
      // patch support 2*int(bl*aspect*rho/2) must equal the coefficient block
      const int64_t patch = 2 * ((int64_t)t->box_length[l] * t->aspect[a] * t->redundancy / 2);
      if (S != patch || S < 2) BH_FAIL(BH_ERR_UNSUPPORTED, "tiling: level %d axis %d support %lld "
                                       "differs from the patch support %lld", l, a,
                                       (long long)S, (long long)patch);
      if (S > t->N[a]) BH_FAIL(BH_ERR_UNSUPPORTED, "tiling: support %lld exceeds N=%lld",
                               (long long)S, (long long)t->N[a]);
      p *= S;
    }
    const int64_t nb = t->boxes_cumsum[l + 1] - t->boxes_cumsum[l];
    if (t->coeffs_cumsum[l + 1] - t->coeffs_cumsum[l] != nb * p)
      BH_FAIL(BH_ERR_ARG, "tiling: coeffs_cumsum inconsistent at level %d", l);
  }
  return BH_OK;
}

TilingDev to_dev(const bh_tiling* t) {
  TilingDev d{};
  d.ndim = t->ndim;
  d.num_levels = t->num_levels;
  d.redundancy = t->redundancy;
  d.windowing = t->windowing;
  for (int a = 0; a < BH_MAX_DIM; ++a) {
    d.N[a] = a < t->ndim ? t->N[a] : 1;
    d.aspect[a] = a < t->ndim ? t->aspect[a] : 1;
  }
  d.gpacked_off[0] = 0;
  for (int l = 0; l < BH_MAX_LEVELS; ++l) {
    d.box_length[l] = t->box_length[l];
    int64_t p = 1;
    for (int a = 0; a < BH_MAX_DIM; ++a) {
      d.support[l][a] = (a < t->ndim) ? t->support[l][a] : 1;
      if (l < t->num_levels && a < t->ndim) p *= t->support[l][a];
    }
    d.gpacked_off[l + 1] = d.gpacked_off[l] + (l < t->num_levels ? p : 0);
  }
  for (int l = 0; l <= BH_MAX_LEVELS; ++l) {
    d.boxes_cumsum[l] = t->boxes_cumsum[l];
    d.coeffs_cumsum[l] = t->coeffs_cumsum[l];
  }
  return d;
}

}  // namespace bh

namespace {

using bh::TilingDev;

// Windowed Gaussian of box `c` at centred Fourier index `fm` (transforms.py:36-96).
__device__ inline double filter_value(const TilingDev& T, int lvl, const int32_t* c,
                                      const long long* fm) {
  const int bl = T.box_length[lvl];
  double expo = 0.0;
  bool inside = true;
  for (int a = 0; a < T.ndim; ++a) {
    const double Nf = (double)T.N[a];
    const double x = (double)(fm[a] - (long long)c[a]) + 0.5 * Nf;
    double md = fmod(x, Nf);
    if (md < 0.0) md += Nf;
    const double diff = md - 0.5 * Nf;
    const double sig = (double)((bl / 2) * T.aspect[a]);
    const double q = diff / sig;
    expo += q * q;
    const double ws = (double)(bl * T.aspect[a]) * ((double)T.redundancy / 2.0);
    if (T.windowing == BH_WINDOW_RECTANGULAR_MIRROR) inside = inside && (fabs(diff) < ws);
    else inside = inside && (-ws <= diff) && (diff < ws);
  }
  return inside ? exp(-expo) : 0.0;
}

__device__ inline void unflatten(long long f, const int64_t* shape, int d, int64_t* idx) {
  for (int a = d - 1; a >= 0; --a) {
    idx[a] = f % shape[a];
    f /= shape[a];
  }
}

// sum_b g_b^2 at every grid point, accumulated in box order (transforms.py:511-555)
__global__ void k_sum_gsquare(TilingDev T, const int32_t* __restrict__ centres,
                              double* __restrict__ out, long long total) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= total) return;
  int64_t m[BH_MAX_DIM]; long long fm[BH_MAX_DIM];
  unflatten(i, T.N, T.ndim, m);
  for (int a = 0; a < T.ndim; ++a) fm[a] = m[a] - T.N[a] / 2;
  double acc = 0.0;
  int lvl = 0;
  const long long nbox = T.boxes_cumsum[T.num_levels];
  for (long long b = 0; b < nbox; ++b) {
    while (b >= T.boxes_cumsum[lvl + 1]) ++lvl;
    const double g = filter_value(T, lvl, centres + b * T.ndim, fm);
    acc += g * g;
  }
  out[i] = acc;
}

// packed first-box filter of each level (transforms.py:557-589)
__global__ void k_gpacked(TilingDev T, const int32_t* __restrict__ centres,
                          double* __restrict__ out) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= T.gpacked_off[T.num_levels]) return;
  int lvl = 0;
  while (i >= T.gpacked_off[lvl + 1]) ++lvl;
  int64_t S[BH_MAX_DIM], j[BH_MAX_DIM]; long long fm[BH_MAX_DIM];
  for (int a = 0; a < T.ndim; ++a) S[a] = T.support[lvl][a];
  unflatten(i - T.gpacked_off[lvl], S, T.ndim, j);
  const int32_t* c = centres + T.boxes_cumsum[lvl] * T.ndim;
  for (int a = 0; a < T.ndim; ++a) {
    const long long cs = c[a] + T.N[a] / 2;
    const long long m = pmod(pmod(cs - S[a] / 2, T.N[a]) + j[a], T.N[a]);
    fm[a] = m - T.N[a] / 2;
  }
  out[i] = filter_value(T, lvl, c, fm);
}

}  // namespace

namespace {

__device__ inline int find_level(const int64_t* cum, int L, long long e) {
  int l = 0;
  while (l + 1 < L && e >= cum[l + 1]) ++l;
  return l;
}

// real spatial input -> complex workspace
__global__ void k_real_to_complex(const double* __restrict__ in, double2* __restrict__ out,
                                  long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_double2(in[i], 0.0);
}

// out[i] = in[(i + N/2) mod N] per axis (ifftshift == fftshift for even N)
__global__ void k_shift_copy(TilingDev T, const double2* __restrict__ in,
                             double2* __restrict__ out, long long total) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= total) return;
  int64_t m[BH_MAX_DIM];
  long long f = i;
  for (int a = T.ndim - 1; a >= 0; --a) { m[a] = f % T.N[a]; f /= T.N[a]; }
  long long src = 0;
  for (int a = 0; a < T.ndim; ++a) src = src * T.N[a] + (m[a] + T.N[a] / 2) % T.N[a];
  out[i] = in[src];
}

// Forward analysis gather (transforms.py:591-694): for coefficient element e of
// box b, support index j:  q = (j + S/2 - r) mod S (ifftshift after roll r),
// grid index m = (cs - S/2 + q) mod N (centred), value g[q] * F_shift[m] / sum_g2[m].
__global__ void k_mswpt_gather(TilingDev T, const int32_t* __restrict__ centres,
                               const double2* __restrict__ F, const double* __restrict__ sg,
                               const double* __restrict__ gp, double2* __restrict__ out) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long total = T.coeffs_cumsum[T.num_levels];
  if (e >= total) return;
  const int lvl = find_level(T.coeffs_cumsum, T.num_levels, e);
  const int d = T.ndim;
  int64_t S[BH_MAX_DIM];
  long long P = 1;
  for (int a = 0; a < d; ++a) { S[a] = T.support[lvl][a]; P *= S[a]; }
  const long long local = e - T.coeffs_cumsum[lvl];
  const long long box = T.boxes_cumsum[lvl] + local / P;
  long long jf = local % P;
  int64_t j[BH_MAX_DIM];
  for (int a = d - 1; a >= 0; --a) { j[a] = jf % S[a]; jf /= S[a]; }
  long long qf = 0, mf = 0, uf = 0;
  for (int a = 0; a < d; ++a) {
    const long long N = T.N[a];
    const long long cs = (long long)centres[box * d + a] + N / 2;
    const long long bs = S[a] / 2;
    const long long ri = fdiv(cs + N / 2, bs);
    const long long r = (1 - 2 * (ri & 1)) * (bs / 2);
    const long long q = pmod(j[a] + S[a] / 2 - r, S[a]);
    const long long m = pmod(pmod(cs - S[a] / 2, N) + q, N);
    qf = qf * S[a] + q;
    mf = mf * N + m;
    uf = uf * N + (m + N / 2) % N;
  }
  const double2 v = F[uf];
  const double s = sg[mf];
  const double g = gp[T.gpacked_off[lvl] + qf];
  out[e] = make_double2(g * (v.x / s), g * (v.y / s));
}

__global__ void k_scale_mask(TilingDev T, double2* __restrict__ c, double scale, int half) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long total = T.coeffs_cumsum[T.num_levels];
  if (e >= total) return;
  const int lvl = find_level(T.coeffs_cumsum, T.num_levels, e);
  const long long seg = T.coeffs_cumsum[lvl + 1] - T.coeffs_cumsum[lvl];
  double2 v = c[e];
  if (half && (e - T.coeffs_cumsum[lvl]) >= seg / 2) v = make_double2(0.0, 0.0);
  c[e] = make_double2(scale * v.x, scale * v.y);
}

// Synthesis gather (transforms.py:735-850): ft[m] = sum over boxes (box order)
// of g[jj] * fftshift(fftn(c_b))[(jj + r) mod S], jj = (m - start_b) mod N < S.
// With sg != nullptr the sum is divided by sum_g2[m]: the adjoint of the
// analysis gather (which divides by it before windowing).
__global__ void k_mswpt_synth(TilingDev T, const int32_t* __restrict__ centres,
                              const double2* __restrict__ Cf, const double* __restrict__ gp,
                              const double* __restrict__ sg, double2* __restrict__ ft,
                              long long total) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int d = T.ndim;
  int64_t m[BH_MAX_DIM];
  long long f = i;
  for (int a = d - 1; a >= 0; --a) { m[a] = f % T.N[a]; f /= T.N[a]; }
  double re = 0.0, im = 0.0;
  for (int lvl = 0; lvl < T.num_levels; ++lvl) {
    int64_t S[BH_MAX_DIM];
    long long P = 1;
    for (int a = 0; a < d; ++a) { S[a] = T.support[lvl][a]; P *= S[a]; }
    for (long long b = T.boxes_cumsum[lvl]; b < T.boxes_cumsum[lvl + 1]; ++b) {
      bool in = true;
      long long jf = 0, hf = 0;
      for (int a = 0; a < d && in; ++a) {
        const long long N = T.N[a];
        const long long c = centres[b * d + a];
        const long long c0 = pmod(c + N / 2, N);
        const long long start = pmod(c0 - S[a] / 2, N);
        const long long jj = pmod(m[a] - start, N);
        if (jj >= S[a]) { in = false; break; }
        const long long bs = S[a] / 2;
        const long long ri = fdiv(c + N / 2, bs);
        const long long r = (1 - 2 * (ri & 1)) * (bs / 2);
        const long long ish = pmod(jj + r, S[a]);
        const long long h = pmod(ish - S[a] / 2, S[a]);
        jf = jf * S[a] + jj;
        hf = hf * S[a] + h;
      }
      if (!in) continue;
      const long long cbase = T.coeffs_cumsum[lvl] + (b - T.boxes_cumsum[lvl]) * P;
      const double2 v = Cf[cbase + hf];
      const double g = gp[T.gpacked_off[lvl] + jf];
      re += g * v.x;
      im += g * v.y;
    }
  }
  if (sg) {
    const double q = sg[i];
    re /= q;
    im /= q;
  }
  ft[i] = make_double2(re, im);
}

__global__ void k_real_part(const double2* __restrict__ in, double* __restrict__ out,
                            long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i].x;
}

// batched fftn over the box axes of every level, in place
int level_ffts(const bh::TilingDev& T, double2* coeffs, int dir, void* ws, size_t ws_bytes,
               hipStream_t s) {
  for (int l = 0; l < T.num_levels; ++l) {
    const int64_t nb = T.boxes_cumsum[l + 1] - T.boxes_cumsum[l];
    if (nb == 0) continue;
    double2* base = coeffs + T.coeffs_cumsum[l];
    int64_t P = 1;
    for (int a = 0; a < T.ndim; ++a) P *= T.support[l][a];
    for (int a = 0; a < T.ndim; ++a) {
      const int64_t n = T.support[l][a];
      int64_t inner = 1;
      for (int b = a + 1; b < T.ndim; ++b) inner *= T.support[l][b];
      const int64_t batch = nb * P / n;
      int rc = bh::fft_strided(base, n, batch, inner, 1, n * inner, inner, dir,
                               1.0 / sqrt((double)n), ws, ws_bytes, s);
      if (rc) return rc;
    }
  }
  return BH_OK;
}

size_t level_fft_ws(const bh_tiling* t) {
  size_t w = 0;
  for (int l = 0; l < t->num_levels; ++l) {
    int64_t nb = t->boxes_cumsum[l + 1] - t->boxes_cumsum[l];
    int64_t P = 1;
    for (int a = 0; a < t->ndim; ++a) P *= t->support[l][a];
    for (int a = 0; a < t->ndim; ++a) {
      size_t x = bh::fft_workspace_bytes(t->support[l][a], nb * P / t->support[l][a]);
      if (x > w) w = x;
    }
  }
  return w;
}

size_t grid_fft_ws(const bh_tiling* t) {
  const int64_t total = bh::prod(t->N, t->ndim);
  size_t w = 0;
  for (int a = 0; a < t->ndim; ++a) {
    size_t x = bh::fft_workspace_bytes(t->N[a], total / t->N[a]);
    if (x > w) w = x;
  }
  return w;
}

}  // namespace

namespace {
// cpos = 0.5 (a + i b / vg), cneg = 0.5 (a - i b / vg), vg = c |2 pi centre / L|
// of the coefficient's box (forward_solver_utils.py:763-785); b == nullptr
// stands for b = 0; optional half mask on cpos ('pos_only')
__global__ void k_combine_velocity(TilingDev T, const int32_t* __restrict__ centres, double L0,
                                   double L1, double L2, double c, const double2* __restrict__ a,
                                   const double2* __restrict__ b, int half,
                                   double2* __restrict__ pos, double2* __restrict__ neg) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long total = T.coeffs_cumsum[T.num_levels];
  if (e >= total) return;
  const int lvl = find_level(T.coeffs_cumsum, T.num_levels, e);
  long long P = 1;
  for (int x = 0; x < T.ndim; ++x) P *= T.support[lvl][x];
  const long long box = T.boxes_cumsum[lvl] + (e - T.coeffs_cumsum[lvl]) / P;
  const double L[3] = {L0, L1, L2};
  double n2 = 0.0;
  for (int x = 0; x < T.ndim; ++x) {
    const double p = 2.0 * bh::kPi * ((double)centres[box * T.ndim + x] / L[x]);
    n2 += p * p;
  }
  const double vg = c * sqrt(n2);
  const double2 av = a[e];
  const double2 bv = b ? b[e] : make_double2(0.0, 0.0);
  // i b / vg = (-b.im / vg, b.re / vg)
  const double ur = (-bv.y) / vg, ui = bv.x / vg;
  double2 r = make_double2(0.5 * (av.x + ur), 0.5 * (av.y + ui));
  const long long seg = T.coeffs_cumsum[lvl + 1] - T.coeffs_cumsum[lvl];
  if (half && (e - T.coeffs_cumsum[lvl]) >= seg / 2) r = make_double2(0.0, 0.0);
  pos[e] = r;
  if (neg) neg[e] = make_double2(0.5 * (av.x - ur), 0.5 * (av.y - ui));
}
}  // namespace

extern "C" int bh_mswpt_combine_velocity(const bh_tiling* t, const double* grid_size, double c,
                                         const double* a, const double* b, int half_mask,
                                         double* cpos, double* cneg, bh_stream s) {
  bh::clear_error();
  if (int rc = bh::check_tiling(t)) return rc;
  if (!grid_size || !a || !cpos) BH_FAIL(BH_ERR_ARG, "bh_mswpt_combine_velocity: null pointer");
  if (!(c > 0.0)) BH_FAIL(BH_ERR_ARG, "bh_mswpt_combine_velocity: c must be positive");
  const bh::TilingDev T = bh::to_dev(t);
  const long long nc = t->coeffs_cumsum[t->num_levels];
  hipLaunchKernelGGL(k_combine_velocity, dim3(bh::grid1d(nc, 256)), dim3(256), 0, (hipStream_t)s,
                     T, t->centres, grid_size[0], t->ndim > 1 ? grid_size[1] : 1.0,
                     t->ndim > 2 ? grid_size[2] : 1.0, c, (const double2*)a, (const double2*)b,
                     half_mask, (double2*)cpos, (double2*)cneg);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

extern "C" int bh_mswpt_setup(const bh_tiling* t, double* sum_gsquare, double* gpacked,
                              bh_stream s) {
  bh::clear_error();
  if (int rc = bh::check_tiling(t)) return rc;
  if (!sum_gsquare || !gpacked) BH_FAIL(BH_ERR_ARG, "bh_mswpt_setup: null output");
  const bh::TilingDev T = bh::to_dev(t);
  const long long total = bh::prod(t->N, t->ndim);
  hipStream_t st = (hipStream_t)s;
  hipLaunchKernelGGL(k_sum_gsquare, dim3(bh::grid1d(total, 256)), dim3(256), 0, st, T,
                     t->centres, sum_gsquare, total);
  hipLaunchKernelGGL(k_gpacked, dim3(bh::grid1d(T.gpacked_off[T.num_levels], 256)), dim3(256),
                     0, st, T, t->centres, gpacked);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

extern "C" size_t bh_mswpt_workspace_size(const bh_tiling* t) {
  if (bh::check_tiling(t)) return 0;
  const int64_t total = bh::prod(t->N, t->ndim);
  size_t fw = grid_fft_ws(t);
  size_t lw = level_fft_ws(t);
  const int64_t nc = t->coeffs_cumsum[t->num_levels];
  const int64_t big = total > nc ? total : nc;
  return (size_t)(2 * big) * sizeof(double2) + (fw > lw ? fw : lw);
}

extern "C" int bh_mswpt_forward(const bh_tiling* t, const double* sum_gsquare,
                                const double* gpacked, const double* input, int input_kind,
                                double scale, int half_mask, double* coeffs, void* ws,
                                size_t ws_bytes, bh_stream s) {
  bh::clear_error();
  if (int rc = bh::check_tiling(t)) return rc;
  if (!sum_gsquare || !gpacked || !input || !coeffs || !ws)
    BH_FAIL(BH_ERR_ARG, "bh_mswpt_forward: null pointer");
  const size_t need = bh_mswpt_workspace_size(t);
  if (ws_bytes < need) BH_FAIL(BH_ERR_WORKSPACE, "bh_mswpt_forward: workspace %zu < %zu",
                               ws_bytes, need);
  const bh::TilingDev T = bh::to_dev(t);
  hipStream_t st = (hipStream_t)s;
  const long long total = bh::prod(t->N, t->ndim);
  const int64_t nc = t->coeffs_cumsum[t->num_levels];
  const int64_t big = total > nc ? total : nc;
  double2* F = (double2*)ws;
  void* fws = (char*)ws + (size_t)(2 * big) * sizeof(double2);
  const size_t fws_bytes = ws_bytes - (size_t)(2 * big) * sizeof(double2);
  const unsigned g = bh::grid1d(total, 256);
  if (input_kind == BH_INPUT_REAL_SPATIAL) {
    hipLaunchKernelGGL(k_real_to_complex, dim3(g), dim3(256), 0, st, input, F, total);
  } else if (input_kind == BH_INPUT_COMPLEX_SPATIAL) {
    BH_HIP(hipMemcpyAsync(F, input, total * sizeof(double2), hipMemcpyDeviceToDevice, st));
  } else if (input_kind == BH_INPUT_COMPLEX_FOURIER) {
    hipLaunchKernelGGL(k_shift_copy, dim3(g), dim3(256), 0, st, T, (const double2*)input, F,
                       total);
  } else {
    BH_FAIL(BH_ERR_ARG, "bh_mswpt_forward: input_kind=%d", input_kind);
  }
  BH_LAUNCH_CHECK();
  if (input_kind != BH_INPUT_COMPLEX_FOURIER) {
    if (int rc = bh::fftn_ortho(F, t->ndim, t->N, -1, fws, fws_bytes, st)) return rc;
  }
  hipLaunchKernelGGL(k_mswpt_gather, dim3(bh::grid1d(nc, 256)), dim3(256), 0, st, T,
                     t->centres, F, sum_gsquare, gpacked, (double2*)coeffs);
  BH_LAUNCH_CHECK();
  if (int rc = level_ffts(T, (double2*)coeffs, +1, fws, fws_bytes, st)) return rc;
  if (scale != 1.0 || half_mask) {
    hipLaunchKernelGGL(k_scale_mask, dim3(bh::grid1d(nc, 256)), dim3(256), 0, st, T,
                       (double2*)coeffs, scale, half_mask);
    BH_LAUNCH_CHECK();
  }
  return BH_OK;
}

extern "C" int bh_mswpt_inverse(const bh_tiling* t, const double* gpacked,
                                const double* coeffs, int output_fourier, double* out,
                                void* ws, size_t ws_bytes, bh_stream s) {
  bh::clear_error();
  if (int rc = bh::check_tiling(t)) return rc;
  if (!gpacked || !coeffs || !out || !ws) BH_FAIL(BH_ERR_ARG, "bh_mswpt_inverse: null pointer");
  const size_t need = bh_mswpt_workspace_size(t);
  if (ws_bytes < need) BH_FAIL(BH_ERR_WORKSPACE, "bh_mswpt_inverse: workspace %zu < %zu",
                               ws_bytes, need);
  const bh::TilingDev T = bh::to_dev(t);
  hipStream_t st = (hipStream_t)s;
  const long long total = bh::prod(t->N, t->ndim);
  const int64_t nc = t->coeffs_cumsum[t->num_levels];
  const int64_t big = total > nc ? total : nc;
  double2* C = (double2*)ws;
  double2* G = C + big;
  void* fws = (char*)ws + (size_t)(2 * big) * sizeof(double2);
  const size_t fws_bytes = ws_bytes - (size_t)(2 * big) * sizeof(double2);
  BH_HIP(hipMemcpyAsync(C, coeffs, nc * sizeof(double2), hipMemcpyDeviceToDevice, st));
  if (int rc = level_ffts(T, C, -1, fws, fws_bytes, st)) return rc;
  double2* dst = output_fourier ? (double2*)out : G;
  hipLaunchKernelGGL(k_mswpt_synth, dim3(bh::grid1d(total, 256)), dim3(256), 0, st, T,
                     t->centres, C, gpacked, (const double*)nullptr, dst, total);
  BH_LAUNCH_CHECK();
  if (!output_fourier) {
    hipLaunchKernelGGL(k_shift_copy, dim3(bh::grid1d(total, 256)), dim3(256), 0, st, T, G,
                       (double2*)out, total);
    BH_LAUNCH_CHECK();
    if (int rc = bh::fftn_ortho((double2*)out, t->ndim, t->N, +1, fws, fws_bytes, st)) return rc;
  }
  return BH_OK;
}

extern "C" int bh_mswpt_adjoint(const bh_tiling* t, const double* sum_gsquare,
                                const double* gpacked, const double* coeffs, int output_real,
                                double* out, void* ws, size_t ws_bytes, bh_stream s) {
  bh::clear_error();
  if (int rc = bh::check_tiling(t)) return rc;
  if (!sum_gsquare || !gpacked || !coeffs || !out || !ws)
    BH_FAIL(BH_ERR_ARG, "bh_mswpt_adjoint: null pointer");
  const size_t need = bh_mswpt_workspace_size(t);
  if (ws_bytes < need) BH_FAIL(BH_ERR_WORKSPACE, "bh_mswpt_adjoint: workspace %zu < %zu",
                               ws_bytes, need);
  const bh::TilingDev T = bh::to_dev(t);
  hipStream_t st = (hipStream_t)s;
  const long long total = bh::prod(t->N, t->ndim);
  const int64_t nc = t->coeffs_cumsum[t->num_levels];
  const int64_t big = total > nc ? total : nc;
  double2* C = (double2*)ws;
  double2* G = C + big;
  void* fws = (char*)ws + (size_t)(2 * big) * sizeof(double2);
  const size_t fws_bytes = ws_bytes - (size_t)(2 * big) * sizeof(double2);
  BH_HIP(hipMemcpyAsync(C, coeffs, nc * sizeof(double2), hipMemcpyDeviceToDevice, st));
  if (int rc = level_ffts(T, C, -1, fws, fws_bytes, st)) return rc;
  hipLaunchKernelGGL(k_mswpt_synth, dim3(bh::grid1d(total, 256)), dim3(256), 0, st, T,
                     t->centres, C, gpacked, sum_gsquare, G, total);
  BH_LAUNCH_CHECK();
  // G (centred) -> C (unshifted) -> unitary inverse FFT in place
  hipLaunchKernelGGL(k_shift_copy, dim3(bh::grid1d(total, 256)), dim3(256), 0, st, T, G, C,
                     total);
  BH_LAUNCH_CHECK();
  if (int rc = bh::fftn_ortho(C, t->ndim, t->N, +1, fws, fws_bytes, st)) return rc;
  if (output_real) {
    hipLaunchKernelGGL(k_real_part, dim3(bh::grid1d(total, 256)), dim3(256), 0, st, C, out, total);
  } else {
    BH_HIP(hipMemcpyAsync(out, C, total * sizeof(double2), hipMemcpyDeviceToDevice, st));
  }
  BH_LAUNCH_CHECK();
  return BH_OK;
}
