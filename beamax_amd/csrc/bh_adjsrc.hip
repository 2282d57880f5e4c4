// Adjoint source (spectral time derivative) and the Fourier crop along time.
#include "bh_internal.hpp"

namespace {

__global__ void k_load_windowed(const double* __restrict__ d, const double* __restrict__ w,
                                int w2d, long long nt, long long ns, double2* __restrict__ X) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= nt * ns) return;
  double v = d[i];
  if (w) v = (w2d ? w[i] : w[i / ns]) * v;
  X[i] = make_double2(v, 0.0);
}

// X[k, s] *= 2 pi i f_k,  f = fftfreq(nt, dt) (msgb_solver.py:98-103)
__global__ void k_deriv_mult(double2* __restrict__ X, long long nt, long long ns, double dt) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= nt * ns) return;
  const long long k = i / ns;
  const long long half = (nt - 1) / 2 + 1;
  const double val = 1.0 / ((double)nt * dt);
  const double f = (double)(k < half ? k : k - nt) * val;
  const double m = 2.0 * bh::kPi * f;
  const double2 x = X[i];
  X[i] = make_double2(-m * x.y, m * x.x);
}

__global__ void k_store_source(const double2* __restrict__ X, const double* __restrict__ c,
                               long long nt, long long ns, double* __restrict__ out) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= nt * ns) return;
  const double cs = c[i % ns];
  out[i] = -(cs * cs) * X[i].x;
}

// gather selected columns (times weight) of an (nt, ns_in) real array
__global__ void k_load_cols(const double* __restrict__ d, long long nt, long long ns_in,
                            const long long* __restrict__ cols, const double* __restrict__ wt,
                            long long ns, double2* __restrict__ X) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= nt * ns) return;
  const long long t = i / ns, j = i - t * ns;
  const long long cj = cols ? cols[j] : j;
  double v = d[t * ns_in + cj];
  if (wt) v *= wt[j];
  X[i] = make_double2(v, 0.0);
}

// keep frequencies f in [-m/2, m - m/2): Y[f mod m] = X[f mod n]
__global__ void k_crop(const double2* __restrict__ X, long long n, long long m, long long ns,
                       double2* __restrict__ Y) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= m * ns) return;
  const long long r = i / ns, j = i - r * ns;
  const long long f = r < (m - m / 2) ? r : r - m;
  const long long src = f >= 0 ? f : f + n;
  Y[i] = X[src * ns + j];
}

__global__ void k_real_part(const double2* __restrict__ Y, long long n, double* __restrict__ out) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) out[i] = Y[i].x;
}

}  // namespace

extern "C" size_t bh_adjoint_source_workspace_size(int64_t nt, int64_t ns) {
  if (nt <= 0 || ns <= 0) return 0;
  return (size_t)(nt * ns) * sizeof(double2) + bh::fft_workspace_bytes(nt, ns);
}

extern "C" int bh_adjoint_source(const double* data, int64_t nt, int64_t ns, double dt,
                                 const double* window, int window_2d, const double* c_src,
                                 double* out, void* ws, size_t ws_bytes, bh_stream s) {
  bh::clear_error();
  if (!data || !c_src || !out || !ws) BH_FAIL(BH_ERR_ARG, "bh_adjoint_source: null pointer");
  if (nt <= 0 || ns <= 0) BH_FAIL(BH_ERR_ARG, "bh_adjoint_source: nt=%lld ns=%lld",
                                  (long long)nt, (long long)ns);
  if (!(dt > 0.0)) BH_FAIL(BH_ERR_ARG, "bh_adjoint_source: dt must be positive");
  const size_t need = bh_adjoint_source_workspace_size(nt, ns);
  if (ws_bytes < need) BH_FAIL(BH_ERR_WORKSPACE, "bh_adjoint_source: workspace %zu < %zu",
                               ws_bytes, need);
  hipStream_t st = (hipStream_t)s;
  double2* X = (double2*)ws;
  void* fws = (char*)ws + (size_t)(nt * ns) * sizeof(double2);
  const size_t fws_bytes = ws_bytes - (size_t)(nt * ns) * sizeof(double2);
  const unsigned g = bh::grid1d(nt * ns, 256);
  hipLaunchKernelGGL(k_load_windowed, dim3(g), dim3(256), 0, st, data, window, window_2d,
                     (long long)nt, (long long)ns, X);
  BH_LAUNCH_CHECK();
  if (int rc = bh::fft_strided(X, nt, ns, ns, 1, nt * ns, ns, -1, 1.0, fws, fws_bytes, st))
    return rc;
  hipLaunchKernelGGL(k_deriv_mult, dim3(g), dim3(256), 0, st, X, (long long)nt, (long long)ns, dt);
  if (int rc = bh::fft_strided(X, nt, ns, ns, 1, nt * ns, ns, +1, 1.0 / (double)nt, fws,
                               fws_bytes, st))
    return rc;
  hipLaunchKernelGGL(k_store_source, dim3(g), dim3(256), 0, st, X, c_src, (long long)nt,
                     (long long)ns, out);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

extern "C" size_t bh_fourier_crop_time_workspace_size(int64_t nt, int64_t nt_out, int64_t ns) {
  if (nt <= 0 || nt_out <= 0 || ns <= 0) return 0;
  size_t a = bh::fft_workspace_bytes(nt, ns), b = bh::fft_workspace_bytes(nt_out, ns);
  return (size_t)((nt + nt_out) * ns) * sizeof(double2) + (a > b ? a : b);
}

extern "C" int bh_fourier_crop_time(const double* data, int64_t nt, int64_t ns_in,
                                    const int64_t* cols, const double* col_weight, int64_t ns,
                                    int64_t nt_out, double* out, void* ws, size_t ws_bytes,
                                    bh_stream s) {
  bh::clear_error();
  if (!data || !out || !ws) BH_FAIL(BH_ERR_ARG, "bh_fourier_crop_time: null pointer");
  if (nt <= 0 || ns <= 0 || ns_in <= 0 || nt_out <= 0 || nt_out > nt)
    BH_FAIL(BH_ERR_ARG, "bh_fourier_crop_time: nt=%lld nt_out=%lld ns=%lld", (long long)nt,
            (long long)nt_out, (long long)ns);
  const size_t need = bh_fourier_crop_time_workspace_size(nt, nt_out, ns);
  if (ws_bytes < need) BH_FAIL(BH_ERR_WORKSPACE, "bh_fourier_crop_time: workspace %zu < %zu",
                               ws_bytes, need);
  hipStream_t st = (hipStream_t)s;
  double2* X = (double2*)ws;
  double2* Y = X + nt * ns;
  void* fws = (char*)(Y + nt_out * ns);
  const size_t fws_bytes = ws_bytes - (size_t)((nt + nt_out) * ns) * sizeof(double2);
  hipLaunchKernelGGL(k_load_cols, dim3(bh::grid1d(nt * ns, 256)), dim3(256), 0, st, data,
                     (long long)nt, (long long)ns_in, (const long long*)cols, col_weight,
                     (long long)ns, X);
  BH_LAUNCH_CHECK();
  if (int rc = bh::fft_strided(X, nt, ns, ns, 1, nt * ns, ns, -1, 1.0 / sqrt((double)nt), fws,
                               fws_bytes, st))
    return rc;
  hipLaunchKernelGGL(k_crop, dim3(bh::grid1d(nt_out * ns, 256)), dim3(256), 0, st, X,
                     (long long)nt, (long long)nt_out, (long long)ns, Y);
  if (int rc = bh::fft_strided(Y, nt_out, ns, ns, 1, nt_out * ns, ns, +1,
                               1.0 / sqrt((double)nt_out), fws, fws_bytes, st))
    return rc;
  hipLaunchKernelGGL(k_real_part, dim3(bh::grid1d(nt_out * ns, 256)), dim3(256), 0, st, Y,
                     (long long)(nt_out * ns), out);
  BH_LAUNCH_CHECK();
  return BH_OK;
}
