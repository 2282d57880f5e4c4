// Adjoint source (spectral time derivative) and the Fourier crop along time.
//
// The (time, sensor) arrays are transformed along time.  Internally they are
// held sensor-major -- one contiguous time series per sensor -- so that the
// FFTs read and write whole rows; the layout changes on the way in and out
// through 32 x 32 tiles in LDS (coalesced on both sides).  Transforming the
// time-major array in place instead read every element 8 KB from the last
// (512 sensors of complex doubles per time step): the 2414-point transforms
// of one 512-sensor face took 236 us that way, 205 us on rows.
#include "bh_internal.hpp"

namespace {

constexpr int kTile = 32;

// X[j nt + t] = w(t, j) d[t ns_in + c_j]  (c_j = cols[j] or j; w = wt[j], or
// the window w1[t] / w2[t ns + j], or 1), as complex
__global__ __launch_bounds__(256) void k_load_t(const double* __restrict__ d, long long nt,
                                                long long ns_in, const long long* __restrict__ cols,
                                                const double* __restrict__ wt,
                                                const double* __restrict__ w1,
                                                const double* __restrict__ w2, long long ns,
                                                double2* __restrict__ X) {
  __shared__ double tile[kTile][kTile + 1];
  const int tx = threadIdx.x & (kTile - 1), ty = threadIdx.x / kTile;
  const long long t0 = blockIdx.x * (long long)kTile, j0 = blockIdx.y * (long long)kTile;
  const long long j = j0 + tx;
  const bool jin = j < ns;
  const long long cj = jin ? (cols ? cols[j] : j) : 0;
  for (int r = ty; r < kTile; r += 256 / kTile) {
    const long long t = t0 + r;
    double v = 0.0;
    if (jin && t < nt) {
      v = d[t * ns_in + cj];
      if (wt) v *= wt[j];
      if (w1) v = w1[t] * v;
      if (w2) v = w2[t * ns + j] * v;
    }
    tile[r][tx] = v;
  }
  __syncthreads();
  for (int r = ty; r < kTile; r += 256 / kTile) {
    const long long jj = j0 + r, t = t0 + tx;
    if (jj < ns && t < nt) X[jj * nt + t] = make_double2(tile[tx][r], 0.0);
  }
}

// out[t ns + j] = scale_j Re Y[j nt + t]  (scale_j = -c_j^2, or 1)
__global__ __launch_bounds__(256) void k_store_t(const double2* __restrict__ Y, long long nt,
                                                 long long ns, const double* __restrict__ c,
                                                 double* __restrict__ out) {
  __shared__ double tile[kTile][kTile + 1];
  const int tx = threadIdx.x & (kTile - 1), ty = threadIdx.x / kTile;
  const long long t0 = blockIdx.x * (long long)kTile, j0 = blockIdx.y * (long long)kTile;
  for (int r = ty; r < kTile; r += 256 / kTile) {
    const long long jj = j0 + r, t = t0 + tx;
    tile[r][tx] = (jj < ns && t < nt) ? Y[jj * nt + t].x : 0.0;
  }
  __syncthreads();
  const long long j = j0 + tx;
  for (int r = ty; r < kTile; r += 256 / kTile) {
    const long long t = t0 + r;
    if (j < ns && t < nt) {
      const double v = tile[tx][r];
      if (c) {
        const double cs = c[j];
        out[t * ns + j] = -(cs * cs) * v;
      } else {
        out[t * ns + j] = v;
      }
    }
  }
}

dim3 tiles(long long nt, long long ns) {
  return dim3((unsigned)((nt + kTile - 1) / kTile), (unsigned)((ns + kTile - 1) / kTile));
}

// X[s, k] *= 2 pi i f_k,  f = fftfreq(nt, dt) (msgb_solver.py:98-103)
__global__ void k_deriv_mult(double2* __restrict__ X, long long nt, long long ns, double dt) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= nt * ns) return;
  const long long k = i % nt;
  const long long half = (nt - 1) / 2 + 1;
  const double val = 1.0 / ((double)nt * dt);
  const double f = (double)(k < half ? k : k - nt) * val;
  const double m = 2.0 * bh::kPi * f;
  const double2 x = X[i];
  X[i] = make_double2(-m * x.y, m * x.x);
}

// keep frequencies f in [-m/2, m - m/2): Y[j, f mod m] = X[j, f mod n]
__global__ void k_crop(const double2* __restrict__ X, long long n, long long m, long long ns,
                       double2* __restrict__ Y) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= m * ns) return;
  const long long j = i / m, r = i - j * m;
  const long long f = r < (m - m / 2) ? r : r - m;
  const long long src = f >= 0 ? f : f + n;
  Y[i] = X[j * n + src];
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
  // X sensor-major: row j holds sensor j's time series
  hipLaunchKernelGGL(k_load_t, tiles(nt, ns), dim3(256), 0, st, data, (long long)nt,
                     (long long)ns, nullptr, nullptr, (window && !window_2d) ? window : nullptr,
                     (window && window_2d) ? window : nullptr, (long long)ns, X);
  BH_LAUNCH_CHECK();
  if (int rc = bh::fft_strided(X, nt, ns, 1, 0, nt, 1, -1, 1.0, fws, fws_bytes, st)) return rc;
  hipLaunchKernelGGL(k_deriv_mult, dim3(g), dim3(256), 0, st, X, (long long)nt, (long long)ns, dt);
  if (int rc = bh::fft_strided(X, nt, ns, 1, 0, nt, 1, +1, 1.0 / (double)nt, fws, fws_bytes, st))
    return rc;
  hipLaunchKernelGGL(k_store_t, tiles(nt, ns), dim3(256), 0, st, X, (long long)nt, (long long)ns,
                     c_src, out);
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
  // X, Y sensor-major: row j holds column j's time series
  hipLaunchKernelGGL(k_load_t, tiles(nt, ns), dim3(256), 0, st, data, (long long)nt,
                     (long long)ns_in, (const long long*)cols, col_weight, nullptr, nullptr,
                     (long long)ns, X);
  BH_LAUNCH_CHECK();
  if (int rc = bh::fft_strided(X, nt, ns, 1, 0, nt, 1, -1, 1.0 / sqrt((double)nt), fws,
                               fws_bytes, st))
    return rc;
  hipLaunchKernelGGL(k_crop, dim3(bh::grid1d(nt_out * ns, 256)), dim3(256), 0, st, X,
                     (long long)nt, (long long)nt_out, (long long)ns, Y);
  if (int rc = bh::fft_strided(Y, nt_out, ns, 1, 0, nt_out, 1, +1, 1.0 / sqrt((double)nt_out),
                               fws, fws_bytes, st))
    return rc;
  hipLaunchKernelGGL(k_store_t, tiles(nt_out, ns), dim3(256), 0, st, Y, (long long)nt_out,
                     (long long)ns, nullptr, out);
  BH_LAUNCH_CHECK();
  return BH_OK;
}
