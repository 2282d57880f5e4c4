// Batched strided complex FFT (FP64), mixed radix Stockham autosort.
//
// n <= 2560: one LDS-resident pass per launch; each workgroup loads tpb whole
// transforms (several up to 1024 points, one above; coalesced along whichever
// axis is contiguous), runs every radix stage between two LDS buffers with a
// per-block twiddle table, and stores.
// n > 2560: the same stages run as one global-memory pass per launch.
//
// Stage s with radix R over sub-length Ns produces output o from
//   k = o mod Ns, q = (o / Ns) mod R, j = (o / (Ns R)) Ns + k,
//   out[o] = sum_r in[j + r n/R] w^(r (n/(Ns R)) (k + q Ns)),  w = e^(dir 2 pi i / n)
// which is the radix-R Stockham butterfly with its twiddle folded in.  For
// n a power of two (radices 4, then one 2) the stages run as butterflies
// instead: thread j of a stage reads in[j + r n/R] (r < R), multiplies
// input r by w^(r m k) (k = j mod Ns, m = n / (Ns R)) and writes the R-point
// DFT of the products to out[(j - k) R + q Ns + k] (q < R) -- the same
// outputs, with the index arithmetic and twiddle reads of one butterfly
// shared by its R outputs (the per-output form spent most of its time on
// divisions by Ns and R).
#include "bh_internal.hpp"

namespace {

constexpr int kMaxStages = 40;

struct FftArgs {
  double2* data;
  long long batch, inner, inner_stride, outer_stride, elem_stride;
  int n, nst, tpb, dir;
  int pow2;   // every radix 4 or 2
  double scale;
  int radix[kMaxStages];
};

bool make_plan(int64_t n, FftArgs& a) {
  a.n = (int)n;
  a.nst = 0;
  int64_t m = n;
  while (m % 4 == 0 && a.nst < kMaxStages) { a.radix[a.nst++] = 4; m /= 4; }
  while (m % 2 == 0 && a.nst < kMaxStages) { a.radix[a.nst++] = 2; m /= 2; }
  a.pow2 = m == 1 ? 1 : 0;
  for (int64_t f = 3; m > 1 && a.nst < kMaxStages; f += 2) {
    if (f * f > m) { a.radix[a.nst++] = (int)m; m = 1; break; }
    while (m % f == 0 && a.nst < kMaxStages) { a.radix[a.nst++] = (int)f; m /= f; }
  }
  return m == 1;
}

__device__ inline long long fft_base(long long t, const FftArgs& a) {
  return (t / a.inner) * a.outer_stride + (t % a.inner) * a.inner_stride;
}

__global__ __launch_bounds__(1024) void k_fft_lds(FftArgs a) {
  extern __shared__ double2 sm[];
  const int n = a.n, tpb = a.tpb, tot = tpb * n;
  double2* b0 = sm;
  double2* b1 = sm + tot;
  double2* tw = sm + 2 * tot;
  // twiddles w^i = (cos, dir sin)(2 pi i / n): sincospi for the first
  // octant (n divisible by 8) or half turn (n even), the rest by the exact
  // symmetries (c, s)(pi/2 - x) = (s, c)(x), (c, s)(x + pi/2) = (-s, c)(x)
  // and (c, s)(x + pi) = (-c, -s)(x); each entry written once
  const double dr = (double)a.dir;
  if ((n & 7) == 0) {
    const int n4 = n >> 2, n8 = n >> 3;
    for (int i = threadIdx.x; i <= n8; i += blockDim.x) {
      double sv, cv;
      sincospi(2.0 * (double)i / (double)n, &sv, &cv);
      double c1 = cv, s1 = sv, c2 = sv, s2 = cv;   // angles x and pi/2 - x
      for (int q = 0; q < 4; ++q) {
        tw[i + q * n4] = make_double2(c1, dr * s1);
        if (i > 0 && i < n8) tw[n4 - i + q * n4] = make_double2(c2, dr * s2);
        double t = c1; c1 = -s1; s1 = t;
        t = c2; c2 = -s2; s2 = t;
      }
    }
  } else if ((n & 1) == 0) {
    const int n2 = n >> 1;
    for (int i = threadIdx.x; i < n2; i += blockDim.x) {
      double sv, cv;
      sincospi(2.0 * (double)i / (double)n, &sv, &cv);
      tw[i] = make_double2(cv, dr * sv);
      tw[i + n2] = make_double2(-cv, -dr * sv);
    }
  } else {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      double sv, cv;
      sincospi(2.0 * (double)i / (double)n, &sv, &cv);
      tw[i] = make_double2(cv, dr * sv);
    }
  }
  const long long t0 = (long long)blockIdx.x * tpb;
  // the block's global loads are all issued before the first is stored to
  // LDS (a loop storing each before issuing the next waited for one memory
  // round trip per iteration); up to kLoads per thread, the rest in a loop
  constexpr int kLoads = 8;
  auto src_of = [&](int idx, int& tr, int& e) {
    if (a.elem_stride == 1) { tr = idx / n; e = idx - tr * n; }
    else { e = idx / tpb; tr = idx - e * tpb; }
  };
  double2 lv[kLoads];
#pragma unroll
  for (int q = 0; q < kLoads; ++q) {
    const int idx = threadIdx.x + q * blockDim.x;
    lv[q] = make_double2(0.0, 0.0);
    if (idx < tot) {
      int tr, e;
      src_of(idx, tr, e);
      if (t0 + tr < a.batch) lv[q] = a.data[fft_base(t0 + tr, a) + (long long)e * a.elem_stride];
    }
  }
#pragma unroll
  for (int q = 0; q < kLoads; ++q) {
    const int idx = threadIdx.x + q * blockDim.x;
    if (idx < tot) {
      int tr, e;
      src_of(idx, tr, e);
      b0[tr * n + e] = lv[q];
    }
  }
  for (int idx = threadIdx.x + kLoads * blockDim.x; idx < tot; idx += blockDim.x) {
    int tr, e;
    src_of(idx, tr, e);
    const long long t = t0 + tr;
    double2 v = make_double2(0.0, 0.0);
    if (t < a.batch) v = a.data[fft_base(t, a) + (long long)e * a.elem_stride];
    b0[tr * n + e] = v;
  }
  __syncthreads();
  int Ns = 1;
  if (a.pow2) {
    for (int st = 0; st < a.nst; ++st) {
      const int R = a.radix[st];
      const int T = n / R;
      const int m = n / (Ns * R);
      for (int bi = threadIdx.x; bi < tpb * T; bi += blockDim.x) {
        const int tr = bi / T;
        const int j = bi - tr * T;
        const int k = j & (Ns - 1);
        const double2* src = b0 + tr * n + j;
        double2* dst = b1 + tr * n + (j - k) * R + k;
        const int e = m * k;   // r e < n for r < R
        auto cmul = [](double2 x, double2 w) {
          return make_double2(x.x * w.x - x.y * w.y, x.x * w.y + x.y * w.x);
        };
        if (R == 4) {
          const double2 v0 = src[0];
          const double2 v1 = cmul(src[T], tw[e]);
          const double2 v2 = cmul(src[2 * T], tw[2 * e]);
          const double2 v3 = cmul(src[3 * T], tw[3 * e]);
          const double2 a0 = make_double2(v0.x + v2.x, v0.y + v2.y);
          const double2 a1 = make_double2(v0.x - v2.x, v0.y - v2.y);
          const double2 a2 = make_double2(v1.x + v3.x, v1.y + v3.y);
          // (v1 - v3) times e^(dir pi i / 2) = dir i
          const double2 a3 = make_double2(-dr * (v1.y - v3.y), dr * (v1.x - v3.x));
          dst[0] = make_double2(a0.x + a2.x, a0.y + a2.y);
          dst[Ns] = make_double2(a1.x + a3.x, a1.y + a3.y);
          dst[2 * Ns] = make_double2(a0.x - a2.x, a0.y - a2.y);
          dst[3 * Ns] = make_double2(a1.x - a3.x, a1.y - a3.y);
        } else {
          const double2 v0 = src[0];
          const double2 v1 = cmul(src[T], tw[e]);
          dst[0] = make_double2(v0.x + v1.x, v0.y + v1.y);
          dst[Ns] = make_double2(v0.x - v1.x, v0.y - v1.y);
        }
      }
      __syncthreads();
      double2* tmp = b0; b0 = b1; b1 = tmp;
      Ns *= R;
    }
  } else
  for (int st = 0; st < a.nst; ++st) {
    const int R = a.radix[st];
    const int T = n / R;
    const int m = n / (Ns * R);
    for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
      const int tr = idx / n;
      const int o = idx - tr * n;
      const int k = o % Ns;
      const int q = (o / Ns) % R;
      const int j = (o / (Ns * R)) * Ns + k;
      const int e = m * (k + q * Ns);
      const double2* src = b0 + tr * n + j;
      double re = 0.0, im = 0.0;
      int acc = 0;
      for (int r = 0; r < R; ++r) {
        const double2 x = src[r * T];
        const double2 w = tw[acc];
        re += x.x * w.x - x.y * w.y;
        im += x.x * w.y + x.y * w.x;
        acc += e;
        if (acc >= n) acc -= n;
      }
      b1[tr * n + o] = make_double2(re, im);
    }
    __syncthreads();
    double2* tmp = b0; b0 = b1; b1 = tmp;
    Ns *= R;
  }
  for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
    int tr, e;
    if (a.elem_stride == 1) { tr = idx / n; e = idx - tr * n; }
    else { e = idx / tpb; tr = idx - e * tpb; }
    const long long t = t0 + tr;
    if (t < a.batch) {
      const double2 v = b0[tr * n + e];
      a.data[fft_base(t, a) + (long long)e * a.elem_stride] =
          make_double2(v.x * a.scale, v.y * a.scale);
    }
  }
}

// ---- global-memory passes for n > 1024 -------------------------------------
__global__ void k_fft_twiddles(double2* tw, int n, int dir) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double sv, cv;
    sincospi(2.0 * (double)i / (double)n, &sv, &cv);
    tw[i] = make_double2(cv, dir * sv);
  }
}

__global__ void k_fft_gather(FftArgs a, double2* dst) {
  const long long tot = a.batch * a.n;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < tot;
       idx += (long long)gridDim.x * blockDim.x) {
    long long t, e;
    if (a.elem_stride == 1) { t = idx / a.n; e = idx - t * a.n; }
    else { e = idx / a.batch; t = idx - e * a.batch; }
    dst[t * a.n + e] = a.data[fft_base(t, a) + e * a.elem_stride];
  }
}

__global__ void k_fft_scatter(FftArgs a, const double2* src) {
  const long long tot = a.batch * a.n;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < tot;
       idx += (long long)gridDim.x * blockDim.x) {
    long long t, e;
    if (a.elem_stride == 1) { t = idx / a.n; e = idx - t * a.n; }
    else { e = idx / a.batch; t = idx - e * a.batch; }
    const double2 v = src[t * a.n + e];
    a.data[fft_base(t, a) + e * a.elem_stride] = make_double2(v.x * a.scale, v.y * a.scale);
  }
}

__global__ void k_fft_stage(const double2* __restrict__ in, double2* __restrict__ out,
                            const double2* __restrict__ tw, long long batch, int n, int R,
                            int Ns) {
  const long long tot = batch * n;
  const int T = n / R;
  const int m = n / (Ns * R);
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < tot;
       idx += (long long)gridDim.x * blockDim.x) {
    const long long tr = idx / n;
    const int o = (int)(idx - tr * n);
    const int k = o % Ns;
    const int q = (o / Ns) % R;
    const int j = (o / (Ns * R)) * Ns + k;
    const long long e = (long long)m * (k + q * Ns);
    const double2* src = in + tr * n + j;
    double re = 0.0, im = 0.0;
    long long acc = 0;
    for (int r = 0; r < R; ++r) {
      const double2 x = src[(long long)r * T];
      const double2 w = tw[acc];
      re += x.x * w.x - x.y * w.y;
      im += x.x * w.y + x.y * w.x;
      acc += e;
      if (acc >= n) acc -= n;
    }
    out[tr * n + o] = make_double2(re, im);
  }
}

}  // namespace

namespace bh {

size_t fft_workspace_bytes(int64_t n, int64_t batch) {
  if (n <= kFftLdsMax) return 0;
  return (size_t)(2 * batch * n + n) * sizeof(double2);
}

int fft_strided(double2* data, int64_t n, int64_t batch, int64_t inner, int64_t inner_stride,
                int64_t outer_stride, int64_t elem_stride, int dir, double scale, void* ws,
                size_t ws_bytes, hipStream_t s) {
  if (n < 1 || batch < 0 || inner < 1) BH_FAIL(BH_ERR_ARG, "fft: bad sizes n=%lld batch=%lld",
                                             (long long)n, (long long)batch);
  if (dir != -1 && dir != 1) BH_FAIL(BH_ERR_ARG, "fft: dir must be -1 or +1");
  if (batch == 0) return BH_OK;
  FftArgs a;
  if (!make_plan(n, a)) BH_FAIL(BH_ERR_UNSUPPORTED, "fft: cannot factor n=%lld", (long long)n);
  a.data = data;
  a.batch = batch;
  a.inner = inner;
  a.inner_stride = inner_stride;
  a.outer_stride = outer_stride;
  a.elem_stride = elem_stride;
  a.dir = dir;
  a.scale = scale;
  if (n <= kFftLdsMax) {
    a.tpb = (int)(kFftLdsShare / n);
    if (a.tpb < 1) a.tpb = 1;
    if (a.tpb > batch) a.tpb = (int)batch;
    const size_t lds = (size_t)(2 * a.tpb * n + n) * sizeof(double2);
    if (lds > 65536) {
      // above the default dynamic LDS limit of a launch
      static bool raised = false;
      if (!raised) {
        if (hipFuncSetAttribute((const void*)k_fft_lds, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)((2 * kFftLdsMax + kFftLdsMax) * sizeof(double2))) != hipSuccess)
          BH_FAIL(BH_ERR_HIP, "fft: cannot raise the LDS limit of k_fft_lds");
        raised = true;
      }
    }
    const unsigned grid = (unsigned)((batch + a.tpb - 1) / a.tpb);
    // a workgroup that holds one long transform takes 16 waves, so that the
    // CU it fills alone still has four waves per SIMD to hide LDS latency
    // (the butterfly form of a power of two has n / 4 butterflies per
    // transform and stage: 256 threads)
    const unsigned threads = (a.tpb == 1 && n >= 1024 && !a.pow2) ? 1024 : 256;
    hipLaunchKernelGGL(k_fft_lds, dim3(grid), dim3(threads), lds, s, a);
    BH_LAUNCH_CHECK();
    return BH_OK;
  }
  const size_t need = fft_workspace_bytes(n, batch);
  if (ws == nullptr || ws_bytes < need)
    BH_FAIL(BH_ERR_WORKSPACE, "fft: workspace of %zu bytes needed, %zu given", need, ws_bytes);
  double2* bufA = (double2*)ws;
  double2* bufB = bufA + batch * n;
  double2* tw = bufB + batch * n;
  hipLaunchKernelGGL(k_fft_twiddles, dim3(grid1d(n, 256)), dim3(256), 0, s, tw, (int)n, dir);
  const unsigned g = grid1d(batch * n, 256) > 65536 ? 65536 : grid1d(batch * n, 256);
  hipLaunchKernelGGL(k_fft_gather, dim3(g), dim3(256), 0, s, a, bufA);
  int Ns = 1;
  for (int st = 0; st < a.nst; ++st) {
    hipLaunchKernelGGL(k_fft_stage, dim3(g), dim3(256), 0, s, bufA, bufB, tw, (long long)batch,
                       (int)n, a.radix[st], Ns);
    double2* tmp = bufA; bufA = bufB; bufB = tmp;
    Ns *= a.radix[st];
  }
  hipLaunchKernelGGL(k_fft_scatter, dim3(g), dim3(256), 0, s, a, bufA);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

int fftn_ortho(double2* data, int ndim, const int64_t* N, int dir, void* ws, size_t ws_bytes,
               hipStream_t s) {
  const int64_t total = prod(N, ndim);
  for (int ax = 0; ax < ndim; ++ax) {
    const int64_t n = N[ax];
    const int64_t inner = prod(N + ax + 1, ndim - ax - 1);
    const int64_t batch = total / n;
    int rc = fft_strided(data, n, batch, inner, 1, n * inner, inner, dir,
                         1.0 / sqrt((double)n), ws, ws_bytes, s);
    if (rc) return rc;
  }
  return BH_OK;
}

}  // namespace bh

extern "C" size_t bh_fft_workspace_size(int64_t n, int64_t batch) {
  return bh::fft_workspace_bytes(n, batch);
}

extern "C" int bh_fft(double* data, int64_t n, int64_t batch, int64_t inner,
                      int64_t inner_stride, int64_t outer_stride, int64_t elem_stride, int dir,
                      double scale, void* ws, size_t ws_bytes, bh_stream s) {
  bh::clear_error();
  if (data == nullptr) BH_FAIL(BH_ERR_ARG, "bh_fft: data is null");
  return bh::fft_strided((double2*)data, n, batch, inner, inner_stride, outer_stride,
                         elem_stride, dir, scale, ws, ws_bytes, (hipStream_t)s);
}
