// Kernels of the learned reconstruction network (a U-Net on 2-D images,
// NHWC bf16 activations): bf16 MFMA matrix products with FP32 accumulation
// (a general GEMM, and 3x3 convolutions with their weight and input
// gradients as implicit GEMMs that gather their operands straight from the
// activations), and the elementwise pieces of forward and backward passes
// and of the Adam update.
// No reference counterpart: the reference has no reconstruction network.
//
// GEMM: C[M][N] = alpha * sum_k A(m, k) B(k, n) (+ bias[n], ReLU), with
//   A stored [M][lda] (ta = 0, k contiguous) or [K][lda] (ta = 1),
//   B stored [K][ldb] (tb = 0, n contiguous) or [N][ldb] (tb = 1),
// the contiguous dimension of each operand a multiple of 8 (16-byte vectors).
// Workgroup tile 64 x 64, k step 32, four waves of 32 x 32 each made of 2 x 2
// v_mfma_f32_16x16x32_bf16 blocks; both operands are staged in LDS with k
// contiguous so every fragment is one 16-byte LDS read.  Split-K (grid z)
// writes FP32 partial tiles that bh_nn_sum_partials adds in order.
#include "bh_internal.hpp"

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float floatx4;
typedef unsigned short bfbits;   // bf16 storage

__device__ inline float bf2f(bfbits b) { return __uint_as_float((unsigned)b << 16); }
__device__ inline bfbits f2bf(float f) {
  const __bf16 h = (__bf16)f;   // round to nearest even
  return __builtin_bit_cast(bfbits, h);
}

constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;

// 16 zero bytes in global memory (never written): a load of an element
// outside its operand reads them instead, a pointer select between two
// global addresses.  Written as "cond ? *p : zero", the compiler selects
// between p and a private copy of the zero and issues a flat load, which
// also counts on the LDS counter, so that every LDS wait then waits for the
// global loads in flight as well (and a const array would be constant
// memory, with the same flat result).
__device__ __attribute__((aligned(16))) unsigned kZero16[4] = {0u, 0u, 0u, 0u};
__device__ inline uint4 load16_or_zero(bool ok, const bfbits* p) {
  return *(ok ? reinterpret_cast<const uint4*>(p) : reinterpret_cast<const uint4*>(kZero16));
}

struct GemmArgs {
  const bfbits* A;
  const bfbits* B;
  void* C;
  long long M, N, K, lda, ldb, ldc;
  int ta, tb;
  int out_bf16;
  const float* bias;
  int relu;
  float alpha;
  long long kchunk;         // k range per split (multiple of BK)
  long long split_stride;   // elements between split partials (FP32 output)
};

__device__ inline void load_k_contig(const bfbits* src, long long rows, long long K, long long ld,
                                     long long r0, long long k0, bfbits (*dst)[LDK]) {
  // 64 rows x 32 k, thread t: row t / 4, k segment (t % 4) * 8
  const int t = threadIdx.x;
  const int row = t >> 2, ks = (t & 3) * 8;
  const long long r = r0 + row, k = k0 + ks;
  uint4 v = make_uint4(0, 0, 0, 0);
  if (r < rows && k < K) v = *reinterpret_cast<const uint4*>(src + r * ld + k);
  *reinterpret_cast<uint4*>(&dst[row][ks]) = v;
}

__device__ inline void load_k_strided(const bfbits* src, long long rows, long long K, long long ld,
                                      long long r0, long long k0, bfbits (*dst)[LDK]) {
  // storage [K][ld] with the 64 rows contiguous: thread t reads 8 rows at
  // k = t / 8 and writes them transposed
  const int t = threadIdx.x;
  const int kk = t >> 3, rs = (t & 7) * 8;
  const long long k = k0 + kk, r = r0 + rs;
  uint4 v = make_uint4(0, 0, 0, 0);
  if (k < K && r < rows) v = *reinterpret_cast<const uint4*>(src + k * ld + r);
  const bfbits* e = reinterpret_cast<const bfbits*>(&v);
  for (int i = 0; i < 8; ++i) dst[rs + i][kk] = e[i];
}

__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) bfbits As[BM][LDK];
  __shared__ __attribute__((aligned(16))) bfbits Bs[BN][LDK];
  const long long m0 = blockIdx.y * (long long)BM, n0 = blockIdx.x * (long long)BN;
  const long long kb = blockIdx.z * g.kchunk;
  const long long ke = (kb + g.kchunk < g.K) ? kb + g.kchunk : g.K;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  floatx4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (long long k0 = kb; k0 < ke; k0 += BK) {
    if (g.ta) load_k_strided(g.A, g.M, ke, g.lda, m0, k0, As);
    else load_k_contig(g.A, g.M, ke, g.lda, m0, k0, As);
    if (g.tb) load_k_contig(g.B, g.N, ke, g.ldb, n0, k0, Bs);
    else load_k_strided(g.B, g.N, ke, g.ldb, n0, k0, Bs);
    __syncthreads();
    bf16x8 a[2], b[2];
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    for (int i = 0; i < 2; ++i)
      a[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&As[wm + i * 16 + fr][fk]));
    for (int j = 0; j < 2; ++j)
      b[j] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&Bs[wn + j * 16 + fr][fk]));
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // C fragment: column lane % 16, rows (lane / 16) * 4 + r
  const int cc = lane & 15, cr = (lane >> 4) * 4;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      const long long n = n0 + wn + j * 16 + cc;
      if (n >= g.N) continue;
      const float bv = g.bias ? g.bias[n] : 0.f;
      for (int r = 0; r < 4; ++r) {
        const long long m = m0 + wm + i * 16 + cr + r;
        if (m >= g.M) continue;
        float v = g.alpha * acc[i][j][r] + bv;
        if (g.relu) v = fmaxf(v, 0.f);
        if (g.out_bf16) ((bfbits*)g.C)[m * g.ldc + n] = f2bf(v);
        else ((float*)g.C)[blockIdx.z * g.split_stride + m * g.ldc + n] = v;
      }
    }
}

__global__ void k_sum_partials(const float* __restrict__ part, int nsplit, long long n,
                               float* __restrict__ out) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int z = 0; z < nsplit; ++z) s += part[z * n + i];
  out[i] = s;
}

// 2x2 average pooling and its adjoint (NHWC)
__global__ void k_pool2(const bfbits* __restrict__ X, long long B, int H, int W, int C,
                        bfbits* __restrict__ Y) {
  const int Ho = H / 2, Wo = W / 2;
  const long long total = B * Ho * Wo * (long long)C;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int c = (int)(i % C);
  const long long p = i / C;
  const int x = (int)(p % Wo), y = (int)((p / Wo) % Ho);
  const long long n = p / ((long long)Wo * Ho);
  float s = 0.f;
  for (int dy = 0; dy < 2; ++dy)
    for (int dx = 0; dx < 2; ++dx)
      s += bf2f(X[((n * H + 2 * y + dy) * W + 2 * x + dx) * C + c]);
  Y[i] = f2bf(0.25f * s);
}

__global__ void k_pool2_bwd(const bfbits* __restrict__ dY, long long B, int H, int W, int C,
                            bfbits* __restrict__ dX) {
  const long long total = B * H * W * (long long)C;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int c = (int)(i % C);
  const long long p = i / C;
  const int x = (int)(p % W), y = (int)((p / W) % H);
  const long long n = p / ((long long)W * H);
  dX[i] = f2bf(0.25f * bf2f(dY[((n * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c]));
}

// nearest 2x upsampling and its adjoint
__global__ void k_up2(const bfbits* __restrict__ X, long long B, int H, int W, int C,
                      bfbits* __restrict__ Y) {
  const long long total = B * 4LL * H * W * C;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int c = (int)(i % C);
  const long long p = i / C;
  const int x = (int)(p % (2 * W)), y = (int)((p / (2 * W)) % (2 * H));
  const long long n = p / (4LL * W * H);
  Y[i] = X[((n * H + y / 2) * W + x / 2) * C + c];
}

__global__ void k_up2_bwd(const bfbits* __restrict__ dY, long long B, int H, int W, int C,
                          bfbits* __restrict__ dX) {
  const long long total = B * H * W * (long long)C;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int c = (int)(i % C);
  const long long p = i / C;
  const int x = (int)(p % W), y = (int)((p / W) % H);
  const long long n = p / ((long long)W * H);
  float s = 0.f;
  for (int dy = 0; dy < 2; ++dy)
    for (int dx = 0; dx < 2; ++dx)
      s += bf2f(dY[((n * 2 * H + 2 * y + dy) * 2 * W + 2 * x + dx) * C + c]);
  dX[i] = f2bf(s);
}

// Z = X + Y (bf16, FP32 add)
__global__ void k_add_bf16(const bfbits* __restrict__ X, const bfbits* __restrict__ Y, long long n,
                           bfbits* __restrict__ Z) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) Z[i] = f2bf(bf2f(X[i]) + bf2f(Y[i]));
}

// dZ = dY where Yact > 0, else 0
__global__ void k_relu_bwd(const bfbits* __restrict__ dY, const bfbits* __restrict__ Yact,
                           long long n, bfbits* __restrict__ dZ) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) dZ[i] = bf2f(Yact[i]) > 0.f ? dY[i] : (bfbits)0;
}

// The same element-wise kernels on 16-byte chunks of 8 channels (C a
// multiple of 8, 16-byte aligned operands; the kernels above otherwise):
// one chunk per thread, the same FP32 operations per value in the same
// order, so the results are bit-identical.  The 2x2 kernels index by the
// small image's pixel p = (n h + y) w + x: its 2x2 block in the large image
// is rows 2 (n h + y) + dy, columns 2 x + dx.
__device__ inline float bf_lo(unsigned u) { return __uint_as_float(u << 16); }
__device__ inline float bf_hi(unsigned u) { return __uint_as_float(u & 0xffff0000u); }
__device__ inline unsigned bf_pack(float lo, float hi) {
  return (unsigned)f2bf(lo) | ((unsigned)f2bf(hi) << 16);
}

// large (2h, 2w) image from the small (h, w) one: nearest upsampling
// (copy), or the adjoint of average pooling (0.25 x, rounded to bf16)
__global__ void k_spread2_v8(const uint4* __restrict__ X, long long n8, int w, int C8, int quarter,
                             uint4* __restrict__ Y) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n8) return;
  const int cg = (int)(i % C8);
  const long long p = i / C8;
  const int x = (int)(p % w);
  const long long q = p / w;
  uint4 v = X[i];
  if (quarter) {
    unsigned* u = reinterpret_cast<unsigned*>(&v);
#pragma unroll
    for (int k = 0; k < 4; ++k) u[k] = bf_pack(0.25f * bf_lo(u[k]), 0.25f * bf_hi(u[k]));
  }
#pragma unroll
  for (int dy = 0; dy < 2; ++dy)
#pragma unroll
    for (int dx = 0; dx < 2; ++dx)
      Y[((2 * q + dy) * 2 * w + 2 * x + dx) * C8 + cg] = v;
}

// small (h, w) image from the large one: 2x2 sums in (dy, dx) order, times
// 0.25 (average pooling) or 1 (adjoint of nearest upsampling)
__global__ void k_gather2_v8(const uint4* __restrict__ X, long long n8, int w, int C8, float scale,
                             uint4* __restrict__ Y) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n8) return;
  const int cg = (int)(i % C8);
  const long long p = i / C8;
  const int x = (int)(p % w);
  const long long q = p / w;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int dy = 0; dy < 2; ++dy)
#pragma unroll
    for (int dx = 0; dx < 2; ++dx) {
      const uint4 v = X[((2 * q + dy) * 2 * w + 2 * x + dx) * C8 + cg];
      const unsigned u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        s[2 * k] += bf_lo(u[k]);
        s[2 * k + 1] += bf_hi(u[k]);
      }
    }
  Y[i] = uint4{bf_pack(scale * s[0], scale * s[1]), bf_pack(scale * s[2], scale * s[3]),
               bf_pack(scale * s[4], scale * s[5]), bf_pack(scale * s[6], scale * s[7])};
}

__global__ void k_add_bf16_v8(const uint4* __restrict__ X, const uint4* __restrict__ Y,
                              long long n8, uint4* __restrict__ Z) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n8) return;
  const uint4 a = X[i], b = Y[i];
  Z[i] = uint4{bf_pack(bf_lo(a.x) + bf_lo(b.x), bf_hi(a.x) + bf_hi(b.x)),
               bf_pack(bf_lo(a.y) + bf_lo(b.y), bf_hi(a.y) + bf_hi(b.y)),
               bf_pack(bf_lo(a.z) + bf_lo(b.z), bf_hi(a.z) + bf_hi(b.z)),
               bf_pack(bf_lo(a.w) + bf_lo(b.w), bf_hi(a.w) + bf_hi(b.w))};
}

__device__ inline unsigned relu_mask2(unsigned d, unsigned y) {
  return (bf_lo(y) > 0.f ? d & 0xffffu : 0u) | (bf_hi(y) > 0.f ? d & 0xffff0000u : 0u);
}

__global__ void k_relu_bwd_v8(const uint4* __restrict__ dY, const uint4* __restrict__ Yact,
                              long long n8, uint4* __restrict__ dZ) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n8) return;
  const uint4 d = dY[i], y = Yact[i];
  dZ[i] = uint4{relu_mask2(d.x, y.x), relu_mask2(d.y, y.y), relu_mask2(d.z, y.z),
                relu_mask2(d.w, y.w)};
}

// column sums of a [P][C] bf16 matrix in two fixed-order passes.  Pass 1:
// block b sums rows [b * chunk, (b + 1) * chunk); its threads form G row
// groups of C columns (coalesced along C), group sums added in group order.
// Pass 2: one block per column adds the block partials by a fixed tree.
__global__ __launch_bounds__(256) void k_colsum_part(const bfbits* __restrict__ X, long long P,
                                                    int C, long long chunk,
                                                    float* __restrict__ part) {
  __shared__ float red[256];
  const long long p0 = blockIdx.x * chunk;
  const long long p1 = (p0 + chunk < P) ? p0 + chunk : P;
  const int t = threadIdx.x;
  for (int cb = 0; cb < C; cb += 256) {
    const int cw = (C - cb < 256) ? C - cb : 256;   // columns in this pass
    const int G = 256 / cw;
    const int c = t % cw, r = t / cw;
    float s = 0.f;
    if (r < G)
      for (long long p = p0 + r; p < p1; p += G) s += bf2f(X[p * C + cb + c]);
    red[t] = s;
    __syncthreads();
    if (t < cw) {
      float tot = 0.f;
      for (int g = 0; g < G; ++g) tot += red[g * cw + t];
      part[blockIdx.x * (long long)C + cb + t] = tot;
    }
    __syncthreads();
  }
}

// pass 1 on 16-byte chunks (C a multiple of 8, at most 2048 columns, X
// 16-byte aligned): G = 256 / (C / 8) row groups of C / 8 threads, each
// summing 8 columns
__global__ __launch_bounds__(256) void k_colsum_part_v8(const uint4* __restrict__ X, long long P,
                                                       int C8, long long chunk,
                                                       float* __restrict__ part) {
  __shared__ float red[256][9];
  const long long p0 = blockIdx.x * chunk;
  const long long p1 = (p0 + chunk < P) ? p0 + chunk : P;
  const int t = threadIdx.x, G = 256 / C8;
  const int cg = t % C8, r = t / C8;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (r < G)
    for (long long p = p0 + r; p < p1; p += G) {
      const uint4 v = X[p * C8 + cg];
      const unsigned u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        s[2 * k] += bf_lo(u[k]);
        s[2 * k + 1] += bf_hi(u[k]);
      }
    }
#pragma unroll
  for (int k = 0; k < 8; ++k) red[t][k] = s[k];
  __syncthreads();
  const int C = 8 * C8;
  for (int c = t; c < C; c += 256) {
    float tot = 0.f;
    for (int g = 0; g < G; ++g) tot += red[g * C8 + c / 8][c % 8];
    part[blockIdx.x * (long long)C + c] = tot;
  }
}

__global__ __launch_bounds__(256) void k_colsum_final(const float* __restrict__ part, long long nb,
                                                     int C, float* __restrict__ out) {
  __shared__ float red[256];
  const int c = blockIdx.x;
  float s = 0.f;
  for (long long b = threadIdx.x; b < nb; b += 256) s += part[b * C + c];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = red[0];
}

inline long long colsum_chunk(long long P) {
  long long ch = (P + 1023) / 1024;   // about 1024 blocks
  return ch < 64 ? 64 : ch;
}

// squared-error loss on channel 0 of pred [P][cs] against target [P]:
// grad[p][c] = (c == 0) * 2 (pred - target) / P (bf16); per-block partial sums of
// the squared error (fixed order) to part[block]
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ pred, int cs,
                                            const float* __restrict__ target, long long P,
                                            bfbits* __restrict__ grad, int gs,
                                            float* __restrict__ part) {
  __shared__ float red[256];
  const long long p = blockIdx.x * 256LL + threadIdx.x;
  float e2 = 0.f;
  if (p < P) {
    const float d = pred[p * cs] - target[p];
    e2 = d * d;
    for (int c = 0; c < gs; ++c) grad[p * gs + c] = c == 0 ? f2bf(2.f * d / (float)P) : (bfbits)0;
  }
  red[threadIdx.x] = e2;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam (Kingma and Ba 2015) on FP32 master weights, bf16 copy refreshed
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long long n, float lr, float b1, float b2, float eps,
                       float c1, float c2, bfbits* __restrict__ wbf) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wbf) wbf[i] = f2bf(wi);
}

__global__ void k_f32_to_bf16(const float* __restrict__ x, long long n, float scale,
                              bfbits* __restrict__ y) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(scale * x[i]);
}

__global__ void k_bf16_to_f32(const bfbits* __restrict__ x, long long n, float scale,
                              float* __restrict__ y) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) y[i] = scale * bf2f(x[i]);
}

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

// ---------------------------------------------------------------------------
// Implicit-GEMM 3x3 convolution (NHWC, same padding):
//   Y[p][o] = act(alpha sum_k A(p, k) Wt[o][k] + bias[o]),
//   A(p, k) = X[n][y + dy][x + dx][c] (zero outside), k = tap C + c,
//   tap = 3 (dy + 1) + dx + 1
// without a column buffer: each k step the workgroup gathers its 128 pixels'
// 32 k values (four 8-channel segments, each inside one tap since C is a
// multiple of 8) straight from X into LDS.  Four waves of 32 pixels x 16 NB
// output channels, v_mfma_f32_16x16x32_bf16 (the GEMM's fragment layout);
// LDS double-buffered with the next step's global loads held in registers
// while the current step multiplies, one barrier per step.
// ---------------------------------------------------------------------------
constexpr int CM = 128, CK = 32, CLD = CK + 8;

struct ConvArgs {
  const bfbits* X;
  const bfbits* Wt;
  const float* bias;
  void* Y;
  long long P;
  int H, Wd, C, Cout;
  float alpha;
  int relu, out_bf16;
};

template <int NB, int KS>
__global__ __launch_bounds__(256 * KS) void k_conv3_igemm(ConvArgs a) {
  // KS k groups of four waves: group g multiplies the k steps g, g + KS, ...
  // into its own accumulators (with its own LDS buffers), and group 0 adds
  // the other groups' tiles at the end in group order.  Two groups give a
  // launch of few workgroups (the 32^2 and 64^2 layers: 256-512) eight
  // waves per workgroup to hide each other's staging latency.
  constexpr int BNc = 16 * NB;
  constexpr int BLD = (BNc * 4 + 255) / 256;   // weight segments per thread
  constexpr int AS_E = KS * 2 * CM * CLD, BS_E = KS * 2 * BNc * CLD;
  constexpr int RED_B = KS > 1 ? (KS - 1) * 4 * 2 * NB * 64 * 16 : 0;   // other groups' tiles
  constexpr int OR = BNc + 8;                                            // staged output row
  constexpr int OUT_B = CM * OR * 2;
  constexpr int SM_B0 = (AS_E + BS_E) * 2 > RED_B ? (AS_E + BS_E) * 2 : RED_B;
  constexpr int SM_B = SM_B0 > OUT_B ? SM_B0 : OUT_B;
  __shared__ __attribute__((aligned(16))) unsigned char smem[SM_B];
  auto As = reinterpret_cast<bfbits(*)[2][CM][CLD]>(smem);
  auto Bs = reinterpret_cast<bfbits(*)[2][BNc][CLD]>(smem + AS_E * 2);
  const int g = KS > 1 ? (int)(threadIdx.x >> 8) : 0;
  const int t = threadIdx.x & 255, lane = t & 63, w = t >> 6;
  const long long p0 = blockIdx.x * (long long)CM;
  const int n0 = blockIdx.y * BNc;
  const int K = 9 * a.C;
  const int seg = (t & 3) * 8;
  int py[2], px[2];
  long long pn[2];
  bool pv[2];
  for (int h = 0; h < 2; ++h) {
    const long long p = p0 + (t >> 2) + 64 * h;
    pv[h] = p < a.P;
    const long long q = pv[h] ? p : 0;
    px[h] = (int)(q % a.Wd);
    py[h] = (int)((q / a.Wd) % a.H);
    pn[h] = q / ((long long)a.Wd * a.H);
  }
  uint4 ra[2], rb[BLD];
  // k steps past the last read zeros (a group may have one step fewer)
  auto gload = [&](int k0) {
    const int k = k0 + seg;
    const bool kin = k < K;
    const int tap = kin ? k / a.C : 0;
    const int ci = kin ? k - tap * a.C : 0;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    for (int h = 0; h < 2; ++h) {
      const int yy = py[h] + dy, xx = px[h] + dx;
      const bool ok = pv[h] && kin && yy >= 0 && yy < a.H && xx >= 0 && xx < a.Wd;
      ra[h] = load16_or_zero(ok, a.X + ((pn[h] * a.H + yy) * a.Wd + xx) * (long long)a.C + ci);
    }
    for (int j = 0; j < BLD; ++j) {
      const int i = t + 256 * j;
      const int row = i >> 2, kk = k0 + (i & 3) * 8;
      const bool ok = i < BNc * 4 && n0 + row < a.Cout && kk < K;
      rb[j] = load16_or_zero(ok, a.Wt + (long long)(n0 + row) * K + kk);
    }
  };
  auto sstore = [&](int buf) {
    for (int h = 0; h < 2; ++h)
      *reinterpret_cast<uint4*>(&As[g][buf][(t >> 2) + 64 * h][seg]) = ra[h];
    for (int j = 0; j < BLD; ++j) {
      const int i = t + 256 * j;
      if (i < BNc * 4) *reinterpret_cast<uint4*>(&Bs[g][buf][i >> 2][(i & 3) * 8]) = rb[j];
    }
  };
  floatx4 acc[2][NB];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < NB; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int nsteps = (K + CK - 1) / CK;
  const int niter = (nsteps + KS - 1) / KS;   // the same for every group
  gload(g * CK);
  sstore(0);
  __syncthreads();
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  for (int it = 0; it < niter; ++it) {
    const int buf = it & 1;
    const bool more = it + 1 < niter;
    if (more) gload(((it + 1) * KS + g) * CK);
    bf16x8 av[2], bv[NB];
    for (int i = 0; i < 2; ++i)
      av[i] = __builtin_bit_cast(
          bf16x8, *reinterpret_cast<const uint4*>(&As[g][buf][w * 32 + i * 16 + fr][fk]));
    for (int j = 0; j < NB; ++j)
      bv[j] = __builtin_bit_cast(bf16x8,
                                 *reinterpret_cast<const uint4*>(&Bs[g][buf][j * 16 + fr][fk]));
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < NB; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
    if (more) sstore(buf ^ 1);
    __syncthreads();
  }
  if constexpr (KS > 1) {
    // the other groups' accumulator tiles through LDS (the operand buffers
    // are free after the loop's last barrier), added by group 0 in order
    float4* red = reinterpret_cast<float4*>(smem);
    if (g > 0)
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < NB; ++j) {
          const floatx4 v = acc[i][j];
          red[((((g - 1) * 4 + w) * 2 + i) * NB + j) * 64 + lane] = make_float4(v[0], v[1], v[2], v[3]);
        }
    __syncthreads();
    if (g == 0)
      for (int q = 1; q < KS; ++q)
        for (int i = 0; i < 2; ++i)
          for (int j = 0; j < NB; ++j) {
            const float4 v = red[((((q - 1) * 4 + w) * 2 + i) * NB + j) * 64 + lane];
            acc[i][j][0] += v.x;
            acc[i][j][1] += v.y;
            acc[i][j][2] += v.z;
            acc[i][j][3] += v.w;
          }
    if (a.out_bf16) __syncthreads();   // the tiles are read before the output is staged
    else if (g > 0) return;
  }
  const int cc = lane & 15, cr = (lane >> 4) * 4;
  if (a.out_bf16) {
    // the bf16 output tile (128 pixels x BNc channels) through LDS, so that
    // it leaves in 16-byte stores of 8 channels (by every thread of the
    // workgroup) instead of one 2-byte store per value
    bfbits* so = reinterpret_cast<bfbits*>(smem);
    if (g == 0)
      for (int j = 0; j < NB; ++j) {
        const int n = n0 + j * 16 + cc;
        const float bv = (a.bias && n < a.Cout) ? a.bias[n] : 0.f;
        for (int i = 0; i < 2; ++i)
          for (int r = 0; r < 4; ++r) {
            float v = a.alpha * acc[i][j][r] + bv;
            if (a.relu) v = fmaxf(v, 0.f);
            so[(w * 32 + i * 16 + cr + r) * OR + j * 16 + cc] = f2bf(v);
          }
      }
    __syncthreads();
    const int nv8 = (min(BNc, a.Cout - n0)) >> 3;
    for (int e = threadIdx.x; e < CM * (BNc / 8); e += 256 * KS) {
      const int p = e / (BNc / 8), c8 = e - p * (BNc / 8);
      if (c8 >= nv8 || p0 + p >= a.P) continue;
      *reinterpret_cast<uint4*>((bfbits*)a.Y + (p0 + p) * a.Cout + n0 + 8 * c8) =
          *reinterpret_cast<const uint4*>(&so[p * OR + 8 * c8]);
    }
    return;
  }
  for (int j = 0; j < NB; ++j) {
    const int n = n0 + j * 16 + cc;
    if (n >= a.Cout) continue;
    const float bv = a.bias ? a.bias[n] : 0.f;
    for (int i = 0; i < 2; ++i)
      for (int r = 0; r < 4; ++r) {
        const long long p = p0 + w * 32 + i * 16 + cr + r;
        if (p >= a.P) continue;
        float v = a.alpha * acc[i][j][r] + bv;
        if (a.relu) v = fmaxf(v, 0.f);
        ((float*)a.Y)[p * a.Cout + n] = v;
      }
  }
}

// Halo-tile form of the same convolution for inputs of up to 96 channels:
// a workgroup owns a tile of TH x 32 output pixels of one image and loads
// the (TH + 2) x 34 input pixels around it into LDS once (coalesced 16-byte
// rows, zero outside the image); every k step then reads its A fragments
// from that halo at the step's tap offset (k = tap C + c, eight channels of
// one tap per lane group), so each input element is fetched from memory
// once instead of once per tap.  The weight tile of the k step is staged
// and double-buffered as in k_conv3_igemm (the whole weight tile loaded
// with the halo, where both fit in 80 KB of LDS, measured no faster on the
// U-Net's layers).  The halo's pixel stride is C or C + 8 channels,
// whichever is an odd number of 16-byte words (16-byte LDS reads of 16
// consecutive pixels then hit distinct banks).
constexpr int HTW = 32;

template <int C8, int TH, int NB>
__global__ __launch_bounds__(256) void k_conv3_halo(ConvArgs a, int tiles_x, int tiles_y) {
  constexpr int C = 8 * C8;
  constexpr int CS = (C8 & 1) ? C : C + 8;
  constexpr int HH = TH + 2, HW = HTW + 2;
  constexpr int MB = TH / 2;                 // 16-pixel blocks per wave (TH x 32 / 4 / 16)
  constexpr int BNc = 16 * NB;
  constexpr int NG = 9 * C8;                 // k groups of 8
  constexpr int NSTEP = (NG + 3) / 4;
  constexpr int BLD = (BNc * 4 + 255) / 256;
  // the halo and the weight tiles, and after the k steps the bf16 output
  // tile (TH x 32 pixels x BNc channels, rows padded by 8 channels), so that
  // it leaves in 16-byte stores of 8 channels instead of one 2-byte store
  // per value
  constexpr int HALO_E = HH * HW * CS, BS_E = 2 * BNc * CLD, OR = BNc + 8;
  constexpr int SM_E = HALO_E + BS_E > TH * HTW * OR ? HALO_E + BS_E : TH * HTW * OR;
  __shared__ __attribute__((aligned(16))) bfbits smem[SM_E];
  bfbits* const halo = smem;
  auto Bs = reinterpret_cast<bfbits(*)[BNc][CLD]>(smem + HALO_E);
  __shared__ uint4 zlds;   // k groups past 9 C read these zeros
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (t == 0) zlds = make_uint4(0, 0, 0, 0);
  int tb = blockIdx.x;
  const int tx = tb % tiles_x;
  tb /= tiles_x;
  const int ty = tb % tiles_y;
  const long long img = tb / tiles_y;
  const int x0 = tx * HTW, y0 = ty * TH;
  const int n0 = blockIdx.y * BNc;
  const int K = 9 * C;
  // the halo's loads are all issued before the first is stored to LDS: a
  // loop that stored each load before issuing the next waited for one
  // memory round trip per iteration ([256, 48, 24] at batch 16: 100 -> 90 us)
  constexpr int NHL = (HH * HW * C8 + 255) / 256;
  uint4 hv[NHL];
#pragma unroll
  for (int q = 0; q < NHL; ++q) {
    const int i = t + 256 * q;
    const int ch = i % C8, pix = i / C8;
    const int hy = pix / HW, hx = pix - hy * HW;
    const int yy = y0 + hy - 1, xx = x0 + hx - 1;
    const bool in = i < HH * HW * C8 && yy >= 0 && yy < a.H && xx >= 0 && xx < a.Wd;
    hv[q] = load16_or_zero(in, a.X + ((img * a.H + (in ? yy : 0)) * a.Wd + (in ? xx : 0)) * C + ch * 8);
  }
  uint4 rb[BLD];
  auto gload = [&](int k0) {
    for (int j = 0; j < BLD; ++j) {
      const int i = t + 256 * j;
      const int row = i >> 2, kk = k0 + (i & 3) * 8;
      const bool ok = i < BNc * 4 && n0 + row < a.Cout && kk < K;
      rb[j] = load16_or_zero(ok, a.Wt + (long long)(n0 + row) * K + kk);
    }
  };
  auto sstore = [&](int buf) {
    for (int j = 0; j < BLD; ++j) {
      const int i = t + 256 * j;
      if (i < BNc * 4) *reinterpret_cast<uint4*>(&Bs[buf][i >> 2][(i & 3) * 8]) = rb[j];
    }
  };
  floatx4 acc[MB][NB];
  for (int i = 0; i < MB; ++i)
    for (int j = 0; j < NB; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  gload(0);
#pragma unroll
  for (int q = 0; q < NHL; ++q) {
    const int i = t + 256 * q;
    if (i < HH * HW * C8) *reinterpret_cast<uint4*>(&halo[(i / C8) * CS + (i % C8) * 8]) = hv[q];
  }
  sstore(0);
  __syncthreads();   // the halo and the first weight tile
  const int fr = lane & 15, fg = lane >> 4;
  int hbase[MB];     // halo offset of this lane's pixel in each block (tap 0)
  for (int i = 0; i < MB; ++i) {
    const int p = w * (MB * 16) + i * 16 + fr;
    hbase[i] = ((p / HTW) * HW + (p % HTW)) * CS;
  }
  for (int st = 0; st < NSTEP; ++st) {
    const int buf = st & 1;
    const bool more = st + 1 < NSTEP;
    if (more) gload((st + 1) * 32);
    const int g = st * 4 + fg;
    const bool gin = g < NG;
    const int tap = gin ? g / C8 : 0;
    const int off = ((tap / 3) * HW + (tap % 3)) * CS + (g - tap * C8) * 8;
    bf16x8 av[MB], bv[NB];
    for (int i = 0; i < MB; ++i)
      av[i] = __builtin_bit_cast(
          bf16x8, *(gin ? reinterpret_cast<const uint4*>(&halo[hbase[i] + off]) : &zlds));
    for (int j = 0; j < NB; ++j)
      bv[j] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&Bs[buf][j * 16 + fr][fg * 8]));
    for (int i = 0; i < MB; ++i)
      for (int j = 0; j < NB; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
    if (more) sstore(buf ^ 1);
    __syncthreads();
  }
  const int cc = lane & 15, cr = (lane >> 4) * 4;
  if (a.out_bf16) {
    // (the loop's last barrier: no wave still reads the halo or weights)
    for (int j = 0; j < NB; ++j) {
      const int n = n0 + j * 16 + cc;
      const float bvv = (a.bias && n < a.Cout) ? a.bias[n] : 0.f;
      for (int i = 0; i < MB; ++i)
        for (int r = 0; r < 4; ++r) {
          float v = a.alpha * acc[i][j][r] + bvv;
          if (a.relu) v = fmaxf(v, 0.f);
          smem[(w * (MB * 16) + i * 16 + cr + r) * OR + j * 16 + cc] = f2bf(v);
        }
    }
    __syncthreads();
    const int nv8 = (min(BNc, a.Cout - n0)) >> 3;   // 8-channel chunks of this tile
    for (int e = t; e < TH * HTW * (BNc / 8); e += 256) {
      const int p = e / (BNc / 8), c8 = e - p * (BNc / 8);
      const int yy = y0 + p / HTW, xx = x0 + p % HTW;
      if (c8 >= nv8 || yy >= a.H || xx >= a.Wd) continue;
      *reinterpret_cast<uint4*>((bfbits*)a.Y + ((img * a.H + yy) * a.Wd + xx) * a.Cout + n0 + 8 * c8) =
          *reinterpret_cast<const uint4*>(&smem[p * OR + 8 * c8]);
    }
    return;
  }
  for (int j = 0; j < NB; ++j) {
    const int n = n0 + j * 16 + cc;
    if (n >= a.Cout) continue;
    const float bvv = a.bias ? a.bias[n] : 0.f;
    for (int i = 0; i < MB; ++i)
      for (int r = 0; r < 4; ++r) {
        const int p = w * (MB * 16) + i * 16 + cr + r;
        const int yy = y0 + p / HTW, xx = x0 + p % HTW;
        if (yy >= a.H || xx >= a.Wd) continue;
        const long long o = ((img * a.H + yy) * a.Wd + xx) * a.Cout + n;
        float v = a.alpha * acc[i][j][r] + bvv;
        if (a.relu) v = fmaxf(v, 0.f);
        ((float*)a.Y)[o] = v;
      }
  }
}

// Weight gradient of the 3x3 convolution without a column buffer:
//   dW[o][k] = sum_p dZ[p][o] A(p, k)   (A as in k_conv3_igemm)
// with the pixels as the MFMA reduction dimension, in 32-pixel steps; grid z
// splits the pixels into chunks whose FP32 partial tiles are added in order
// afterwards (deterministic).  Both operands lie pixel-major in memory; the
// kernels below stage them that way and let the LDS read transpose them.
constexpr int DWK = 128, DWP = 32;

struct ConvDwArgs {
  const bfbits* X;
  const bfbits* dZ;
  float* out;
  long long P;
  int H, Wd, C, Cout;
  long long chunk;          // per split: pixels (a multiple of DWP), or 4x8 tiles
  long long ntiles;         // 4x8-pixel tiles (the whole-tile kernel)
  long long split_stride;   // floats between split partials
};

// out[i] = sum_z part[z][i] in a fixed order for many splits: a block owns
// 16 outputs; its 16 phases each add a contiguous run of splits in order,
// then the 16 phase sums are added in order
__global__ __launch_bounds__(256) void k_sum_splits(const float* __restrict__ part, int ns,
                                                    long long n, float* __restrict__ out) {
  __shared__ float red[16][17];
  const int j = threadIdx.x & 15, q = threadIdx.x >> 4;
  const long long i = blockIdx.x * 16LL + j;
  const int per = (ns + 15) / 16;
  const int z0 = min(ns, q * per), z1 = min(ns, z0 + per);
  float s = 0.f;
  if (i < n) {
#pragma unroll 8
    for (int z = z0; z < z1; ++z) s += part[z * n + i];
  }
  red[q][j] = s;
  __syncthreads();
  if (q == 0 && i < n) {
    float t = 0.f;
    for (int r = 0; r < 16; ++r) t += red[r][j];
    out[i] = t;
  }
}

// Weight gradient of a slice of 8, 24 or 48 input channels (grid x; the
// launch uses 8 and 24) into a tile of up to 64 output channels (grid y),
// image width a multiple of 8, with
// gfx950's transposing LDS reads.  A step
// is a tile of 4 image rows x 8 pixels; the operands are staged as they lie
// in memory -- the tile's 6 x 10 input pixels (rows y - 1 .. y + 4, pixels
// x - 1 .. x + 8, all C channels: 1.9 staged pixels per pixel, 3.75 with
// one-row steps of 32 pixels) and its 32 dZ rows -- with plain 16-byte
// copies, and each MFMA operand (8 consecutive pixels of one channel per
// lane) is read with two ds_read_b64_tr_b16: in a 16-lane group, source lane
// 4 r + q points at 4 channels (quad q) of pixel r, and output lane i gets
// channel i of the four pixels.  For the X operand the quad's address holds
// the tap: 16 k columns (tap, channel) of a block span up to two taps, but
// a 4-channel quad never straddles one (C is a multiple of 8).  No register
// transposes, no transposed LDS writes.  One workgroup holds the whole dW
// tile of its slice (9 taps x C8 channel octets, as k columns tap * C + c;
// waves own the 16-column k blocks w, w + 4, ...); grid z splits the
// tiles (image-major, then row group, then 8-pixel column) into chunks; LDS
// double-buffered with later steps' loads issued before the multiplies.
typedef short v4s __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));

__device__ inline v4s lds_tr16(const bfbits* p) {
  return __builtin_amdgcn_ds_read_tr16_b64_v4i16((__attribute__((address_space(3))) v4s*)(p));
}

// 8 consecutive rows (pixels) of one column: rows p .. p + 3 at lo, p + 4 .. p + 7 at hi
__device__ inline bf16x8 tr_operand(const bfbits* lo, const bfbits* hi) {
  const v8s v = __builtin_shufflevector(lds_tr16(lo), lds_tr16(hi), 0, 1, 2, 3, 4, 5, 6, 7);
  return __builtin_bit_cast(bf16x8, v);
}

template <int C8, int NB>
__global__ __launch_bounds__(256) void k_conv3_dw_tr(ConvDwArgs a) {
  constexpr int C = 8 * C8;
  constexpr int K = 9 * C;
  constexpr int KB = (K + 15) / 16;          // 16-column k blocks
  constexpr int KBW = (KB + 3) / 4;          // per wave
  constexpr int BO = 16 * NB;
  constexpr int XR = C + 8;                  // staged pixel row (elements): 16 B of padding
  constexpr int ZR = BO + 8;
  constexpr int NXT = 6 * 10 * C8;           // X copies per step (16 B each)
  constexpr int NZT = 32 * (BO / 8);         // dZ copies per step
  constexpr int NXL = (NXT + 255) / 256, NZL = (NZT + 255) / 256;
  __shared__ __attribute__((aligned(16))) bfbits Xs[2][6][10][XR];
  __shared__ __attribute__((aligned(16))) bfbits Zs[2][32][ZR];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const long long tbeg = blockIdx.z * a.chunk;
  const long long tend = min(a.ntiles, tbeg + a.chunk);
  const int c0 = blockIdx.x * C, o0 = blockIdx.y * BO;
  const bfbits* Xc = a.X + c0;
  const bfbits* dZo = a.dZ + o0;
  const unsigned WB = (unsigned)a.Wd >> 3, HT = ((unsigned)a.H + 3) >> 2;
  // 8 or 24 input channels: two register sets, the loads of step st + 2 issued
  // while step st multiplies (every load is unconditional -- past the chunk
  // it reads the zero block -- so the waits before the LDS stores count
  // older loads only; [256, 24, 24] at batch 16: 83 -> 68 us); 48 channels:
  // one step ahead (the second set's registers cost more than it hides)
  constexpr bool AHEAD2 = C8 <= 3;
  uint4 xr0[NXL], zr0[NZL], xr1[AHEAD2 ? NXL : 1], zr1[AHEAD2 ? NZL : 1];
  auto gload = [&](long long tile, auto& xr, auto& zr) {
    const bool live = tile < tend;
    const unsigned tq = (unsigned)(live ? tile : 0);
    const unsigned rest = tq / WB;
    const int x0 = 8 * (int)(tq - rest * WB), y0 = 4 * (int)(rest % HT);
    const long long img = (long long)(rest / HT) * a.H;
#pragma unroll
    for (int q = 0; q < NXL; ++q) {
      const int i = t + 256 * q;
      const int cg = i % C8, px = (i / C8) % 10, sr = i / (10 * C8);
      const int y = y0 - 1 + sr, x = x0 - 1 + px;
      const bool ok = live && i < NXT && y >= 0 && y < a.H && x >= 0 && x < a.Wd;
      xr[q] = load16_or_zero(ok, Xc + ((img + (ok ? y : 0)) * a.Wd + (ok ? x : 0)) * (long long)a.C + 8 * cg);
    }
#pragma unroll
    for (int q = 0; q < NZL; ++q) {
      const int i = t + 256 * q;
      const int og = i % (BO / 8), pi = i / (BO / 8);
      const int y = y0 + (pi >> 3), x = x0 + (pi & 7);
      const bool ok = live && i < NZT && y < a.H && o0 + 8 * og < a.Cout;
      zr[q] = load16_or_zero(ok, dZo + (ok ? ((img + y) * a.Wd + x) * a.Cout + 8 * og : 0));
    }
  };
  auto sstore = [&](int buf, const auto& xr, const auto& zr) {
#pragma unroll
    for (int q = 0; q < NXL; ++q) {
      const int i = t + 256 * q;
      if (i < NXT) {
        const int cg = i % C8, px = (i / C8) % 10, sr = i / (10 * C8);
        *reinterpret_cast<uint4*>(&Xs[buf][sr][px][8 * cg]) = xr[q];
      }
    }
#pragma unroll
    for (int q = 0; q < NZL; ++q) {
      const int i = t + 256 * q;
      if (i < NZT) *reinterpret_cast<uint4*>(&Zs[buf][i / (BO / 8)][8 * (i % (BO / 8))]) = zr[q];
    }
  };
  floatx4 acc[KBW][NB];
  for (int i = 0; i < KBW; ++i)
    for (int j = 0; j < NB; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  // this lane's source position in its 16-lane group: pixel r of the
  // group's 8-pixel block (g = tile row), 4-column quad qd
  const int g = lane >> 4, r = (lane & 15) >> 2, qd = lane & 3;
  // per owned k block: the staged (dy, pixel shift, channel) of this lane's
  // quad (columns past 9 C read tap 8, channel 0: their outputs are dropped)
  int xoff[KBW];
  for (int i = 0; i < KBW; ++i) {
    int col = 16 * (w + 4 * i) + 4 * qd;
    if (col >= K) col = 8 * C;
    const int tap = col / C, c = col - tap * C;
    xoff[i] = ((g + tap / 3) * 10 + r + tap % 3) * XR + c;   // staged row g + dy + 1, px = r + dx + 1
  }
  const int nsteps = (int)(tend - tbeg);
  auto mul = [&](int buf) {
    const bfbits* xb = &Xs[buf][0][0][0];
    bf16x8 zf[NB];
    for (int j = 0; j < NB; ++j)
      zf[j] = tr_operand(&Zs[buf][8 * g + r][16 * j + 4 * qd], &Zs[buf][8 * g + 4 + r][16 * j + 4 * qd]);
    for (int i = 0; i < KBW; ++i) {
      const int kb = w + 4 * i;
      if (kb >= KB) break;
      const bf16x8 xf = tr_operand(xb + xoff[i], xb + xoff[i] + 4 * XR);
      for (int j = 0; j < NB; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(zf[j], xf, acc[i][j], 0, 0, 0);
    }
  };
  if constexpr (AHEAD2) {
    if (nsteps > 0) {
      gload(tbeg, xr0, zr0);
      gload(tbeg + 1, xr1, zr1);
      sstore(0, xr0, zr0);
    }
    __syncthreads();
    for (int st = 0; st < nsteps; st += 2) {
      // step st in buffer 0 (set 0 free, set 1 holds step st + 1)
      gload(tbeg + st + 2, xr0, zr0);
      mul(0);
      sstore(1, xr1, zr1);
      __syncthreads();
      if (st + 1 >= nsteps) break;
      // step st + 1 in buffer 1 (set 0 holds step st + 2)
      gload(tbeg + st + 3, xr1, zr1);
      mul(1);
      sstore(0, xr0, zr0);
      __syncthreads();
    }
  } else {
    if (nsteps > 0) {
      gload(tbeg, xr0, zr0);
      sstore(0, xr0, zr0);
    }
    __syncthreads();
    for (int st = 0; st < nsteps; ++st) {
      const int buf = st & 1;
      const bool more = st + 1 < nsteps;
      if (more) gload(tbeg + st + 1, xr0, zr0);
      mul(buf);
      if (more) sstore(buf ^ 1, xr0, zr0);
      __syncthreads();
    }
  }
  float* out = a.out + blockIdx.z * a.split_stride;
  const int cc = lane & 15, cr = (lane >> 4) * 4, KF = 9 * a.C;
  for (int i = 0; i < KBW; ++i) {
    const int k = 16 * (w + 4 * i) + cc;
    if (k >= K) continue;
    const int tap = k / C, kf = tap * a.C + c0 + (k - tap * C);
    for (int j = 0; j < NB; ++j)
      for (int q = 0; q < 4; ++q) {
        const int o = o0 + 16 * j + cr + q;
        if (o < a.Cout) out[(long long)o * KF + kf] = acc[i][j][q];
      }
  }
}

// The k-tiled weight gradient, for the shapes the whole-tile kernel does not
// take (widths not a multiple of 8, other channel counts; tile: 16 MO
// output channels x 128 k columns,
// four waves of 16 MO x 32) with the same transposing LDS reads: per 32-pixel
// step the tile's 16 column octets (tap, 8 channels) are staged pixel-major
// -- octet o holds the 32 pixels' 8 channels at its tap, one 16-byte copy
// per (pixel, octet), consecutive threads taking consecutive octets of one
// pixel -- and the dZ rows likewise; an operand lane group reads its 16
// columns as 4 quads of the two octets of its k block.
template <int MO>
__global__ __launch_bounds__(256) void k_conv3_dwk_tr(ConvDwArgs a) {
  constexpr int BO = 16 * MO;
  constexpr int XO = 32 * 8 + 32;            // one octet's 32 pixels (+ 64 B between octets)
  constexpr int ZR = BO + 8;
  constexpr int NXT = 16 * 32, NZT = 32 * (BO / 8);
  constexpr int NXL = NXT / 256, NZL = (NZT + 255) / 256;
  __shared__ __attribute__((aligned(16))) bfbits Xs[2][16][XO];
  __shared__ __attribute__((aligned(16))) bfbits Zs[2][32][ZR];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int k0 = blockIdx.x * DWK, o0 = blockIdx.y * BO;
  const int K = 9 * a.C;
  const long long pbeg = blockIdx.z * a.chunk;
  const long long pend = min(a.P, pbeg + a.chunk);
  // this thread's X octets (i % 16 for i = t + 256 q: the same for every q)
  // and their taps
  const int oc = t & 15;
  const int kx = k0 + 8 * oc;
  const bool kin = kx < K;
  const int tap = kin ? kx / a.C : 0;
  const int ci = kin ? kx - tap * a.C : 0;
  const int dy = tap / 3 - 1, dx = tap % 3 - 1;
  uint4 xr[NXL], zr[NZL];
  auto gload = [&](long long p0) {
#pragma unroll
    for (int q = 0; q < NXL; ++q) {
      const int px = (t + 256 * q) >> 4;
      const long long p = p0 + px;
      bool ok = kin && p < pend;
      const unsigned qq = (unsigned)(ok ? p : 0);
      const int x = (int)(qq % (unsigned)a.Wd) + dx;
      const unsigned rr = qq / (unsigned)a.Wd;
      const int y = (int)(rr % (unsigned)a.H) + dy;
      const long long n = rr / (unsigned)a.H;
      ok = ok && y >= 0 && y < a.H && x >= 0 && x < a.Wd;
      xr[q] = load16_or_zero(
          ok, a.X + ((n * a.H + (ok ? y : 0)) * a.Wd + (ok ? x : 0)) * (long long)a.C + ci);
    }
#pragma unroll
    for (int q = 0; q < NZL; ++q) {
      const int i = t + 256 * q;
      const int og = i % (BO / 8), pi = i / (BO / 8);
      const bool ok = i < NZT && p0 + pi < pend && o0 + 8 * og < a.Cout;
      zr[q] = load16_or_zero(ok, a.dZ + (ok ? (p0 + pi) * a.Cout + o0 + 8 * og : 0));
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < NXL; ++q)
      *reinterpret_cast<uint4*>(&Xs[buf][oc][((t + 256 * q) >> 4) * 8]) = xr[q];
#pragma unroll
    for (int q = 0; q < NZL; ++q) {
      const int i = t + 256 * q;
      if (i < NZT) *reinterpret_cast<uint4*>(&Zs[buf][i / (BO / 8)][8 * (i % (BO / 8))]) = zr[q];
    }
  };
  floatx4 acc[MO][2];
  for (int i = 0; i < MO; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int g = lane >> 4, r = (lane & 15) >> 2, qd = lane & 3;
  const int nsteps = (int)((pend - pbeg + DWP - 1) / DWP);
  if (nsteps > 0) {
    gload(pbeg);
    sstore(0);
  }
  __syncthreads();
  for (int st = 0; st < nsteps; ++st) {
    const int buf = st & 1;
    const bool more = st + 1 < nsteps;
    if (more) gload(pbeg + (long long)(st + 1) * DWP);
    bf16x8 za[MO], xb[2];
    for (int i = 0; i < MO; ++i)
      za[i] = tr_operand(&Zs[buf][8 * g + r][16 * i + 4 * qd], &Zs[buf][8 * g + 4 + r][16 * i + 4 * qd]);
    for (int j = 0; j < 2; ++j) {
      // k block 2 w + j: octets 4 w + 2 j and 4 w + 2 j + 1
      const bfbits* xo = &Xs[buf][4 * w + 2 * j + (qd >> 1)][0];
      xb[j] = tr_operand(xo + (8 * g + r) * 8 + 4 * (qd & 1), xo + (8 * g + 4 + r) * 8 + 4 * (qd & 1));
    }
    for (int i = 0; i < MO; ++i)
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(za[i], xb[j], acc[i][j], 0, 0, 0);
    if (more) sstore(buf ^ 1);
    __syncthreads();
  }
  float* out = a.out + blockIdx.z * a.split_stride;
  const int cc = lane & 15, cr = (lane >> 4) * 4;
  for (int j = 0; j < 2; ++j) {
    const int k = k0 + 32 * w + 16 * j + cc;
    if (k >= K) continue;
    for (int i = 0; i < MO; ++i)
      for (int q = 0; q < 4; ++q) {
        const int o = o0 + 16 * i + cr + q;
        if (o < a.Cout) out[(long long)o * K + k] = acc[i][j][q];
      }
  }
}

// Weights of the input gradient as a convolution of dY:
// Wf[c][tap' Cout + o] = Wt[o][(8 - tap') C + c]
__global__ void k_conv3_wflip(const bfbits* __restrict__ Wt, int Cout, int C,
                              bfbits* __restrict__ Wf) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long n = (long long)Cout * 9 * C;
  if (i >= n) return;
  const int o = (int)(i % Cout);
  const int tp = (int)((i / Cout) % 9);
  const int c = (int)(i / (9LL * Cout));
  Wf[i] = Wt[(long long)o * 9 * C + (8 - tp) * (long long)C + c];
}

}  // namespace

// output-channel tile of the convolution kernels: 16, 32, 48, 64 or 96
// channels (wider layers take several tiles of 96)
static int conv_tile(int Cout) {
  return Cout <= 16 ? 1 : Cout <= 32 ? 2 : Cout <= 48 ? 3 : Cout <= 64 ? 4 : 6;
}

extern "C" int bh_nn_conv3(const void* X, int64_t B, int H, int W, int C, const void* Wt,
                           int Cout, const float* bias, int relu, float alpha, void* Y,
                           int out_bf16, bh_stream s) {
  bh::clear_error();
  if (B <= 0 || H <= 0 || W <= 0 || C <= 0 || Cout <= 0)
    BH_FAIL(BH_ERR_ARG, "bh_nn_conv3: shape (%lld, %d, %d, %d) -> %d", (long long)B, H, W, C,
            Cout);
  if (!X || !Wt || !Y) BH_FAIL(BH_ERR_ARG, "bh_nn_conv3: null pointer");
  if (C % 8 || Cout % 8 || !aligned16(X) || !aligned16(Wt))
    BH_FAIL(BH_ERR_ARG, "bh_nn_conv3: channels (%d, %d) must be multiples of 8 and the "
                        "operands 16-byte aligned", C, Cout);
  ConvArgs a;
  a.X = (const bfbits*)X;
  a.Wt = (const bfbits*)Wt;
  a.bias = bias;
  a.Y = Y;
  a.P = B * (long long)H * W;
  a.H = H;
  a.Wd = W;
  a.C = C;
  a.Cout = Cout;
  a.alpha = alpha;
  a.relu = relu ? 1 : 0;
  a.out_bf16 = out_bf16 ? 1 : 0;
  hipStream_t st = (hipStream_t)s;
  const int nb = conv_tile(Cout);
  const unsigned gy = (unsigned)((Cout + 16 * nb - 1) / (16 * nb));
  // inputs of 8, 24 or 48 channels into up to 64 outputs, or 96 channels
  // into up to 48: halo tiles of 8 x 32 pixels, or 4 x 32 where 8-row tiles
  // would give fewer than 2048 workgroups.  Wider outputs keep the gathered
  // form (their accumulators leave the halo kernel two waves per SIMD;
  // measured on the U-Net's layers at batch 16)
  const int c8 = C / 8;
  if (((c8 == 1 || c8 == 3 || c8 == 6) && nb <= 4) || (c8 == 12 && nb <= 3)) {
    const long long tiles_x = (W + HTW - 1) / HTW;
    const int th = ((H + 7) / 8) * tiles_x * B * gy >= 2048 ? 8 : 4;
    const long long tiles_y = (H + th - 1) / th;
    const long long nt = tiles_x * tiles_y * B;
    if (nt > 0x7fffffffLL) BH_FAIL(BH_ERR_ARG, "bh_nn_conv3: %lld tiles", nt);
    const dim3 grid((unsigned)nt, gy);
    const int tx = (int)tiles_x, ty = (int)tiles_y;
#define BH_HALO(C8V, THV)                                                                      \
  do {                                                                                        \
    if (nb == 1) hipLaunchKernelGGL((k_conv3_halo<C8V, THV, 1>), grid, dim3(256), 0, st, a, tx, ty); \
    else if (nb == 2) hipLaunchKernelGGL((k_conv3_halo<C8V, THV, 2>), grid, dim3(256), 0, st, a, tx, ty); \
    else if (nb == 3) hipLaunchKernelGGL((k_conv3_halo<C8V, THV, 3>), grid, dim3(256), 0, st, a, tx, ty); \
    else if (nb == 4) hipLaunchKernelGGL((k_conv3_halo<C8V, THV, 4>), grid, dim3(256), 0, st, a, tx, ty); \
    else hipLaunchKernelGGL((k_conv3_halo<C8V, THV, 6>), grid, dim3(256), 0, st, a, tx, ty); \
  } while (0)
    if (th == 8) {
      if (c8 == 1) BH_HALO(1, 8);
      else if (c8 == 3) BH_HALO(3, 8);
      else if (c8 == 6) BH_HALO(6, 8);
      else BH_HALO(12, 8);
    } else {
      if (c8 == 1) BH_HALO(1, 4);
      else if (c8 == 3) BH_HALO(3, 4);
      else if (c8 == 6) BH_HALO(6, 4);
      else BH_HALO(12, 4);
    }
#undef BH_HALO
    BH_LAUNCH_CHECK();
    return BH_OK;
  }
  const unsigned gx = (unsigned)((a.P + CM - 1) / CM);
  const dim3 grid(gx, gy);
  // fewer than four workgroups per CU: two k groups per workgroup
  if ((long long)gx * gy < 1024) {
    if (nb == 1) hipLaunchKernelGGL((k_conv3_igemm<1, 2>), grid, dim3(512), 0, st, a);
    else if (nb == 2) hipLaunchKernelGGL((k_conv3_igemm<2, 2>), grid, dim3(512), 0, st, a);
    else if (nb == 3) hipLaunchKernelGGL((k_conv3_igemm<3, 2>), grid, dim3(512), 0, st, a);
    else if (nb == 4) hipLaunchKernelGGL((k_conv3_igemm<4, 2>), grid, dim3(512), 0, st, a);
    else hipLaunchKernelGGL((k_conv3_igemm<6, 2>), grid, dim3(512), 0, st, a);
  } else {
    if (nb == 1) hipLaunchKernelGGL((k_conv3_igemm<1, 1>), grid, dim3(256), 0, st, a);
    else if (nb == 2) hipLaunchKernelGGL((k_conv3_igemm<2, 1>), grid, dim3(256), 0, st, a);
    else if (nb == 3) hipLaunchKernelGGL((k_conv3_igemm<3, 1>), grid, dim3(256), 0, st, a);
    else if (nb == 4) hipLaunchKernelGGL((k_conv3_igemm<4, 1>), grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL((k_conv3_igemm<6, 1>), grid, dim3(256), 0, st, a);
  }
  BH_LAUNCH_CHECK();
  return BH_OK;
}

// the whole-tile weight gradient: 8 input channels (one slice: [256, 8, 24]
// at batch 16 55 -> 40 us against the k-tiled kernel) or a multiple of 24
// (slices of 24), widths a multiple of 8; output tiles of up to 64 channels
// (48 where 64 do not divide Cout: 96 outputs as 2 x 48, not 64 + 32).
// Slices of 24 also where 48 divide C: the 48-channel kernel's registers
// (160-208 against 92-112) cost more occupancy than the second slice's dZ
// re-read costs bandwidth.  Batch 16, slices of 48 -> 24: [128, 96, 48] 70
// -> 67 us, [32, 192, 192] 64 -> 44, [256, 48, 24] 95 -> 81 (1024
// workgroups, below).
static bool conv_dw_whole(int W, int C, int Cout) {
  (void)Cout;
  return (C == 8 || C % 24 == 0) && W % 8 == 0;
}
static int conv_dw_slice(int C) { return C == 8 ? 8 : 24; }
static int conv_dw_otile(int Cout) {
  return Cout <= 64 ? conv_tile(Cout) : Cout % 64 == 0 ? 4 : 3;
}

// pixel split of the weight gradient: about 1024 workgroups (the kernel is
// bound by the latency of its staging loads, so it needs several resident
// workgroups per CU; 512 or 1024 of the whole-tile kernel, which does more
// per step), chunks of at least 512 pixels (shorter chunks spend more on their
// partial tiles than they gain in parallelism), a multiple of 32 (of the
// whole-tile kernel: whole 4x8 tiles, *ntiles of them in all).  Measured
// on the U-Net's twelve layer shapes at batch 16 against 2048 workgroups
// and 256-pixel chunks: 1150 -> 1040 us of weight gradients in all.
static void conv_dw_split(long long B, int H, int W, int C, int Cout, long long* chunk,
                          int* nsplit, long long* ntiles) {
  if (conv_dw_whole(W, C, Cout)) {
    const long long nt = B * ((H + 3) / 4) * (W / 8);
    const int nb = conv_dw_otile(Cout);
    const long long blocks = (long long)(C / conv_dw_slice(C)) * ((Cout + 16 * nb - 1) / (16 * nb));
    // 1024 workgroups for one or two dW tiles per split (<= 112 registers: 4
    // resident per CU hide the staging loads better; [256, 24, 24] 57 -> 49
    // us, [256, 8, 24] 48 -> 40), else
    // 512 (more splits of the wider kernels add more partial-sum traffic
    // than they hide: [128, 96, 48] 68 -> 76 us at 1024; fewer hide less
    // latency: 80 us at 384, 92 at 256)
    const long long target = blocks <= 2 ? 1024 : 512;
    const long long ns = std::max(1LL, std::min((target + blocks - 1) / blocks, (nt + 15) / 16));
    const long long ch = (nt + ns - 1) / ns;
    *chunk = ch;
    *nsplit = (int)((nt + ch - 1) / ch);
    *ntiles = nt;
    return;
  }
  const long long P = B * (long long)H * W;
  const long long tiles = (long long)((9 * C + DWK - 1) / DWK) *
                          ((Cout + 16 * conv_tile(Cout) - 1) / (16 * conv_tile(Cout)));
  long long ns = (1024 + tiles - 1) / tiles;
  ns = std::max(1LL, std::min(ns, (P + 511) / 512));
  long long ch = (P + ns - 1) / ns;
  ch = (ch + DWP - 1) / DWP * DWP;
  *chunk = ch;
  *nsplit = (int)((P + ch - 1) / ch);
  *ntiles = 0;
}

extern "C" size_t bh_nn_conv3_dw_workspace_size(int64_t B, int H, int W, int C, int Cout) {
  if (B <= 0 || H <= 0 || W <= 0 || C <= 0 || Cout <= 0) return 0;
  long long chunk, nt;
  int ns;
  conv_dw_split(B, H, W, C, Cout, &chunk, &ns, &nt);
  return ns > 1 ? (size_t)ns * Cout * 9 * C * sizeof(float) : 0;
}

extern "C" int bh_nn_conv3_dw(const void* X, const void* dZ, int64_t B, int H, int W, int C,
                              int Cout, float* dW, void* ws, size_t ws_bytes, bh_stream s) {
  bh::clear_error();
  if (B <= 0 || H <= 0 || W <= 0 || C <= 0 || Cout <= 0)
    BH_FAIL(BH_ERR_ARG, "bh_nn_conv3_dw: shape (%lld, %d, %d, %d) -> %d", (long long)B, H, W, C,
            Cout);
  if (!X || !dZ || !dW) BH_FAIL(BH_ERR_ARG, "bh_nn_conv3_dw: null pointer");
  if (C % 8 || Cout % 8 || !aligned16(X) || !aligned16(dZ))
    BH_FAIL(BH_ERR_ARG, "bh_nn_conv3_dw: channels (%d, %d) must be multiples of 8 and the "
                        "operands 16-byte aligned", C, Cout);
  const long long P = B * (long long)H * W;
  if (P >= (1LL << 32)) BH_FAIL(BH_ERR_ARG, "bh_nn_conv3_dw: %lld pixels (limit 2^32)", P);
  ConvDwArgs a;
  a.X = (const bfbits*)X;
  a.dZ = (const bfbits*)dZ;
  a.P = P;
  a.H = H;
  a.Wd = W;
  a.C = C;
  a.Cout = Cout;
  int ns;
  conv_dw_split(B, H, W, C, Cout, &a.chunk, &ns, &a.ntiles);
  const long long n = (long long)Cout * 9 * C;
  a.split_stride = n;
  if (ns > 1) {
    if (!ws || ws_bytes < (size_t)ns * n * sizeof(float))
      BH_FAIL(BH_ERR_ARG, "bh_nn_conv3_dw: workspace of %zu bytes, %zu needed", ws_bytes,
              (size_t)ns * n * sizeof(float));
    a.out = (float*)ws;
  } else {
    a.out = dW;
  }
  hipStream_t st = (hipStream_t)s;
  const int mo = conv_tile(Cout);
  if (conv_dw_whole(W, C, Cout)) {
    const int cs = conv_dw_slice(C), nb = conv_dw_otile(Cout);
    const dim3 g1((unsigned)(C / cs), (unsigned)((Cout + 16 * nb - 1) / (16 * nb)), (unsigned)ns);
#define BH_DWF(C8V)                                                                            \
  do {                                                                                        \
    if (nb == 1) hipLaunchKernelGGL((k_conv3_dw_tr<C8V, 1>), g1, dim3(256), 0, st, a);         \
    else if (nb == 2) hipLaunchKernelGGL((k_conv3_dw_tr<C8V, 2>), g1, dim3(256), 0, st, a);    \
    else if (nb == 3) hipLaunchKernelGGL((k_conv3_dw_tr<C8V, 3>), g1, dim3(256), 0, st, a);    \
    else hipLaunchKernelGGL((k_conv3_dw_tr<C8V, 4>), g1, dim3(256), 0, st, a);                 \
  } while (0)
    if (cs == 8) BH_DWF(1);
    else BH_DWF(3);
#undef BH_DWF
  } else {
  const dim3 grid((unsigned)((9 * C + DWK - 1) / DWK), (unsigned)((Cout + 16 * mo - 1) / (16 * mo)),
                  (unsigned)ns);
  if (mo == 1) hipLaunchKernelGGL(k_conv3_dwk_tr<1>, grid, dim3(256), 0, st, a);
  else if (mo == 2) hipLaunchKernelGGL(k_conv3_dwk_tr<2>, grid, dim3(256), 0, st, a);
  else if (mo == 3) hipLaunchKernelGGL(k_conv3_dwk_tr<3>, grid, dim3(256), 0, st, a);
  else if (mo == 4) hipLaunchKernelGGL(k_conv3_dwk_tr<4>, grid, dim3(256), 0, st, a);
  else hipLaunchKernelGGL(k_conv3_dwk_tr<6>, grid, dim3(256), 0, st, a);
  }
  BH_LAUNCH_CHECK();
  if (ns > 1) {
    hipLaunchKernelGGL(k_sum_splits, dim3((unsigned)((n + 15) / 16)), dim3(256), 0, st,
                       (const float*)ws, ns, n, dW);
    BH_LAUNCH_CHECK();
  }
  return BH_OK;
}

extern "C" int bh_nn_conv3_wflip(const void* Wt, int Cout, int C, void* Wf, bh_stream s) {
  bh::clear_error();
  if (Cout <= 0 || C <= 0 || !Wt || !Wf) BH_FAIL(BH_ERR_ARG, "bh_nn_conv3_wflip: arguments");
  const long long n = (long long)Cout * 9 * C;
  hipLaunchKernelGGL(k_conv3_wflip, dim3(bh::grid1d(n, 256)), dim3(256), 0, (hipStream_t)s,
                     (const bfbits*)Wt, Cout, C, (bfbits*)Wf);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

extern "C" size_t bh_nn_gemm_workspace_size(int64_t M, int64_t N, int64_t K, int nsplit) {
  if (nsplit <= 1 || M <= 0 || N <= 0 || K <= 0) return 0;
  return (size_t)nsplit * (size_t)M * (size_t)N * sizeof(float);
}

extern "C" int bh_nn_gemm(const void* A, int ta, int64_t lda, const void* B, int tb, int64_t ldb,
                          int64_t M, int64_t N, int64_t K, float alpha, const float* bias,
                          int relu, void* C, int out_bf16, int64_t ldc, int nsplit, void* ws,
                          size_t ws_bytes, bh_stream s) {
  bh::clear_error();
  if (M <= 0 || N <= 0) return BH_OK;
  if (!A || !B || !C) BH_FAIL(BH_ERR_ARG, "bh_nn_gemm: null pointer");
  if (K <= 0) BH_FAIL(BH_ERR_ARG, "bh_nn_gemm: K=%lld", (long long)K);
  const int64_t ca = ta ? M : K, cb = tb ? K : N;   // contiguous extents
  if (ca % 8 || cb % 8 || lda % 8 || ldb % 8 || !aligned16(A) || !aligned16(B))
    BH_FAIL(BH_ERR_ARG, "bh_nn_gemm: contiguous extents and leading dimensions must be "
                        "multiples of 8 and the operands 16-byte aligned");
  if (lda < ca || ldb < cb || ldc < N) BH_FAIL(BH_ERR_ARG, "bh_nn_gemm: leading dimension");
  if (nsplit < 1) nsplit = 1;
  if (nsplit > 1 && (out_bf16 || bias || relu))
    BH_FAIL(BH_ERR_ARG, "bh_nn_gemm: split-K gives FP32 output without epilogue");
  long long kchunk = ((K + nsplit - 1) / nsplit + BK - 1) / BK * BK;
  nsplit = (int)((K + kchunk - 1) / kchunk);
  GemmArgs g;
  g.A = (const bfbits*)A;
  g.B = (const bfbits*)B;
  g.M = M;
  g.N = N;
  g.K = K;
  g.lda = lda;
  g.ldb = ldb;
  g.ldc = ldc;
  g.ta = ta ? 1 : 0;
  g.tb = tb ? 1 : 0;
  g.out_bf16 = out_bf16 ? 1 : 0;
  g.bias = bias;
  g.relu = relu ? 1 : 0;
  g.alpha = alpha;
  g.kchunk = kchunk;
  hipStream_t st = (hipStream_t)s;
  const dim3 grid((unsigned)((N + BN - 1) / BN), (unsigned)((M + BM - 1) / BM), (unsigned)nsplit);
  if (nsplit == 1) {
    g.C = C;
    g.split_stride = 0;
    hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, st, g);
    BH_LAUNCH_CHECK();
    return BH_OK;
  }
  if (ldc != N) BH_FAIL(BH_ERR_ARG, "bh_nn_gemm: split-K needs ldc == N");
  const size_t need = (size_t)nsplit * (size_t)M * (size_t)N * sizeof(float);
  if (!ws || ws_bytes < need)
    BH_FAIL(BH_ERR_ARG, "bh_nn_gemm: workspace of %zu bytes needed", need);
  g.C = ws;
  g.split_stride = M * N;
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, st, g);
  BH_LAUNCH_CHECK();
  const long long n = M * N;
  hipLaunchKernelGGL(k_sum_partials, dim3(bh::grid1d(n, 256)), dim3(256), 0, st,
                     (const float*)ws, nsplit, n, (float*)C);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

static int check_img(const char* name, int64_t B, int H, int W, int C) {
  if (B <= 0 || H <= 0 || W <= 0 || C <= 0)
    BH_FAIL(BH_ERR_ARG, "%s: shape (%lld, %d, %d, %d)", name, (long long)B, H, W, C);
  return BH_OK;
}

extern "C" int bh_nn_pool2(const void* X, int64_t B, int H, int W, int C, void* Y, int backward,
                           bh_stream s) {
  bh::clear_error();
  if (int rc = check_img("bh_nn_pool2", B, H, W, C)) return rc;
  if (!X || !Y) BH_FAIL(BH_ERR_ARG, "bh_nn_pool2: null pointer");
  if (H % 2 || W % 2) BH_FAIL(BH_ERR_ARG, "bh_nn_pool2: H and W must be even");
  hipStream_t st = (hipStream_t)s;
  if (C % 8 == 0 && aligned16(X) && aligned16(Y)) {
    const long long n8 = B * (long long)(H / 2) * (W / 2) * (C / 8);
    if (!backward)
      hipLaunchKernelGGL(k_gather2_v8, dim3(bh::grid1d(n8, 256)), dim3(256), 0, st,
                         (const uint4*)X, n8, W / 2, C / 8, 0.25f, (uint4*)Y);
    else
      hipLaunchKernelGGL(k_spread2_v8, dim3(bh::grid1d(n8, 256)), dim3(256), 0, st,
                         (const uint4*)X, n8, W / 2, C / 8, 1, (uint4*)Y);
  } else if (!backward) {
    const long long n = B * (long long)(H / 2) * (W / 2) * C;
    hipLaunchKernelGGL(k_pool2, dim3(bh::grid1d(n, 256)), dim3(256), 0, st, (const bfbits*)X,
                       (long long)B, H, W, C, (bfbits*)Y);
  } else {
    // X = dY (B, H/2, W/2, C), Y = dX (B, H, W, C)
    const long long n = B * (long long)H * W * C;
    hipLaunchKernelGGL(k_pool2_bwd, dim3(bh::grid1d(n, 256)), dim3(256), 0, st,
                       (const bfbits*)X, (long long)B, H, W, C, (bfbits*)Y);
  }
  BH_LAUNCH_CHECK();
  return BH_OK;
}

extern "C" int bh_nn_up2(const void* X, int64_t B, int H, int W, int C, void* Y, int backward,
                         bh_stream s) {
  bh::clear_error();
  if (int rc = check_img("bh_nn_up2", B, H, W, C)) return rc;
  if (!X || !Y) BH_FAIL(BH_ERR_ARG, "bh_nn_up2: null pointer");
  hipStream_t st = (hipStream_t)s;
  if (C % 8 == 0 && aligned16(X) && aligned16(Y)) {
    const long long n8 = B * (long long)H * W * (C / 8);
    if (!backward)
      hipLaunchKernelGGL(k_spread2_v8, dim3(bh::grid1d(n8, 256)), dim3(256), 0, st,
                         (const uint4*)X, n8, W, C / 8, 0, (uint4*)Y);
    else
      hipLaunchKernelGGL(k_gather2_v8, dim3(bh::grid1d(n8, 256)), dim3(256), 0, st,
                         (const uint4*)X, n8, W, C / 8, 1.0f, (uint4*)Y);
  } else if (!backward) {
    // X (B, H, W, C) -> Y (B, 2H, 2W, C)
    const long long n = B * 4LL * H * W * C;
    hipLaunchKernelGGL(k_up2, dim3(bh::grid1d(n, 256)), dim3(256), 0, st, (const bfbits*)X,
                       (long long)B, H, W, C, (bfbits*)Y);
  } else {
    // X = dY (B, 2H, 2W, C) -> Y = dX (B, H, W, C)
    const long long n = B * (long long)H * W * C;
    hipLaunchKernelGGL(k_up2_bwd, dim3(bh::grid1d(n, 256)), dim3(256), 0, st, (const bfbits*)X,
                       (long long)B, H, W, C, (bfbits*)Y);
  }
  BH_LAUNCH_CHECK();
  return BH_OK;
}

extern "C" int bh_nn_add(const void* X, const void* Y, int64_t n, void* Z, bh_stream s) {
  bh::clear_error();
  if (n <= 0) return BH_OK;
  if (!X || !Y || !Z) BH_FAIL(BH_ERR_ARG, "bh_nn_add: null pointer");
  if (n % 8 == 0 && aligned16(X) && aligned16(Y) && aligned16(Z))
    hipLaunchKernelGGL(k_add_bf16_v8, dim3(bh::grid1d(n / 8, 256)), dim3(256), 0, (hipStream_t)s,
                       (const uint4*)X, (const uint4*)Y, (long long)n / 8, (uint4*)Z);
  else
    hipLaunchKernelGGL(k_add_bf16, dim3(bh::grid1d(n, 256)), dim3(256), 0, (hipStream_t)s,
                       (const bfbits*)X, (const bfbits*)Y, (long long)n, (bfbits*)Z);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

extern "C" int bh_nn_relu_bwd(const void* dY, const void* Yact, int64_t n, void* dZ, bh_stream s) {
  bh::clear_error();
  if (n <= 0) return BH_OK;
  if (!dY || !Yact || !dZ) BH_FAIL(BH_ERR_ARG, "bh_nn_relu_bwd: null pointer");
  if (n % 8 == 0 && aligned16(dY) && aligned16(Yact) && aligned16(dZ))
    hipLaunchKernelGGL(k_relu_bwd_v8, dim3(bh::grid1d(n / 8, 256)), dim3(256), 0, (hipStream_t)s,
                       (const uint4*)dY, (const uint4*)Yact, (long long)n / 8, (uint4*)dZ);
  else
    hipLaunchKernelGGL(k_relu_bwd, dim3(bh::grid1d(n, 256)), dim3(256), 0, (hipStream_t)s,
                       (const bfbits*)dY, (const bfbits*)Yact, (long long)n, (bfbits*)dZ);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

extern "C" size_t bh_nn_colsum_workspace_size(int64_t P, int C) {
  if (P <= 0 || C <= 0) return 0;
  const long long ch = colsum_chunk(P);
  return (size_t)((P + ch - 1) / ch) * (size_t)C * sizeof(float);
}

extern "C" int bh_nn_colsum(const void* X, int64_t P, int C, float* out, void* ws, size_t ws_bytes,
                            bh_stream s) {
  bh::clear_error();
  if (C <= 0) return BH_OK;
  if (!X || !out || P < 0) BH_FAIL(BH_ERR_ARG, "bh_nn_colsum: bad arguments");
  hipStream_t st = (hipStream_t)s;
  const long long ch = colsum_chunk(P);
  const long long nb = (P + ch - 1) / ch;
  const size_t need = (size_t)nb * (size_t)C * sizeof(float);
  if (nb > 0 && (!ws || ws_bytes < need))
    BH_FAIL(BH_ERR_ARG, "bh_nn_colsum: workspace of %zu bytes needed", need);
  if (nb > 0 && C % 8 == 0 && C <= 2048 && aligned16(X))
    hipLaunchKernelGGL(k_colsum_part_v8, dim3((unsigned)nb), dim3(256), 0, st, (const uint4*)X,
                       (long long)P, C / 8, ch, (float*)ws);
  else if (nb > 0)
    hipLaunchKernelGGL(k_colsum_part, dim3((unsigned)nb), dim3(256), 0, st, (const bfbits*)X,
                       (long long)P, C, ch, (float*)ws);
  hipLaunchKernelGGL(k_colsum_final, dim3((unsigned)C), dim3(256), 0, st, (const float*)ws, nb, C,
                     out);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

extern "C" size_t bh_nn_mse_workspace_size(int64_t P) {
  return P <= 0 ? 0 : (size_t)((P + 255) / 256) * sizeof(float);
}

extern "C" int bh_nn_mse(const float* pred, int pred_stride, const float* target, int64_t P,
                         void* grad, int grad_stride, float* part, size_t part_bytes,
                         bh_stream s) {
  bh::clear_error();
  if (P <= 0) return BH_OK;
  if (!pred || !target || !grad || !part || pred_stride < 1 || grad_stride < 1)
    BH_FAIL(BH_ERR_ARG, "bh_nn_mse: bad arguments");
  const long long nb = (P + 255) / 256;
  if (part_bytes < (size_t)nb * sizeof(float))
    BH_FAIL(BH_ERR_ARG, "bh_nn_mse: %lld partials need %zu bytes", nb, (size_t)nb * sizeof(float));
  hipLaunchKernelGGL(k_mse, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)s, pred, pred_stride,
                     target, (long long)P, (bfbits*)grad, grad_stride, part);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

extern "C" int bh_nn_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr,
                          float beta1, float beta2, float eps, int64_t step, void* w_bf16,
                          bh_stream s) {
  bh::clear_error();
  if (n <= 0) return BH_OK;
  if (!w || !g || !m || !v || step < 1) BH_FAIL(BH_ERR_ARG, "bh_nn_adam: bad arguments");
  const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(bh::grid1d(n, 256)), dim3(256), 0, (hipStream_t)s, w, g, m, v,
                     (long long)n, lr, beta1, beta2, eps, c1, c2, (bfbits*)w_bf16);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

extern "C" int bh_nn_cast(const void* x, int from_bf16, int64_t n, float scale, void* y,
                          bh_stream s) {
  bh::clear_error();
  if (n <= 0) return BH_OK;
  if (!x || !y) BH_FAIL(BH_ERR_ARG, "bh_nn_cast: null pointer");
  hipStream_t st = (hipStream_t)s;
  if (from_bf16)
    hipLaunchKernelGGL(k_bf16_to_f32, dim3(bh::grid1d(n, 256)), dim3(256), 0, st,
                       (const bfbits*)x, (long long)n, scale, (float*)y);
  else
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(bh::grid1d(n, 256)), dim3(256), 0, st,
                       (const float*)x, (long long)n, scale, (bfbits*)y);
  BH_LAUNCH_CHECK();
  return BH_OK;
}
