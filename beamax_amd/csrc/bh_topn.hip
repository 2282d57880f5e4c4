// top_n selection: the k largest |c| under the order (|c|, index), ascending.
//
// 1. key[i] = bit pattern of |c_i| = sqrt(re*re + im*im) (IEEE, each op
//    correctly rounded; for non-negative doubles the bit pattern orders like
//    the value).
// 2. MSB-first radix select, 8 bits per pass: a grid-wide LDS histogram of the
//    keys that match the prefix fixed so far, then one small block picks the
//    digit.  After 8 passes the k-th largest key T is known, together with how
//    many of the entries equal to T are needed (the largest indices win).
// 3. Stable compaction of the winners (per-block counts, one scan block,
//    per-block ordered writes).
// 4. Sort of the winners on (key, index): bitonic within 512-element LDS
//    tiles, then stable merge levels (one thread per element, its output slot
//    from a binary search in the partner run).
// 5. Gather of indices and values.
#include <mutex>

#include "bh_internal.hpp"

namespace {

constexpr int kSelBlock = 256;
constexpr int kItems = 16;                       // keys per thread per block pass
constexpr int kChunk = kSelBlock * kItems;       // keys per block
constexpr int kSortTile = 512;

struct SelState {
  unsigned long long prefix;   // fixed high bits of T
  unsigned long long mask;     // which bits are fixed
  long long need;              // entries still to take at or below the prefix
  long long eq_total;          // number of keys == T (after the last pass)
};

__global__ void k_keys(const double2* __restrict__ c, unsigned long long* __restrict__ key,
                       long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 v = c[i];
  const double m = __dsqrt_rn(__dadd_rn(__dmul_rn(v.x, v.x), __dmul_rn(v.y, v.y)));
  key[i] = (unsigned long long)__double_as_longlong(m);
}

__global__ void k_init_state(SelState* st, unsigned int* hist, long long k) {
  if (threadIdx.x == 0) {
    st->prefix = 0ull;
    st->mask = 0ull;
    st->need = k;
    st->eq_total = 0;
  }
  for (int i = threadIdx.x; i < 256; i += blockDim.x) hist[i] = 0u;
}

__global__ __launch_bounds__(kSelBlock) void k_hist(const unsigned long long* __restrict__ key,
                                                    long long n, const SelState* st, int shift,
                                                    unsigned int* __restrict__ hist) {
  __shared__ unsigned int h[256];
  for (int i = threadIdx.x; i < 256; i += blockDim.x) h[i] = 0u;
  __syncthreads();
  const unsigned long long prefix = st->prefix, mask = st->mask;
  const long long base = (long long)blockIdx.x * kChunk;
  for (int it = 0; it < kItems; ++it) {
    const long long i = base + (long long)it * kSelBlock + threadIdx.x;
    if (i < n) {
      const unsigned long long kv = key[i];
      if ((kv & mask) == prefix) atomicAdd(&h[(kv >> shift) & 0xffu], 1u);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 256; i += blockDim.x)
    if (h[i]) atomicAdd(&hist[i], h[i]);
}

// one block of 256: choose the digit of the k-th largest key (the largest
// digit whose suffix count reaches `need`, from a parallel suffix scan),
// clear the histogram
__global__ __launch_bounds__(256) void k_pick(SelState* st, unsigned int* hist, int shift) {
  __shared__ long long sfx[256];
  const int t = threadIdx.x;
  const long long hv = hist[t];
  const long long need = st->need;
  sfx[t] = hv;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const long long v = t + off < 256 ? sfx[t + off] : 0;
    __syncthreads();
    sfx[t] += v;
    __syncthreads();
  }
  // sfx is non-increasing in the digit: exactly one thread holds the digit
  const long long above = sfx[t] - hv;   // keys with a larger digit
  if (sfx[t] >= need && above < need) {
    st->prefix |= ((unsigned long long)t) << shift;
    st->mask |= 0xffull << shift;
    st->need = need - above;
    st->eq_total = hv;
  }
  hist[t] = 0u;
}

// per-block counts of keys > T and keys == T
__global__ __launch_bounds__(kSelBlock) void k_count(const unsigned long long* __restrict__ key,
                                                     long long n, const SelState* st,
                                                     long long* __restrict__ cnt) {
  __shared__ int sgt, seq;
  if (threadIdx.x == 0) { sgt = 0; seq = 0; }
  __syncthreads();
  const unsigned long long T = st->prefix;
  int gt = 0, eq = 0;
  const long long base = (long long)blockIdx.x * kChunk;
  for (int it = 0; it < kItems; ++it) {
    const long long i = base + (long long)it * kSelBlock + threadIdx.x;
    if (i < n) {
      const unsigned long long kv = key[i];
      gt += kv > T;
      eq += kv == T;
    }
  }
  atomicAdd(&sgt, gt);
  atomicAdd(&seq, eq);
  __syncthreads();
  if (threadIdx.x == 0) {
    cnt[2 * blockIdx.x] = sgt;
    cnt[2 * blockIdx.x + 1] = seq;
  }
}

// one block of 1024 threads: exclusive scans of the per-block counts (gt and
// eq) in place; thread t owns a contiguous slice of the blocks, the slice
// sums are scanned in LDS
__global__ __launch_bounds__(1024) void k_scan_counts(long long* cnt, long long nblk) {
  __shared__ long long sg[1024], se[1024];
  const int t = threadIdx.x;
  const long long per = (nblk + 1023) / 1024;
  const long long b0 = t * per;
  const long long b1 = b0 + per < nblk ? b0 + per : nblk;
  long long g = 0, e = 0;
  for (long long b = b0; b < b1; ++b) {
    g += cnt[2 * b];
    e += cnt[2 * b + 1];
  }
  sg[t] = g;
  se[t] = e;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const long long ag = t >= off ? sg[t - off] : 0, ae = t >= off ? se[t - off] : 0;
    __syncthreads();
    sg[t] += ag;
    se[t] += ae;
    __syncthreads();
  }
  long long rg = sg[t] - g, re = se[t] - e;
  for (long long b = b0; b < b1; ++b) {
    const long long gg = cnt[2 * b], ee = cnt[2 * b + 1];
    cnt[2 * b] = rg;
    cnt[2 * b + 1] = re;
    rg += gg;
    re += ee;
  }
}

// ordered compaction: winners are keys > T and the last `need` keys == T in
// index order.  Output slot = (#winners before i).
__global__ __launch_bounds__(kSelBlock) void k_compact(const unsigned long long* __restrict__ key,
                                                       long long n, const SelState* st,
                                                       const long long* __restrict__ cnt,
                                                       unsigned long long* __restrict__ okey,
                                                       long long* __restrict__ oidx) {
  __shared__ int wgt[kSelBlock / 64], weq[kSelBlock / 64];
  const unsigned long long T = st->prefix;
  const long long skip_eq = st->eq_total - st->need;   // equal keys not taken (lowest indices)
  long long gt_base = cnt[2 * blockIdx.x];
  long long eq_base = cnt[2 * blockIdx.x + 1];
  const long long base = (long long)blockIdx.x * kChunk;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int it = 0; it < kItems; ++it) {
    const long long i = base + (long long)it * kSelBlock + threadIdx.x;
    unsigned long long kv = 0;
    bool g = false, e = false;
    if (i < n) {
      kv = key[i];
      g = kv > T;
      e = kv == T;
    }
    const unsigned long long bg = __ballot(g), be = __ballot(e);
    const unsigned long long lower = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int rg = __popcll(bg & lower), re = __popcll(be & lower);
    if (lane == 0) { wgt[wave] = __popcll(bg); weq[wave] = __popcll(be); }
    __syncthreads();
    int og = 0, oe = 0, tg = 0, te = 0;
    for (int w = 0; w < kSelBlock / 64; ++w) {
      if (w < wave) { og += wgt[w]; oe += weq[w]; }
      tg += wgt[w];
      te += weq[w];
    }
    const long long my_gt = gt_base + og + rg;   // rank among keys > T
    const long long my_eq = eq_base + oe + re;   // rank among keys == T
    if (g) {
      const long long taken_eq_before = my_eq - skip_eq > 0 ? my_eq - skip_eq : 0;
      const long long slot = my_gt + taken_eq_before;
      okey[slot] = kv;
      oidx[slot] = i;
    } else if (e && my_eq >= skip_eq) {
      const long long slot = my_gt + (my_eq - skip_eq);
      okey[slot] = kv;
      oidx[slot] = i;
    }
    gt_base += tg;
    eq_base += te;
    __syncthreads();
  }
}

__device__ inline bool pair_less(unsigned long long ka, long long ia, unsigned long long kb,
                                 long long ib) {
  return ka < kb || (ka == kb && ia < ib);
}

__global__ void k_pad(unsigned long long* key, long long* idx, long long k, long long P) {
  const long long i = k + blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < P) {
    key[i] = ~0ull;
    idx[i] = 0x7fffffffffffffffll;
  }
}

// Bitonic sort of each 512-element tile in LDS (ascending by pair_less).
__global__ __launch_bounds__(256) void k_bitonic_tile(unsigned long long* key, long long* idx,
                                                      long long P) {
  __shared__ unsigned long long sk[kSortTile];
  __shared__ long long si[kSortTile];
  const long long base = (long long)blockIdx.x * kSortTile;
  const int tile = (int)((P < kSortTile) ? P : kSortTile);
  for (int i = threadIdx.x; i < tile; i += blockDim.x) {
    sk[i] = key[base + i];
    si[i] = idx[base + i];
  }
  __syncthreads();
  for (int kk = 2; kk <= tile; kk <<= 1) {
    for (int j = kk >> 1; j > 0; j >>= 1) {
      for (int t = threadIdx.x; t < tile / 2; t += blockDim.x) {
        const int lo = ((t / j) * 2 * j) + (t % j);
        const int hi = lo + j;
        const bool up = (lo & kk) == 0;
        const bool sw = up ? pair_less(sk[hi], si[hi], sk[lo], si[lo])
                           : pair_less(sk[lo], si[lo], sk[hi], si[hi]);
        if (sw) {
          unsigned long long tk = sk[lo]; sk[lo] = sk[hi]; sk[hi] = tk;
          long long ti = si[lo]; si[lo] = si[hi]; si[hi] = ti;
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < tile; i += blockDim.x) {
    key[base + i] = sk[i];
    idx[base + i] = si[i];
  }
}

// One merge level: adjacent sorted runs of length L into runs of 2L, one
// thread per element.  An element of the first run of a pair lands at its
// index plus the number of second-run elements below it; an element of the
// second run at its index plus the number of first-run elements not above it
// (a stable merge, so the equal padding pairs keep distinct slots).
__global__ void k_merge_runs(const unsigned long long* __restrict__ key,
                             const long long* __restrict__ idx,
                             unsigned long long* __restrict__ okey, long long* __restrict__ oidx,
                             long long P, long long L) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= P) return;
  const long long run = t / L;
  const long long pos = t - run * L;
  const long long a0 = (run & ~1ll) * L;
  const bool first = (run & 1) == 0;
  const long long o = first ? a0 + L : a0;
  const unsigned long long k = key[t];
  const long long ix = idx[t];
  long long lo = 0, hi = L;
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    const unsigned long long km = key[o + mid];
    const long long im = idx[o + mid];
    const bool before = first ? pair_less(km, im, k, ix) : !pair_less(k, ix, km, im);
    if (before) lo = mid + 1;
    else hi = mid;
  }
  okey[a0 + pos + lo] = k;
  oidx[a0 + pos + lo] = ix;
}

__global__ void k_gather_out(const double2* __restrict__ c, const long long* __restrict__ sidx,
                             long long k, long long* __restrict__ idx_out,
                             double2* __restrict__ val_out) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= k) return;
  const long long j = sidx[i];
  idx_out[i] = j;
  val_out[i] = c[j];
}

long long pow2_at_least(long long x) {
  long long p = 1;
  while (p < x) p <<= 1;
  return p;
}

struct TopnWs {
  unsigned long long* key;
  unsigned long long* okey;
  long long* oidx;
  unsigned long long* okey2;   // merge ping-pong
  long long* oidx2;
  long long* cnt;
  unsigned int* hist;
  SelState* st;
  size_t bytes;
};

TopnWs carve(void* ws, long long n, long long k) {
  const long long P = pow2_at_least(k < kSortTile ? kSortTile : k);
  const long long nblk = (n + kChunk - 1) / kChunk;
  char* p = (char*)ws;
  TopnWs w;
  auto take = [&](size_t bytes) {
    char* r = p;
    p += (bytes + 255) / 256 * 256;
    return (void*)r;
  };
  w.key = (unsigned long long*)take(n * 8);
  w.okey = (unsigned long long*)take(P * 8);
  w.oidx = (long long*)take(P * 8);
  w.okey2 = (unsigned long long*)take(P * 8);
  w.oidx2 = (long long*)take(P * 8);
  w.cnt = (long long*)take(2 * nblk * 8);
  w.hist = (unsigned int*)take(256 * 4);
  w.st = (SelState*)take(sizeof(SelState));
  w.bytes = (size_t)(p - (char*)ws);
  return w;
}

}  // namespace

namespace {
__global__ void k_emit_key(const SelState* st, unsigned long long* out) { *out = st->prefix; }
}  // namespace

namespace bh {
long long sort_pairs_capacity(long long n) {
  return pow2_at_least(n < kSortTile ? kSortTile : n);
}

int sort_pairs(unsigned long long* key, long long* idx, unsigned long long* key2, long long* idx2,
               long long n, unsigned long long** key_out, long long** idx_out, hipStream_t s) {
  const long long P = sort_pairs_capacity(n);
  if (P > n)
    hipLaunchKernelGGL(k_pad, dim3(bh::grid1d(P - n, 256)), dim3(256), 0, s, key, idx, n, P);
  hipLaunchKernelGGL(k_bitonic_tile, dim3((unsigned)(P / kSortTile)), dim3(256), 0, s, key, idx, P);
  unsigned long long* ka = key;
  long long* ia = idx;
  unsigned long long* kb = key2;
  long long* ib = idx2;
  for (long long L = kSortTile; L < P; L <<= 1) {
    hipLaunchKernelGGL(k_merge_runs, dim3(bh::grid1d(P, 256)), dim3(256), 0, s, ka, ia, kb, ib, P,
                       L);
    unsigned long long* tk = ka; ka = kb; kb = tk;
    long long* ti = ia; ia = ib; ib = ti;
  }
  *key_out = ka;
  *idx_out = ia;
  BH_LAUNCH_CHECK();
  return BH_OK;
}

size_t kth_key_ws_bytes() { return 256 * 4 + 256; }

void kth_largest_key(const unsigned long long* keys, long long n, long long k,
                     unsigned long long* out, void* ws, hipStream_t s) {
  unsigned int* hist = (unsigned int*)ws;
  SelState* st = (SelState*)((char*)ws + 256 * 4);
  const long long nblk = (n + kChunk - 1) / kChunk;
  hipLaunchKernelGGL(k_init_state, dim3(1), dim3(256), 0, s, st, hist, k);
  for (int shift = 56; shift >= 0; shift -= 8) {
    hipLaunchKernelGGL(k_hist, dim3((unsigned)nblk), dim3(kSelBlock), 0, s, keys, n, st, shift,
                       hist);
    hipLaunchKernelGGL(k_pick, dim3(1), dim3(256), 0, s, st, hist, shift);
  }
  hipLaunchKernelGGL(k_emit_key, dim3(1), dim3(1), 0, s, st, out);
}
}  // namespace bh

extern "C" size_t bh_topn_workspace_size(int64_t n, int64_t k) {
  if (n <= 0 || k <= 0) return 0;
  if (k > n) k = n;
  TopnWs w = carve(nullptr, n, k);
  return w.bytes;
}

namespace {
// The selection on a workspace the caller has checked (k <= n, sized by
// bh_topn_workspace_size).
int topn_run(const double* c, int64_t n, int64_t k, int64_t* idx_out, double* val_out, void* ws,
             bh_stream s) {
  hipStream_t st = (hipStream_t)s;
  TopnWs w = carve(ws, n, k);
  const long long nblk = (n + kChunk - 1) / kChunk;
  hipLaunchKernelGGL(k_keys, dim3(bh::grid1d(n, 256)), dim3(256), 0, st, (const double2*)c,
                     w.key, (long long)n);
  hipLaunchKernelGGL(k_init_state, dim3(1), dim3(256), 0, st, w.st, w.hist, (long long)k);
  for (int shift = 56; shift >= 0; shift -= 8) {
    hipLaunchKernelGGL(k_hist, dim3((unsigned)nblk), dim3(kSelBlock), 0, st, w.key,
                       (long long)n, w.st, shift, w.hist);
    hipLaunchKernelGGL(k_pick, dim3(1), dim3(256), 0, st, w.st, w.hist, shift);
  }
  hipLaunchKernelGGL(k_count, dim3((unsigned)nblk), dim3(kSelBlock), 0, st, w.key, (long long)n,
                     w.st, w.cnt);
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, w.cnt, nblk);
  const long long P = pow2_at_least(k < kSortTile ? kSortTile : k);
  hipLaunchKernelGGL(k_pad, dim3(bh::grid1d(P - k, 256)), dim3(256), 0, st, w.okey, w.oidx,
                     (long long)k, P);
  hipLaunchKernelGGL(k_compact, dim3((unsigned)nblk), dim3(kSelBlock), 0, st, w.key,
                     (long long)n, w.st, w.cnt, w.okey, w.oidx);
  BH_LAUNCH_CHECK();
  // sort of P (power of two >= 512) pairs, ascending: bitonic within
  // 512-element tiles, then log2(P / 512) merge levels
  const unsigned tiles = (unsigned)(P / kSortTile);
  hipLaunchKernelGGL(k_bitonic_tile, dim3(tiles), dim3(256), 0, st, w.okey, w.oidx, P);
  unsigned long long* ka = w.okey;
  long long* ia = w.oidx;
  unsigned long long* kb = w.okey2;
  long long* ib = w.oidx2;
  for (long long L = kSortTile; L < P; L <<= 1) {
    hipLaunchKernelGGL(k_merge_runs, dim3(bh::grid1d(P, 256)), dim3(256), 0, st, ka, ia, kb, ib,
                       P, L);
    unsigned long long* tk = ka; ka = kb; kb = tk;
    long long* ti = ia; ia = ib; ib = ti;
  }
  hipLaunchKernelGGL(k_gather_out, dim3(bh::grid1d(k, 256)), dim3(256), 0, st,
                     (const double2*)c, ia, (long long)k, (long long*)idx_out,
                     (double2*)val_out);
  BH_LAUNCH_CHECK();
  return BH_OK;
}
}  // namespace

extern "C" int bh_topn(const double* c, int64_t n, int64_t k, int64_t* idx_out,
                       double* val_out, void* ws, size_t ws_bytes, bh_stream s) {
  bh::clear_error();
  if (n <= 0) BH_FAIL(BH_ERR_ARG, "bh_topn: empty input (n=%lld)", (long long)n);
  if (k <= 0) BH_FAIL(BH_ERR_ARG, "top_n threshold must be a positive integer.");
  if (k > n) k = n;
  if (!c || !idx_out || !val_out || !ws) BH_FAIL(BH_ERR_ARG, "bh_topn: null pointer");
  const size_t need = bh_topn_workspace_size(n, k);
  if (ws_bytes < need) BH_FAIL(BH_ERR_WORKSPACE, "bh_topn: workspace %zu < %zu", ws_bytes, need);
  return topn_run(c, n, k, idx_out, val_out, ws, s);
}

// ---------------------------------------------------------------------------
// bh_topn_noalloc: the same selection with its workspace in a fixed device
// array of this library (no allocation by the library or the caller).
// ---------------------------------------------------------------------------
namespace {
constexpr long long kNoallocMaxN = 1ll << 22;   // coefficients
constexpr long long kNoallocMaxK = 1ll << 18;   // selected entries

constexpr size_t round256(size_t x) { return (x + 255) / 256 * 256; }
// carve()'s layout for n keys and a sort of P (a power of two) entries
constexpr size_t topn_layout_bytes(long long n, long long P) {
  return round256((size_t)n * 8) + 4 * round256((size_t)P * 8) +
         round256((size_t)2 * ((n + kChunk - 1) / kChunk) * 8) + round256(256 * 4) +
         round256(sizeof(SelState));
}
constexpr size_t kScratchBytes = topn_layout_bytes(kNoallocMaxN, kNoallocMaxK);

// One copy per device, held while the library stays loaded.  Host code never
// uses this name as an address: the device address comes from
// hipGetSymbolAddress, once per device (scratch_address).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};

int scratch_address(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0)
    BH_FAIL(BH_ERR_HIP, "bh_topn_noalloc: no current device");
  if (dev >= kMaxDevices)
    BH_FAIL(BH_ERR_UNSUPPORTED, "bh_topn_noalloc: device %d >= %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      BH_FAIL(BH_ERR_HIP, "bh_topn_noalloc: hipGetSymbolAddress failed on device %d: %s", dev,
              hipGetErrorString(e));
    g_scratch_addr[dev] = p;
  }
  *out = g_scratch_addr[dev];
  return BH_OK;
}
}  // namespace

extern "C" int bh_topn_noalloc(const double* c, int64_t n, int64_t k, int64_t* idx_out,
                               double* val_out, bh_stream s) {
  bh::clear_error();
  if (n <= 0) BH_FAIL(BH_ERR_ARG, "bh_topn_noalloc: empty input (n=%lld)", (long long)n);
  if (k <= 0) BH_FAIL(BH_ERR_ARG, "top_n threshold must be a positive integer.");
  if (k > n) k = n;
  if (!c || !idx_out || !val_out) BH_FAIL(BH_ERR_ARG, "bh_topn_noalloc: null pointer");
  if (n > kNoallocMaxN || k > kNoallocMaxK || bh_topn_workspace_size(n, k) > kScratchBytes)
    BH_FAIL(BH_ERR_UNSUPPORTED,
            "bh_topn_noalloc: n=%lld, k=%lld exceed its fixed scratch (n <= %lld, k <= %lld); "
            "use bh_topn with a workspace",
            (long long)n, (long long)k, kNoallocMaxN, kNoallocMaxK);
  void* ws = nullptr;
  if (int rc = scratch_address(&ws)) return rc;
  return topn_run(c, n, k, idx_out, val_out, ws, s);
}
