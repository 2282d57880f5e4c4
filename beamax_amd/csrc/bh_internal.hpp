// Internal (non-exported) interfaces shared by the library's translation units.
#pragma once

#include "bh_common.hpp"

namespace bh {

// Batched strided 1-D complex FFT in place (see bh_fft in the public header).
// ws may be null when n <= kFftLdsMax (LDS-resident path needs no workspace):
// up to 1024 points several transforms share a workgroup; above, one
// transform per workgroup with its two buffers and twiddles in up to 120 KiB
// of the CU's 160 KiB of LDS.
constexpr int64_t kFftLdsMax = 2560;
constexpr int64_t kFftLdsShare = 1024;
size_t fft_workspace_bytes(int64_t n, int64_t batch);
int fft_strided(double2* data, int64_t n, int64_t batch, int64_t inner, int64_t inner_stride,
                int64_t outer_stride, int64_t elem_stride, int dir, double scale, void* ws,
                size_t ws_bytes, hipStream_t s);

// Multi-dimensional unitary FFT of a C-order array of shape N[0..d) in place:
// dir -1 = fftn(ortho), +1 = ifftn(ortho).  No shifts.
int fftn_ortho(double2* data, int ndim, const int64_t* N, int dir, void* ws, size_t ws_bytes,
               hipStream_t s);

inline int64_t prod(const int64_t* v, int n) {
  int64_t p = 1;
  for (int i = 0; i < n; ++i) p *= v[i];
  return p;
}

inline unsigned grid1d(int64_t work, int block) {
  int64_t g = (work + block - 1) / block;
  if (g < 1) g = 1;
  if (g > 0x7fffffff) g = 0x7fffffff;
  return (unsigned)g;
}

// device-side copy of the tiling scalars (everything but the centre pointer)
struct TilingDev {
  int32_t ndim, num_levels, redundancy, windowing;
  int64_t N[BH_MAX_DIM];
  int32_t aspect[BH_MAX_DIM];
  int32_t box_length[BH_MAX_LEVELS];
  int32_t support[BH_MAX_LEVELS][BH_MAX_DIM];
  int64_t boxes_cumsum[BH_MAX_LEVELS + 1];
  int64_t coeffs_cumsum[BH_MAX_LEVELS + 1];
  int64_t gpacked_off[BH_MAX_LEVELS + 1];
};

int check_tiling(const bh_tiling* t);
TilingDev to_dev(const bh_tiling* t);

// The k-th largest (1 <= k <= n) of n keys that are bit patterns of
// non-negative doubles (so they order like the values), by the top_n radix
// select; written to *out (device).  ws: kth_key_ws_bytes() bytes.
size_t kth_key_ws_bytes();
void kth_largest_key(const unsigned long long* keys, long long n, long long k,
                     unsigned long long* out, void* ws, hipStream_t s);

// Ascending sort of n (key, index) pairs by (key, index): bitonic 512-pair
// tiles, then stable merge levels.  key/idx and key2/idx2 each hold
// sort_pairs_capacity(n) entries (a power of two, padded here); the sorted
// pairs end in one of the two buffers, returned in *key_out / *idx_out.
long long sort_pairs_capacity(long long n);
int sort_pairs(unsigned long long* key, long long* idx, unsigned long long* key2, long long* idx2,
               long long n, unsigned long long** key_out, long long** idx_out, hipStream_t s);

// Activity windows of a closed-form forward beam table (entries [0, count)
// with mode +1 and, if mirror, [count, 2 count) with mode -1): per entry the
// first and last time row on which its envelope can reach a point run at
// cut E (rows outside hold no term above 2^-E of the beam).  Only for
// imaginary alpha and increasing ts; other entries get all rows.  ws:
// beam_windows_bytes(total) bytes; *win points into it.
size_t beam_windows_bytes(long long total);
int beam_windows(int ndim, const bh_fwd_beams* b, int mirror, double c, const double* ts,
                 long long nt, const bh_segments* seg, double E, void* ws, const int2** win,
                 hipStream_t st);

// Status words of the kernels that guard beam / coefficient indices (one per
// translation unit): OR the bits into *bits, clear them if asked.  Both
// synchronise the stream.
int beams_status(int clear, unsigned* bits, hipStream_t s);
int transpose_status(int clear, unsigned* bits, hipStream_t s);

}  // namespace bh
