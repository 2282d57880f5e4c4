// Exact transpose of the frozen-index forward beam sum.
//
// With the beam set frozen, the real forward sum is linear in the beam
// coefficients c_k:  u(t, s) = sum_k sum_{mode} 2 Re(c_k G_k(t, s)), where
// G_k = A_k(t) exp(i w_k (p.delta + delta^T M delta / 2)) is the Gaussian beam
// of unit coefficient (A_k(t) from the closed-form propagator with a0 of the
// beam table).  Its transpose for the real inner products <u, r> and
// Re<c, g> is
//     g_k = sum_{mode} sum_{t, s} 2 r(t, s) conj(G_k(t, s)),
// which this file evaluates.  Work unit: a tile of 64 time rows (one per
// lane) x up to 32 point groups; a wave takes one beam at a time, builds the
// 64 records of its tile (lane = time row), finds per row the groups inside
// the certified envelope (the same closed-form run intervals as the culled
// forward sum), then walks the active (row, group) pairs with the lanes over
// the group's 64 points, and reduces its lanes once per (beam, tile).  Every
// (tile, beam) slot is written exactly once and the slots are summed in a
// fixed order, so the result is deterministic without atomics.
#include <type_traits>

#include "bh_segcull.hpp"

namespace {

constexpr int kTileT = 64;   // time rows per tile (one per lane)
constexpr int kTileG = 32;   // 64-point groups per tile (bits of a mask word)

struct TArgs {
  bh_fwd_beams b;
  long long count;
  int modes;             // 2: mirrored pair (+c, -c) per beam
  double c;
  const double* ts;
  long long nt;
  SegDev G;
  double E;              // cull exponent (log2 units); <= 0: every pair
  const void* resid;     // (nt, S) row-major, float or double
  int resid_f64;
  double* slots;         // [ntt * ntg][count] complex
  long long ntg;
};

template <int D, bool F64>
__global__ __launch_bounds__(256) void k_fwd_transpose(TArgs A) {
  using Rec = typename RecOf<D, F64>::type;
  using Part = typename std::conditional<F64, double, float>::type;
  __shared__ Rec rec[4][kTileT];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const long long t0 = (long long)blockIdx.y * kTileT;
  const long long g0 = (long long)blockIdx.z * kTileG;
  const long long ngl = A.G.ngroups - g0;
  const int ng = (int)(ngl < kTileG ? ngl : kTileG);
  const long long tl = t0 + lane;
  const bool tvalid = tl < A.nt;
  const double tv = tvalid ? A.ts[tl] : 0.0;
  const long long tile = (long long)blockIdx.y * A.ntg + blockIdx.z;
  const long long nw = (long long)gridDim.x * 4;
  const long long S = A.G.S;
  for (long long k = (long long)blockIdx.x * 4 + wv; k < A.count; k += nw) {
    double x0[D], p0[D];
    cd al[D];
    for (int a = 0; a < D; ++a) {
      x0[a] = A.b.x0[k * D + a];
      p0[a] = A.b.p0[k * D + a];
      al[a] = mkc(A.b.alpha[2 * (k * D + a)], A.b.alpha[2 * (k * D + a) + 1]);
    }
    const cd a0 = mkc(A.b.a0[2 * k], A.b.a0[2 * k + 1]);
    const double w = A.b.omega[k];
    double acc_re = 0.0, acc_im = 0.0;
    for (int md = 0; md < A.modes; ++md) {
      Rec R;
      bool live = tvalid;
      if (live) {
        State<D> St;
        hom_diag_state_fast<D>(x0, p0, al, a0, md ? -A.c : A.c, tv, St);
        make_rec<D>(St, w, R);
        live = !rec_zero<D>(R);
      }
      if (!live) zero_rec<D>(R);
      // groups of this tile inside the envelope at this lane's time row
      unsigned msk = 0;
      if (live) {
        double Q[D][D];
        if (A.E > 0.0) env_matrix<D>(R, Q);
        long long cur = -1;
        int lo = 1, hi = 0;
        for (int gl = 0; gl < ng; ++gl) {
          bool act = true;
          if (A.E > 0.0) {
            const long long g = g0 + gl;
            const int ks = A.G.grp_seg[g];
            if (ks != cur) {
              seg_interval<D>(A.G, ks, R.x, Q, A.E, lo, hi);
              cur = ks;
            }
            const int j0 = (int)((g - A.G.g0[ks]) * 64);
            act = (hi >= lo) && (hi >= j0) && (lo <= j0 + 63);
          }
          if (act) msk |= 1u << gl;
        }
      }
      wave_lds_fence();   // the previous mode's rows are consumed
      rec[wv][lane] = R;
      wave_lds_fence();   // every lane's record is visible to the wave
      for (int j = 0; j < kTileT; ++j) {
        unsigned m = __builtin_amdgcn_readlane(msk, j);
        if (!m) continue;
        const Rec Rj = rec[wv][j];
        const long long t = t0 + j;
        Part pr = 0, pi = 0;
        while (m) {
          const int gl = __builtin_ctz(m);
          m &= m - 1;
          const long long slot = (g0 + gl) * 64 + lane;
          double dl[D];
          for (int a = 0; a < D; ++a) dl[a] = A.G.pts[slot * D + a] - Rj.x[a];
          const int col = A.G.col[slot];
          double r = 0.0;
          if (col >= 0)
            r = A.resid_f64 ? ((const double*)A.resid)[t * S + col]
                            : (double)((const float*)A.resid)[t * S + col];
          if constexpr (F64) eval_f64_conj<D>(Rj, dl, r, pr, pi);
          else eval_split_conj<D>(Rj, dl, (float)r, pr, pi);
        }
        acc_re += (double)pr;
        acc_im += (double)pi;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      acc_re += __shfl_xor(acc_re, off);
      acc_im += __shfl_xor(acc_im, off);
    }
    if (lane == 0) {
      A.slots[2 * (tile * A.count + k)] = acc_re;
      A.slots[2 * (tile * A.count + k) + 1] = acc_im;
    }
  }
}

// ---------------------------------------------------------------------------
// FP32-split form by per-pair quadratics (the forward's evaluation turned
// around).  Work unit: 64 time rows (one per lane) x up to 32 units of 128
// points (two adjacent groups of a run, the forward's culling unit).  A wave
// takes one beam at a time: lane = row builds the row's record and a 32-bit
// mask of the units inside the certified envelope (the forward's intervals,
// so both sums hold the same terms); the active (row, unit) pairs are
// compacted into an LDS list in (row, unit) order; then in passes of 64 each
// lane forms one pair's phase/envelope quadratic in the point index
// (pair_quad) and the wave walks the pass, lane l taking the unit's points l
// and l + 64, weighting each term by the residual (gathered into group order
// once per call, zero on padding slots) and accumulating r 2|A| 2^-e (cos,
// -sin)(2 pi cycles) -- r times the conjugate of the forward's complex term.
// Per lane the pass partials are FP32, the beam's totals FP64; the wave
// reduces them once per (beam, tile) into its slot.
// ---------------------------------------------------------------------------
// 128-point units per tile (bits of a row's mask): 8, 16 or 32, the fewest
// that hold the points' units (fewer units, less LDS for the pair list, more
// resident blocks)
constexpr int kTileU = 32;
inline int tile_units(long long nunits) { return nunits <= 8 ? 8 : nunits <= 16 ? 16 : kTileU; }

struct TQArgs {
  bh_fwd_beams b;
  long long count;
  int modes;
  double c;
  const double* ts;
  long long nt;
  SegDev G;
  double E;
  const float* rg;       // residual in group order: [nt][ngroups * 64]
  double* slots;         // [ntt * ntu][count] complex
  long long ntu;         // unit tiles
  long long nunits;      // ngroups / 2
  // activity windows per entry (mode +1: beam k, mode -1: count + k), or
  // null: a wave skips a (beam, mode) whose window misses its 64 rows
  const int2* win;
};

// r at (t, slot) for every 64-point group slot, zero where the slot is padding
__global__ void k_gather_resid(const void* __restrict__ resid, int f64, long long nt, long long S,
                               const int32_t* __restrict__ col, long long nslot,
                               float* __restrict__ rg) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= nt * nslot) return;
  const long long t = i / nslot, q = i - t * nslot;
  const int c = col[q];
  float v = 0.f;
  if (c >= 0)
    v = f64 ? (float)((const double*)resid)[t * S + c] : ((const float*)resid)[t * S + c];
  rg[i] = v;
}

// Up to 16 units per tile the 1-D and 2-D forms run at four waves per SIMD
// (the LDS allows four blocks per CU; a few registers spill outside the
// evaluation loop); otherwise the compiler's choice.
template <int D, int TU>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((D <= 2 && TU <= 16) ? 4 : 1)))
void k_fwd_transpose_q(TQArgs A) {
  __shared__ RecS<D> rec[4][64];
  __shared__ unsigned short plist[4][64 * TU];   // row << 5 | unit
  __shared__ PairQ coef[4][64];
  __shared__ double upt[TU][D], ustep[TU][D];
  __shared__ int useg[TU], uj0[TU];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const long long t0 = (long long)blockIdx.y * 64;
  const long long ub = (long long)blockIdx.z * TU;
  const long long nul = A.nunits - ub;
  const int nu = (int)(nul < TU ? nul : TU);
  if (threadIdx.x < nu) {
    const long long g = 2 * (ub + threadIdx.x);
    const int k = A.G.grp_seg[g];
    for (int a = 0; a < D; ++a) {
      upt[threadIdx.x][a] = A.G.pts[g * 64 * D + a];
      ustep[threadIdx.x][a] = A.G.v[(long long)k * D + a];
    }
    useg[threadIdx.x] = k;
    uj0[threadIdx.x] = (int)((g - A.G.g0[k]) * 64);
  }
  __syncthreads();
  const long long tl = t0 + lane;
  const bool tvalid = tl < A.nt;
  const double tv = tvalid ? A.ts[tl] : 0.0;
  const long long tile = (long long)blockIdx.y * A.ntu + blockIdx.z;
  const long long nw = (long long)gridDim.x * 4;
  const long long S64 = A.G.ngroups * 64;
  const unsigned all_units = nu == 32 ? 0xffffffffu : ((1u << nu) - 1u);
  const double ld = (double)lane;
  const float lf = (float)lane;
  for (long long k = (long long)blockIdx.x * 4 + wv; k < A.count; k += nw) {
    double acc_re = 0.0, acc_im = 0.0;
    for (int md = 0; md < A.modes; ++md) {
      if (A.win) {
        // no term of this (beam, mode) on rows [t0, t0 + 63] reaches the cut
        const int2 w = A.win[md ? A.count + k : k];
        if (w.y < t0 || w.x > t0 + 63) continue;
      }
      RecS<D> R;
      bool live = tvalid;
      if (live) {
        double x0[D], p0[D];
        cd al[D];
        for (int a = 0; a < D; ++a) {
          x0[a] = A.b.x0[k * D + a];
          p0[a] = A.b.p0[k * D + a];
          al[a] = mkc(A.b.alpha[2 * (k * D + a)], A.b.alpha[2 * (k * D + a) + 1]);
        }
        const cd a0 = mkc(A.b.a0[2 * k], A.b.a0[2 * k + 1]);
        const double w = A.b.omega[k];
        // 2-D beams with alpha = i b: the forward's records in real
        // arithmetic (fast_rec2, the same records as the windowed forward
        // sum); the test is the same for the whole wave (one beam)
        bool done = false;
        if constexpr (D == 2) {
          const FastEnt2 F = fast_consts2(x0, p0, al, a0, w);
          if (F.iP > 0.0) {
            fast_rec2(F, md ? -A.c : A.c, tv, R);
            done = true;
          }
        }
        if (!done) {
          State<D> St;
          hom_diag_state_fast<D, true>(x0, p0, al, a0, md ? -A.c : A.c, tv, St);
          make_rec<D>(St, w, R);
        }
        live = !rec_zero<D>(R);
      }
      if (!live) zero_rec<D>(R);
      unsigned msk = 0u;
      if (live) {
        if (A.E > 0.0) {
          double Q[D][D];
          env_matrix<D>(R, Q);
          int cur = -1, lo = 1, hi = 0;
          for (int u = 0; u < nu; ++u) {
            const int ks = useg[u];
            if (ks != cur) {
              seg_interval<D>(A.G, ks, R.x, Q, A.E, lo, hi);
              cur = ks;
            }
            const int j0 = uj0[u];
            if (hi >= lo && hi >= j0 && lo <= j0 + 127) msk |= 1u << u;
          }
        } else {
          msk = all_units;
        }
      }
      wave_lds_fence();   // the previous mode's list, records and pass are consumed
      rec[wv][lane] = R;
      // compaction of the active (row, unit) pairs, row-major
      const int cnt = __popc(msk);
      int incl = cnt;
      for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
      }
      const int total = __shfl(incl, 63);
      int pos = incl - cnt;
      for (unsigned m = msk; m; m &= m - 1)
        plist[wv][pos++] = (unsigned short)((lane << 5) | __builtin_ctz(m));
      wave_lds_fence();
      for (int b = 0; b < total; b += 64) {
        const int e = b + lane;
        if (e < total) {
          const int pr = plist[wv][e];
          coef[wv][lane] = pair_quad<D>(rec[wv][pr >> 5], upt[pr & 31], ustep[pr & 31]);
        }
        wave_lds_fence();
        const int n = __builtin_amdgcn_readfirstlane(total - b < 64 ? total - b : 64);
        float sr = 0.f, si = 0.f;
        for (int sl = 0; sl < n; ++sl) {
          const int pr = __builtin_amdgcn_readfirstlane((int)plist[wv][b + sl]);
          const PairQ c = coef[wv][sl];
          const float* rr = A.rg + (t0 + (pr >> 5)) * S64 + (2 * (ub + (pr & 31))) * 64;
          const float r0 = rr[lane], r1 = rr[64 + lane];
          const double cyc0 = fma(ld, fma(ld, c.C, c.B), c.A);
          const double ld1 = ld + 64.0;
          const double cyc1 = fma(ld1, fma(ld1, c.C, c.B), c.A);
          const float f0 = __uint_as_float(
              __builtin_amdgcn_alignbit(0x7Fu, (unsigned)__double_as_longlong(cyc0), 9u));
          const float f1 = __uint_as_float(
              __builtin_amdgcn_alignbit(0x7Fu, (unsigned)__double_as_longlong(cyc1), 9u));
          const float u0 = fmaf(lf, c.s, c.nl0s);
          const float u1 = fmaf(64.f, c.s, u0);
          const float m0 = r0 * __builtin_amdgcn_exp2f(fmaf(-u0, u0, c.nlog));
          const float m1 = r1 * __builtin_amdgcn_exp2f(fmaf(-u1, u1, c.nlog));
          sr = fmaf(m0, __builtin_amdgcn_cosf(f0), sr);
          sr = fmaf(m1, __builtin_amdgcn_cosf(f1), sr);
          si = fmaf(-m0, __builtin_amdgcn_sinf(f0), si);
          si = fmaf(-m1, __builtin_amdgcn_sinf(f1), si);
        }
        acc_re += (double)sr;
        acc_im += (double)si;
        wave_lds_fence();   // the pass's coefficients are consumed
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      acc_re += __shfl_xor(acc_re, off);
      acc_im += __shfl_xor(acc_im, off);
    }
    if (lane == 0) {
      A.slots[2 * (tile * A.count + k)] = acc_re;
      A.slots[2 * (tile * A.count + k) + 1] = acc_im;
    }
  }
}

// grad[k] = sum over tiles (fixed order) of slots[tile][k]
__global__ void k_slot_reduce(const double* __restrict__ slots, long long tiles, long long count,
                              double* __restrict__ grad) {
  const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= count) return;
  double re = 0.0, im = 0.0;
  for (long long q = 0; q < tiles; ++q) {
    re += slots[2 * (q * count + k)];
    im += slots[2 * (q * count + k) + 1];
  }
  grad[2 * k] = re;
  grad[2 * k + 1] = im;
}

__global__ void k_fill_zero(double* __restrict__ p, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// status word of this translation unit (BH_STATUS_COEFF_INDEX), read and
// cleared by bh_device_status
__device__ unsigned int g_status = 0u;

// out[idx[i]] = scale * vals[i]; the padding index -1 is skipped (the
// parameter kernels make such beams inert), any other index outside
// [0, total) is skipped and recorded in the status word
__global__ void k_scatter(const int64_t* __restrict__ idx, const double* __restrict__ vals,
                          long long k, double scale, long long total, double* __restrict__ out) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= k) return;
  const int64_t j = idx[i];
  if (j < 0 || j >= total) {
    if (j != -1) atomicOr(&g_status, (unsigned)BH_STATUS_COEFF_INDEX);
    return;
  }
  out[2 * j] = scale * vals[2 * i];
  out[2 * j + 1] = scale * vals[2 * i + 1];
}

// vals[i] = coeffs[idx[i]] (0 for an index outside [0, total))
__global__ void k_gather(const double* __restrict__ coeffs, long long total,
                         const int64_t* __restrict__ idx, long long k, double* __restrict__ vals) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= k) return;
  const int64_t j = idx[i];
  const bool ok = j >= 0 && j < total;
  if (!ok && j != -1) atomicOr(&g_status, (unsigned)BH_STATUS_COEFF_INDEX);
  vals[2 * i] = ok ? coeffs[2 * j] : 0.0;
  vals[2 * i + 1] = ok ? coeffs[2 * j + 1] : 0.0;
}

struct TShape {
  long long ntt, ntg, gx;
};

inline TShape transpose_shape(long long count, long long nt, long long ngroups) {
  TShape s;
  s.ntt = (nt + kTileT - 1) / kTileT;
  s.ntg = (ngroups + kTileG - 1) / kTileG;
  const long long tiles = s.ntt * s.ntg;
  long long gx = (4096 + tiles - 1) / tiles;       // ~16 block rounds of 256 CUs
  const long long maxgx = (count + 3) / 4;          // at least one beam per wave
  if (gx > maxgx) gx = maxgx;
  if (gx > 65535) gx = 65535;
  s.gx = gx < 1 ? 1 : gx;
  return s;
}

// launch shape of the FP32-split form: unit tiles instead of group tiles
inline TShape transpose_q_shape(long long count, long long nt, long long ngroups) {
  TShape s;
  s.ntt = (nt + 63) / 64;
  const int tu = tile_units(ngroups / 2);
  s.ntg = (ngroups / 2 + tu - 1) / tu;
  const long long tiles = s.ntt * s.ntg;
  long long gx = (4096 + tiles - 1) / tiles;
  const long long maxgx = (count + 3) / 4;
  if (gx > maxgx) gx = maxgx;
  if (gx > 65535) gx = 65535;
  s.gx = gx < 1 ? 1 : gx;
  return s;
}

size_t round256(size_t x) { return (x + 255) / 256 * 256; }

}  // namespace

namespace {
// the FP64 form's slots, or the FP32-split form's slots followed by the
// residual in group order; the larger of the two
size_t transpose_base_bytes(int64_t count, int64_t nt, int64_t ngroups) {
  const TShape sh = transpose_shape(count, nt, ngroups);
  const size_t f64 = (size_t)sh.ntt * sh.ntg * count * 2 * sizeof(double);
  const TShape sq = transpose_q_shape(count, nt, ngroups);
  const size_t f32 = round256((size_t)sq.ntt * sq.ntg * count * 2 * sizeof(double)) +
                     (size_t)nt * ngroups * 64 * sizeof(float);
  return round256(f64 > f32 ? f64 : f32);
}
}  // namespace

// The slots (and the FP32-split form's residual in group order), then room
// for the activity windows of both modes; with less than that (but at least
// the slots) the FP32-split form runs without windows.
extern "C" size_t bh_forward_transpose_workspace_size(int64_t count, int64_t nt,
                                                      int64_t ngroups) {
  if (count <= 0 || nt <= 0 || ngroups <= 0) return 0;
  return transpose_base_bytes(count, nt, ngroups) + bh::beam_windows_bytes(2 * count);
}

extern "C" int bh_forward_transpose(const bh_fwd_beams* b, int ndim, int mirror, double c,
                                    const double* ts, int64_t nt, const bh_segments* seg,
                                    double cull_log2, int precision, const void* resid,
                                    int resid_f64, double* grad, void* ws, size_t ws_bytes,
                                    bh_stream s) {
  bh::clear_error();
  if (int rc = check_segments(seg, ndim)) return rc;
  if (!b || !ts || !resid || !grad) BH_FAIL(BH_ERR_ARG, "bh_forward_transpose: null pointer");
  if (nt <= 0 || nt > 65535 * (int64_t)kTileT)
    BH_FAIL(BH_ERR_ARG, "bh_forward_transpose: nt=%lld", (long long)nt);
  if (precision != BH_PRECISION_FP64 && precision != BH_PRECISION_FP32_SPLIT)
    BH_FAIL(BH_ERR_ARG, "bh_forward_transpose: precision=%d", precision);
  hipStream_t st = (hipStream_t)s;
  if (b->count == 0) return BH_OK;
  if (!b->x0 || !b->p0 || !b->omega || !b->alpha || !b->a0)
    BH_FAIL(BH_ERR_ARG, "bh_forward_transpose: null beam array");
  const TShape sh = transpose_shape(b->count, nt, seg->ngroups);
  if (sh.ntg > 65535) BH_FAIL(BH_ERR_ARG, "bh_forward_transpose: %lld point groups",
                              (long long)seg->ngroups);
  const size_t need = transpose_base_bytes(b->count, nt, seg->ngroups);
  if (!ws || ws_bytes < need)
    BH_FAIL(BH_ERR_WORKSPACE, "bh_forward_transpose: workspace %zu < %zu", ws_bytes, need);
  if (precision == BH_PRECISION_FP32_SPLIT) {
    if (seg->ngroups % 2)
      BH_FAIL(BH_ERR_ARG, "bh_forward_transpose: %lld groups (runs hold an even number)",
              (long long)seg->ngroups);
    const TShape sq = transpose_q_shape(b->count, nt, seg->ngroups);
    TQArgs q{};
    q.b = *b;
    q.count = b->count;
    q.modes = mirror ? 2 : 1;
    q.c = c;
    q.ts = ts;
    q.nt = nt;
    q.G = to_segdev(seg);
    q.E = cull_log2;
    q.slots = (double*)ws;
    float* rg = (float*)((char*)ws + round256((size_t)sq.ntt * sq.ntg * b->count * 2 *
                                              sizeof(double)));
    q.rg = rg;
    q.ntu = sq.ntg;
    q.nunits = seg->ngroups / 2;
    q.win = nullptr;
    const long long entries = b->count * (mirror ? 2 : 1);
    if (cull_log2 > 0.0 && nt < 65535 && ws_bytes >= need + bh::beam_windows_bytes(entries)) {
      if (int rc = bh::beam_windows(ndim, b, mirror, c, ts, nt, seg, cull_log2, (char*)ws + need,
                                    &q.win, st))
        return rc;
    }
    const long long nslot = seg->ngroups * 64;
    hipLaunchKernelGGL(k_gather_resid, dim3(bh::grid1d(nt * nslot, 256)), dim3(256), 0, st, resid,
                       resid_f64, (long long)nt, (long long)seg->S, seg->col, nslot, rg);
    dim3 grid((unsigned)sq.gx, (unsigned)sq.ntt, (unsigned)sq.ntg);
    const int tu = tile_units(q.nunits);
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, grid, dim3(256), 0, st, q); };
    if (ndim == 1) {
      if (tu == 8) go(k_fwd_transpose_q<1, 8>);
      else if (tu == 16) go(k_fwd_transpose_q<1, 16>);
      else go(k_fwd_transpose_q<1, kTileU>);
    } else if (ndim == 2) {
      if (tu == 8) go(k_fwd_transpose_q<2, 8>);
      else if (tu == 16) go(k_fwd_transpose_q<2, 16>);
      else go(k_fwd_transpose_q<2, kTileU>);
    } else {
      if (tu == 8) go(k_fwd_transpose_q<3, 8>);
      else if (tu == 16) go(k_fwd_transpose_q<3, 16>);
      else go(k_fwd_transpose_q<3, kTileU>);
    }
    BH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_slot_reduce, dim3(bh::grid1d(b->count, 256)), dim3(256), 0, st,
                       (const double*)ws, sq.ntt * sq.ntg, (long long)b->count, grad);
    BH_LAUNCH_CHECK();
    return BH_OK;
  }
  TArgs a{};
  a.b = *b;
  a.count = b->count;
  a.modes = mirror ? 2 : 1;
  a.c = c;
  a.ts = ts;
  a.nt = nt;
  a.G = to_segdev(seg);
  a.E = cull_log2;
  a.resid = resid;
  a.resid_f64 = resid_f64;
  a.slots = (double*)ws;
  a.ntg = sh.ntg;
  dim3 grid((unsigned)sh.gx, (unsigned)sh.ntt, (unsigned)sh.ntg);
  const bool f64 = precision == BH_PRECISION_FP64;
  if (ndim == 1) {
    if (f64) hipLaunchKernelGGL((k_fwd_transpose<1, true>), grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL((k_fwd_transpose<1, false>), grid, dim3(256), 0, st, a);
  } else if (ndim == 2) {
    if (f64) hipLaunchKernelGGL((k_fwd_transpose<2, true>), grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL((k_fwd_transpose<2, false>), grid, dim3(256), 0, st, a);
  } else {
    if (f64) hipLaunchKernelGGL((k_fwd_transpose<3, true>), grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL((k_fwd_transpose<3, false>), grid, dim3(256), 0, st, a);
  }
  BH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_slot_reduce, dim3(bh::grid1d(b->count, 256)), dim3(256), 0, st,
                     (const double*)ws, sh.ntt * sh.ntg, (long long)b->count, grad);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

extern "C" int bh_coeff_scatter(const int64_t* idx, const double* vals, int64_t k, double scale,
                                int64_t total, double* coeffs, bh_stream s) {
  bh::clear_error();
  if (total <= 0 || k < 0 || k > total)
    BH_FAIL(BH_ERR_ARG, "bh_coeff_scatter: k=%lld total=%lld", (long long)k, (long long)total);
  if (!coeffs || (k > 0 && (!idx || !vals)))
    BH_FAIL(BH_ERR_ARG, "bh_coeff_scatter: null pointer");
  hipStream_t st = (hipStream_t)s;
  hipLaunchKernelGGL(k_fill_zero, dim3(bh::grid1d(2 * total, 256)), dim3(256), 0, st, coeffs,
                     (long long)(2 * total));
  BH_LAUNCH_CHECK();
  if (k > 0) {
    hipLaunchKernelGGL(k_scatter, dim3(bh::grid1d(k, 256)), dim3(256), 0, st, idx, vals,
                       (long long)k, scale, (long long)total, coeffs);
    BH_LAUNCH_CHECK();
  }
  return BH_OK;
}

extern "C" int bh_coeff_gather(const double* coeffs, int64_t total, const int64_t* idx,
                               int64_t k, double* vals, bh_stream s) {
  bh::clear_error();
  if (total <= 0 || k < 0) BH_FAIL(BH_ERR_ARG, "bh_coeff_gather: k=%lld total=%lld",
                                   (long long)k, (long long)total);
  if (k == 0) return BH_OK;
  if (!coeffs || !idx || !vals) BH_FAIL(BH_ERR_ARG, "bh_coeff_gather: null pointer");
  hipLaunchKernelGGL(k_gather, dim3(bh::grid1d(k, 256)), dim3(256), 0, (hipStream_t)s, coeffs,
                     (long long)total, idx, (long long)k, vals);
  BH_LAUNCH_CHECK();
  return BH_OK;
}

namespace bh {
int transpose_status(int clear, unsigned* bits, hipStream_t s) {
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
