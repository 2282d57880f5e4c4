// Culled real beam summations over segment-decomposed points.
//
// Workgroup = 4 waves; it owns GPB = 4*SPT groups of 64 points (wave w owns
// local groups w, w+4, ...; lane l holds point l of each) and, in the forward
// form, one time level.  Beams stream in chunks of 256: lane i of wave q
// builds (forward: closed form at time t) or loads (static) the record of beam
// chunk + 64q + i and solves, for every segment of the block, the index
// interval on which the beam's envelope 2^-(delta^T Q delta) stays above
// 2^-cull_log2 (Q = w log2(e) ImM / 2; along a segment the exponent is a
// quadratic in the point index).  A wave ballot per group turns those
// intervals into a 64-bit mask per (group, wave q); each wave then evaluates
// exactly the active (record, group) pairs, in record order -- no atomics, so
// the result is deterministic.  Every evaluated term is computed exactly as
// in the dense kernels.
#include <type_traits>
#include <vector>

#include "bh_segcull.hpp"

namespace {

constexpr int kRecChunk = 256;
constexpr int kStaticSplit = 64;   // beam ranges of a precomputed (image) sum
// blocks a culled launch aims at: about 20 rounds of 4 blocks on each of 256
// CUs, so that the blocks still running at the end are a small share
constexpr long long kTargetBlocks = 20480;

struct CulledArgs {
  // forward source: mirror != 0 -> beams [0, count) with mode +1 and their
  // copies [count, 2 count) with mode -1; mirror == 0 -> beam j has mode +1
  // for j < npos, -1 otherwise
  bh_fwd_beams b;
  long long count, total, npos;
  int mirror;
  double phase;   // added to every beam's arg(A) (cycles)
  const double* traj;   // non-null: beam states per (time row, beam) from bh_ode_forward
  double c;
  const double* ts;
  // static source
  const void* recs;
  // common
  SegDev G;
  double E;       // cull exponent (log2 units); <= 0: evaluate all pairs
  double scale;
  void* out;
  unsigned long long* pairs;
  void* part;          // split > 1: partial sums [split][rows][S] (FP64; FP32 in the
                       // FP32-split form, whose output is FP32)
  int split;
  // forward entries in window order (null: entries in table order): position
  // bi holds entry perm[bi] (its beam and mode as above), which can only be
  // active on time rows [row_on[bi], row_off[bi]]
  const int32_t* perm;
  const int2* rows_act;
  // with perm: per chunk of kRecChunk positions the hull [first, last] of
  // its entries' windows; a block skips a chunk outside it (null: no skip)
  const int2* chunk_rows;
  // with perm, 2-D: per position the time-independent constants of its beam
  // (fast_consts2); null: formed from the beam table when needed
  const FastEnt2* fast;
  long long rows;      // output rows (nt for the forward, 1 for the static form)
  int ostride;         // output element o goes to out[o * ostride + ooff]
  int ooff;
};

template <int D>
__device__ inline void shift_phase(RecS<D>& R, double cyc) { R.ph0 += cyc; }
template <int D>
__device__ inline void shift_phase(RecD<D>& R, double cyc) { R.ang += bh::kTwoPi * cyc; }

// One work unit: point-group block bx, time row by, beam range bz of A.split.
template <int D, int SPT, bool F64, bool STATIC>
__device__ __forceinline__ void culled_body(const CulledArgs& A, long long bx, long long by,
                                            long long bz) {
  using Rec = typename RecOf<D, F64>::type;
  constexpr int GPB = 4 * SPT;
  // culling and work unit: U adjacent groups of one run (runs hold an even
  // number of groups, so a unit never straddles two); the FP32-split form
  // evaluates a unit's two groups from one set of pair coefficients
  constexpr int U = F64 ? 1 : 2;
  static_assert(F64 || SPT >= 2, "the FP32-split form needs at least 8 groups per block");
  constexpr int NU = GPB / U;    // units per block
  constexpr int UPW = SPT / U;   // units per wave
  // FP32-split: wave wv owns units 4 i + ((wv + i) mod 4), i < UPW -- in
  // every four consecutive units one, at a position that rotates with i, so
  // that a wave does not get the same place (say, the corner units) of every
  // run; the F64 form: groups wv + 4 i
  auto unit_of = [wv = (int)(threadIdx.x >> 6)](int i) { return 4 * i + ((wv + i) & 3); };
  // precomputed FP32-split records stay in global memory (L2): the masks
  // use each thread's own record from registers, the pair coefficients read
  // theirs from the table, and the block's LDS drops to 13 KB (measured on
  // the 512^2 adjoint: 4.44 against 4.84 ms with the records in LDS)
  constexpr bool GREC = STATIC && !F64;
  __shared__ Rec rec[GREC ? 1 : kRecChunk];
  // masks of the chunk being evaluated and of the next one (a wave builds
  // the next chunk's masks while the other waves still evaluate)
  __shared__ unsigned long long mask[2][NU][4];
  // the block's units grouped by run: per piece its run, first unit, number
  // of units and the index of its first point within the run
  __shared__ int4 runs[NU];
  __shared__ int nruns;
  // FP32-split: per unit its first point and step; per wave the compacted
  // list of active (record, unit) pairs of the chunk and one pass of 64
  // pair coefficients
  __shared__ double gpt[F64 ? 1 : NU][D], gstep[F64 ? 1 : NU][D];
  __shared__ unsigned char plist[F64 ? 1 : 4][F64 ? 1 : 256 * UPW];
  __shared__ PairQ coef[F64 ? 1 : 4][F64 ? 1 : 64];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const long long gbase = bx * GPB;
  const long long t = STATIC ? 0 : by;
  const double tv = STATIC ? 0.0 : A.ts[t];
  const int ng = (int)((A.G.ngroups - gbase) < GPB ? (A.G.ngroups - gbase) : GPB);
  const int nu = (ng + U - 1) / U;
  // group tables in LDS; a wave's accumulators (its SPT groups: unit i,
  // group h at i * U + h) in registers
  double acc[SPT];
  for (int i = 0; i < SPT; ++i) acc[i] = 0.0;
  const double ld = (double)lane;
  const float lf = (float)lane;
  if (wv == 0) {
    // lane u < nu: unit u; a piece starts where the run changes
    int k = -1, kp = -1;
    const bool in = lane < nu;
    if (in) {
      const long long g = gbase + (long long)lane * U;
      k = A.G.grp_seg[g];
      kp = lane > 0 ? A.G.grp_seg[g - U] : -1;
      if constexpr (!F64) {
        for (int a = 0; a < D; ++a) {
          gpt[lane][a] = A.G.pts[g * 64 * D + a];
          gstep[lane][a] = A.G.v[(long long)k * D + a];
        }
      }
    }
    const bool first = in && k != kp;
    const unsigned long long st = __ballot(first);
    if (first) {
      const unsigned long long above = lane < 63 ? st >> (lane + 1) : 0ull;
      const int next = above ? lane + 1 + __builtin_ctzll(above) : nu;
      const long long g = gbase + (long long)lane * U;
      runs[__popcll(st & ((1ull << lane) - 1ull))] =
          make_int4(k, lane, next - lane, (int)((g - A.G.g0[k]) * 64));
    }
    if (lane == 0) nruns = __popcll(st);
  }
  __syncthreads();   // the runs table and the group tables
  unsigned long long npairs = 0;
  const long long total = STATIC ? A.count : A.total;
  const long long nchunk = (total + kRecChunk - 1) / kRecChunk;
  const long long c_lo = nchunk * bz / A.split;
  const long long c_hi = nchunk * (bz + 1) / A.split;
  // Per chunk: each thread builds its record in registers and its wave the
  // chunk's masks (into the mask buffer of the chunk's parity), both before
  // the barrier that waits for the previous chunk's evaluation, so a wave
  // that finished early fills the wait with this work; then the records go
  // to LDS and, after a second barrier, the waves evaluate.  Two chunks
  // apart the mask buffer is rewritten only after every wave has passed the
  // second barrier of the chunk in between, so no wave still reads it.
  // The mask buffers alternate over the chunks the block processes (skipped
  // chunks pass no barrier, so the parity counts processed chunks only).
  int npro = 0;
  for (long long cb = c_lo * kRecChunk; cb < c_hi * kRecChunk && cb < total; cb += kRecChunk) {
    if constexpr (!STATIC) {
      if (A.chunk_rows) {
        // no entry of the chunk can reach a run on this row: the whole
        // block skips it (the test is block-uniform)
        const int2 h = A.chunk_rows[cb / kRecChunk];
        if (t < h.x || t > h.y) continue;
      }
    }
    const int par = npro & 1;
    ++npro;
    const long long bi = cb + threadIdx.x;
    Rec R;
    bool live = bi < total;
    if constexpr (!STATIC) {
      // outside its window the entry's envelope reaches no run: a wave whose
      // entries are all outside skips the records (the branch is uniform)
      if (live && A.rows_act) {
        const int2 w = A.rows_act[bi];
        live = t >= w.x && t <= w.y;
      }
    }
    if (live) {
      if constexpr (STATIC) {
        R = ((const Rec*)A.recs)[bi];
      } else {
        const long long e = A.perm ? (long long)A.perm[bi] : bi;
        const bool neg = A.mirror ? e >= A.count : e >= A.npos;
        const long long j = (A.mirror && neg) ? e - A.count : e;
        // FP32-split 2-D records of alpha = i b beams from the window-order
        // constants (fast_rec2, real arithmetic); other beams, and every beam
        // of a launch without window tables, by the general closed form
        // (the same record to rounding)
        bool fastf = false;
        if constexpr (D == 2 && !F64) {
          if (!A.traj) {
            FastEnt2 F;
            F.iP = -1.0;
            if (A.fast) F = A.fast[bi];
            fastf = F.iP > 0.0;
            if (fastf) fast_rec2(F, neg ? -A.c : A.c, tv, R);
          }
        }
        if (!fastf) {
          double x0[D], p0[D];
          cd al[D];
          for (int a = 0; a < D; ++a) {
            x0[a] = A.b.x0[j * D + a];
            p0[a] = A.b.p0[j * D + a];
            al[a] = mkc(A.b.alpha[2 * (j * D + a)], A.b.alpha[2 * (j * D + a) + 1]);
          }
          State<D> St;
          if (A.traj) {
            // x, p, M (complex), A (complex) of beam bi at this time row
            const double* y = A.traj + (t * A.total + e) * (2 * D + 2 * D * D + 2);
            for (int a = 0; a < D; ++a) {
              St.x[a] = y[a];
              St.p[a] = y[D + a];
              for (int e = 0; e < D; ++e)
                St.M[a][e] = mkc(y[2 * D + 2 * (a * D + e)], y[2 * D + 2 * (a * D + e) + 1]);
            }
            St.A = mkc(y[2 * D + 2 * D * D], y[2 * D + 2 * D * D + 1]);
          } else {
            hom_diag_state_fast<D, !F64>(x0, p0, al, mkc(A.b.a0[2 * j], A.b.a0[2 * j + 1]),
                                         neg ? -A.c : A.c, tv, St);
          }
          make_rec<D>(St, A.b.omega[j], R);
        }
        if (A.phase != 0.0) shift_phase<D>(R, A.phase);
      }
      live = !rec_zero<D>(R);   // zero amplitude contributes exactly 0
    } else {
      zero_rec<D>(R);
    }
    // per-unit activity masks by ballot (word = this wave's 64 records): one
    // closed-form interval per run piece, then two compares per unit
    double Q[D][D];
    const bool cull = A.E > 0.0;
    if (live && cull) env_matrix<D>(R, Q);
    const int nr = __ballot(live) ? nruns : 0;
    if (nr == 0 && lane < nu) mask[par][lane][wv] = 0ull;   // no live record in the wave
    for (int r = 0; r < nr; ++r) {
      const int4 ru = runs[r];
      const int rk = __builtin_amdgcn_readfirstlane(ru.x);
      const int ru0 = __builtin_amdgcn_readfirstlane(ru.y);
      const int run = __builtin_amdgcn_readfirstlane(ru.z);
      const int rj0 = __builtin_amdgcn_readfirstlane(ru.w);
      int lo = 0x3fffffff, hi = -1;   // empty
      if (live) {
        if (cull) {
          int l2, h2;
          seg_interval<D>(A.G, rk, R.x, Q, A.E, l2, h2);
          if (h2 >= l2) { lo = l2; hi = h2; }
        } else {
          lo = 0;
          hi = 0x3fffffff;
        }
      }
      for (int i = 0; i < run; ++i) {
        const int j0 = rj0 + i * 64 * U;
        // hi >= j0 and lo <= j0 + 64 U - 1 as lane masks (v_cmp into SGPRs)
        const unsigned long long m = __builtin_amdgcn_sicmp(hi, j0, 39) &
                                     __builtin_amdgcn_sicmp(lo, j0 + 64 * U - 1, 41);
        if (lane == 0) mask[par][ru0 + i][wv] = m;
        npairs += __popcll(m);
      }
    }
    if constexpr (!GREC) {
      __syncthreads();   // every wave is done with the previous chunk's records
      rec[threadIdx.x] = R;
    }
    __syncthreads();   // this chunk's records and masks are complete
    if constexpr (!F64) {
      // 1. compact the wave's active pairs: unit by unit (its units
      //    unit_of(i)), record order within a unit
      int gs[UPW + 1];
      int base = 0;
      for (int i = 0; i < UPW; ++i) {
        gs[i] = base;
        const int ul = unit_of(i);
        if (ul >= nu) continue;
        for (int q = 0; q < 4; ++q) {
          unsigned long long m = mask[par][ul][q];
          const unsigned lo32 = __builtin_amdgcn_readfirstlane((unsigned)m);
          const unsigned hi32 = __builtin_amdgcn_readfirstlane((unsigned)(m >> 32));
          m = ((unsigned long long)hi32 << 32) | lo32;
          if (!m) continue;
          if (__builtin_amdgcn_inverse_ballot_w64(m)) {   // this lane's bit of m
            const int pos = __builtin_amdgcn_mbcnt_hi(hi32, __builtin_amdgcn_mbcnt_lo(lo32, 0u));
            plist[wv][base + pos] = (unsigned char)(q * 64 + lane);
          }
          base += __popcll(m);
        }
      }
      gs[UPW] = base;
      wave_lds_fence();
      // 2. passes of up to 64 pairs: coefficients lane-parallel, then each
      //    unit's slots in order; lane l of a unit evaluates its points
      //    gpt + l gstep (pa) and gpt + (l + 64) gstep (pb).  A unit's terms
      //    of the chunk form one FP32 chain in record order whatever the pass
      //    boundaries, so every tile shape gives the same sums.
      float pa[UPW], pb[UPW];
      for (int i = 0; i < UPW; ++i) {
        pa[i] = 0.f;
        pb[i] = 0.f;
      }
      for (int b = 0; b < base; b += 64) {
        const int e = b + lane;
        if (e < base) {
          int gi = 0;
          for (int i = 1; i < UPW; ++i) gi = (e >= gs[i]) ? i : gi;
          const int ul = unit_of(gi);
          if constexpr (GREC)
            coef[wv][lane] = pair_quad<D>(((const Rec*)A.recs)[cb + plist[wv][e]], gpt[ul],
                                          gstep[ul]);
          else
            coef[wv][lane] = pair_quad<D>(rec[plist[wv][e]], gpt[ul], gstep[ul]);
        }
        wave_lds_fence();
        for (int i = 0; i < UPW; ++i) {
          // wave-uniform bounds: a scalar loop
          const int lo = __builtin_amdgcn_readfirstlane((gs[i] > b ? gs[i] : b) - b);
          const int hi = __builtin_amdgcn_readfirstlane((gs[i + 1] < b + 64 ? gs[i + 1] : b + 64) - b);
          if (lo >= hi) continue;
          // two pairs per step, each at lanes l and l + 64
          int sl = lo;
          for (; sl + 1 < hi; sl += 2) {
            const PairQ c1 = coef[wv][sl], c2 = coef[wv][sl + 1];
            eval_quad2(c1, ld, lf, pa[i], pb[i]);
            eval_quad2(c2, ld, lf, pa[i], pb[i]);
          }
          if (sl < hi) eval_quad2(coef[wv][sl], ld, lf, pa[i], pb[i]);
        }
        wave_lds_fence();   // the slots are consumed before the next pass
      }
      for (int i = 0; i < UPW; ++i) {
        acc[i * 2] += (double)pa[i];
        acc[i * 2 + 1] += (double)pb[i];
      }
    } else
    for (int i = 0; i < SPT; ++i) {
      const int gl = wv + 4 * i;
      if (gl >= ng) continue;
      using Part = typename std::conditional<F64, double, float>::type;
      Part part = 0;
      double pti[D];
      {
        const long long slot = (gbase + gl) * 64 + lane;
        for (int a = 0; a < D; ++a) pti[a] = A.G.pts[slot * D + a];
      }
      for (int q = 0; q < 4; ++q) {
        unsigned long long m = mask[par][gl][q];
        const unsigned lo32 = __builtin_amdgcn_readfirstlane((unsigned)m);
        const unsigned hi32 = __builtin_amdgcn_readfirstlane((unsigned)(m >> 32));
        m = ((unsigned long long)hi32 << 32) | lo32;
        while (m) {
          const int r1 = __builtin_ctzll(m);
          m &= m - 1;
          const Rec R1 = rec[q * 64 + r1];
          double d1[D];
          for (int a = 0; a < D; ++a) d1[a] = pti[a] - R1.x[a];
          if (m) {
            const int r2 = __builtin_ctzll(m);
            m &= m - 1;
            const Rec R2 = rec[q * 64 + r2];
            double d2[D];
            for (int a = 0; a < D; ++a) d2[a] = pti[a] - R2.x[a];
            if constexpr (F64) {
              const double e1 = eval_f64<D>(R1, d1), e2 = eval_f64<D>(R2, d2);
              part += e1;
              part += e2;
            } else {
              eval_split<D>(R1, d1, part);
              eval_split<D>(R2, d2, part);
            }
          } else {
            if constexpr (F64) part += eval_f64<D>(R1, d1);
            else eval_split<D>(R1, d1, part);
          }
        }
      }
      acc[i] += (double)part;
    }
  }
  if (A.pairs && lane == 0) atomicAdd(A.pairs, npairs * U);
  for (int i = 0; i < SPT; ++i) {
    // unit unit_of(i / U), group i % U (F64: group wv + 4 i)
    const int gl = F64 ? wv + 4 * i : unit_of(i / U) * U + i % U;
    if (gl >= ng) continue;
    const int col = A.G.col[(gbase + gl) * 64 + lane];
    if (col < 0) continue;
    const long long o = STATIC ? col : t * A.G.S + col;
    if (A.split > 1) {
      if constexpr (F64) ((double*)A.part)[bz * A.rows * A.G.S + o] = acc[i];
      else ((float*)A.part)[bz * A.rows * A.G.S + o] = (float)acc[i];
    } else {
      const long long oo = o * A.ostride + A.ooff;
      if constexpr (F64) ((double*)A.out)[oo] = A.scale * acc[i];
      else ((float*)A.out)[oo] = (float)(A.scale * acc[i]);
    }
  }
}

// The FP32-split form runs at four waves per SIMD (its 128-VGPR budget costs
// a few spilled registers outside the evaluation loop; the extra wave hides
// more latency than the spills cost); the FP64 form keeps the compiler's
// choice.
template <int D, int SPT, bool F64, bool STATIC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(F64 ? 1 : 4)))
void k_culled_sum(CulledArgs A) {
  culled_body<D, SPT, F64, STATIC>(A, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Many forward sums in one launch: solve s owns blocks [start[s], start[s+1])
// of the 1-D grid, laid out as (time row, point-group block).
struct BatchDesc {
  CulledArgs a;
  long long ngx;   // point-group blocks per time row
};

template <int D, int SPT, bool F64>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(F64 ? 1 : 4)))
void k_culled_batch(const BatchDesc* __restrict__ d,
                                                      const long long* __restrict__ start,
                                                      int nsolve) {
  const long long b = blockIdx.x;
  int lo = 0, hi = nsolve - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const CulledArgs A = d[lo].a;
  const long long local = b - start[lo];
  culled_body<D, SPT, F64, false>(A, local % d[lo].ngx, local / d[lo].ngx, 0);
}

// ordered sum of the split partials (deterministic)
template <typename P>
__global__ void k_combine(const P* __restrict__ part, int split, long long n, double scale,
                          int f64, void* out, int ostride, int ooff) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int z = 0; z < split; ++z) s += (double)part[(long long)z * n + i];
  const long long o = i * ostride + ooff;
  if (f64) ((double*)out)[o] = scale * s;
  else ((float*)out)[o] = (float)(scale * s);
}

__global__ void k_zero_out(void* out, long long nbytes4) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < nbytes4) ((unsigned*)out)[i] = 0u;
}

// static records (same as the dense static path)
template <int D, bool F64>
__global__ void k_static_records_c(bh_static_beams b, void* recs) {
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

// ---------------------------------------------------------------------------
// Activity windows of the forward entries.  For a closed-form beam with
// alpha = i b (b > 0, the forward beams' form), 1/M(t) = 1/M0 + k t P_perp
// (k = c0 / |p|), so Im M(t) = (B^-1 + k^2 t^2 P_perp B P_perp)^-1 decreases
// (in the positive semidefinite order) with |t|: the envelope metric
// Q(t) = w log2(e) Im M(t) / 2 dominates Q(tmax) for every |t| <= tmax.  An
// entry can therefore be active on a run only while its centre x(t) = x0 +
// c0 p^ t lies within sqrt(E) of the run in the metric of Q(tmax): a capsule
// around the run, whose intersection with the line x(t) is one time interval
// (a convex condition).  The runs are lengthened by 2.01 points at each end
// and E by 1e-6 relative, covering the one-point widening of the culled
// sums' intervals and the FP32 rounding of their envelope coefficients, so
// every (entry, row) whose masks the kernel could set lies inside its
// window.  Any other beam (Re alpha != 0, a metric that is not positive
// definite, a time grid that is not increasing) gets every row.
// ---------------------------------------------------------------------------
struct WinArgs {
  bh_fwd_beams b;
  long long count, total, npos;
  int mirror;
  double c;
  const double* ts;
  long long nt;
  SegDev G;
  double E;
  const double* trange;      // [0] max |t|, [1] 1 if ts is not increasing
  unsigned long long* key;   // (row_on << 32) | entry
  long long* idx;            // entry
  int2* win;                 // per entry: first and last row
};

__global__ void k_time_range(const double* __restrict__ ts, long long nt, double* out) {
  __shared__ double smax[256];
  __shared__ int sbad[256];
  double m = 0.0;
  int bad = 0;
  for (long long i = threadIdx.x; i < nt; i += blockDim.x) {
    m = fmax(m, fabs(ts[i]));
    if (i > 0 && !(ts[i] > ts[i - 1])) bad = 1;
  }
  smax[threadIdx.x] = m;
  sbad[threadIdx.x] = bad;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + o]);
      sbad[threadIdx.x] |= sbad[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = smax[0];
    out[1] = sbad[0] ? 1.0 : 0.0;
  }
}

// times t where |y0 + vy t - s| <= R for a point s of the segment [ya, yb]
// (Euclidean, D dimensions): the union of the two end balls and the slab
// part, one interval (lo > hi when empty)
template <int D>
__device__ inline void capsule_times(const double* y0, const double* vy, const double* ya,
                                     const double* yb, double R, double& lo, double& hi) {
  lo = 1e300;
  hi = -1e300;
  double vv = 0.0;
  for (int a = 0; a < D; ++a) vv += vy[a] * vy[a];
  const double R2 = R * R;
  auto quad = [&](double a2, double a1, double a0, double& l, double& h) {
    // a2 t^2 + a1 t + a0 <= 0 with a2 >= 0
    if (a2 > 0.0) {
      const double disc = a1 * a1 - 4.0 * a2 * a0;
      if (disc < 0.0) { l = 1e300; h = -1e300; return; }
      const double sq = sqrt(disc);
      l = (-a1 - sq) / (2.0 * a2);
      h = (-a1 + sq) / (2.0 * a2);
    } else if (a1 != 0.0) {
      const double tt = -a0 / a1;
      if (a1 > 0.0) { l = -1e300; h = tt; } else { l = tt; h = 1e300; }
    } else if (a0 <= 0.0) {
      l = -1e300; h = 1e300;
    } else {
      l = 1e300; h = -1e300;
    }
  };
  for (int end = 0; end < 2; ++end) {
    const double* cc = end ? yb : ya;
    double B = 0.0, C = -R2;
    for (int a = 0; a < D; ++a) {
      const double d = y0[a] - cc[a];
      B += d * vy[a];
      C += d * d;
    }
    double l, h;
    quad(vv, 2.0 * B, C, l, h);
    if (l <= h) { lo = fmin(lo, l); hi = fmax(hi, h); }
  }
  double u[D], uu = 0.0;
  for (int a = 0; a < D; ++a) { u[a] = yb[a] - ya[a]; uu += u[a] * u[a]; }
  if (uu > 0.0) {
    double w0u = 0.0, vyu = 0.0, w0w0 = 0.0, w0vy = 0.0;
    for (int a = 0; a < D; ++a) {
      const double w0 = y0[a] - ya[a];
      w0u += w0 * u[a];
      vyu += vy[a] * u[a];
      w0w0 += w0 * w0;
      w0vy += w0 * vy[a];
    }
    // squared distance to the line through the segment, times |u|^2
    double l, h;
    quad(fmax(vv * uu - vyu * vyu, 0.0), 2.0 * (w0vy * uu - w0u * vyu), w0w0 * uu - w0u * w0u - R2 * uu,
         l, h);
    // the foot of the perpendicular inside the segment: 0 <= w.u <= |u|^2
    double sl, sh;
    if (vyu != 0.0) {
      const double t1 = -w0u / vyu, t2 = (uu - w0u) / vyu;
      sl = fmin(t1, t2);
      sh = fmax(t1, t2);
    } else if (w0u >= 0.0 && w0u <= uu) {
      sl = -1e300; sh = 1e300;
    } else {
      sl = 1e300; sh = -1e300;
    }
    l = fmax(l, sl);
    h = fmin(h, sh);
    if (l <= h) { lo = fmin(lo, l); hi = fmax(hi, h); }
  }
}

template <int D>
__global__ __launch_bounds__(128) void k_beam_windows(WinArgs W) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= W.total) return;
  const bool neg = W.mirror ? e >= W.count : e >= W.npos;
  const long long j = (W.mirror && neg) ? e - W.count : e;
  int on = 0, off = (int)W.nt - 1;
  double x0[D], p0[D];
  cd al[D];
  bool pure = true;
  for (int a = 0; a < D; ++a) {
    x0[a] = W.b.x0[j * D + a];
    p0[a] = W.b.p0[j * D + a];
    al[a] = mkc(W.b.alpha[2 * (j * D + a)], W.b.alpha[2 * (j * D + a) + 1]);
    pure = pure && al[a].re == 0.0 && al[a].im > 0.0;
  }
  const cd a0 = mkc(W.b.a0[2 * j], W.b.a0[2 * j + 1]);
  const double w = W.b.omega[j];
  const double tmax = W.trange[0];
  if (W.E > 0.0 && W.trange[1] == 0.0 && pure) {
    if (a0.re == 0.0 && a0.im == 0.0) {
      on = (int)W.nt;   // zero amplitude: never evaluated
      off = -1;
    } else {
      const double c0 = neg ? -W.c : W.c;
      State<D> St;
      hom_diag_state_fast<D>(x0, p0, al, a0, c0, tmax, St);
      // metric Q(tmax) and its Cholesky factor (lower): Q = L L^T
      double Q[D][D], L[D][D] = {};
      const double sc = 0.5 * w * bh::kLog2e;
      for (int a = 0; a < D; ++a)
        for (int b = 0; b < D; ++b) Q[a][b] = sc * 0.5 * (St.M[a][b].im + St.M[b][a].im);
      bool pd = true;
      for (int a = 0; a < D && pd; ++a) {
        for (int b = 0; b <= a; ++b) {
          double v = Q[a][b];
          for (int k = 0; k < b; ++k) v -= L[a][k] * L[b][k];
          if (a == b) {
            if (!(v > 0.0)) { pd = false; break; }
            L[a][a] = sqrt(v);
          } else {
            L[a][b] = v / L[b][b];
          }
        }
      }
      if (pd) {
        double np2 = 0.0;
        for (int a = 0; a < D; ++a) np2 += p0[a] * p0[a];
        const double inp = 1.0 / sqrt(np2);
        // y = L^T x
        auto ty = [&](const double* x, double* y) {
          for (int a = 0; a < D; ++a) {
            double v = 0.0;
            for (int b = a; b < D; ++b) v += L[b][a] * x[b];
            y[a] = v;
          }
        };
        double vel[D], y0[D], vy[D];
        for (int a = 0; a < D; ++a) vel[a] = c0 * p0[a] * inp;
        ty(x0, y0);
        ty(vel, vy);
        const double R = sqrt(W.E * (1.0 + 1e-6));
        double tlo = 1e300, thi = -1e300;
        for (long long k = 0; k < W.G.nseg; ++k) {
          const double n = (double)W.G.n[k];
          double pa[D], pb[D], ya[D], yb[D];
          for (int a = 0; a < D; ++a) {
            const double v = W.G.v[k * D + a];
            pa[a] = W.G.p0[k * D + a] - 2.01 * v;
            pb[a] = W.G.p0[k * D + a] + (n - 1.0 + 2.01) * v;
          }
          ty(pa, ya);
          ty(pb, yb);
          double l, h;
          capsule_times<D>(y0, vy, ya, yb, R, l, h);
          // only the part inside the time grid counts (a run the beam's line
          // crossed before the first time must not stretch the window)
          l = fmax(l, W.ts[0]);
          h = fmin(h, W.ts[W.nt - 1]);
          if (l <= h) { tlo = fmin(tlo, l); thi = fmax(thi, h); }
        }
        if (!(tlo <= thi)) {
          on = (int)W.nt;
          off = -1;
        } else {
          // first row with ts >= tlo, last row with ts <= thi (ts increasing)
          long long a = 0, b = W.nt;
          while (a < b) { const long long m = (a + b) >> 1; if (W.ts[m] < tlo) a = m + 1; else b = m; }
          on = (int)a;
          a = 0; b = W.nt;
          while (a < b) { const long long m = (a + b) >> 1; if (W.ts[m] <= thi) a = m + 1; else b = m; }
          off = (int)a - 1;
        }
      }
    }
  }
  W.win[e] = make_int2(on, off);
  if (!W.key) return;   // windows only, no sort keys
  // order: eight ranges of the table (so that every beam range of the launch
  // draws from all of them alike), then window start and end, then entry
  const unsigned long long rg = (unsigned long long)((e * 8) / W.total);
  W.key[e] = (rg << 61) | ((unsigned long long)(unsigned)on << 45) |
             ((unsigned long long)(unsigned)(off + 1) << 29) | (unsigned long long)e;
  W.idx[e] = e;
}

// entries in window order: perm[bi] = sorted entry, rows_act[bi] its window
// and, in 2-D (fast != null), fast[bi] its beam's fast_consts2
__global__ void k_window_order(const long long* __restrict__ sidx, const int2* __restrict__ win,
                               long long total, int32_t* __restrict__ perm,
                               int2* __restrict__ rows_act, bh_fwd_beams b, long long count,
                               int mirror, FastEnt2* __restrict__ fast) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long long e = sidx[i];
  perm[i] = (int32_t)e;
  rows_act[i] = win[e];
  if (fast) {
    const long long j = (mirror && e >= count) ? e - count : e;
    double x0[2], p0[2];
    cd al[2];
    for (int a = 0; a < 2; ++a) {
      x0[a] = b.x0[j * 2 + a];
      p0[a] = b.p0[j * 2 + a];
      al[a] = mkc(b.alpha[2 * (j * 2 + a)], b.alpha[2 * (j * 2 + a) + 1]);
    }
    fast[i] = fast_consts2(x0, p0, al, mkc(b.a0[2 * j], b.a0[2 * j + 1]), b.omega[j]);
  }
}

// per chunk of kRecChunk positions, the hull of its entries' windows
static_assert(kRecChunk == 256, "one block of four waves per chunk");
__global__ __launch_bounds__(256) void k_chunk_rows(const int2* __restrict__ rows_act,
                                                    long long total, int2* __restrict__ hull) {
  __shared__ int smin[4], smax[4];
  const long long i = blockIdx.x * (long long)kRecChunk + threadIdx.x;
  int lo = 0x7fffffff, hi = -1;
  if (i < total) {
    const int2 w = rows_act[i];
    lo = w.x;
    hi = w.y;
  }
  for (int o = 32; o > 0; o >>= 1) {
    lo = min(lo, __shfl_xor(lo, o));
    hi = max(hi, __shfl_xor(hi, o));
  }
  if ((threadIdx.x & 63) == 0) {
    smin[threadIdx.x >> 6] = lo;
    smax[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    hull[blockIdx.x] = make_int2(min(min(smin[0], smin[1]), min(smin[2], smin[3])),
                                 max(max(smax[0], smax[1]), max(smax[2], smax[3])));
}

size_t round256(size_t x) { return (x + 255) / 256 * 256; }

// bytes of the window tables of `total` entries
size_t window_bytes(long long total) {
  const long long P = bh::sort_pairs_capacity(total);
  const long long nchunk = (total + kRecChunk - 1) / kRecChunk;
  return round256(16) + 4 * round256((size_t)P * 8) + round256((size_t)total * 8) +
         round256((size_t)total * 4) + round256((size_t)total * 8) +
         round256((size_t)nchunk * 8) + round256((size_t)total * sizeof(FastEnt2));
}

// the window tables of `total` entries carved from ws, in this order
struct WinTables {
  double* trange;
  unsigned long long *key, *key2;
  long long *idx, *idx2;
  int2* win;
  int32_t* perm;
  int2* rows_act;
  int2* hull;
  FastEnt2* fast;
};
WinTables window_tables(void* ws, long long total) {
  char* p = (char*)ws;
  auto take = [&](size_t bytes) { char* r = p; p += round256(bytes); return (void*)r; };
  const long long P = bh::sort_pairs_capacity(total);
  const long long nchunk = (total + kRecChunk - 1) / kRecChunk;
  WinTables w;
  w.trange = (double*)take(16);
  w.key = (unsigned long long*)take((size_t)P * 8);
  w.idx = (long long*)take((size_t)P * 8);
  w.key2 = (unsigned long long*)take((size_t)P * 8);
  w.idx2 = (long long*)take((size_t)P * 8);
  w.win = (int2*)take((size_t)total * 8);
  w.perm = (int32_t*)take((size_t)total * 4);
  w.rows_act = (int2*)take((size_t)total * 8);
  w.hull = (int2*)take((size_t)nchunk * 8);
  w.fast = (FastEnt2*)take((size_t)total * sizeof(FastEnt2));
  return w;
}

// the windows and the entry order of a forward launch, carved from ws
int build_windows(int ndim, const CulledArgs& a, const double* ts, long long nt, void* ws,
                  const int32_t** perm, const int2** rows_act, const int2** chunk_rows,
                  const FastEnt2** fast, hipStream_t st) {
  const long long total = a.total;
  const WinTables wt = window_tables(ws, total);
  double* trange = wt.trange;
  unsigned long long* key = wt.key;
  long long* idx = wt.idx;
  unsigned long long* key2 = wt.key2;
  long long* idx2 = wt.idx2;
  int2* win = wt.win;
  int32_t* pm = wt.perm;
  int2* ra = wt.rows_act;
  const long long nchunk = (total + kRecChunk - 1) / kRecChunk;
  int2* hull = wt.hull;
  hipLaunchKernelGGL(k_time_range, dim3(1), dim3(256), 0, st, ts, nt, trange);
  WinArgs W{};
  W.b = a.b;
  W.count = a.count;
  W.total = total;
  W.npos = a.npos;
  W.mirror = a.mirror;
  W.c = a.c;
  W.ts = ts;
  W.nt = nt;
  W.G = a.G;
  W.E = a.E;
  W.trange = trange;
  W.key = key;
  W.idx = idx;
  W.win = win;
  const unsigned g = bh::grid1d(total, 128);
  if (ndim == 1) hipLaunchKernelGGL(k_beam_windows<1>, dim3(g), dim3(128), 0, st, W);
  else if (ndim == 2) hipLaunchKernelGGL(k_beam_windows<2>, dim3(g), dim3(128), 0, st, W);
  else hipLaunchKernelGGL(k_beam_windows<3>, dim3(g), dim3(128), 0, st, W);
  unsigned long long* ko;
  long long* io;
  if (int rc = bh::sort_pairs(key, idx, key2, idx2, total, &ko, &io, st)) return rc;
  FastEnt2* fe = ndim == 2 ? wt.fast : nullptr;
  hipLaunchKernelGGL(k_window_order, dim3(bh::grid1d(total, 256)), dim3(256), 0, st, io, win,
                     total, pm, ra, a.b, a.count, a.mirror, fe);
  BH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_chunk_rows, dim3((unsigned)nchunk), dim3(kRecChunk), 0, st, ra, total,
                     hull);
  BH_LAUNCH_CHECK();
  *perm = pm;
  *rows_act = ra;
  *chunk_rows = hull;
  *fast = fe;
  return BH_OK;
}

}  // namespace

namespace bh {

size_t beam_windows_bytes(long long total) {
  return round256(16) + round256((size_t)total * 8);
}

int beam_windows(int ndim, const bh_fwd_beams* b, int mirror, double c, const double* ts,
                 long long nt, const bh_segments* seg, double E, void* ws, const int2** win,
                 hipStream_t st) {
  char* p = (char*)ws;
  double* trange = (double*)p;
  int2* w = (int2*)(p + round256(16));
  WinArgs W{};
  W.b = *b;
  W.count = b->count;
  W.total = b->count * (mirror ? 2 : 1);
  W.npos = b->count;
  W.mirror = mirror ? 1 : 0;
  W.c = c;
  W.ts = ts;
  W.nt = nt;
  W.G = to_segdev(seg);
  W.E = E;
  W.trange = trange;
  W.key = nullptr;
  W.idx = nullptr;
  W.win = w;
  hipLaunchKernelGGL(k_time_range, dim3(1), dim3(256), 0, st, ts, nt, trange);
  const unsigned g = grid1d(W.total, 128);
  if (ndim == 1) hipLaunchKernelGGL(k_beam_windows<1>, dim3(g), dim3(128), 0, st, W);
  else if (ndim == 2) hipLaunchKernelGGL(k_beam_windows<2>, dim3(g), dim3(128), 0, st, W);
  else hipLaunchKernelGGL(k_beam_windows<3>, dim3(g), dim3(128), 0, st, W);
  BH_LAUNCH_CHECK();
  *win = w;
  return BH_OK;
}

}  // namespace bh

namespace {

// Launch shape: groups per block GPB in {32, 16, 8, 4} (fewer groups per
// block = more blocks but more record work per group), then a beam split so
// that the grid covers the chip several times over (blocks resident per CU
// are limited to ~4 by LDS).
struct Shape {
  int gpb;
  int split;
};

inline Shape culled_shape(long long G, long long ny, long long nchunk, bool have_part,
                          int min_gpb, int max_split) {
  // kTargetBlocks (2414 rows, 2^-40 cut: split 4 gives 30.4 ms, split 2
  // 31.2 ms, split 1 32.5 ms on the 512^2 forward; at 2^-32, split 8 26.65
  // ms against 26.83 ms at 4)
  const long long target = kTargetBlocks;
  Shape sh{min_gpb, 1};
  const int opts[4] = {32, 16, 8, 4};
  for (int o : opts) {
    if ((G + o - 1) / o * ny >= 1024 || o <= min_gpb) {
      sh.gpb = o;
      break;
    }
  }
  const long long blocks = (G + sh.gpb - 1) / sh.gpb * ny;
  if (have_part && blocks < target) {
    long long sp = (target + blocks - 1) / blocks;
    if (sp > max_split) sp = max_split;
    if (sp > nchunk) sp = nchunk;
    sh.split = (int)(sp < 1 ? 1 : sp);
  }
  return sh;
}

template <int D, bool F64, bool STATIC>
int launch_culled(CulledArgs a, long long ny, long long nchunk, void* part_ws, size_t part_bytes,
                  hipStream_t s) {
  const long long G = a.G.ngroups;
  const size_t per_split = (size_t)ny * a.G.S * sizeof(double);
  // precomputed (image) sums: 16 groups per block and up to 64 beam ranges
  // (the 512^2 adjoint: 2.46 ms against 2.74 ms with 16 ranges, 2.56 ms
  // with 32; earlier, with the records in LDS, 16 groups and 16 ranges took
  // 4.76 ms against 5.00 ms with 8 and 8)
  const Shape sh = culled_shape(G, ny, nchunk, part_ws != nullptr && part_bytes >= 2 * per_split,
                                F64 ? 4 : (STATIC ? 16 : 8), STATIC ? kStaticSplit : 8);
  a.split = sh.split;
  if (sh.split > 1 && (size_t)sh.split * per_split > part_bytes)
    a.split = (int)(part_bytes / per_split);
  a.part = (double*)part_ws;
  a.rows = ny;
  dim3 grid((unsigned)((G + sh.gpb - 1) / sh.gpb), (unsigned)ny, (unsigned)a.split);
  if (sh.gpb == 32) hipLaunchKernelGGL((k_culled_sum<D, 8, F64, STATIC>), grid, dim3(256), 0, s, a);
  else if (sh.gpb == 16) hipLaunchKernelGGL((k_culled_sum<D, 4, F64, STATIC>), grid, dim3(256), 0, s, a);
  else if (sh.gpb == 8) hipLaunchKernelGGL((k_culled_sum<D, 2, F64, STATIC>), grid, dim3(256), 0, s, a);
  else if constexpr (F64) hipLaunchKernelGGL((k_culled_sum<D, 1, F64, STATIC>), grid, dim3(256), 0, s, a);
  if (a.split > 1) {
    const long long n = ny * a.G.S;
    using P = typename std::conditional<F64, double, float>::type;
    hipLaunchKernelGGL(k_combine<P>, dim3(bh::grid1d(n, 256)), dim3(256), 0, s, (const P*)part_ws,
                       a.split, n, a.scale, (int)F64, a.out, a.ostride, a.ooff);
  }
  return BH_OK;
}

}  // namespace

extern "C" size_t bh_forward_sum_culled_workspace_size(int64_t nt, int64_t S) {
  if (nt <= 0 || S <= 0) return 0;
  return (size_t)8 * nt * S * sizeof(double);
}

// The partial sums of the beam ranges, then the activity windows and the
// window order of the forward entries (mirror: 2 count of them).
extern "C" size_t bh_forward_culled_workspace_size(int64_t count, int mirror, int64_t nt,
                                                   int64_t S) {
  if (nt <= 0 || S <= 0 || count <= 0) return bh_forward_sum_culled_workspace_size(nt, S);
  return round256(bh_forward_sum_culled_workspace_size(nt, S)) +
         window_bytes(count * (mirror ? 2 : 1));
}

namespace {
int forward_culled(const char* name, const bh_fwd_beams* b, int ndim, long long npos, int mirror,
                   double c, const double* ts, int64_t nt, const bh_segments* seg,
                   double cull_log2, int precision, int complex_out, void* out,
                   unsigned long long* pairs, void* ws, size_t ws_bytes, bh_stream s,
                   const double* traj = nullptr) {
  bh::clear_error();
  if (int rc = check_segments(seg, ndim)) return rc;
  if (!b || !ts || !out) BH_FAIL(BH_ERR_ARG, "%s: null pointer", name);
  if (nt <= 0 || nt > 65535) BH_FAIL(BH_ERR_ARG, "%s: nt=%lld", name, (long long)nt);
  if (precision != BH_PRECISION_FP64 && precision != BH_PRECISION_FP32_SPLIT)
    BH_FAIL(BH_ERR_ARG, "%s: precision=%d", name, precision);
  if (!mirror && (npos < 0 || npos > b->count))
    BH_FAIL(BH_ERR_ARG, "%s: npos=%lld outside [0, %lld]", name, npos, (long long)b->count);
  hipStream_t st = (hipStream_t)s;
  const bool f64 = precision == BH_PRECISION_FP64;
  const int ostride = complex_out ? 2 : 1;
  if (b->count == 0) {
    const long long n4 = (long long)nt * seg->S * ostride * (f64 ? 2 : 1);
    hipLaunchKernelGGL(k_zero_out, dim3(bh::grid1d(n4, 256)), dim3(256), 0, st, out, n4);
    BH_LAUNCH_CHECK();
    return BH_OK;
  }
  if (!b->x0 || !b->p0 || !b->omega || !b->alpha || !b->a0)
    BH_FAIL(BH_ERR_ARG, "%s: null beam array", name);
  CulledArgs a{};
  a.b = *b;
  a.count = b->count;
  a.total = b->count * (mirror ? 2 : 1);
  a.npos = mirror ? b->count : npos;
  a.mirror = mirror ? 1 : 0;
  a.c = c;
  a.ts = ts;
  a.traj = traj;
  a.recs = nullptr;
  a.G = to_segdev(seg);
  a.E = cull_log2;
  a.out = out;
  a.pairs = pairs;
  a.ostride = ostride;
  const long long nchunk = (a.total + kRecChunk - 1) / kRecChunk;
  // activity windows (closed-form beams, a cut, room for the tables after
  // the partial sums): the entries in window order
  size_t part_bytes = ws_bytes;
  a.perm = nullptr;
  a.rows_act = nullptr;
  a.chunk_rows = nullptr;
  const size_t part_std = round256(bh_forward_sum_culled_workspace_size(nt, seg->S));
  if (!traj && cull_log2 > 0.0 && ws && ws_bytes >= part_std + window_bytes(a.total) &&
      a.total < (1ll << 29) && nt < 65535) {
    if (int rc = build_windows(ndim, a, ts, nt, (char*)ws + part_std, &a.perm, &a.rows_act,
                               &a.chunk_rows, &a.fast, st))
      return rc;
    part_bytes = part_std;
  }
  // real sum: 2 Re sum A G.  Complex sum: Re = half of it, Im = half of the
  // real sum with every arg(A) lowered by a quarter cycle (cos(x - pi/2) = sin x)
  for (int part = 0; part < (complex_out ? 2 : 1); ++part) {
    a.scale = complex_out ? 0.5 : 1.0;
    a.phase = part ? -0.25 : 0.0;
    a.ooff = part;
    if (ndim == 1) { if (f64) launch_culled<1, true, false>(a, nt, nchunk, ws, part_bytes, st); else launch_culled<1, false, false>(a, nt, nchunk, ws, part_bytes, st); }
    else if (ndim == 2) { if (f64) launch_culled<2, true, false>(a, nt, nchunk, ws, part_bytes, st); else launch_culled<2, false, false>(a, nt, nchunk, ws, part_bytes, st); }
    else { if (f64) launch_culled<3, true, false>(a, nt, nchunk, ws, part_bytes, st); else launch_culled<3, false, false>(a, nt, nchunk, ws, part_bytes, st); }
    BH_LAUNCH_CHECK();
  }
  return BH_OK;
}
}  // namespace

extern "C" int bh_forward_sum_culled(const bh_fwd_beams* b, int ndim, int mirror, double c,
                                     const double* ts, int64_t nt, const bh_segments* seg,
                                     double cull_log2, int precision, void* out,
                                     unsigned long long* pairs, void* ws, size_t ws_bytes,
                                     bh_stream s) {
  return forward_culled("bh_forward_sum_culled", b, ndim, b ? (long long)b->count : 0, mirror, c,
                        ts, nt, seg, cull_log2, precision, 0, out, pairs, ws, ws_bytes, s);
}

extern "C" int bh_forward_sum_traj(const bh_fwd_beams* b, int ndim, int mirror,
                                   const double* traj, const double* ts, int64_t nt,
                                   const bh_segments* seg, double cull_log2, int precision,
                                   void* out, unsigned long long* pairs, void* ws,
                                   size_t ws_bytes, bh_stream s) {
  if (!traj) {
    bh::clear_error();
    BH_FAIL(BH_ERR_ARG, "bh_forward_sum_traj: null trajectory table");
  }
  return forward_culled("bh_forward_sum_traj", b, ndim, b ? (long long)b->count : 0, mirror, 0.0,
                        ts, nt, seg, cull_log2, precision, 0, out, pairs, ws, ws_bytes, s, traj);
}

extern "C" int bh_forward_sum_modes(const bh_fwd_beams* b, int ndim, int64_t npos, int mirror,
                                    double c, const double* ts, int64_t nt,
                                    const bh_segments* seg, double cull_log2, int precision,
                                    int complex_out, void* out, unsigned long long* pairs,
                                    void* ws, size_t ws_bytes, bh_stream s) {
  return forward_culled("bh_forward_sum_modes", b, ndim, npos, mirror, c, ts, nt, seg, cull_log2,
                        precision, complex_out, out, pairs, ws, ws_bytes, s);
}

extern "C" size_t bh_static_sum_culled_workspace_size(int64_t count, int ndim, int64_t P) {
  if (count <= 0 || P <= 0) return 0;
  const size_t rec = (bh_static_sum_workspace_size(count, ndim) + 255) / 256 * 256;
  // the launch has at least P / 2048 blocks per beam range (at most 32
  // groups of 64 points per block), so it splits at most
  // ceil(kTargetBlocks / (P / 2048)) ways
  const long long blocks = P / 2048 > 1 ? P / 2048 : 1;
  long long split = (kTargetBlocks + blocks - 1) / blocks;
  if (split > kStaticSplit) split = kStaticSplit;
  return rec + (size_t)split * P * sizeof(double);
}

extern "C" int bh_static_sum_culled(const bh_static_beams* b, int ndim, const bh_segments* seg,
                                    double scale, double cull_log2, int precision, void* out,
                                    unsigned long long* pairs, void* ws, size_t ws_bytes,
                                    bh_stream s) {
  bh::clear_error();
  if (int rc = check_segments(seg, ndim)) return rc;
  if (!b || !out) BH_FAIL(BH_ERR_ARG, "bh_static_sum_culled: null pointer");
  if (precision != BH_PRECISION_FP64 && precision != BH_PRECISION_FP32_SPLIT)
    BH_FAIL(BH_ERR_ARG, "bh_static_sum_culled: precision=%d", precision);
  hipStream_t st = (hipStream_t)s;
  const bool f64 = precision == BH_PRECISION_FP64;
  if (b->count == 0) {
    const long long n4 = (long long)seg->S * (f64 ? 2 : 1);
    hipLaunchKernelGGL(k_zero_out, dim3(bh::grid1d(n4, 256)), dim3(256), 0, st, out, n4);
    BH_LAUNCH_CHECK();
    return BH_OK;
  }
  const size_t rec = (bh_static_sum_workspace_size(b->count, ndim) + 255) / 256 * 256;
  if (!ws || ws_bytes < rec)
    BH_FAIL(BH_ERR_WORKSPACE, "bh_static_sum_culled: workspace %zu < %zu", ws_bytes, rec);
  CulledArgs a{};
  a.count = b->count;
  a.total = b->count;
  a.ostride = 1;
  a.ooff = 0;
  a.recs = ws;
  a.G = to_segdev(seg);
  a.E = cull_log2;
  a.scale = scale;
  a.out = out;
  a.pairs = pairs;
  void* part = ws_bytes > rec ? (char*)ws + rec : nullptr;
  const size_t part_bytes = ws_bytes > rec ? ws_bytes - rec : 0;
  const long long nchunk = (a.count + kRecChunk - 1) / kRecChunk;
  const unsigned gr = bh::grid1d(b->count, 128);
#define BH_SC(D, F)                                                                           \
  do {                                                                                        \
    hipLaunchKernelGGL((k_static_records_c<D, F>), dim3(gr), dim3(128), 0, st, *b, ws);        \
    launch_culled<D, F, true>(a, 1, nchunk, part, part_bytes, st);                            \
  } while (0)
  if (ndim == 1) { if (f64) BH_SC(1, true); else BH_SC(1, false); }
  else if (ndim == 2) { if (f64) BH_SC(2, true); else BH_SC(2, false); }
  else { if (f64) BH_SC(3, true); else BH_SC(3, false); }
#undef BH_SC
  BH_LAUNCH_CHECK();
  return BH_OK;
}

namespace {
// a solve of the batch visits its entries in activity-window order when it
// brings a window workspace and the windows apply (a cut, closed-form beams)
bool batch_windowed(const bh_batch_solve& q, double cull_log2) {
  return q.win && cull_log2 > 0.0 && q.beams.count > 0 && 2 * q.beams.count < (1ll << 29) &&
         q.nt > 0 && q.nt < 65535;
}
}  // namespace

extern "C" size_t bh_forward_windows_workspace_size(int64_t count, int mirror) {
  if (count <= 0) return 0;
  return window_bytes(count * (mirror ? 2 : 1));
}

extern "C" int bh_forward_batch_windows(const bh_batch_solve* solves, int nsolve, int ndim,
                                        double cull_log2, bh_stream s) {
  bh::clear_error();
  if (!solves || nsolve <= 0)
    BH_FAIL(BH_ERR_ARG, "bh_forward_batch_windows: null pointer or nsolve=%d", nsolve);
  if (ndim < 1 || ndim > 3) BH_FAIL(BH_ERR_ARG, "bh_forward_batch_windows: ndim=%d", ndim);
  hipStream_t st = (hipStream_t)s;
  for (int i = 0; i < nsolve; ++i) {
    const bh_batch_solve& q = solves[i];
    if (!batch_windowed(q, cull_log2)) continue;
    if (int rc = check_segments(&q.seg, ndim)) return rc;
    CulledArgs a{};
    a.b = q.beams;
    a.count = q.beams.count;
    a.total = 2 * q.beams.count;
    a.npos = q.beams.count;
    a.mirror = 1;
    a.c = q.c;
    a.G = to_segdev(&q.seg);
    a.E = cull_log2;
    const int32_t* pm;
    const int2* ra;
    const int2* hull;
    const FastEnt2* fe;
    if (int rc = build_windows(ndim, a, q.ts, q.nt, q.win, &pm, &ra, &hull, &fe, st)) return rc;
  }
  return BH_OK;
}

extern "C" size_t bh_forward_batch_plan_size(int nsolve) {
  if (nsolve <= 0) return 0;
  return ((size_t)nsolve * sizeof(BatchDesc) + 255) / 256 * 256 + (size_t)(nsolve + 1) * 8;
}

extern "C" int bh_forward_batch_plan(const bh_batch_solve* solves, int nsolve, int ndim,
                                     double cull_log2, int gpb, void* plan, size_t plan_bytes,
                                     int64_t* nblocks, bh_stream s) {
  bh::clear_error();
  if (!solves || nsolve <= 0 || !plan || !nblocks)
    BH_FAIL(BH_ERR_ARG, "bh_forward_batch_plan: null pointer or nsolve=%d", nsolve);
  if (gpb != 4 && gpb != 8 && gpb != 16 && gpb != 32)
    BH_FAIL(BH_ERR_ARG, "bh_forward_batch_plan: gpb=%d (4, 8, 16 or 32)", gpb);
  if (ndim < 1 || ndim > 3) BH_FAIL(BH_ERR_ARG, "bh_forward_batch_plan: ndim=%d", ndim);
  const size_t need = bh_forward_batch_plan_size(nsolve);
  if (plan_bytes < need)
    BH_FAIL(BH_ERR_WORKSPACE, "bh_forward_batch_plan: plan %zu < %zu bytes", plan_bytes, need);
  std::vector<BatchDesc> d((size_t)nsolve);
  std::vector<long long> start((size_t)nsolve + 1);
  long long nb = 0;
  for (int i = 0; i < nsolve; ++i) {
    const bh_batch_solve& q = solves[i];
    if (int rc = check_segments(&q.seg, ndim)) return rc;
    if (q.nt < 0 || q.nt > 65535 || !q.out || (q.nt > 0 && !q.ts))
      BH_FAIL(BH_ERR_ARG, "bh_forward_batch_plan: solve %d: nt=%lld or null pointer", i,
              (long long)q.nt);
    if (q.beams.count > 0 && (!q.beams.x0 || !q.beams.p0 || !q.beams.omega ||
                              !q.beams.alpha || !q.beams.a0))
      BH_FAIL(BH_ERR_ARG, "bh_forward_batch_plan: solve %d: null beam array", i);
    CulledArgs a{};
    a.b = q.beams;
    a.count = q.beams.count;
    a.total = 2 * q.beams.count;
    a.npos = q.beams.count;
    a.mirror = 1;
    a.phase = 0.0;
    a.c = q.c;
    a.ts = q.ts;
    a.recs = nullptr;
    a.G = to_segdev(&q.seg);
    a.E = cull_log2;
    a.scale = 1.0;
    a.out = q.out;
    a.pairs = nullptr;
    a.part = nullptr;
    a.split = 1;
    a.rows = q.nt;
    a.ostride = 1;
    a.ooff = 0;
    if (batch_windowed(q, cull_log2)) {
      // the tables bh_forward_batch_windows fills before every run
      const WinTables wt = window_tables(q.win, a.total);
      a.perm = wt.perm;
      a.rows_act = wt.rows_act;
      a.chunk_rows = wt.hull;
      a.fast = ndim == 2 ? wt.fast : nullptr;
    }
    d[i].a = a;
    d[i].ngx = (q.seg.ngroups + gpb - 1) / gpb;
    start[i] = nb;
    nb += d[i].ngx * q.nt;
  }
  start[nsolve] = nb;
  if (nb > 0x7fffffffLL) BH_FAIL(BH_ERR_ARG, "bh_forward_batch_plan: %lld blocks", nb);
  const size_t dbytes = ((size_t)nsolve * sizeof(BatchDesc) + 255) / 256 * 256;
  hipStream_t st = (hipStream_t)s;
  BH_HIP(hipMemcpyAsync(plan, d.data(), (size_t)nsolve * sizeof(BatchDesc),
                        hipMemcpyHostToDevice, st));
  BH_HIP(hipMemcpyAsync((char*)plan + dbytes, start.data(), (size_t)(nsolve + 1) * 8,
                        hipMemcpyHostToDevice, st));
  BH_HIP(hipStreamSynchronize(st));   // the host staging vectors end here
  *nblocks = nb;
  return BH_OK;
}

extern "C" int bh_forward_batch_run(const void* plan, int nsolve, int64_t nblocks, int ndim,
                                    int precision, int gpb, unsigned long long* pairs,
                                    bh_stream s) {
  bh::clear_error();
  if (!plan || nsolve <= 0) BH_FAIL(BH_ERR_ARG, "bh_forward_batch_run: null plan");
  if (precision != BH_PRECISION_FP64 && precision != BH_PRECISION_FP32_SPLIT)
    BH_FAIL(BH_ERR_ARG, "bh_forward_batch_run: precision=%d", precision);
  if (pairs) BH_FAIL(BH_ERR_UNSUPPORTED, "bh_forward_batch_run: pair counting is not wired");
  if (nblocks <= 0) return BH_OK;
  const bool f64 = precision == BH_PRECISION_FP64;
  if (!f64 && gpb == 4)
    BH_FAIL(BH_ERR_ARG, "bh_forward_batch_run: the FP32-split form takes 8, 16 or 32 groups "
                        "per block (gpb=4 is FP64 only)");
  const size_t dbytes = ((size_t)nsolve * sizeof(BatchDesc) + 255) / 256 * 256;
  const BatchDesc* d = (const BatchDesc*)plan;
  const long long* start = (const long long*)((const char*)plan + dbytes);
  hipStream_t st = (hipStream_t)s;
  dim3 grid((unsigned)nblocks);
#define BH_BATCH(DD, SS)                                                                       \
  do {                                                                                        \
    if (f64) hipLaunchKernelGGL((k_culled_batch<DD, SS, true>), grid, dim3(256), 0, st, d,     \
                                start, nsolve);                                               \
    else hipLaunchKernelGGL((k_culled_batch<DD, SS, false>), grid, dim3(256), 0, st, d, start, \
                            nsolve);                                                          \
  } while (0)
#define BH_BATCH_D(DD)                                                                         \
  do {                                                                                        \
    if (gpb == 32) BH_BATCH(DD, 8);                                                           \
    else if (gpb == 16) BH_BATCH(DD, 4);                                                      \
    else if (gpb == 8) BH_BATCH(DD, 2);                                                       \
    else hipLaunchKernelGGL((k_culled_batch<DD, 1, true>), grid, dim3(256), 0, st, d, start,   \
                            nsolve);                                                          \
  } while (0)
  if (ndim == 1) BH_BATCH_D(1);
  else if (ndim == 2) BH_BATCH_D(2);
  else BH_BATCH_D(3);
#undef BH_BATCH_D
#undef BH_BATCH
  BH_LAUNCH_CHECK();
  return BH_OK;
}
