/*
 * beamax_amd: C-ABI of the MI355X (gfx950) MSGB forward/adjoint operator.
 *
 * Every entry point takes plain pointers and sizes; device pointers are HIP
 * device memory owned by the caller (the Python host layer allocates them as
 * PyTorch-ROCm tensors).  Work is enqueued on the caller's stream `s`
 * (hipStream_t; NULL = the legacy default stream); nothing synchronises the
 * device except where a function says so.  Every function returns 0 on
 * success or a negative BH_ERR_* code; bh_last_error() returns the message of
 * the last failure on the calling thread.  Complex arrays are interleaved
 * (re, im) FP64.  The library allocates no device memory except inside
 * bh_topn / bh_*_sum / bh_mswpt_* through the workspace the caller passes.
 *
 * Reference interfaces replaced (elma16/beamax 0.2.2, file:line):
 *   bh_mswpt_setup      MSWPT.__init__ sum_gsquare + packed filters
 *                       (beamax/transforms.py:423-589)
 *   bh_mswpt_forward    MSWPT.forward (beamax/transforms.py:696-733, 591-694)
 *                       + compute_coefficients 'pos_only' epilogue
 *                       (beamax/solvers/msgb_solvers/forward_solver_utils.py:716-786)
 *   bh_mswpt_inverse    MSWPT.inverse (beamax/transforms.py:735-850)
 *   bh_fft              utils.unitary_fft/unitary_ifft building block
 *                       (beamax/utils/fft.py:5-38)
 *   bh_topn             _threshold_top_n (forward_solver_utils.py:65-89)
 *   bh_threshold        _threshold_hard / _percentile / _hard_reassign /
 *                       _bao_energy / _perc_max_abs (forward_solver_utils.py:14-178)
 *   bh_forward_params   compute_forward_parameters (forward_solver_utils.py:255-373)
 *   bh_forward_sum      compute_forward_result with solve_hom_diag and
 *                       compute_gaussian_beam_real (forward_solver_utils.py:677-713,
 *                       beamax/gb/gb_solvers.py:356-422, beamax/gb/core.py:194-329)
 *   bh_solve_hom_diag   gb_solvers.solve_hom_diag (beamax/gb/gb_solvers.py:356-422)
 *   bh_solve_hom_TR     gb_solvers.solve_hom_TR (beamax/gb/gb_solvers.py:484-594)
 *   bh_adjoint_source   _form_adjoint_source (msgb_solver.py:68-120)
 *   bh_adjoint_params   compute_adj_parameters + compute_TR_parameters +
 *                       solve_hom_TR at t = 0 (adjoint_solver_utils.py:75-220,
 *                       tr_solver_utils.py:205-435, gb_solvers.py:484-594)
 *   bh_tr_params        compute_TR_parameters + solve_hom_TR at t = 0 for
 *                       MSGBSolver.time_reversal (msgb_solver.py:691-746, 984-1111)
 *   bh_mswpt_combine_velocity  compute_coefficients with dpdt != 0
 *                       (forward_solver_utils.py:755-786)
 *   bh_fourier_crop_time  prepare_data_domain_for_msgb Fourier crop
 *                       (examples/reconstruction/2d_time_reversal_and_adjoint.py:111-165)
 *   bh_speed_eval, bh_ode_forward, bh_forward_sum_traj: heterogeneous media
 *                       (utils/interp.py:158-270, gb_solvers.py:787-1016 with
 *                       compute_gaussian_beam_real on the saved states)
 *   bh_boundary_terminal, bh_ode_static: the adjoint / time reversal in a
 *                       heterogeneous medium (compute_TR_parameters with c(x_T),
 *                       tr_solver_utils.py:205-435, adjoint_solver_utils.py:75-220;
 *                       solve_ODE_batch_t from t_start to 0, gb_solvers.py:1017-1113)
 *   bh_forward_batch_plan, bh_forward_batch_run: a batch of forward sums in
 *                       one launch (no reference counterpart)
 *   bh_forward_sum_modes  the complex-input forward (msgb_solver.py:635-689)
 *                       and the complex beam sum (gb/core.py:105-168)
 *   bh_forward_sum_culled, bh_static_sum_culled: the same sums with a certified
 *                       envelope cut (no reference counterpart; same seams)
 *   bh_forward_transpose, bh_coeff_scatter, bh_coeff_gather, bh_mswpt_adjoint: the exact
 *                       transpose of the frozen-index forward (beam sum ->
 *                       coefficients -> MSWPT analysis); the reference gets
 *                       it only through JAX autodiff of forward
 *                       (msgb_solver.py:858-982), it has no entry point of its own
 *   bh_static_sum       compute_TR_result + compute_gaussian_beam_real_TR at
 *                       t = 0 + c^-2 weight (tr_solver_utils.py:677-750,
 *                       beamax/gb/core.py:332-425, msgb_solver.py:123-127)
 */
#ifndef BEAMAX_AMD_H
#define BEAMAX_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* bh_stream;

#define BH_OK 0
#define BH_ERR_ARG -1
#define BH_ERR_HIP -2
#define BH_ERR_UNSUPPORTED -3
#define BH_ERR_WORKSPACE -4

/* Bits of bh_device_status: a kernel met a beam or coefficient index outside
 * [-1, total) (-1 is the reference's padding index of an empty threshold slot,
 * whose value is 0: forward_solver_utils.py:58-61, 174-177; it stays an inert
 * zero-amplitude beam).
 * Such a beam gets a NaN amplitude, so the outputs it reaches are NaN. */
#define BH_STATUS_FWD_INDEX 1u   /* bh_forward_params */
#define BH_STATUS_ADJ_INDEX 2u   /* bh_adjoint_params, bh_tr_params, bh_boundary_terminal */
#define BH_STATUS_COEFF_INDEX 4u /* bh_coeff_scatter, bh_coeff_gather */

#define BH_MAX_DIM 3
#define BH_MAX_LEVELS 12

#define BH_WINDOW_RECTANGULAR 1
#define BH_WINDOW_RECTANGULAR_MIRROR 2

#define BH_INPUT_REAL_SPATIAL 0
#define BH_INPUT_COMPLEX_SPATIAL 1
#define BH_INPUT_COMPLEX_FOURIER 2

#define BH_PRECISION_FP64 0
#define BH_PRECISION_FP32_SPLIT 1

/* Dyadic Fourier tiling + MSWPT layout (host struct; `centres` is device). */
typedef struct {
  int32_t ndim;                           /* 1..3 */
  int32_t num_levels;                     /* 1..BH_MAX_LEVELS */
  int32_t redundancy;                     /* 1 or 2 */
  int32_t windowing;                      /* BH_WINDOW_* */
  int64_t N[BH_MAX_DIM];                  /* grid shape (even) */
  int32_t aspect[BH_MAX_DIM];             /* box aspect ratio per axis */
  int32_t box_length[BH_MAX_LEVELS];      /* smallest-axis box length per level */
  int32_t support[BH_MAX_LEVELS][BH_MAX_DIM]; /* coefficient block shape per level */
  int64_t boxes_cumsum[BH_MAX_LEVELS + 1];
  int64_t coeffs_cumsum[BH_MAX_LEVELS + 1];
  const int32_t* centres;                 /* device, total_boxes x ndim, raw centres */
} bh_tiling;

/* Forward beams (device SoA, FP64). Beam i: x0[i*d..], p0[i*d..], omega[i],
 * alpha[2*(i*d+a)..] (diagonal of M0), a0[2*i..] (amplitude incl. the
 * coefficient).  bh_forward_sum mirrors every beam to mode -1 when asked. */
typedef struct {
  int64_t count;
  double* x0;
  double* p0;
  double* omega;
  double* alpha;
  double* a0;
} bh_fwd_beams;

/* Beams frozen at one time level (device SoA, FP64): adjoint beams at t = 0.
 * M is the full d x d complex Hessian, row-major, interleaved. */
typedef struct {
  int64_t count;
  double* x;
  double* p;
  double* M;
  double* A;
  double* omega;
} bh_static_beams;

const char* bh_last_error(void);
int bh_version(void);

/* The status bits the index guards of the current device's kernels have set
 * since the last clear (BH_STATUS_*), OR-ed into *status; clear != 0 resets
 * them.  Synchronises stream s (not capturable in a graph).  No reference
 * counterpart: it reports caller errors (an index the reference's thresholds
 * never produce) that the asynchronous kernels cannot return as a status. */
int bh_device_status(int clear, uint32_t* status, bh_stream s);

/* Batched 1-D complex FFT (FP64) along one strided axis, in place.
 * Transform t in [0,batch) starts at (t/inner)*outer_stride +
 * (t%inner)*inner_stride, element e at + e*elem_stride (units: complex
 * elements).  dir = -1 forward (exp(-2 pi i jk/n)), +1 inverse; result
 * multiplied by `scale`.  `ws` needs bh_fft_workspace_size bytes. */
size_t bh_fft_workspace_size(int64_t n, int64_t batch);
int bh_fft(double* data, int64_t n, int64_t batch, int64_t inner, int64_t inner_stride,
           int64_t outer_stride, int64_t elem_stride, int dir, double scale, void* ws,
           size_t ws_bytes, bh_stream s);

/* MSWPT plan tables: sum_gsquare (prod N doubles, centred index order) and
 * the packed first-box filter of every level (sum_l prod support[l] doubles,
 * levels concatenated). */
int bh_mswpt_setup(const bh_tiling* t, double* sum_gsquare, double* gpacked, bh_stream s);
size_t bh_mswpt_workspace_size(const bh_tiling* t);
/* coeffs[total] = scale * MSWPT(input); if half_mask != 0 the second half
 * (flat entries >= len_l/2) of every level is set to zero. */
int bh_mswpt_forward(const bh_tiling* t, const double* sum_gsquare, const double* gpacked,
                     const double* input, int input_kind, double scale, int half_mask,
                     double* coeffs, void* ws, size_t ws_bytes, bh_stream s);
/* cpos = 0.5 (a + i b / vg) and (cneg nullable) cneg = 0.5 (a - i b / vg)
 * for the coefficients a = MSWPT(p0), b = MSWPT(dpdt) (b NULL: dpdt = 0);
 * vg = c |2 pi centre / L| of each coefficient's box; with half_mask the
 * second half of every level of cpos is zeroed ('pos_only'). */
int bh_mswpt_combine_velocity(const bh_tiling* t, const double* grid_size, double c,
                              const double* a, const double* b, int half_mask, double* cpos,
                              double* cneg, bh_stream s);
/* out[prod N] complex = MSWPT^-1(coeffs) in the spatial (output_fourier = 0)
 * or centred Fourier (1) domain. */
int bh_mswpt_inverse(const bh_tiling* t, const double* gpacked, const double* coeffs,
                     int output_fourier, double* out, void* ws, size_t ws_bytes,
                     bh_stream s);

/* top_n: the k largest |c| under the order (|c|, index), returned ascending.
 * |c| = sqrt(re*re + im*im), each operation correctly rounded. */
size_t bh_topn_workspace_size(int64_t n, int64_t k);
int bh_topn(const double* c, int64_t n, int64_t k, int64_t* idx_out, double* val_out,
            void* ws, size_t ws_bytes, bh_stream s);
/* bh_topn without a workspace argument and without any allocation: its
 * workspace is a fixed array of this library in each device's memory
 * (about 40 MiB, held while the library is loaded), for callers that cannot
 * allocate device memory (code under HIP graph capture, allocation-free
 * serving loops).  No counterpart in the reference.  Limits: n <= 4194304,
 * k <= 262144 (BH_ERR_UNSUPPORTED beyond).  Every call on a device shares
 * that device's array: calls that overlap on one device (other streams,
 * other threads) race and must be serialised by the caller.  The first call
 * on a device looks its array up (hipGetSymbolAddress). */
int bh_topn_noalloc(const double* c, int64_t n, int64_t k, int64_t* idx_out, double* val_out,
                    bh_stream s);

/* The other thresholding strategies (forward_solver_utils.py:14-178):
 * strategy 0 hard (|c| > val), 1 percentile (|c| > the val-th percentile of
 * |c|, linear interpolation), 2 hard_reassign (|c| >= val max|c|, values
 * scaled by sqrt(total / retained energy)), 3 bao_energy (|c| |centre|^d >
 * val; needs the tiling t), 4 perc_max_abs (|c| > val max|c|).  Writes the
 * selected indices (ascending) and values to idx_out[0..count) and
 * val_out[0..count) (capacity n each) and count to *count_out (device). */
size_t bh_threshold_workspace_size(int64_t n);
int bh_threshold(const double* c, int64_t n, int strategy, double val, const bh_tiling* t,
                 int64_t* idx_out, double* val_out, int64_t* count_out, void* ws,
                 size_t ws_bytes, bh_stream s);

/* Forward beam parameters for flat coefficient indices idx[k] with values
 * vals[k] (complex); grid_size = N*dx (host, ndim). */
int bh_forward_params(const bh_tiling* t, const double* grid_size, const int64_t* idx,
                      const double* vals, int64_t k, bh_fwd_beams out, bh_stream s);

/* Real beam sum at sensors:  out[t*S + s] = sum_b 2|A_b(t)| cos(...) exp(...).
 * ts: device, nt values; sensors: device S x d; mirror != 0 adds every beam
 * again with mode -1 (after the +1 block).  precision BH_PRECISION_FP64 writes
 * double, BH_PRECISION_FP32_SPLIT writes float. periodic (host, d) applies
 * the minimum-image offset with period domain_size (host, d). */
int bh_forward_sum(const bh_fwd_beams* b, int ndim, int mirror, double c, const double* ts,
                   int64_t nt, const double* sensors, int64_t S, const double* domain_size,
                   const int32_t* periodic, int precision, void* out, bh_stream s);

/* Closed-form trajectories (SolverFn): x0,p0 (b x d), alpha (b x d complex),
 * a0 (b complex), mode (b); outputs xt,pt (b,nt,d), Mt (b,nt,d,d complex),
 * At (b,nt complex). */
int bh_solve_hom_diag(int64_t b, int ndim, const double* x0, const double* p0,
                      const double* alpha, const double* a0, const double* mode, double c,
                      const double* ts, int64_t nt, double* xt, double* pt, double* Mt,
                      double* At, bh_stream s);
/* solve_hom_TR as a SolverFn (gb_solvers.py:484-594): terminal states xT, pT
 * (b x d), mT (b x d x d complex), aT (b complex), mode (b); per-beam times
 * ts (b x nt, row-major), each measured from ts[beam][0]; outputs xt, pt
 * (b,nt,d), Mt (b,nt,d,d complex), At (b,nt complex). */
int bh_solve_hom_TR(int64_t b, int ndim, const double* xT, const double* pT, const double* mT,
                    const double* aT, const double* mode, double c, const double* ts, int64_t nt,
                    double* xt, double* pt, double* Mt, double* At, bh_stream s);

/* Adjoint source -c^2 Re IFFT_t(2 pi i f FFT_t(w r)) along axis 0 of an
 * (nt, ns) FP64 array; window (nullable) has nt entries (window_2d = 0) or
 * nt*ns (window_2d = 1); c_src device ns. */
size_t bh_adjoint_source_workspace_size(int64_t nt, int64_t ns);
int bh_adjoint_source(const double* data, int64_t nt, int64_t ns, double dt,
                      const double* window, int window_2d, const double* c_src, double* out,
                      void* ws, size_t ws_bytes, bh_stream s);

/* Fourier crop along time (the multi-face adjoint glue): columns cols[ns]
 * (nullable = 0..ns-1) of an (nt, ns_in) real array, times col_weight[ns]
 * (nullable), are resampled to nt_out samples by keeping the centred nt_out
 * temporal frequencies of the unitary FFT; out is (nt_out, ns). */
size_t bh_fourier_crop_time_workspace_size(int64_t nt, int64_t nt_out, int64_t ns);
int bh_fourier_crop_time(const double* data, int64_t nt, int64_t ns_in, const int64_t* cols,
                         const double* col_weight, int64_t ns, int64_t nt_out, double* out,
                         void* ws, size_t ws_bytes, bh_stream s);

/* Adjoint beams: data-domain tiling t, data grid size L (host, d), selected
 * indices/values (k), detector plane normal axis / value / inward sign (host
 * geometry), homogeneous speed c, relative guard; beams are propagated to t = 0.
 * `swap_axes` permutes spatial axes 0 and 1 of the result (face rotation). */
int bh_adjoint_params(const bh_tiling* t, const double* data_grid_size, const int64_t* idx,
                      const double* vals, int64_t k, int normal_axis, double normal_value,
                      double inward_sign, double c, double guard, int swap_axes,
                      bh_static_beams out, bh_stream s);

/* Time-reversal beams: compute_TR_parameters + solve_hom_TR to t = 0
 * (no B^-1 factor and no evanescent exclusion); same arguments as
 * bh_adjoint_params without the guard. */
int bh_tr_params(const bh_tiling* t, const double* data_grid_size, const int64_t* idx,
                 const double* vals, int64_t k, int normal_axis, double normal_value,
                 double inward_sign, double c, int swap_axes, bh_static_beams out, bh_stream s);

/* out[P] = scale * sum_b field_b(points[p]) for beams frozen at one time. */
size_t bh_static_sum_workspace_size(int64_t count, int ndim);
int bh_static_sum(const bh_static_beams* b, int ndim, const double* points, int64_t P,
                  double scale, int precision, void* out, void* ws, size_t ws_bytes,
                  bh_stream s);

/* Evaluation points decomposed into straight, equally spaced runs
 * ("segments") packed into groups of 64 points (one point per lane).
 * Segment k: points p0[k] + j*v[k], j < n[k]; its groups are g0[k] ...
 * Group g holds pts[g*64 .. g*64+63] (d doubles each, padded slots repeat a
 * valid point) and writes them to output columns col[g*64 + lane] (-1 =
 * padding).  All arrays are device memory. */
typedef struct {
  int64_t nseg;
  int64_t ngroups;
  int64_t S;                 /* number of output columns */
  const double* seg_p0;      /* nseg x d */
  const double* seg_v;       /* nseg x d */
  const int32_t* seg_n;      /* nseg */
  const int32_t* seg_g0;     /* nseg */
  const int32_t* grp_seg;    /* ngroups */
  const double* pts;         /* ngroups*64 x d */
  const int32_t* col;        /* ngroups*64 */
} bh_segments;

/* Culled forms of bh_forward_sum / bh_static_sum: a (beam, 64-point group)
 * pair is evaluated only if some point of the group has envelope
 * 2^-exponent >= 2^-cull_log2 relative to the beam's own 2|A| (exact closed
 * form along each segment, widened by one point); every evaluated term is
 * the same as in the dense form.  cull_log2 <= 0 evaluates every pair.
 * Dropped terms are each below 2^-cull_log2 * 2|A_b| (in absolute value).
 * `pairs` (nullable, device) is incremented by the number of evaluated
 * (beam, group) pairs; each pair is 64 lane evaluations.  Workspace: the
 * forward form may run without (ws NULL); with the size returned by
 * bh_forward_sum_culled_workspace_size the beam range is split over more
 * workgroups and the partial sums are added in a fixed order.  The static
 * form needs at least bh_static_sum_workspace_size bytes (records) and uses
 * bh_static_sum_culled_workspace_size for the split.
 * With the larger size of bh_forward_culled_workspace_size (count = beams of
 * the table, mirror as passed to the sum) the forward forms of closed-form
 * beams also compute each (beam, mode)'s time window -- the rows on which
 * its envelope can reach a run, from the metric of its envelope at the
 * latest time, which bounds every earlier one for alpha = i b, b > 0 -- and
 * visit the beams in window order, so that workgroups skip the records of
 * beams outside their window; the evaluated pairs are the same, the order in
 * which they are added differs. */
size_t bh_forward_sum_culled_workspace_size(int64_t nt, int64_t S);
size_t bh_forward_culled_workspace_size(int64_t count, int mirror, int64_t nt, int64_t S);
int bh_forward_sum_culled(const bh_fwd_beams* b, int ndim, int mirror, double c,
                          const double* ts, int64_t nt, const bh_segments* seg, double cull_log2,
                          int precision, void* out, unsigned long long* pairs, void* ws,
                          size_t ws_bytes, bh_stream s);
/* bh_forward_sum_culled with per-beam modes and an optional complex sum:
 * mirror = 0: beams [0, npos) propagate with mode +1 and [npos, count) with
 * mode -1 (the complex-input path of msgb_solver.py:635-689, whose positive
 * and negative coefficient sets are thresholded separately); mirror != 0:
 * every beam twice, modes +1 and -1 (npos ignored), as bh_forward_sum_culled.  complex_out = 0
 * writes the real sum 2 Re sum A G (compute_gaussian_beam_real,
 * gb/core.py:194-329); complex_out = 1 writes sum A G (compute_gaussian_beam
 * summed over beams, gb/core.py:105-168) as interleaved complex (nt, S), FP32
 * pairs in FP32-split, FP64 pairs in FP64.  Same workspace as
 * bh_forward_sum_culled. */
int bh_forward_sum_modes(const bh_fwd_beams* b, int ndim, int64_t npos, int mirror, double c,
                         const double* ts, int64_t nt, const bh_segments* seg, double cull_log2,
                         int precision, int complex_out, void* out, unsigned long long* pairs,
                         void* ws, size_t ws_bytes, bh_stream s);
/* Many independent forward sums in one launch (a batch of solves of
 * different grids, beam counts, time grids and sensors).  Each solve: a
 * mirrored beam table, its speed c, device time grid ts[nt], segment tables
 * and output (nt, seg.S) (float in FP32-split, double in FP64).
 * bh_forward_batch_plan writes the device plan (bh_forward_batch_plan_size
 * bytes; it copies from the host and synchronises the stream, so call it
 * outside graph capture); bh_forward_batch_run is one kernel launch over all
 * solves (capturable), same terms and summation order per point as
 * bh_forward_sum_culled without a beam split.  gpb = point groups per block
 * (4, 8, 16, 32; the tile shape an autotuner picks).  No reference
 * counterpart: the reference runs solves one by one. */
typedef struct bh_batch_solve {
  bh_fwd_beams beams;
  double c;
  const double* ts;
  int64_t nt;
  bh_segments seg;
  void* out;
  /* optional (null: none): bh_forward_windows_workspace_size(beams.count, 1)
   * bytes of device memory for this solve's activity windows */
  void* win;
} bh_batch_solve;
/* Activity windows (the forward entries' time rows of possible activity, as
 * bh_forward_sum_culled builds them) of every solve that brings a window
 * workspace, for a cut > 0; enqueues kernels only (capturable).  A plan
 * built from the same solves then visits each solve's entries in window
 * order and skips their inactive rows; call it again whenever the beams
 * change (before every bh_forward_batch_run). */
size_t bh_forward_windows_workspace_size(int64_t count, int mirror);
int bh_forward_batch_windows(const bh_batch_solve* solves, int nsolve, int ndim,
                             double cull_log2, bh_stream s);
size_t bh_forward_batch_plan_size(int nsolve);
int bh_forward_batch_plan(const bh_batch_solve* solves, int nsolve, int ndim, double cull_log2,
                          int gpb, void* plan, size_t plan_bytes, int64_t* nblocks, bh_stream s);
int bh_forward_batch_run(const void* plan, int nsolve, int64_t nblocks, int ndim, int precision,
                         int gpb, unsigned long long* pairs, bh_stream s);
size_t bh_static_sum_culled_workspace_size(int64_t count, int ndim, int64_t P);
int bh_static_sum_culled(const bh_static_beams* b, int ndim, const bh_segments* seg,
                         double scale, double cull_log2, int precision, void* out,
                         unsigned long long* pairs, void* ws, size_t ws_bytes, bh_stream s);

/* Grid-valued sound speed (utils/interp.py:158-270 make_c_function_from_grid):
 * method 0 n-linear on grid values, 1 cubic B-spline on prefiltered
 * coefficients; boundary 0 nearest (clamp), 1 wrap, 2 reflect, 3 mirror;
 * values: device array of shape N (C order); lam: absorption coefficient of
 * the amplitude equation (Domain.lam). */
typedef struct {
  int32_t ndim, method, boundary;
  int64_t N[3];
  double spacing[3], origin[3];
  const double* values;
  double lam;
} bh_speed;
/* c, grad c (n x ndim) and Hessian (n x ndim x ndim) at points x (n x ndim). */
int bh_speed_eval(const bh_speed* sp, const double* x, int64_t n, double* c, double* grad,
                  double* hess, bh_stream s);
/* Beam ODEs in the medium (coupled_rhs_absorption + solve_ODE_base,
 * gb_solvers.py:787-1016): each beam of b (and its mode -1 copy when mirror
 * != 0, appended after the count originals) from M0 = diag(alpha), A0 = a0 at
 * ts[0] (mirror < 0: every beam with mode -1 instead), written at every ts[k]
 * to traj[(k * total + beam) * NS], NS = 2d +
 * 2d^2 + 2 doubles: x, p, M (complex, row-major), A (complex).  Adaptive
 * Dormand-Prince 5(4) with rtol/atol; status[beam] = 0, 1 (more than
 * max_steps steps), 2 (non-finite state). */
int bh_ode_forward(const bh_fwd_beams* b, int ndim, int mirror, const bh_speed* sp,
                   const double* ts, int64_t nt, double rtol, double atol, int64_t max_steps,
                   double* traj, int* status, bh_stream s);
/* bh_adjoint_params (time_reversal = 0) or bh_tr_params (1) with the sound
 * speed read from a speed model at each beam's detector point x_T (its value
 * sets the momentum, B^-1 and xdot; its gradient the pdot terms of
 * mT_inverse), leaving each beam at its terminal state instead of
 * propagating it: out holds (x_T, p_T, M_T, a_T, omega), t_start[i] the
 * terminal time and mode[i] the branch sign (device arrays of k doubles). */
int bh_boundary_terminal(const bh_tiling* t, const double* data_grid_size, const int64_t* idx,
                         const double* vals, int64_t k, int normal_axis, double normal_value,
                         double inward_sign, const bh_speed* sp, double guard, int time_reversal,
                         int swap_axes, bh_static_beams out, double* t_start, double* mode,
                         bh_stream s);
/* Beams taken from t_start[i] to t_end[i] (t_end NULL: t = 0, the adjoint's
 * terminal-value beams) in place through the beam ODEs, full complex M:
 * solve_ODE_batch_t with ts_i = (t_start_i, t_end_i) (gb_solvers.py:1017-1113,
 * coupled_rhs 870-908) when lam = 0, the coupled_rhs_absorption amplitude
 * term (820-867) otherwise.  Adaptive Dormand-Prince 5(4); first step dt0
 * (<= 0: the whole interval).  Beams with zero amplitude, t_end = t_start or
 * mode 0 are left as they are.  status[i] = 0, 1 (more than max_steps
 * steps) or 2 (non-finite state; the beam is left at its start state). */
int bh_ode_static(const bh_static_beams* b, int ndim, const bh_speed* sp, const double* t_start,
                  const double* t_end, const double* mode, double lam, double rtol, double atol,
                  double dt0, int64_t max_steps, int* status, bh_stream s);
/* bh_forward_sum_culled with the beam states read from a bh_ode_forward
 * trajectory table (beams [0, total) with the table's modes) instead of the
 * homogeneous closed form; omega from b (beam j and its mirror share it). */
int bh_forward_sum_traj(const bh_fwd_beams* b, int ndim, int mirror, const double* traj,
                        const double* ts, int64_t nt, const bh_segments* seg, double cull_log2,
                        int precision, void* out, unsigned long long* pairs, void* ws,
                        size_t ws_bytes, bh_stream s);

/* Exact transpose of the frozen-index forward beam sum.  For the beam table b
 * with unit coefficients (a0 without the MSWPT coefficient) and a residual
 * resid (nt, S) (float, or double if resid_f64), grad[k] (complex) =
 * sum_{mode} sum_{t,s} 2 resid(t,s) conj(G_k(t,s)), G_k the complex Gaussian
 * beam of the forward sum (both mirrored modes when mirror != 0), so that
 * <forward(c), resid> = Re sum_k conj(grad[k]) c[k].  Terms outside the
 * envelope bound of bh_forward_sum_culled (cull_log2) are skipped the same
 * way; cull_log2 <= 0 evaluates every term.  The workspace holds one partial
 * per (beam, tile of 64 time rows x a tile of point groups), added in a fixed
 * order (deterministic), the FP32-split form's residual in group order, and
 * then the beams' activity windows (the FP32-split form skips a beam's time
 * tiles outside its window); a workspace with room for the partials but not
 * the windows gives the same result without skipping. */
size_t bh_forward_transpose_workspace_size(int64_t count, int64_t nt, int64_t ngroups);
int bh_forward_transpose(const bh_fwd_beams* b, int ndim, int mirror, double c,
                         const double* ts, int64_t nt, const bh_segments* seg, double cull_log2,
                         int precision, const void* resid, int resid_f64, double* grad,
                         void* ws, size_t ws_bytes, bh_stream s);
/* coeffs[total] complex = 0, then coeffs[idx[i]] = scale * vals[i] (indices
 * outside [0, total) are skipped): the transpose of gathering the selected
 * coefficients. */
int bh_coeff_scatter(const int64_t* idx, const double* vals, int64_t k, double scale,
                     int64_t total, double* coeffs, bh_stream s);
/* vals[i] = coeffs[idx[i]] (complex; 0 for an index outside [0, total)):
 * the selected coefficients of a frozen index set. */
int bh_coeff_gather(const double* coeffs, int64_t total, const int64_t* idx, int64_t k,
                    double* vals, bh_stream s);
/* Adjoint of bh_mswpt_forward (no scale, no mask): out = W^H coeffs, i.e.
 * the synthesis of bh_mswpt_inverse divided by sum_g2 before the inverse
 * FFT; output_real != 0 writes its real part (prod N doubles, the transpose
 * for real spatial input), else prod N complex. */
int bh_mswpt_adjoint(const bh_tiling* t, const double* sum_gsquare, const double* gpacked,
                     const double* coeffs, int output_real, double* out, void* ws,
                     size_t ws_bytes, bh_stream s);

/* ---- learned reconstruction network (no reference counterpart: the
 * reference has no reconstruction solver; these serve the U-Net that gives
 * CG on the normal equations its starting image).  Tensors are NHWC bf16
 * (uint16 bit patterns) unless stated; device pointers, caller-allocated. */

/* C[M][N] = alpha * sum_k A(m,k) B(k,n) (+ bias[n]) (ReLU if relu) with bf16
 * MFMA and FP32 accumulation.  A stored [M][lda] (ta = 0) or [K][lda]
 * (ta = 1); B stored [K][ldb] (tb = 0) or [N][ldb] (tb = 1); the contiguous
 * extent and leading dimension of each operand multiples of 8, operands
 * 16-byte aligned.  C is bf16 (out_bf16) or FP32 with leading dimension ldc.
 * nsplit > 1 splits K (FP32 output, no bias/ReLU, ldc = N) through a
 * workspace of bh_nn_gemm_workspace_size bytes, partials added in order. */
size_t bh_nn_gemm_workspace_size(int64_t M, int64_t N, int64_t K, int nsplit);
int bh_nn_gemm(const void* A, int ta, int64_t lda, const void* B, int tb, int64_t ldb, int64_t M,
               int64_t N, int64_t K, float alpha, const float* bias, int relu, void* C,
               int out_bf16, int64_t ldc, int nsplit, void* ws, size_t ws_bytes, bh_stream s);
/* 3x3 same-padded convolution as an implicit GEMM (no column buffer):
 * Y[p][o] = act(alpha sum_{tap, c} X[n][y+dy][x+dx][c] Wt[o][tap*C + c] + bias[o]),
 * tap = 3(dy+1) + dx+1, X zero outside the image; X (B, H, W, C) and Wt
 * [Cout][9C] bf16, C and Cout multiples of 8, 16-byte aligned; bias FP32
 * (optional); ReLU if relu; Y (B, H, W, Cout) bf16 (out_bf16) or FP32. */
int bh_nn_conv3(const void* X, int64_t B, int H, int W, int C, const void* Wt, int Cout,
                const float* bias, int relu, float alpha, void* Y, int out_bf16, bh_stream s);
/* Wf[c][tap*Cout + o] = Wt[o][(8 - tap)*C + c]: with these weights
 * bh_nn_conv3 of dZ (B, H, W, Cout) gives the input gradient dX (B, H, W, C). */
int bh_nn_conv3_wflip(const void* Wt, int Cout, int C, void* Wf, bh_stream s);
/* Weight gradient dW[o][tap*C + c] = sum_p dZ[p][o] X(p shifted by tap)[c]
 * (FP32), implicit as bh_nn_conv3; the pixels are split into chunks whose
 * partial sums are added in order through a workspace of
 * bh_nn_conv3_dw_workspace_size bytes (0: none needed). */
size_t bh_nn_conv3_dw_workspace_size(int64_t B, int H, int W, int C, int Cout);
int bh_nn_conv3_dw(const void* X, const void* dZ, int64_t B, int H, int W, int C, int Cout,
                   float* dW, void* ws, size_t ws_bytes, bh_stream s);
/* 2x2 average pooling of X (B, H, W, C) into Y (B, H/2, W/2, C); backward:
 * X = dY (B, H/2, W/2, C) into Y = dX (B, H, W, C). */
int bh_nn_pool2(const void* X, int64_t B, int H, int W, int C, void* Y, int backward, bh_stream s);
/* nearest 2x upsampling of X (B, H, W, C) into Y (B, 2H, 2W, C); backward:
 * X = dY (B, 2H, 2W, C) summed into Y = dX (B, H, W, C). */
int bh_nn_up2(const void* X, int64_t B, int H, int W, int C, void* Y, int backward, bh_stream s);
/* Z = X + Y (bf16, n elements). */
int bh_nn_add(const void* X, const void* Y, int64_t n, void* Z, bh_stream s);
/* dZ = dY where Yact > 0, else 0 (ReLU backward). */
int bh_nn_relu_bwd(const void* dY, const void* Yact, int64_t n, void* dZ, bh_stream s);
/* out[c] = sum_p X[p][c] (FP32, fixed order: row-chunk partials, then a
 * fixed tree): bias gradients.  ws: bh_nn_colsum_workspace_size bytes. */
size_t bh_nn_colsum_workspace_size(int64_t P, int C);
int bh_nn_colsum(const void* X, int64_t P, int C, float* out, void* ws, size_t ws_bytes,
                 bh_stream s);
/* Squared error of channel 0 of pred (FP32, stride pred_stride) against
 * target (P): grad (bf16, stride grad_stride, channel 0 = 2(pred - target)/P,
 * other channels 0) and per-256-point partial sums of the squared error. */
size_t bh_nn_mse_workspace_size(int64_t P);
int bh_nn_mse(const float* pred, int pred_stride, const float* target, int64_t P, void* grad,
              int grad_stride, float* part, size_t part_bytes, bh_stream s);
/* One Adam step (Kingma and Ba 2015, bias-corrected, step >= 1) on FP32
 * weights w with gradient g and moments m, v; w_bf16 (optional) receives the
 * bf16 copy used by the GEMMs. */
int bh_nn_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float beta1,
               float beta2, float eps, int64_t step, void* w_bf16, bh_stream s);
/* y = scale * x, FP32 -> bf16 (from_bf16 = 0) or bf16 -> FP32 (1). */
int bh_nn_cast(const void* x, int from_bf16, int64_t n, float scale, void* y, bh_stream s);

#ifdef __cplusplus
}
#endif

#endif /* BEAMAX_AMD_H */
