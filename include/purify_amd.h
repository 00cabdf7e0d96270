/*
 * purify_amd.h -- C-ABI of the MI355X-native PURIFY measurement operator Phi = W.G.F.Z.S.
 *
 * Drop-in boundary: the entry points below are what the reference's operator factory would bind
 * to build and apply Phi.  Plain pointers and sizes only; no C++ or torch types cross this line.
 * References are file:line in lao19881213/purify-1 (PURIFY 2.0.0).
 *
 * Conventions
 *   - Complex data is interleaved (re, im).  A PF_C64 operator works on float pairs on the
 *     device, a PF_C128 operator on double pairs.  Host-side buffers (PF_HOST) are always
 *     complex128, the type of the reference's Vector<t_complex>; the library converts.
 *   - Image vectors have imsizey*imsizex entries, index j*imsizex + i (operators.h:163).
 *     Visibility vectors have M entries in the caller's order.
 *   - Every function returns PF_OK (0) or a negative pf_status; pf_last_error() gives the
 *     message of the last failure on the calling thread.  Construction errors carry the
 *     reference's runtime_error texts.
 *   - An operator is immutable after creation except for its norm.  Applies on one handle must
 *     be externally serialised (one workspace per handle); different handles may be used
 *     concurrently from different threads/streams.
 */
#ifndef PURIFY_AMD_H
#define PURIFY_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pf_op pf_op;
typedef void *pf_stream; /* a hipStream_t; NULL = the null stream, except for an apply whose input and output
                            are both on the host (a synchronous call): that runs on the calling thread's
                            own non-blocking stream, as do constructions and power methods */

enum pf_status {
  PF_OK = 0,
  PF_ERR_ARG = -1,      /* invalid argument */
  PF_ERR_HIP = -2,      /* HIP runtime failure */
  PF_ERR_UNSUPPORTED = -3,
  PF_ERR_NOMEM = -4,
  PF_ERR_STATE = -5
};

enum pf_precision { PF_C64 = 0, PF_C128 = 1 };
enum pf_location { PF_HOST = 0, PF_DEVICE = 1 };
/* kernels::kernel, cpp/purify/kernels.h:15 */
enum pf_kernel { PF_KB = 0, PF_GAUSS = 1, PF_BOX = 2, PF_PSWF = 3, PF_KBMIN = 4, PF_GAUSS_ALT = 5 };
/* utilities::vis_units, cpp/purify/utilities.h:14 */
enum pf_units { PF_LAMBDA = 0, PF_RADIANS = 1, PF_PIXELS = 2 };

/* Operator description.  Replaces the argument list of
 *   measurementoperator::init_degrid_operator_2d(u, v, w, weights, imsizey, imsizex,
 *       oversample_ratio, power_iters, power_tol, kernel, Ju, Jv, w_term, Jw, cellx, celly)
 *   (cpp/purify/operators.h:373-396) and of its vis_params overload (operators.h:398-414),
 * selected by `units` (PF_PIXELS = the raw-vector overload). */
typedef struct pf_op_desc {
  int32_t imsizey, imsizex;
  double oversample_ratio; /* grid is floor(os*imsizey) x floor(os*imsizex) */
  int32_t kernel;          /* enum pf_kernel */
  int32_t Ju, Jv;
  int32_t w_term; /* 0: plain KB gridding (operators.cc:8-46); 1: projection rows (operators.cc:49-99);
                     2: chirp rows, wproj_utilities::wprojection_matrix of the plain G (wproj_utilities.cc:
                     132-188) with energy_chirp / energy_wproj, built on the device (square grids, Ju = Jv) */
  int32_t Jw;
  double cellx, celly; /* arcsec; used by unit conversion and the w-kernel */
  int32_t units;       /* enum pf_units of u, v (w is always in wavelengths) */
  int32_t precision;   /* enum pf_precision */
  int64_t M;
  const double *u, *v, *w; /* host, length M (w may be NULL when w_term == 0) */
  const double *weights;   /* host complex128, length M; NULL = unit weights */
  int32_t power_iters;     /* power-method iterations for the normalisation (operators.h:390) */
  double power_tol;
  int32_t tile_u, tile_v; /* gridding tile shape; 0 = library default; either < 0 (pf_op_create only) = chosen by
                           * timing Phi + Phi^dagger pairs on one build per candidate shape (32 x 32, 64 x 32,
                           * 32 x 64, 64 x 64) and, complex64, per permutation bucket (read back with
                           * pf_op_get_tuning "tile_u", "tile_v", "perm_bucket") */
  /* w_term = 2: the energy fractions the chirp and the convolved row keep (energy_fraction_chirp,
   * energy_fraction_wproj of wprojection_matrix); energy_chirp must be below 1 */
  double energy_chirp, energy_wproj;
} pf_op_desc;

/* Fills a descriptor with the reference's defaults: os 2, kb, Ju = Jv = 4, w_term 0, Jw 6,
 * cells 1, units pixels, power_iters 100, power_tol 1e-4, complex64; energy_chirp = energy_wproj = 0.999. */
void pf_op_desc_init(pf_op_desc *d);

/* Build Phi: copies and sorts the coordinates, uploads them, builds twiddles and the correction
 * vectors, then (power_iters > 0) normalises by the power-method norm computed on the GPU in FP64 (a
 * complex64 operator takes the norm of its complex128 twin, built from the same inputs for the iteration
 * and released, so both precisions get the same norm).  Replaces measurementoperator::init_degrid_operator_2d (operators.h:373-414)
 * and factory::measurement_operator_factory's gpu_serial branch (measurement_operator_factory.h:
 * 38-96). */
int pf_op_create(const pf_op_desc *desc, pf_op **out);
int pf_op_destroy(pf_op *op);

/* y = Phi x (operators.h:393, the LinearTransform's direct operator). */
int pf_op_forward(pf_op *op, const void *x, void *y, int x_loc, int y_loc, pf_stream stream);
/* x = Phi^dagger y (operators.h:394, the adjoint operator). */
int pf_op_adjoint(pf_op *op, const void *y, void *x, int y_loc, int x_loc, pf_stream stream);

/* out = Phi^dagger Phi x: operators::init_grid_degrid_operator_2d (compact_operators.h:150-171, built
 * by base_grid_degrid_operator_2d, 118-147), the self-adjoint grid-degrid operator divided by the
 * square of the measurement operator's norm (that function runs the same power iteration on
 * Phi^dagger Phi).  The reference applies an explicit G^H G pruned of entries with |value| <= 1e-7
 * (compact_operators.h:109-111); this applies G then G^H exactly.  x and out: imsizey*imsizex
 * (complex128 on the host, or the operator's precision on the device); out may alias x. */
int pf_op_apply_normal(pf_op *op, const void *x, void *out, int x_loc, int out_loc, pf_stream stream);

/* ---- PSF convolution (operators::init_psf_convolve_2d, compact_operators.h:279-306) ----
 * psf = Phi^dagger Phi delta (the given measurement operator, delta at row floor(ny/2) - 1, column
 * floor(nx/2) - 1); the operator is x -> IFFT(FFT(psf) . FFT(x)) with unitary n x n FFTs (a circular
 * convolution on the unpadded image), divided by the square of its own power-method norm.  start:
 * host complex128 power-method start vector, or NULL for a fixed pseudo-random one (uniform in
 * [-1, 1] per component, mt19937_64 seed 0x5eed; the reference draws T::Random).  Image sides must be
 * powers of two.  op must outlive the handle; applies share op's host staging buffers. */
typedef struct pf_psf pf_psf;
int pf_psf_create(pf_op *op, int32_t power_iters, double power_tol, const double *start, pf_psf **out);
int pf_psf_apply(pf_psf *psf, const void *x, void *out, int x_loc, int out_loc, pf_stream stream);
int pf_psf_norm(const pf_psf *psf, double *norm);
int pf_psf_destroy(pf_psf *psf);

/* ---- Sparsity dictionary Psi (factory::wavelet_operator_factory, wavelet_operator_factory.h:20-45:
 * sopt::wavelets::SARA wrapped by sopt::linear_transform(sara, imsizey, imsizex)) ----
 * nwav entries: db[i] = 0 for the Dirac basis or the Daubechies order 1..PF_DB_MAX ("DB1".."DB10"),
 * levels[i] levels of the periodic 2-D transform (image sides divisible by 2^levels).  As a
 * LinearTransform: the direct operator is the synthesis Psi (nwav stacked coefficient images ->
 * image), the adjoint the analysis Psi^dagger; SARA divides each entry by sqrt(nwav).  SOPT is not
 * in the reference tree: the filters are Daubechies' extremal-phase filters (pf_wavelet_filter), the
 * transform the periodic down-convolution lo[i] = sum_k h[k] x[(2i + k) mod n], low-pass half first,
 * rows then columns per level. */
#define PF_DB_MAX 10
typedef struct pf_wavelet pf_wavelet;
/* the Daubechies scaling filter with db vanishing moments (2 * db taps, sum sqrt(2)) into lo */
int pf_wavelet_filter(int32_t db, double *lo, int32_t *len);
int pf_wavelet_create(int32_t imsizey, int32_t imsizex, int32_t nwav, const int32_t *db, const int32_t *levels,
                      int32_t precision, pf_wavelet **out);
int pf_wavelet_forward(pf_wavelet *wv, const void *coeffs, void *image, int c_loc, int x_loc, pf_stream stream);
int pf_wavelet_adjoint(pf_wavelet *wv, const void *image, void *coeffs, int x_loc, int c_loc, pf_stream stream);
int pf_wavelet_destroy(pf_wavelet *wv);

/* ---- Imaging PADMM on the device (sopt::algorithm::ImagingProximalADMM with the parameters of
 * padmm_factory, algorithm_factory.h:36-81) ----
 * Iteration (SOPT variable names): z = P_eps(-lambda - r); x~ = x - Phi^dagger(r + lambda + z) / nu;
 * x = prox(x~); r = Phi x - y; lambda += lagrange_update_scale (r + z); from x = Phi^dagger y / nu,
 * r = Phi x - y, lambda = 0.  The l1 proximal is the closed form x~ + Psi(soft(Psi^dagger x~, gamma l1_nu)
 * - Psi^dagger x~) / l1_nu (l1_nu: the proximal's frame bound; nu is the ADMM step) for a tight frame (one orthonormal wavelet: main.cc:121-126), or, for SARA, the
 * iterative constrained one: argmin 1/2 ||x - x~||^2 + gamma ||Psi^dagger x||_1 over the positivity /
 * real constraints, by the dual forward-backward iteration (u = P_gamma(u + Psi^dagger x / l1_nu),
 * x = P_C(x~ - Psi u), from u = 0) until the objective's relative change is below l1_tolerance.
 * SOPT is not in the reference tree: this is a restatement of the published algorithms and its
 * agreement with SOPT's iterates is unpinned.  Norms are FP64 reductions on the device; a tight-frame
 * iteration runs without host synchronisation and can be captured once into a hipGraph and replayed;
 * the iterative proximal reads its objective on the host every step, so it always runs eagerly
 * (use_graph is ignored, and pf_padmm_enqueue returns only when its iterations are done).  phi and
 * psi must outlive the solver, and no other call may use them while it runs (their work buffers are
 * shared). */
typedef struct pf_padmm pf_padmm;
typedef struct pf_padmm_params {
  double gamma;                 /* l1 weight; <= 0: max |Psi^dagger Phi^dagger y| * 1e-3 (algorithm_factory.h:70-73) */
  double epsilon;               /* l2-ball radius (utilities::calculate_l2_radius), > 0 */
  double nu;                    /* 1 (algorithm_factory.h:62; the l1 proximal's nu is also 1) */
  double lagrange_update_scale; /* 0.9 */
  double relative_variation;    /* 1e-3 */
  double residual_tolerance;    /* <= 0: epsilon (algorithm_factory.h:74) */
  int32_t itermax;              /* the most iterations the solver will run (sizes its statistics) */
  /* the l1 proximal (algorithm_factory.h:60-66) */
  int32_t tight_frame;          /* 1: closed form (one basis); 0: iterative, constrained; -1: 1 iff one basis */
  int32_t l1_itermax;           /* 50: iterative proximal steps at most */
  double l1_tolerance;          /* 1e-2: its objective's relative change that stops it */
  double l1_nu;                 /* 1: its frame bound (step 1 / l1_nu) */
  int32_t positivity;           /* 1: Re x >= 0 and Im x >= 0 in the iterative proximal */
  int32_t real_constraint;      /* 1: Im x = 0 in the iterative proximal */
} pf_padmm_params;
typedef struct pf_padmm_result {
  int32_t niters, converged;
  double relative_variation, residual_norm, gamma;
  int32_t l1_iters_last;        /* iterative proximal: steps of the last call, */
  int64_t l1_iters_total;       /* and of all calls since the last reset */
} pf_padmm_result;
enum { PF_PADMM_IMAGE = 0, PF_PADMM_RESIDUAL = 1 };
int pf_padmm_params_init(pf_padmm_params *p);
/* y: the visibilities (host complex128, or device in the operator's precision), M of them. */
int pf_padmm_create(pf_op *phi, pf_wavelet *psi, const void *y, int y_loc, const pf_padmm_params *params,
                    pf_padmm **out);
/* Runs iters more iterations (use_graph: replay one captured iteration; while the capture lasts HIP
 * refuses legacy-stream work from every thread of the process, so a multi-threaded caller should not
 * use it while other threads issue such work).  check_every > 0: every that
 * many iterations, stop once ||x - x_prev|| / ||x|| < relative_variation and ||r|| <
 * residual_tolerance (check_every = 1 stops exactly at the first converged iteration, as SOPT does);
 * 0: run all iters. */
int pf_padmm_run(pf_padmm *h, int32_t iters, int32_t use_graph, int32_t check_every, pf_padmm_result *res);
/* Enqueues iters iterations on the solver's own stream and returns without waiting (no convergence
 * checks); pf_padmm_sync waits and reports.  Solvers on different handles then run concurrently: the
 * batched multi-channel path.  The first graph enqueue captures (see use_graph above). */
int pf_padmm_enqueue(pf_padmm *h, int32_t iters, int32_t use_graph);
int pf_padmm_sync(pf_padmm *h, pf_padmm_result *res);
int pf_padmm_reset(pf_padmm *h);

/* ---- Forward-backward on the device (sopt::algorithm::ImagingForwardBackward as
 * cpp/example/fb_random_coverage.cc:89-106 configures it: PhiTPhi the compact operator, the target the
 * dirty image Phi^dagger y) ----
 * x~ = x - (beta / sigma^2)(Phi^dagger Phi x - Phi^dagger y); x = prox_(mu beta ||Psi^dagger .||_1)(x~) with the
 * l1 proximal of the PADMM above (closed form for one basis, else iterative and constrained); from
 * x = 0; converged when ||x - x_prev|| / ||x|| < relative_variation.  Restated from the published
 * forward-backward splitting (SOPT is not in the reference tree): agreement with SOPT is unpinned.  The
 * handle is a pf_padmm: pf_padmm_run / enqueue / sync / reset / get / stats / destroy drive it;
 * PF_PADMM_RESIDUAL returns the image-domain residual Phi^dagger Phi x_prev - Phi^dagger y (imsizey*imsizex),
 * and the third statistic is its squared norm. */
typedef struct pf_fb_params {
  double mu;                    /* <= 0: 1e-5 imsizey imsizex (fb_random_coverage.cc:92) */
  double beta;                  /* <= 0: 0.5 / (imsizey imsizex) (fb_random_coverage.cc:93) */
  double sigma;                 /* noise standard deviation, > 0 */
  double relative_variation;    /* 1e-3 */
  int32_t itermax;              /* 500 */
  int32_t tight_frame;          /* as pf_padmm_params; the example passes 0 with SARA */
  int32_t l1_itermax;           /* 50 */
  double l1_tolerance;          /* 1e-2 */
  double l1_nu;                 /* 1 */
  int32_t positivity;           /* 1 */
  int32_t real_constraint;      /* 1 */
} pf_fb_params;
int pf_fb_params_init(pf_fb_params *p);
int pf_fb_create(pf_op *phi, pf_wavelet *psi, const void *y, int y_loc, const pf_fb_params *params,
                 pf_padmm **out);
/* what: PF_PADMM_IMAGE (x, imsizey*imsizex) or PF_PADMM_RESIDUAL (r = Phi x - y, M) */
int pf_padmm_get(pf_padmm *h, int32_t what, void *out, int out_loc);
/* per-iteration FP64 statistics (||x - x_prev||^2, ||x||^2, ||r||^2) of up to cap iterations */
int pf_padmm_stats(pf_padmm *h, double *out, int32_t cap, int32_t *count);
int pf_padmm_destroy(pf_padmm *h);

/* Power method on Phi^dagger Phi (SOPT power_method as called at operators.h:390-391) over the
 * operator as it currently is (including its norm).  start: host complex128 of imsizey*imsizex,
 * or NULL for all ones.  Iterates in the operator's precision, reduces norms in FP64. */
int pf_op_power_method(pf_op *op, int32_t iters, double tol, const double *start, double *norm_out);
/* init_normalise (operators.h:185-189): throws (PF_ERR_ARG) unless norm > 0. */
int pf_op_set_norm(pf_op *op, double norm);
int pf_op_get_norm(const pf_op *op, double *norm);

typedef struct pf_op_info {
  int32_t imsizey, imsizex, Nv, Nu, Ju, Jv, precision, w_term;
  int64_t M, nnz_w;   /* nnz_w: stored projection-row entries when w_term */
  int32_t tile_u, tile_v, n_tiles;
  int64_t tile_list_len; /* total entries of the gridding tile lists */
  double norm;
  int64_t device_bytes;  /* HBM the handle holds */
} pf_op_info;
int pf_op_get_info(const pf_op *op, pf_op_info *info);

/* ---- stage access (device pointers, op precision), for tests and per-kernel timing ---- */
/* grid = F Z S x without the norm (operators.h:263-296, base_padding_and_FFT_2d direct). */
int pf_op_apply_FZ(pf_op *op, const void *x_dev, void *grid_dev, pf_stream stream);
/* x = S^H Z^H F^H grid without the norm. */
int pf_op_apply_FZ_adjoint(pf_op *op, const void *grid_dev, void *x_dev, pf_stream stream);
/* y = G grid and grid = G^H y without the norm (operators.h:123-145). */
int pf_op_apply_G(pf_op *op, const void *grid_dev, void *y_dev, pf_stream stream);
int pf_op_apply_GH(pf_op *op, const void *y_dev, void *grid_dev, pf_stream stream);

/* ---- Super operator over visibility blocks (measurementoperator::init_combine_operators /
 * init_super_operator, operators.h:513-578): y = [Phi_1 x; Phi_2 x; ...] (blocks stacked in order),
 * x = sum_b Phi_b^dagger y_b, each Phi_b normalised by its own norm.  The operators must share precision,
 * image and grid sizes, kernel and supports; F Z S then runs once per apply (the blocks' G_b read one
 * grid; the G_b^H accumulate into one).  Device pointers in the operators' precision.  The operators must
 * outlive the handle and are not used by other calls meanwhile (their buffers are shared). ---- */
typedef struct pf_super pf_super;
int pf_super_create(pf_op *const *ops, int32_t nblocks, pf_super **out);
int pf_super_size(const pf_super *sp, int64_t *M);
int pf_super_forward(pf_super *sp, const void *x_dev, void *y_dev, pf_stream stream);
int pf_super_adjoint(pf_super *sp, const void *y_dev, void *x_dev, pf_stream stream);
int pf_super_destroy(pf_super *sp);

/* Unitary 2-D FFT of an Nv x Nu grid (operators.h:209-261): sign -1 is FFTW_FORWARD / sqrt(N),
 * +1 is FFTW_BACKWARD / sqrt(N).  Powers of two from 2 to 16384 per axis.  Device pointers of the
 * given precision; in == out is allowed.  workspace: NULL or a device buffer of Nv*Nu complex. */
int pf_fft2d(int32_t Nv, int32_t Nu, int32_t precision, int32_t sign, const void *in_dev, void *out_dev,
             void *workspace_dev, pf_stream stream);

/* The same transform as pf_fft2d, allocating nothing: no counterpart in the reference.  It is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The twiddle table (computed on the device at each call) and the row pass's output
 * live in a static device array of the library, PF_FFT2D_NOALLOC_MAX_BYTES of workspace plus 128 KiB of
 * twiddles, which each GPU holds from the first call on it until the process ends; grids larger than
 * the workspace (Nv * Nu * 8 bytes for PF_C64, * 16 for PF_C128) are rejected with PF_ERR_ARG.  Calls on
 * one GPU share that array: calls that overlap on one GPU (different streams or threads) race, calls on
 * one stream are ordered.  Asynchronous on `stream`; in == out is allowed.  Results agree with pf_fft2d
 * to rounding (the twiddles are FP64 device sincospi values instead of long-double host ones). */
#define PF_FFT2D_NOALLOC_MAX_BYTES (32u << 20)
int pf_fft2d_noalloc(int32_t Nv, int32_t Nu, int32_t precision, int32_t sign, const void *in_dev, void *out_dev,
                     pf_stream stream);

/* Evaluates a gridding kernel (which = 0) or its correction FT (which = 1) on the GPU exactly as the
 * operator's kernels do (kernels.cc:233-323), for n host points; results to host.  For which = 1
 * the argument is the pixel position r and the function is FT(r / ftsize - 0.5). */
int pf_kernel_eval(int32_t kind, int32_t which, int32_t J, double os, int32_t ftsize, int32_t precision,
                   const double *x_host, double *out_host, int64_t n);

/* ---- tuning (no counterpart in the reference: FFTW_MEASURE planning, operators.h:215-223, is its
 *      closest analogue) ---- */
/* Keys: "kernel_poly" (complex64, Ju = Jv <= 8, KB / KB-min / Gaussian / PSWF: 1 = the taps of both
 * axes from per-tap degree-9 polynomials in the fractional offset, checked at build to 1e-7 of the
 * kernel's peak, the default where they apply and ahead of the tables; 0 = off); "grid_kernel" (0 = residue-class lanes, 2 = rank-4 updates of 16 x 16 tiles on the matrix
 * cores, the default where the grid sides are at least 16 + J - 1: both sum every grid cell in a fixed
 * order, so they are bit-identical run to run; 1 = LDS atomics, whose in-wave summation order is not
 * architecturally fixed: set only on explicit request, never chosen by pf_op_autotune);
 * "kernel_table" (complex64, J <= 8: interpolated tap tables, about 1e-7 relative error: 1 = in gridding,
 * the default, 2 = in gridding and degridding; 0 = evaluate the kernel for every tap); "fft_columns"
 * (0 = the column FFT in one pass up to 8192 rows, the default; 1 = two passes of N1 x N2 beyond 512
 * rows); "fft_row_blocks" (row FFT grid: that many resident blocks per compute unit, each looping over
 * lines with the next line prefetched; 0 = one line per block; default 3 for complex64, 2 for
 * complex128); "grid_waves" (matrix-core gridding with that many waves, each taking every
 * grid_waves-th 16 x 16 tile with its batches pipelined across tiles; 0 = one wave per tile, the
 * default); read-only "tile_u", "tile_v", "perm_bucket" (complex64: the caller-order
 * permutation's bucket of visibilities, else 0), "w_footprint" (side of the stored rows' footprint with
 * w-projection rows, else 0) and "split_tiles" (tiles whose stored-row gridding list is cut into parts
 * that are summed in a fixed order afterwards, so that a dense tile does not form the launch's tail);
 * "w_rows_packed" (chirp w-projection rows of at most 8 x 8 taps: 1 = stored as their kept entries, a
 * 64-bit keep mask and offset per row and the kept values, the default; 0 = dense rows, rebuilt from the
 * kept entries on first use; results are bit-identical either way at one gridding entry per step, and equal
 * to rounding at more); "w_grid_entries" (gridding with kept-entry rows: the list entries one wave step
 * takes, 1, 2 or 4, each over 64 / entries lanes that take its kept values in stored order and with its own
 * accumulator copy of the tile, the copies summed in order at the end; allowed while the operator's largest
 * row keeps at most 64 / entries taps and the copies fit 160 KiB of LDS; the default is the most allowed
 * within 16 KiB of copies; 0 without kept-entry rows); "w_grid_owner" (dense stored rows of 65 to 576
 * taps whose K - 1 fits in the tile: 1 = each row read once into its owner tile extended by K - 1, then
 * a combine pass adds each cell's owner slots and the left, upper and upper-left extensions in that
 * order, the default; 0 = the tile lists, a row read for every tile it touches; equal to rounding; 0 for
 * other operators); "w_grid_depth" (gridding with kept-entry
 * rows at one entry per step: the entries whose tap values are requested ahead of their use, 8, 16 (the
 * default) or 32); read-only
 * "w_degrid_slots" (kept-entry rows: the kept values each degridding lane takes, 2, 4 or 8, the fewest that
 * hold the operator's largest row over 8 lanes; else 0).
 * Results do not depend on grid_kernel or fft_columns beyond floating-point summation order. */
/* pf_op_set_norm, pf_op_set_tuning and pf_op_autotune change what an apply launches: a solver that
 * captured the operator's launches into a graph captures them again on its next run. */
int pf_op_set_tuning(pf_op *op, const char *key, int64_t value);
int pf_op_get_tuning(const pf_op *op, const char *key, int64_t *value);
/* Times the fixed-order gridding-kernel variants (residue-class lanes and matrix cores, not the LDS-atomic
 * one) on this operator (reps adjoint gridding passes each, HIP events on the calling thread's stream) and keeps the
 * fastest, then does the same for the two "fft_columns" modes (reps
 * Phi + Phi^dagger pairs each, the best of three trials).  Writes the chosen gridding variant to *chosen if non-NULL;
 * the column mode is read back with pf_op_get_tuning.  Operators with w-projection rows or a given
 * gridding matrix have one gridding kernel: only the column mode is chosen, and with kept-entry rows the
 * gridding's "w_grid_entries" (reps gridding passes each, after as many warm-up passes, the settings
 * alternating over three trials and each keeping its best).  Last, "fft_row_blocks" is chosen among 0, 2, 3
 * and 4 the same way as the column mode (it replaces a value set before; results do not depend on it). */
int pf_op_autotune(pf_op *op, int32_t reps, int64_t *chosen);

/* ---- per-kernel timing (no counterpart in the reference, whose operators log at construction only,
 *      operators.h:274-290) ---- */
/* on = 1 enables (and clears) HIP-event timing of every kernel the operator launches, recorded on the
 * launching stream; on = 0 disables it. */
int pf_op_profile_enable(pf_op *op, int32_t on);
/* Entry idx of the timing table (kernel name, summed milliseconds, launch count).  Synchronises on the
 * recorded events.  Returns PF_ERR_ARG past the last entry. */
int pf_op_profile_get(pf_op *op, int32_t idx, char *name, int32_t name_len, double *total_ms, int64_t *count);

/* ---- synthetic inputs (the seeded generator of the benchmark and tests) ---- */
/* mt19937_64 seeded with `seed`, Box-Muller normals.  uv ~ N(0, sigma) rad clipped to [-pi, pi]. */
int pf_synth_uv(uint64_t seed, int64_t M, double sigma, double *u, double *v);
/* Sum of nsrc Gaussian sources: centres U[0,n)^2, amplitude U[0.1,1], sigma U[1,8] px. */
int pf_synth_image(uint64_t seed, int32_t ny, int32_t nx, int32_t nsrc, double *img_complex128);
/* Complex standard normal vector (re, im each N(0, 1/2) * scale). */
int pf_synth_cnormal(uint64_t seed, int64_t n, double scale, double *out_complex128);

/* ---- sparse gridding matrices and chirp-based w-projection ---- */
/* A CSR matrix of complex128 values: row r holds entries rowptr[r] .. rowptr[r + 1] - 1, columns
 * ascending.  For a gridding matrix, rows are visibilities and columns grid cells p * Nu + q. */
typedef struct pf_csr {
  int64_t rows, cols, nnz;
  int64_t *rowptr; /* rows + 1 */
  int32_t *colidx; /* nnz */
  double *vals;    /* nnz complex128 (re, im) */
} pf_csr;

/* wproj_utilities::wprojection_matrix (cpp/purify/wproj_utilities.cc:132-188): every row m of G is
 * convolved (row_wise_sparse_convolution, wproj_utilities.h:222-287) with the chirp of w[m]
 * (create_chirp_row / generate_chirp, wproj_utilities.cc:51-130: FFT of the chirp image times sqrt(N),
 * hard-thresholded to keep energy_chirp of its l2 energy), and the result is hard-thresholded to keep
 * energy_wproj (sparsify_row_thres, wproj_utilities.h:101-145).  Runs on the GPU in FP64.  The grid
 * must be square (x_size == y_size), as the reference's chirp generation assumes
 * (wproj_utilities.cc:67-71).  The library allocates GW's arrays; release them with pf_csr_free. */
int pf_wprojection_matrix(const pf_csr *G, int32_t x_size, int32_t y_size, const double *w, double cell_x,
                          double cell_y, double energy_chirp, double energy_wproj, pf_csr *GW);
void pf_csr_free(pf_csr *m);

/* An operator Phi = G F Z S / norm whose gridding matrix G is given (e.g. the output of
 * pf_wprojection_matrix): G->rows visibilities, G->cols = Nv * Nu.  The descriptor supplies the image
 * size, oversampling, kernel (for the correction S), precision and power-method settings; its u, v, w
 * and weights are not used.  G is applied as a row gather and G^H as a gather over the transposed
 * matrix: both deterministic.  No counterpart entry point in the reference, whose operators build G
 * themselves (operators.h:123-145); this one carries a precomputed G into the same operator. */
int pf_op_create_csr(const pf_op_desc *desc, const pf_csr *G, pf_op **out);

/* ---- channel batches (no counterpart in the reference, whose channels are separate runs; the closest
 *      precedent is the stacked visibility blocks of operators.h:513-578) ----
 * n tight-frame PADMM solvers (pf_padmm_create; not forward-backward, not the iterative proximal) run in
 * lock-step: every kernel of the iteration is launched once per group of alike channels, its grid's z
 * index selecting the channel.  Channels form a group when their iterations launch the same kernels with
 * the same shapes (the same image and grid sizes, kernel support, wavelet, precision and tuning); each group
 * runs on its own stream.  A channel computes exactly what it computes alone, bit for bit.  The solvers
 * must outlive the batch; each run synchronises the members' own streams first, adds iters to each
 * member's iteration count (its pf_padmm_get / stats / sync then report as after pf_padmm_run), and
 * returns when the iterations are done.  A member whose operator norm, tuning or l1 weight changed is
 * recorded again.  pf_padmm_batch_info: the groups and the launches one batched iteration makes. */
typedef struct pf_padmm_batch pf_padmm_batch;
int pf_padmm_batch_create(pf_padmm *const *solvers, int32_t n, pf_padmm_batch **out);
int pf_padmm_batch_run(pf_padmm_batch *b, int32_t iters, int32_t use_graph);
int pf_padmm_batch_info(const pf_padmm_batch *b, int32_t *groups, int32_t *launches_per_iteration);
int pf_padmm_batch_destroy(pf_padmm_batch *b);

/* ---- device memory ----
 * Handles (operators, solvers, wavelets, PSFs, FFT plans) and their construction temporaries take device
 * memory from the library's per-device block cache, and their destroy functions return it there without a
 * device-wide synchronisation (hipFree would wait for every stream on the device).  Hence: before
 * destroying a handle, synchronise the streams you passed to it (the library synchronises its own).  The
 * cache keeps at most 16 GiB per device and is released on allocation failure.  Neither function has a
 * counterpart in the reference, whose Eigen/FFTW operators live in host memory.
 * pf_cache_stats: bytes in blocks held by live handles, and bytes cached for reuse.
 * pf_cache_release: hipFree every cached block (this waits for the device); released_bytes may be NULL. */
int pf_cache_stats(int64_t *live_bytes, int64_t *cached_bytes);
int pf_cache_release(int64_t *released_bytes);

/* ---- errors ---- */
const char *pf_last_error(void);
/* Library version string. */
const char *pf_version(void);

#ifdef __cplusplus
}
#endif
#endif
