/*
 * purify_oracle.h -- FP64 CPU restatement of PURIFY's degridding measurement operator
 * Phi = W.G.F.Z.S and its adjoint, used ONLY as test infrastructure (the checker).
 *
 * Nothing in the product library links, loads or calls this code.  It is compiled
 * into oracle/liboracle.so and called from tests/, from __graft_entry__.smoke() and
 * from bench.py's cpu_baseline leg.
 *
 * Each function cites the reference file:line it restates (paths relative to the
 * root of lao19881213/purify-1).  Layouts follow the reference:
 *   image index  j*nx + i    (cpp/purify/operators.h:163)
 *   grid index   p*Nu + q    (cpp/purify/operators.cc:38, utilities.cc:321-333)
 * Complex data is interleaved (re, im) double pairs.
 */
#ifndef PURIFY_ORACLE_H
#define PURIFY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel ids follow kernels::kernel in cpp/purify/kernels.h:15. */
enum orc_kernel { ORC_KB = 0, ORC_GAUSS = 1, ORC_BOX = 2, ORC_PSWF = 3, ORC_KBMIN = 4, ORC_GAUSS_ALT = 5 };
/* w-kernels of cpp/purify/projection_kernels.cc:8-58. */
enum orc_wkernel { ORC_W_APPROX = 0, ORC_W_BOX = 1, ORC_W_SPHERE = 2 };

/* ---- scalar kernels (cpp/purify/kernels.cc) ---- */
double orc_bessel_i0(double x);
double orc_kaiser_bessel(double x, int J);
double orc_kaiser_bessel_general(double x, int J, double alpha);
double orc_ft_kaiser_bessel(double x, int J);
double orc_ft_kaiser_bessel_general(double x, int J, double alpha);
double orc_gaussian(double x, int J);
double orc_ft_gaussian(double x, int J);
double orc_gaussian_general(double x, int J, double sigma);
double orc_ft_gaussian_general(double x, int J, double sigma);
double orc_pswf(double x, int J);
double orc_ft_pswf(double x, int J);
double orc_pill_box(double x, int J);
double orc_ft_pill_box(double x, int J);
/* create_kernels lambdas: which=0 gridding kernel, which=1 correction (FT) kernel; J is the
 * support of that axis, ftsize the oversampled size of that axis. Returns NaN for a bad id. */
double orc_kernel_eval(int kind, int which, double x, int J, int ftsize, double os);

/* ---- operator pieces ---- */
/* S, ny x nx real correction image (cpp/purify/operators.cc:101-121). */
int orc_correction2d(double os, int ny, int nx, int kind, int Ju, int Jv, double *S);
/* Z and Z^dagger (cpp/purify/operators.h:148-183); S is real, length ny*nx. */
void orc_zero_pad(const double *S, int ny, int nx, double os, const double *x, double *grid);
void orc_zero_pad_adjoint(const double *S, int ny, int nx, double os, const double *grid, double *x);
/* Unitary 2-D DFT of an Nv x Nu grid (cpp/purify/operators.h:209-261).  sign = -1 is
 * FFTW_FORWARD followed by /sqrt(N), sign = +1 is FFTW_BACKWARD followed by /sqrt(N). */
void orc_fft2d(int Nv, int Nu, const double *in, double *out, int sign);
/* Unnormalised DFT (no 1/sqrt(N)), the same transform without scaling. */
void orc_fft2d_raw(int Nv, int Nu, const double *in, double *out, int sign);

/* G as CSR, M rows, Ju*Jv entries per row, columns ascending (Eigen RowMajor storage)
 * (cpp/purify/operators.cc:8-46).  rowptr: M+1, colidx/vals: nnz.  Returns nnz or <0. */
int64_t orc_gridding_matrix_nnz(int64_t M, int ny, int nx, double os, int Ju, int Jv, int w_term, int Jw);
/* The u-slot kernel is (kind, Jeval_u), the v-slot kernel (kind, Jeval_v); Ju x Jv taps.
 * literal_phase=1 evaluates exp(-2 pi i (a/2 + b/2)) literally, 0 uses the exact (-1)^(a+b). */
int64_t orc_gridding_matrix(const double *u, const double *v, const double *weights, int64_t M,
                            int ny, int nx, double os, int kind, int Jeval_u, int Jeval_v, int Ju, int Jv,
                            int literal_phase, int64_t *rowptr, int32_t *colidx, double *vals);
/* w-projection G (cpp/purify/operators.cc:49-99 with projection_kernels.cc:16-85 and
 * convolution.h:42-82).  kernel slots follow details::init_gridding_matrix_2d(kernelu, kernelv). */
int64_t orc_gridding_matrix_w(const double *u, const double *v, const double *w, const double *weights,
                              int64_t M, int ny, int nx, double os, int kind, int Jeval_u, int Jeval_v,
                              int Ju, int Jv, int Jw, int wkind, double cellx, double celly,
                              int64_t *rowptr, int32_t *colidx, double *vals);
/* One (Jv+Jw-1) x (Ju+Jw-1) projection kernel, row-major (projection_kernels.cc:59-85). */
void orc_projection(int kind, int Jeval_u, int Ju, int Jv, int Jw, int wkind, double cellx, double celly,
                    int nx, int ny, double os, double u, double v, double w, double *out);
/* Conjugate transpose of a CSR matrix into CSR (operators.h:135-136). */
void orc_csr_adjoint(int64_t rows, int64_t cols, const int64_t *rowptr, const int32_t *colidx,
                     const double *vals, int64_t *t_rowptr, int32_t *t_colidx, double *t_vals);
/* OpenMP row SpMV y = A x (cpp/purify/utilities.h:144-159). */
void orc_spmv(int64_t rows, const int64_t *rowptr, const int32_t *colidx, const double *vals,
              const double *x, double *y);

/* ---- convolution helpers (cpp/purify/convolution.h) ---- */
/* periodic 2-D Daubechies analysis / synthesis of an ny x nx complex image (levels levels, filter h of
 * length L <= 64): the SARA dictionary's bases (wavelet_operator_factory.h:20-45) */
void orc_dwt2(const double *img, int ny, int nx, const double *h, int L, int levels, double *out);
void orc_idwt2(const double *coef, int ny, int nx, const double *h, int L, int levels, double *out);
/* one w-kernel sample (projection_kernels.cc:8-58) at pixel offsets (u, v), w in wavelengths; the pixel ->
 * lambda scales are pixel_to_lambda(cell, imsize, os) (projection_kernels.cc:28-30) */
void orc_wkernel_eval(int wkind, double u, double v, double w, double cellx, double celly, int nx, int ny,
                      double os, double *out);
double orc_pixel_to_lambda(double cell, int imsize, double os);
/* linear_convol_1d: f of length Jf, g already zero padded of length Lg; out length Lg-Jf+1. */
void orc_linear_convol_1d(const double *f, int Jf, const double *g, int Lg, double *out);
/* linear_convol_2d(kernelfu, kernelfv, kernelg): g is Jgv x Jgu row-major; out is
 * (Jfv+Jgv-1) x (Jfu+Jgu-1) row-major.  All complex. */
void orc_linear_convol_2d(const double *fu, int Jfu, const double *fv, int Jfv, const double *g, int Jgv,
                          int Jgu, double *out);

/* ---- chirp / w-projection rows (cpp/purify/wproj_utilities.{h,cc}) ---- */
void orc_fov_cosines(double cell_x, double cell_y, int x_size, int y_size, double *L, double *Mc);
/* Chirp image as the FFT sees it: flat index l + m*x_size (column-major Eigen map) */
void orc_generate_chirp(double w_rate, double cell_x, double cell_y, int x_size, int y_size, double *out);
double orc_sparsify_row_dense_thres(const double *row, int64_t n, double energy);
double orc_sparsify_row_thres_real(const double *absvals, int64_t n, double energy);
/* Sparse row-wise convolution of a G row with a chirp row on an x_size*y_size grid
 * (wproj_utilities.h:222-287).  Dense output row of length x_size*y_size (complex). */
void orc_row_wise_convolution(const int32_t *g_idx, const double *g_val, int g_n, const double *chirp_dense,
                              int x_size, int y_size, double *out_dense);
/* wprojection_matrix (wproj_utilities.cc:132-188) for a CSR G with G.cols == x_size*y_size.
 * Output is written densely per row (rows x x_size*y_size complex) followed by pruning:
 * out_mask[k] = 1 where the entry is kept.  Caller allocates rows*x_size*y_size. */
int orc_wprojection_matrix_dense(int64_t rows, const int64_t *rowptr, const int32_t *colidx,
                                 const double *vals, int x_size, int y_size, const double *w,
                                 double cell_x, double cell_y, double e_chirp, double e_wproj,
                                 double *out_dense, uint8_t *out_mask);

/* ---- utilities ---- */
/* set_cell_size + uv_scale: lambda (units=0) or radians (units=1) -> pixels
 * (cpp/purify/utilities.cc:208-278, operators.h:405-410). */
void orc_uv_to_pixels(const double *u, const double *v, int64_t M, int units, double cell_x, double cell_y,
                      int nx, int ny, double os, double *u_px, double *v_px);
double orc_calculate_l2_radius(int64_t y_size, double sigma, double n_sigma, int chi2);
double orc_snr_to_standard_deviation(const double *y0, int64_t M, double snr);

/* ---- full operator (operators.h:298-414) ---- */
typedef struct orc_op orc_op;
/* Builds Phi with the explicit CSR G and G^H.  power_iters < 0 means "do not normalise"
 * (norm = 1).  w_term selects the projection-kernel G (operators.cc:49-99). */
orc_op *orc_op_create(const double *u_px, const double *v_px, const double *w, const double *weights,
                      int64_t M, int ny, int nx, double os, int kind, int Ju, int Jv, int w_term, int Jw,
                      double cellx, double celly, int literal_phase);
void orc_op_destroy(orc_op *op);
void orc_op_set_norm(orc_op *op, double norm);
double orc_op_get_norm(const orc_op *op);
int64_t orc_op_nnz(const orc_op *op);
void orc_op_forward(const orc_op *op, const double *x, double *y);
void orc_op_adjoint(const orc_op *op, const double *y, double *x);
/* Power method on Phi^dagger Phi returning ||Phi||_2 (SOPT power_method semantics, see
 * the .c file).  start may be NULL for the all-ones vector. */
double orc_op_power_method(const orc_op *op, int iters, double tol, const double *start);
/* Stage access for tests: grid = Z.S x ; grid2 = F grid ; y = G grid2 and adjoints. */
void orc_op_apply_G(const orc_op *op, const double *grid, double *y);
void orc_op_apply_GH(const orc_op *op, const double *y, double *grid);
int orc_op_grid_dims(const orc_op *op, int *Nv, int *Nu);
const double *orc_op_correction(const orc_op *op);

int orc_omp_threads(void);
void orc_set_omp_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
