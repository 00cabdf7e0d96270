/*
 * purify_oracle.c -- FP64 CPU restatement of PURIFY's measurement operator (test infrastructure).
 *
 * This file is the CHECKER.  The product (purify-1_amd/) never links or loads it.
 * Citations are file:line in lao19881213/purify-1.  Arithmetic is IEEE double; complex
 * numbers use C99 double complex.  OpenMP parallelises the same loops the reference does.
 *
 * Deliberate deviation: the gridding phase
 * exp(-2 pi i ((k_u+ju)/2 + (k_v+jv)/2)) (operators.cc:40,93) is evaluated as the exact
 * parity (-1)^((k_u+ju)+(k_v+jv)) unless literal_phase=1 is requested.
 */
#include "purify_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;
static const double PI = 3.14159265358979323846; /* constant::pi, types.h:55 */

/* ------------------------------------------------------------------------------------------ */
/* Scalar kernels                                                                               */
/* ------------------------------------------------------------------------------------------ */

/* Modified Bessel function I0 by its power series sum_k (x^2/4)^k / (k!)^2.  All terms are
 * positive, so the relative error is a few ulp.  Stands in for boost::math::cyl_bessel_i(0, .)
 * (kernels.cc:20-21). */
double orc_bessel_i0(double x) {
  const double q = 0.25 * x * x;
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 500; ++k) {
    term *= q / ((double)k * (double)k);
    sum += term;
    if (term < 1e-17 * sum) break;
  }
  return sum;
}

/* kernels.cc:7-13 */
double orc_kaiser_bessel(double x, int J) { return orc_kaiser_bessel_general(x, J, 2.34 * J); }

/* kernels.cc:15-22: I0(alpha sqrt(1 - (2x/J)^2)) / I0(alpha).  The reference takes the real
 * sqrt of a real argument, so |2x/J| > 1 gives NaN there as well. */
double orc_kaiser_bessel_general(double x, int J, double alpha) {
  const double a = 2.0 * x / J;
  return orc_bessel_i0(alpha * sqrt(1.0 - a * a)) / orc_bessel_i0(alpha);
}

/* kernels.cc:24-38: Re[sin(eta)/eta] / 38828.11016883 with eta = sqrt((pi x J)^2 - alpha^2)
 * taken in the complex plane. */
double orc_ft_kaiser_bessel_general(double x, int J, double alpha) {
  const double t = (PI * x * J) * (PI * x * J) - alpha * alpha;
  const cplx eta = csqrt(t + 0.0 * I);
  const double normalisation = 38828.11016883;
  if (cabs(eta) == 0.0) return 1.0 / normalisation; /* limit sin(eta)/eta -> 1 */
  return creal(csin(eta) / eta) / normalisation;
}

/* kernels.cc:40-47 */
double orc_ft_kaiser_bessel(double x, int J) { return orc_ft_kaiser_bessel_general(x, J, 2.34 * J); }

/* kernels.cc:206-217 */
double orc_gaussian_general(double x, int J, double sigma) {
  (void)J;
  const double a = x / sigma;
  return exp(-a * a * 0.5);
}
/* kernels.cc:219-230 */
double orc_ft_gaussian_general(double x, int J, double sigma) {
  (void)J;
  const double a = x * sigma * PI;
  return sqrt(PI / 2) / sigma * exp(-a * a * 2);
}
/* kernels.cc:49-58 */
double orc_gaussian(double x, int J) { return orc_gaussian_general(x, J, 0.31 * pow(J, 0.52)); }
/* kernels.cc:60-69 */
double orc_ft_gaussian(double x, int J) { return orc_ft_gaussian_general(x, J, 0.31 * pow(J, 0.52)); }

/* kernels.cc:71-105: Schwab's rational approximation, only for J = 6 and alpha = 1. */
static double pswf_fraction(double eta, const double *p, const double *q) {
  /* numerator degree 5 (p[5] = 0), denominator degree 2, Horner's rule as in kernels.cc:86-99 */
  double num = p[5];
  for (int i = 1; i <= 5; ++i) num = eta * num + p[5 - i];
  double den = q[2];
  for (int i = 1; i <= 2; ++i) den = eta * den + q[2 - i];
  return num / den;
}
static double calc_for_pswf(double eta0, int J, double alpha) {
  static const double p1[6] = {8.203343e-2, -3.644705e-1, 6.278660e-1, -5.335581e-1, 2.312756e-1, 0.0};
  static const double p2[6] = {4.028559e-3, -3.697768e-2, 1.021332e-1, -1.201436e-1, 6.412774e-2, 0.0};
  static const double q1[3] = {1., 8.212018e-1, 2.078043e-1};
  static const double q2[3] = {1., 9.599102e-1, 2.918724e-1};
  if (J != 6 || alpha != 1) return 0;
  const double a = fabs(eta0);
  if (0 <= a && a <= 0.75) return pswf_fraction(eta0 * eta0 - 0.75 * 0.75, p1, q1);
  if (0.75 < a && a <= 1) return pswf_fraction(eta0 * eta0 - 1.0, p2, q2);
  return 0;
}
/* kernels.cc:107-122 */
double orc_pswf(double x, int J) {
  const double eta0 = 2 * x / J;
  const double alpha = 1;
  return calc_for_pswf(eta0, J, alpha) * pow(1 - eta0 * eta0, alpha);
}
/* kernels.cc:124-142 */
double orc_ft_pswf(double x, int J) { return calc_for_pswf(2 * x, J, 1); }
/* kernels.cc:188-196 */
double orc_pill_box(double x, int J) {
  (void)x;
  return 1.0 / (double)J;
}
/* kernels.cc:197-205: boost::math::sinc_pi(J x pi) = sin(t)/t with sinc(0) = 1 */
double orc_ft_pill_box(double x, int J) {
  const double t = J * x * PI;
  if (t == 0.0) return 1.0;
  return sin(t) / t;
}

static double kbmin_alpha(int J, double os) {
  /* kernels.cc:259-266 */
  return PI * sqrt((double)J * J / (os * os) * (os - 0.5) * (os - 0.5) - 0.8);
}

/* The lambdas built by create_kernels (kernels.cc:233-323). */
double orc_kernel_eval(int kind, int which, double x, int J, int ftsize, double os) {
  if (which == 1) x = x / ftsize - 0.5;
  switch (kind) {
  case ORC_KB:
    return which ? orc_ft_kaiser_bessel(x, J) : orc_kaiser_bessel(x, J);
  case ORC_KBMIN: {
    const double a = kbmin_alpha(J, os);
    return which ? orc_ft_kaiser_bessel_general(x, J, a) : orc_kaiser_bessel_general(x, J, a);
  }
  case ORC_PSWF:
    return which ? orc_ft_pswf(x, J) : orc_pswf(x, J);
  case ORC_GAUSS:
    return which ? orc_ft_gaussian(x, J) : orc_gaussian(x, J);
  case ORC_BOX:
    return which ? orc_ft_pill_box(x, J) : orc_pill_box(x, J);
  case ORC_GAUSS_ALT:
    return which ? orc_ft_gaussian_general(x, J, 1.0) : orc_gaussian_general(x, J, 1.0);
  default:
    return NAN;
  }
}

/* ------------------------------------------------------------------------------------------ */
/* Correction S and zero padding Z                                                             */
/* ------------------------------------------------------------------------------------------ */

/* operators.cc:101-121.  S(j,i) = 1/FTv(y0+j+1/2) * 1/FTu(x0+i+1/2) * chirp * nx * ny, where
 * the w = 0 chirp is 1/(nx ny) (wproj_utilities.cc:76-98), so S is the outer product. */
int orc_correction2d(double os, int ny, int nx, int kind, int Ju, int Jv, double *S) {
  const int Nu = (int)floor(nx * os), Nv = (int)floor(ny * os);
  const int x0 = (int)floor(Nu * 0.5 - nx * 0.5), y0 = (int)floor(Nv * 0.5 - ny * 0.5);
  if (kind == ORC_PSWF && (Ju != 6 || Jv != 6)) return -1; /* kernels.cc:241-244 */
  for (int j = 0; j < ny; ++j) {
    const double sv = 1.0 / orc_kernel_eval(kind, 1, y0 + j + 0.5, Jv, Nv, os);
    for (int i = 0; i < nx; ++i) {
      const double su = 1.0 / orc_kernel_eval(kind, 1, x0 + i + 0.5, Ju, Nu, os);
      S[(int64_t)j * nx + i] = sv * su;
    }
  }
  return 0;
}

/* operators.h:157-169 */
void orc_zero_pad(const double *S, int ny, int nx, double os, const double *x, double *grid) {
  const int Nu = (int)floor(nx * os), Nv = (int)floor(ny * os);
  const int x0 = (int)floor(Nu * 0.5 - nx * 0.5), y0 = (int)floor(Nv * 0.5 - ny * 0.5);
  const cplx *xc = (const cplx *)x;
  cplx *g = (cplx *)grid;
  memset(grid, 0, sizeof(cplx) * (size_t)Nu * Nv);
#pragma omp parallel for collapse(2)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i)
      g[(int64_t)(y0 + j) * Nu + x0 + i] = S[(int64_t)j * nx + i] * xc[(int64_t)j * nx + i];
}

/* operators.h:170-181 (S is real, so conj(S) = S) */
void orc_zero_pad_adjoint(const double *S, int ny, int nx, double os, const double *grid, double *x) {
  const int Nu = (int)floor(nx * os), Nv = (int)floor(ny * os);
  const int x0 = (int)floor(Nu * 0.5 - nx * 0.5), y0 = (int)floor(Nv * 0.5 - ny * 0.5);
  cplx *xc = (cplx *)x;
  const cplx *g = (const cplx *)grid;
#pragma omp parallel for collapse(2)
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i)
      xc[(int64_t)j * nx + i] = S[(int64_t)j * nx + i] * g[(int64_t)(y0 + j) * Nu + x0 + i];
}

/* ------------------------------------------------------------------------------------------ */
/* FFT: iterative radix-2 for powers of two, direct DFT otherwise (sizes in tests only).        */
/* Stands in for FFTW's plan_dft_2d (operators.h:234-259).                                      */
/* ------------------------------------------------------------------------------------------ */

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

static void twiddles(int n, int sign, cplx *w) {
  /* w[k] = exp(sign 2 pi i k / n), k < n */
  for (int k = 0; k < n; ++k) {
    const double a = 2.0 * PI * (double)k / (double)n;
    w[k] = cos(a) + sign * sin(a) * I;
  }
}

/* In-place transform of one contiguous line of length n with a precomputed table w (length n). */
static void fft_line(cplx *a, int n, const cplx *w, cplx *scratch) {
  if (is_pow2(n)) {
    /* bit reversal */
    for (int i = 1, j = 0; i < n; ++i) {
      int bit = n >> 1;
      for (; j & bit; bit >>= 1) j ^= bit;
      j ^= bit;
      if (i < j) {
        cplx t = a[i];
        a[i] = a[j];
        a[j] = t;
      }
    }
    for (int len = 2; len <= n; len <<= 1) {
      const int half = len >> 1, step = n / len;
      for (int i = 0; i < n; i += len)
        for (int k = 0; k < half; ++k) {
          const cplx t = a[i + k + half] * w[k * step];
          const cplx s = a[i + k];
          a[i + k] = s + t;
          a[i + k + half] = s - t;
        }
    }
  } else {
    for (int k = 0; k < n; ++k) {
      cplx s = 0;
      for (int t = 0; t < n; ++t) s += a[t] * w[(int)(((int64_t)k * t) % n)];
      scratch[k] = s;
    }
    memcpy(a, scratch, sizeof(cplx) * n);
  }
}

void orc_fft2d_raw(int Nv, int Nu, const double *in, double *out, int sign) {
  const cplx *src = (const cplx *)in;
  cplx *dst = (cplx *)out;
  if (dst != src) memcpy(dst, src, sizeof(cplx) * (size_t)Nu * Nv);
  cplx *wu = malloc(sizeof(cplx) * Nu), *wv = malloc(sizeof(cplx) * Nv);
  twiddles(Nu, sign, wu);
  twiddles(Nv, sign, wv);
#pragma omp parallel
  {
    cplx *scratch = malloc(sizeof(cplx) * (Nu > Nv ? Nu : Nv));
#pragma omp for schedule(static)
    for (int r = 0; r < Nv; ++r) fft_line(dst + (int64_t)r * Nu, Nu, wu, scratch);
    /* columns in blocks of 8 for cache locality */
    const int B = 8;
    cplx *col = malloc(sizeof(cplx) * (size_t)Nv * B);
#pragma omp for schedule(static)
    for (int c0 = 0; c0 < Nu; c0 += B) {
      const int nb = (Nu - c0) < B ? (Nu - c0) : B;
      for (int r = 0; r < Nv; ++r)
        for (int b = 0; b < nb; ++b) col[(int64_t)b * Nv + r] = dst[(int64_t)r * Nu + c0 + b];
      for (int b = 0; b < nb; ++b) fft_line(col + (int64_t)b * Nv, Nv, wv, scratch);
      for (int r = 0; r < Nv; ++r)
        for (int b = 0; b < nb; ++b) dst[(int64_t)r * Nu + c0 + b] = col[(int64_t)b * Nv + r];
    }
    free(col);
    free(scratch);
  }
  free(wu);
  free(wv);
}

/* operators.h:242-259: FFTW then output /= sqrt(output.size()) */
void orc_fft2d(int Nv, int Nu, const double *in, double *out, int sign) {
  orc_fft2d_raw(Nv, Nu, in, out, sign);
  const double s = 1.0 / sqrt((double)Nu * (double)Nv);
  const int64_t n = (int64_t)Nu * Nv * 2;
#pragma omp parallel for
  for (int64_t k = 0; k < n; ++k) out[k] *= s;
}

/* ------------------------------------------------------------------------------------------ */
/* Gridding matrix G                                                                           */
/* ------------------------------------------------------------------------------------------ */

/* utilities.cc:350-357 */
static double ref_mod(double x, double y) {
  double r = fmod(x, y);
  if (r < 0) r = y + r;
  return r;
}

/* exp(-2 pi i (a/2 + b/2)) of operators.cc:40,93 */
static cplx phase_ab(double a, double b, int literal) {
  if (literal) return cexp(-2.0 * PI * I * (a * 0.5 + b * 0.5));
  const int64_t s = (int64_t)a + (int64_t)b;
  return (s & 1) ? -1.0 : 1.0;
}

int64_t orc_gridding_matrix_nnz(int64_t M, int ny, int nx, double os, int Ju, int Jv, int w_term, int Jw) {
  const int Nu = (int)floor(nx * os), Nv = (int)floor(ny * os);
  if (w_term) return M * (int64_t)(Ju + Jw - 1) * (Jv + Jw - 1);
  const int ju = Ju < Nu ? Ju : Nu, jv = Jv < Nv ? Jv : Nv;
  return M * (int64_t)ju * jv;
}

/* insertion sort of one CSR row by column (Eigen keeps inner indices ordered) */
static void sort_row(int32_t *c, cplx *v, int n) {
  for (int a = 1; a < n; ++a) {
    int32_t ck = c[a];
    cplx vk = v[a];
    int b = a - 1;
    while (b >= 0 && c[b] > ck) {
      c[b + 1] = c[b];
      v[b + 1] = v[b];
      --b;
    }
    c[b + 1] = ck;
    v[b + 1] = vk;
  }
}

/* operators.cc:8-46 */
int64_t orc_gridding_matrix(const double *u, const double *v, const double *weights, int64_t M, int ny,
                            int nx, double os, int kind, int Jeval_u, int Jeval_v, int Ju, int Jv,
                            int literal_phase, int64_t *rowptr, int32_t *colidx, double *vals) {
  const int Nv = (int)floor(ny * os), Nu = (int)floor(nx * os);
  const int ju_max = Ju < Nu ? Ju : Nu, jv_max = Jv < Nv ? Jv : Nv;
  const int per = ju_max * jv_max;
  const cplx *w = (const cplx *)weights;
  cplx *val = (cplx *)vals;
  for (int64_t m = 0; m <= M; ++m) rowptr[m] = m * per;
#pragma omp parallel for schedule(static)
  for (int64_t m = 0; m < M; ++m) {
    int32_t *c = colidx + m * per;
    cplx *vv = val + m * per;
    int n = 0;
    const double k_u = floor(u[m] - ju_max * 0.5);
    const double k_v = floor(v[m] - jv_max * 0.5);
    for (int ju = 1; ju < ju_max + 1; ++ju) {
      const double ku = orc_kernel_eval(kind, 0, u[m] - (k_u + ju), Jeval_u, Nu, os);
      for (int jv = 1; jv < jv_max + 1; ++jv) {
        const int q = (int)ref_mod(k_u + ju, Nu);
        const int p = (int)ref_mod(k_v + jv, Nv);
        c[n] = p * Nu + q;
        vv[n] = phase_ab(k_u + ju, k_v + jv, literal_phase) * ku *
                orc_kernel_eval(kind, 0, v[m] - (k_v + jv), Jeval_v, Nv, os) * w[m];
        ++n;
      }
    }
    sort_row(c, vv, n);
  }
  return M * per;
}

/* ---- w-projection ---- */

/* projection_kernels.cc:28-30 */
static double pixel_to_lambda(double cell, int imsize, double os) {
  return 60. * 60. * 180. / cell / floor(os * imsize) / PI;
}

/* projection_kernels.cc:8-58 evaluated at (u, v) pixels with du, dv scales */
static cplx wkernel(int wkind, double u, double v, double w, double du, double dv) {
  switch (wkind) {
  case ORC_W_BOX:
    return 1.0; /* box_proj, projection_kernels.cc:8-10 */
  case ORC_W_SPHERE: {
    /* w_proj_sphere, projection_kernels.cc:23-27 */
    const double a = u * du, b = v * dv;
    const double t = 2 * PI * sqrt(a * a + b * b + w * w);
    const double s = (t == 0.0) ? 1.0 : sin(t) / t;
    return 2 * PI * s * cexp(2 * PI * I * w);
  }
  default: {
    /* w_projection_kernel_approx, projection_kernels.cc:31-47; w_proj_approx 16-22 */
    if ((fabs(u * du / w) < 1) && (fabs(v * dv / w) < 1)) {
      const double a = u * du, b = v * dv;
      return cexp(-PI * I * (a * a + b * b) / w) / w / I;
    }
    return 1.0;
  }
  }
}

/* the w-kernel sample at pixel offsets (u, v) (projection_kernels.cc:8-58) with the pixel -> lambda
 * scales of an imsizex x imsizey image at oversampling os (pixel_to_lambda, projection_kernels.cc:28-30) */
void orc_wkernel_eval(int wkind, double u, double v, double w, double cellx, double celly, int nx, int ny,
                      double os, double *out) {
  const cplx c = wkernel(wkind, u, v, w, pixel_to_lambda(cellx, nx, os), pixel_to_lambda(celly, ny, os));
  out[0] = creal(c);
  out[1] = cimag(c);
}
double orc_pixel_to_lambda(double cell, int imsize, double os) { return pixel_to_lambda(cell, imsize, os); }

/* convolution.h:43-52 */
void orc_linear_convol_1d(const double *f, int Jf, const double *g, int Lg, double *out) {
  const cplx *fc = (const cplx *)f, *gc = (const cplx *)g;
  cplx *o = (cplx *)out;
  const int L = Lg - Jf + 1;
  for (int j = 0; j < L; ++j) {
    cplx s = 0;
    for (int t = 0; t < Jf; ++t) s += fc[t] * gc[L - j - 1 + t];
    o[j] = s;
  }
}

/* convolution.h:54-67 */
void orc_linear_convol_2d(const double *fu, int Jfu, const double *fv, int Jfv, const double *g, int Jgv,
                          int Jgu, double *out) {
  const cplx *gc = (const cplx *)g;
  const int R = Jfv + Jgv - 1, C = Jfu + Jgu - 1;
  cplx *buffer = calloc((size_t)R * Jgu, sizeof(cplx)); /* R x Jgu, column-wise convolution */
  cplx *pad = calloc((size_t)(Jgv + 2 * (Jfv - 1) + Jgu + 2 * (Jfu - 1) + 8), sizeof(cplx));
  cplx *tmp = calloc((size_t)(R > C ? R : C) + 8, sizeof(cplx));
  for (int i = 0; i < Jgu; ++i) {
    const int P = Jfv - 1, Lg = Jgv + 2 * P;
    for (int k = 0; k < Lg; ++k) pad[k] = 0;
    for (int k = 0; k < Jgv; ++k) pad[P + k] = gc[(int64_t)k * Jgu + i];
    orc_linear_convol_1d(fv, Jfv, (const double *)pad, Lg, (double *)tmp);
    for (int r = 0; r < R; ++r) buffer[(int64_t)r * Jgu + i] = tmp[r];
  }
  cplx *o = (cplx *)out;
  for (int r = 0; r < R; ++r) {
    const int P = Jfu - 1, Lg = Jgu + 2 * P;
    for (int k = 0; k < Lg; ++k) pad[k] = 0;
    for (int k = 0; k < Jgu; ++k) pad[P + k] = buffer[(int64_t)r * Jgu + k];
    orc_linear_convol_1d(fu, Jfu, (const double *)pad, Lg, (double *)tmp);
    for (int c = 0; c < C; ++c) o[(int64_t)r * C + c] = tmp[c];
  }
  free(buffer);
  free(pad);
  free(tmp);
}

/* projection_kernels.cc:59-85.  Both separable factors use the u-slot kernel (the reference's
 * convol_kernelv calls kernelu, projection_kernels.cc:72-73). */
void orc_projection(int kind, int Jeval_u, int Ju, int Jv, int Jw, int wkind, double cellx, double celly,
                    int nx, int ny, double os, double u, double v, double w, double *out) {
  const int Nu = (int)floor(nx * os), Nv = (int)floor(ny * os);
  const double dul = pixel_to_lambda(cellx, nx, os), dvl = pixel_to_lambda(celly, ny, os);
  cplx fu[64], fv[64];
  cplx *g = malloc(sizeof(cplx) * (size_t)Jw * Jw);
  const double du = u - floor(u - Ju * 0.5);
  for (int ju = 0; ju < Ju; ++ju) fu[ju] = orc_kernel_eval(kind, 0, du - ju - 1, Jeval_u, Nu, os);
  const double dv = v - floor(v - Jv * 0.5);
  for (int jv = 0; jv < Jv; ++jv) fv[jv] = orc_kernel_eval(kind, 0, dv - jv - 1, Jeval_u, Nu, os);
  const double dwu = u - floor(u - Jw * 0.5), dwv = v - floor(v - Jw * 0.5);
  for (int jv = 0; jv < Jw; ++jv)
    for (int ju = 0; ju < Jw; ++ju) g[jv * Jw + ju] = wkernel(wkind, dwu - ju - 1, dwv - jv - 1, w, dul, dvl);
  (void)Nv;
  orc_linear_convol_2d((const double *)fu, Ju, (const double *)fv, Jv, (const double *)g, Jw, Jw, out);
  free(g);
}

/* operators.cc:49-99 */
int64_t orc_gridding_matrix_w(const double *u, const double *v, const double *w, const double *weights,
                              int64_t M, int ny, int nx, double os, int kind, int Jeval_u, int Jeval_v,
                              int Ju, int Jv, int Jw, int wkind, double cellx, double celly,
                              int64_t *rowptr, int32_t *colidx, double *vals) {
  (void)Jeval_v;
  const int Nv = (int)floor(ny * os), Nu = (int)floor(nx * os);
  const int Jwu = Jw + Ju - 1, Jwv = Jw + Jv - 1;
  const int per = Jwu * Jwv;
  const cplx *wt = (const cplx *)weights;
  cplx *val = (cplx *)vals;
  for (int64_t m = 0; m <= M; ++m) rowptr[m] = m * per;
#pragma omp parallel
  {
    cplx *proj = malloc(sizeof(cplx) * (size_t)per);
#pragma omp for schedule(static)
    for (int64_t m = 0; m < M; ++m) {
      orc_projection(kind, Jeval_u, Ju, Jv, Jw, wkind, cellx, celly, nx, ny, os, u[m], v[m], w[m],
                     (double *)proj);
      int32_t *c = colidx + m * per;
      cplx *vv = val + m * per;
      int n = 0;
      const int64_t kwu = (int64_t)floor(u[m] - Jwu * 0.5);
      const int64_t kwv = (int64_t)floor(v[m] - Jwv * 0.5);
      for (int ju = 1; ju < Jwu + 1; ++ju)
        for (int jv = 1; jv < Jwv + 1; ++jv) {
          const int q = (int)ref_mod((double)(kwu + ju), Nu);
          const int p = (int)ref_mod((double)(kwv + jv), Nv);
          c[n] = p * Nu + q;
          vv[n] = phase_ab((double)(kwu + ju), (double)(kwv + jv), 0) * proj[(jv - 1) * Jwu + (ju - 1)] * wt[m];
          ++n;
        }
      sort_row(c, vv, n);
    }
    free(proj);
  }
  return M * per;
}

/* ------------------------------------------------------------------------------------------ */
/* CSR adjoint and SpMV                                                                         */
/* ------------------------------------------------------------------------------------------ */

/* operators.h:135-136: G.adjoint() stored RowMajor; entries of a row ordered by vis index */
void orc_csr_adjoint(int64_t rows, int64_t cols, const int64_t *rowptr, const int32_t *colidx,
                     const double *vals, int64_t *t_rowptr, int32_t *t_colidx, double *t_vals) {
  const cplx *v = (const cplx *)vals;
  cplx *tv = (cplx *)t_vals;
  for (int64_t c = 0; c <= cols; ++c) t_rowptr[c] = 0;
  for (int64_t k = 0; k < rowptr[rows]; ++k) t_rowptr[colidx[k] + 1]++;
  for (int64_t c = 0; c < cols; ++c) t_rowptr[c + 1] += t_rowptr[c];
  int64_t *fill = malloc(sizeof(int64_t) * (size_t)(cols > 0 ? cols : 1));
  for (int64_t c = 0; c < cols; ++c) fill[c] = t_rowptr[c];
  for (int64_t r = 0; r < rows; ++r)
    for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
      const int64_t pos = fill[colidx[k]]++;
      t_colidx[pos] = (int32_t)r;
      tv[pos] = conj(v[k]);
    }
  free(fill);
}

/* utilities.h:144-159 */
void orc_spmv(int64_t rows, const int64_t *rowptr, const int32_t *colidx, const double *vals, const double *x,
              double *y) {
  const cplx *v = (const cplx *)vals, *xc = (const cplx *)x;
  cplx *yc = (cplx *)y;
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < rows; ++k) {
    cplx s = 0;
    for (int64_t t = rowptr[k]; t < rowptr[k + 1]; ++t) s += v[t] * xc[colidx[t]];
    yc[k] = s;
  }
}

/* ------------------------------------------------------------------------------------------ */
/* Chirp rows (wproj_utilities)                                                                 */
/* ------------------------------------------------------------------------------------------ */

/* wproj_utilities.cc:37-49 */
void orc_fov_cosines(double cell_x, double cell_y, int x_size, int y_size, double *L, double *Mc) {
  const double tl = cell_x * x_size, tm = cell_y * y_size;
  *L = 2 * sin(PI / 180. * tl / (60. * 60.) * 0.5);
  *Mc = 2 * sin(PI / 180. * tm / (60. * 60.) * 0.5);
}

/* wproj_utilities.cc:51-109 (generate_dde + generate_chirp with dde = 1).  The reference fills
 * chirp(l, m) of a column-major (y_size x x_size) image, so the flat vector the FFT operator
 * sees holds element (l, m) at l + m * y_size. */
void orc_generate_chirp(double w_rate, double cell_x, double cell_y, int x_size, int y_size, double *out) {
  double L, Mc;
  orc_fov_cosines(cell_x, cell_y, x_size, y_size, &L, &Mc);
  const double dx = L / x_size, dy = Mc / y_size;
  const double nz = (double)y_size * x_size;
  cplx *o = (cplx *)out;
  for (int l = 0; l < x_size; ++l)
    for (int m = 0; m < y_size; ++m) {
      const double x = (l + 0.5 - x_size * 0.5) * dx;
      const double y = (m + 0.5 - y_size * 0.5) * dy;
      o[(int64_t)l + (int64_t)m * y_size] = cexp(-2 * PI * I * w_rate * (sqrt(1 - x * x - y * y) - 1)) / nz;
    }
}

/* wproj_utilities.h:146-185 (bisection for the hard threshold keeping `energy` of the l2 energy) */
static double thres_bisect(const double *absv, int64_t n, double energy) {
  if (energy >= 1) return 0;
  double mx = 0, total = 0;
  for (int64_t j = 0; j < n; ++j) {
    if (absv[j] > mx) mx = absv[j];
    total += absv[j] * absv[j];
  }
  const double target = total * energy;
  double lower = 0, upper = mx, cur = upper * 0.5, old = upper;
  int converged = 0;
  for (int i = 0; i < 200; ++i) {
    double s = 0;
    for (int64_t j = 0; j < n; ++j)
      if (absv[j] > cur) s += absv[j] * absv[j];
    if ((fabs(cur - old) / fabs(old) < 1e-3) && ((s >= target) && (((s / target) - 1) < 0.01))) {
      converged = 1;
      break;
    }
    if (s < target)
      upper = cur;
    else
      lower = cur;
    old = cur;
    cur = (lower + upper) * 0.5;
  }
  if (!converged) cur = lower;
  return cur;
}

double orc_sparsify_row_dense_thres(const double *row, int64_t n, double energy) {
  const cplx *r = (const cplx *)row;
  double *a = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int64_t j = 0; j < n; ++j) a[j] = cabs(r[j]);
  const double t = thres_bisect(a, n, energy);
  free(a);
  return t;
}

/* wproj_utilities.h:101-145 on the stored (nonzero) magnitudes */
double orc_sparsify_row_thres_real(const double *absvals, int64_t n, double energy) {
  return thres_bisect(absvals, n, energy);
}

/* utilities.cc:335-348 ind2sub(sub, cols, rows) = (floor(sub / cols), sub % cols) */
static void ind2sub(int64_t sub, int cols, int *r, int *c) {
  *r = (int)(sub / cols);
  *c = (int)(sub % cols);
}

/* wproj_utilities.h:259-285: output(k) = sum over G entries of G * chirp(pos) with the
 * fftshift-centred bounds check.  The chirp row is given densely. */
void orc_row_wise_convolution(const int32_t *g_idx, const double *g_val, int g_n, const double *chirp_dense,
                              int x_size, int y_size, double *out_dense) {
  const cplx *gv = (const cplx *)g_val, *ch = (const cplx *)chirp_dense;
  cplx *o = (cplx *)out_dense;
  const int xc = (int)floor(x_size * 0.5), yc = (int)floor(y_size * 0.5);
  const int xcu = (int)ceil(x_size * 0.5), ycu = (int)ceil(y_size * 0.5);
  const int64_t N = (int64_t)x_size * y_size;
  for (int64_t k = 0; k < N; ++k) {
    int js, is;
    ind2sub(k, y_size, &js, &is);
    const int iW = (int)ref_mod(is + xc, x_size), jW = (int)ref_mod(js + yc, y_size);
    cplx s = 0;
    for (int t = 0; t < g_n; ++t) {
      int jsg, isg;
      ind2sub(g_idx[t], y_size, &jsg, &isg);
      const int iG = (int)ref_mod(isg + xc, x_size), jG = (int)ref_mod(jsg + yc, y_size);
      const int iC = iW - iG, jC = jW - jG;
      if (-xc <= iC && iC < xcu && -yc <= jC && jC < ycu) {
        const int isC = (int)ref_mod(iC, x_size), jsC = (int)ref_mod(jC, y_size);
        const int64_t pos = (int64_t)jsC * x_size + isC;
        s += gv[t] * ch[pos];
      }
    }
    o[k] = s;
  }
}

/* wproj_utilities.cc:110-188 for a square grid.  The chirp row is FFT(chirp)*sqrt(N), hard
 * thresholded at e_chirp (convert_sparse also drops |c| <= 1e-10 max|c|), convolved with the G
 * row, then thresholded at e_wproj and pruned (|value| > thres kept). */
int orc_wprojection_matrix_dense(int64_t rows, const int64_t *rowptr, const int32_t *colidx,
                                 const double *vals, int x_size, int y_size, const double *w,
                                 double cell_x, double cell_y, double e_chirp, double e_wproj,
                                 double *out_dense, uint8_t *out_mask) {
  const int64_t N = (int64_t)x_size * y_size;
  cplx *img = malloc(sizeof(cplx) * (size_t)N), *ch = malloc(sizeof(cplx) * (size_t)N);
  double *absv = malloc(sizeof(double) * (size_t)N);
  for (int64_t m = 0; m < rows; ++m) {
    /* create_chirp_row(w, cell_x, cell_y, ftsizev = x_size, ftsizeu = y_size) ->
       generate_chirp(w, cell_x, cell_y, x_size = y_size, y_size = x_size) */
    orc_generate_chirp(w[m], cell_x, cell_y, y_size, x_size, (double *)img);
    orc_fft2d(y_size, x_size, (const double *)img, (double *)ch, -1);
    const double sq = sqrt((double)N);
    double mx = 0;
    for (int64_t k = 0; k < N; ++k) {
      ch[k] *= sq;
      absv[k] = cabs(ch[k]);
      if (absv[k] > mx) mx = absv[k];
    }
    const double thr = thres_bisect(absv, N, e_chirp);
    const double keep = (mx * 1e-10 > thr) ? mx * 1e-10 : thr; /* convert_sparse, wproj_utilities.h:90-100 */
    for (int64_t k = 0; k < N; ++k)
      if (!(absv[k] > keep)) ch[k] = 0; /* sparseView keeps |x| > reference */
    cplx *row = (cplx *)out_dense + m * N;
    const int gn = (int)(rowptr[m + 1] - rowptr[m]);
    orc_row_wise_convolution(colidx + rowptr[m], vals + 2 * rowptr[m], gn, (const double *)ch, x_size, y_size,
                             (double *)row);
    /* sparsify_row_thres over the stored nonzeros of the output row */
    int64_t nz = 0;
    for (int64_t k = 0; k < N; ++k)
      if (row[k] != 0) absv[nz++] = cabs(row[k]);
    const double t2 = thres_bisect(absv, nz, e_wproj);
    for (int64_t k = 0; k < N; ++k) {
      const int kept = (row[k] != 0) && (cabs(row[k]) > t2);
      out_mask[m * N + k] = (uint8_t)kept;
      if (!kept) row[k] = 0;
    }
  }
  free(img);
  free(ch);
  free(absv);
  return 0;
}

/* ---- periodic 2-D wavelet transform (the SARA dictionary's Daubechies bases) ----
 * One line: lo[i] = sum_k h[k] x[(2i + k) mod n], hi[i] = sum_k g[k] x[(2i + k) mod n] with
 * g[k] = (-1)^k h[L-1-k]; low-pass half first.  The inverse is the transpose.  A level filters the rows
 * of the top-left (ny >> l) x (nx >> l) block, then its columns; the inverse undoes the columns first. */
static void dwt_line(const cplx *x, int n, int stride, const double *h, const double *g, int L, cplx *out) {
  const int half = n / 2;
  for (int i = 0; i < half; ++i) {
    cplx lo = 0, hi = 0;
    for (int k = 0; k < L; ++k) {
      const cplx a = x[(int64_t)((2 * i + k) % n) * stride];
      lo += h[k] * a;
      hi += g[k] * a;
    }
    out[i] = lo;
    out[half + i] = hi;
  }
}
static void idwt_line(const cplx *c, int n, int stride, const double *h, const double *g, int L, cplx *out) {
  const int half = n / 2;
  for (int j = 0; j < n; ++j) out[j] = 0;
  for (int i = 0; i < half; ++i) {
    const cplx lo = c[(int64_t)i * stride], hi = c[(int64_t)(half + i) * stride];
    for (int k = 0; k < L; ++k) out[(2 * i + k) % n] += h[k] * lo + g[k] * hi;
  }
}

void orc_dwt2(const double *img, int ny, int nx, const double *h, int L, int levels, double *out) {
  cplx *c = (cplx *)out;
  const cplx *x = (const cplx *)img;
  double g[64];
  for (int k = 0; k < L; ++k) g[k] = ((k % 2) ? -1.0 : 1.0) * h[L - 1 - k];
  for (int64_t t = 0; t < (int64_t)ny * nx; ++t) c[t] = x[t];
  for (int l = 0; l < levels; ++l) {
    const int hh = ny >> l, ww = nx >> l;
#pragma omp parallel
    {
      cplx *tmp = malloc(sizeof(cplx) * (size_t)(hh > ww ? hh : ww));
#pragma omp for schedule(static)
      for (int r = 0; r < hh; ++r) {
        dwt_line(c + (int64_t)r * nx, ww, 1, h, g, L, tmp);
        for (int i = 0; i < ww; ++i) c[(int64_t)r * nx + i] = tmp[i];
      }
#pragma omp for schedule(static)
      for (int q = 0; q < ww; ++q) {
        dwt_line(c + q, hh, nx, h, g, L, tmp);
        for (int i = 0; i < hh; ++i) c[(int64_t)i * nx + q] = tmp[i];
      }
      free(tmp);
    }
  }
}

void orc_idwt2(const double *coef, int ny, int nx, const double *h, int L, int levels, double *out) {
  cplx *c = (cplx *)out;
  const cplx *x = (const cplx *)coef;
  double g[64];
  for (int k = 0; k < L; ++k) g[k] = ((k % 2) ? -1.0 : 1.0) * h[L - 1 - k];
  for (int64_t t = 0; t < (int64_t)ny * nx; ++t) c[t] = x[t];
  for (int l = levels - 1; l >= 0; --l) {
    const int hh = ny >> l, ww = nx >> l;
#pragma omp parallel
    {
      cplx *tmp = malloc(sizeof(cplx) * (size_t)(hh > ww ? hh : ww));
#pragma omp for schedule(static)
      for (int q = 0; q < ww; ++q) {
        idwt_line(c + q, hh, nx, h, g, L, tmp);
        for (int i = 0; i < hh; ++i) c[(int64_t)i * nx + q] = tmp[i];
      }
#pragma omp for schedule(static)
      for (int r = 0; r < hh; ++r) {
        idwt_line(c + (int64_t)r * nx, ww, 1, h, g, L, tmp);
        for (int i = 0; i < ww; ++i) c[(int64_t)r * nx + i] = tmp[i];
      }
      free(tmp);
    }
  }
}

/* ------------------------------------------------------------------------------------------ */
/* Utilities                                                                                   */
/* ------------------------------------------------------------------------------------------ */

/* operators.h:405-410: lambda -> set_cell_size (utilities.cc:208-250) -> radians -> uv_scale
 * (utilities.cc:258-278) -> pixels; radians -> uv_scale only (set_cell_size's own radians branch,
 * utilities.cc:240-244, is not on this path) */
void orc_uv_to_pixels(const double *u, const double *v, int64_t M, int units, double cell_x, double cell_y,
                      int nx, int ny, double os, double *u_px, double *v_px) {
  double cu = cell_x, cv = cell_y;
  if (units == 0) {
    double mu = 0, mv = 0;
    for (int64_t m = 0; m < M; ++m) {
      if (u[m] * u[m] > mu) mu = u[m] * u[m];
      if (v[m] * v[m] > mv) mv = v[m] * v[m];
    }
    mu = sqrt(mu);
    mv = sqrt(mv);
    if (cu == 0 && cv == 0) {
      cu = (180 * 3600) / mu / PI / 3;
      cv = (180 * 3600) / mv / PI / 3;
    }
    if (cv == 0) cv = cu;
  }
  double su = 1, sv = 1;
  if (units == 0) {
    su = 180 * 3600 / cu / PI;
    sv = 180 * 3600 / cv / PI;
  }
  const int Nu = (int)floor(os * nx), Nv = (int)floor(os * ny);
  for (int64_t m = 0; m < M; ++m) {
    double ur = u[m], vr = v[m];
    if (units == 0) {
      ur = u[m] / su * 2 * PI;
      vr = v[m] / sv * 2 * PI;
    }
    if (units == 2) {
      u_px[m] = ur;
      v_px[m] = vr;
    } else {
      u_px[m] = ur / (2 * PI) * Nu;
      v_px[m] = vr / (2 * PI) * Nv;
    }
  }
}

/* utilities.cc:396-421 */
double orc_calculate_l2_radius(int64_t y_size, double sigma, double n_sigma, int chi2) {
  if (chi2) {
    const double r = sqrt(2.0 * y_size + n_sigma * sqrt(4.0 * y_size));
    return sigma == 0 ? r : r * sigma;
  }
  const double ys = (double)y_size;
  const double alpha = 1. / (8 * ys) - 1. / (128 * ys * ys);
  const double mean = sqrt(2) * sqrt(ys) * (1 - alpha);
  const double sd = sqrt(2 * ys * alpha * (2 - alpha));
  return sigma == 0 ? mean + n_sigma * sd : (mean + n_sigma * sd) * sigma;
}

/* utilities.cc:422-432 */
double orc_snr_to_standard_deviation(const double *y0, int64_t M, double snr) {
  const cplx *y = (const cplx *)y0;
  double s = 0;
  for (int64_t m = 0; m < M; ++m) s += creal(y[m]) * creal(y[m]) + cimag(y[m]) * cimag(y[m]);
  return sqrt(s) / sqrt((double)M) * pow(10.0, -(snr / 20.0));
}

/* ------------------------------------------------------------------------------------------ */
/* Full operator                                                                               */
/* ------------------------------------------------------------------------------------------ */

struct orc_op {
  int ny, nx, Nv, Nu;
  double os;
  int64_t M, nnz;
  int64_t *g_rowptr, *h_rowptr;
  int32_t *g_col, *h_col;
  double *g_val, *h_val;
  double *S;
  double norm;
};

/* operators.h:298-329 and 373-396.  The gridding kernels are passed as (kernelv, kernelu) into
 * the (kernelu, kernelv) slots (operators.h:324): the u-slot evaluates with Jv, the v-slot with Ju. */
orc_op *orc_op_create(const double *u_px, const double *v_px, const double *w, const double *weights,
                      int64_t M, int ny, int nx, double os, int kind, int Ju, int Jv, int w_term, int Jw,
                      double cellx, double celly, int literal_phase) {
  if (kind == ORC_PSWF && (Ju != 6 || Jv != 6)) return NULL;
  orc_op *op = calloc(1, sizeof(orc_op));
  op->ny = ny;
  op->nx = nx;
  op->os = os;
  op->Nu = (int)floor(nx * os);
  op->Nv = (int)floor(ny * os);
  op->M = M;
  op->norm = 1.0;
  op->S = malloc(sizeof(double) * (size_t)nx * ny);
  orc_correction2d(os, ny, nx, kind, Ju, Jv, op->S);
  const int64_t nnz = orc_gridding_matrix_nnz(M, ny, nx, os, Ju, Jv, w_term, Jw);
  op->nnz = nnz;
  op->g_rowptr = malloc(sizeof(int64_t) * (size_t)(M + 1));
  op->g_col = malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
  op->g_val = malloc(sizeof(double) * 2 * (size_t)(nnz > 0 ? nnz : 1));
  if (w_term) {
    double mean = 0;
    for (int64_t m = 0; m < M; ++m) mean += w[m];
    if (M > 0) mean /= (double)M;
    double *wc = malloc(sizeof(double) * (size_t)(M > 0 ? M : 1));
    for (int64_t m = 0; m < M; ++m) wc[m] = w[m] - mean; /* operators.h:323 */
    orc_gridding_matrix_w(u_px, v_px, wc, weights, M, ny, nx, os, kind, Jv, Ju, Ju, Jv, Jw, ORC_W_APPROX, cellx,
                          celly, op->g_rowptr, op->g_col, op->g_val);
    free(wc);
  } else {
    orc_gridding_matrix(u_px, v_px, weights, M, ny, nx, os, kind, Jv, Ju, Ju, Jv, literal_phase, op->g_rowptr,
                        op->g_col, op->g_val);
  }
  const int64_t Ng = (int64_t)op->Nu * op->Nv;
  op->h_rowptr = malloc(sizeof(int64_t) * (size_t)(Ng + 1));
  op->h_col = malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
  op->h_val = malloc(sizeof(double) * 2 * (size_t)(nnz > 0 ? nnz : 1));
  orc_csr_adjoint(M, Ng, op->g_rowptr, op->g_col, op->g_val, op->h_rowptr, op->h_col, op->h_val);
  return op;
}

void orc_op_destroy(orc_op *op) {
  if (!op) return;
  free(op->g_rowptr);
  free(op->g_col);
  free(op->g_val);
  free(op->h_rowptr);
  free(op->h_col);
  free(op->h_val);
  free(op->S);
  free(op);
}

void orc_op_set_norm(orc_op *op, double norm) { op->norm = norm; }
double orc_op_get_norm(const orc_op *op) { return op->norm; }
int64_t orc_op_nnz(const orc_op *op) { return op->nnz; }
int orc_op_grid_dims(const orc_op *op, int *Nv, int *Nu) {
  *Nv = op->Nv;
  *Nu = op->Nu;
  return 0;
}
const double *orc_op_correction(const orc_op *op) { return op->S; }

void orc_op_apply_G(const orc_op *op, const double *grid, double *y) {
  orc_spmv(op->M, op->g_rowptr, op->g_col, op->g_val, grid, y);
}
void orc_op_apply_GH(const orc_op *op, const double *y, double *grid) {
  orc_spmv((int64_t)op->Nu * op->Nv, op->h_rowptr, op->h_col, op->h_val, y, grid);
}

/* y = G F Z S x / norm (operators.h:293, 325, 393) */
void orc_op_forward(const orc_op *op, const double *x, double *y) {
  const int64_t Ng = (int64_t)op->Nu * op->Nv;
  double *g1 = malloc(sizeof(double) * 2 * (size_t)Ng), *g2 = malloc(sizeof(double) * 2 * (size_t)Ng);
  orc_zero_pad(op->S, op->ny, op->nx, op->os, x, g1);
  orc_fft2d(op->Nv, op->Nu, g1, g2, -1);
  orc_op_apply_G(op, g2, y);
  const double s = 1.0 / op->norm;
  for (int64_t m = 0; m < 2 * op->M; ++m) y[m] *= s;
  free(g1);
  free(g2);
}

/* x = S^H Z^H F^H G^H y / norm (operators.h:294, 326, 394) */
void orc_op_adjoint(const orc_op *op, const double *y, double *x) {
  const int64_t Ng = (int64_t)op->Nu * op->Nv;
  double *g1 = malloc(sizeof(double) * 2 * (size_t)Ng), *g2 = malloc(sizeof(double) * 2 * (size_t)Ng);
  orc_op_apply_GH(op, y, g1);
  orc_fft2d(op->Nv, op->Nu, g1, g2, +1);
  orc_zero_pad_adjoint(op->S, op->ny, op->nx, op->os, g2, x);
  const double s = 1.0 / op->norm;
  for (int64_t k = 0; k < 2 * (int64_t)op->nx * op->ny; ++k) x[k] *= s;
  free(g1);
  free(g2);
}

static double l2norm(const double *a, int64_t n2) {
  /* scaled sum of squares, the stableNorm idea */
  double scale = 0, ssq = 1;
  for (int64_t k = 0; k < n2; ++k) {
    const double v = fabs(a[k]);
    if (v == 0) continue;
    if (scale < v) {
      ssq = 1 + ssq * (scale / v) * (scale / v);
      scale = v;
    } else {
      ssq += (v / scale) * (v / scale);
    }
  }
  return scale * sqrt(ssq);
}

/* SOPT power_method (called at operators.h:390-391; SOPT is not vendored).  Iterates
 * x <- Phi^H Phi x / ||Phi^H Phi x|| from start/||start|| and stops when
 * |old - new| / old < tol^2; returns sqrt(eigenvalue estimate).  niters <= 0 returns 1. */
double orc_op_power_method(const orc_op *op, int iters, double tol, const double *start) {
  if (iters <= 0) return 1.0;
  const int64_t n = (int64_t)op->nx * op->ny;
  double *x = malloc(sizeof(double) * 2 * (size_t)n), *y = malloc(sizeof(double) * 2 * (size_t)op->M);
  for (int64_t k = 0; k < n; ++k) {
    x[2 * k] = start ? start[2 * k] : 1.0;
    x[2 * k + 1] = start ? start[2 * k + 1] : 0.0;
  }
  double nrm = l2norm(x, 2 * n);
  for (int64_t k = 0; k < 2 * n; ++k) x[k] /= nrm;
  double old = 0, val = 1;
  for (int i = 0; i < iters; ++i) {
    orc_op_forward(op, x, y);
    orc_op_adjoint(op, y, x);
    val = l2norm(x, 2 * n);
    if (!(val > 0)) {
      old = val;
      break;
    }
    for (int64_t k = 0; k < 2 * n; ++k) x[k] /= val;
    if (tol * tol > fabs(old - val) / old) {
      old = val;
      break;
    }
    old = val;
  }
  free(x);
  free(y);
  return sqrt(old);
}

int orc_omp_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void orc_set_omp_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}
