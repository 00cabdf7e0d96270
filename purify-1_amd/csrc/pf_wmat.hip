// pf_wmat.hip -- sparse gridding matrices: the chirp-based w-projection matrix of
// wproj_utilities::wprojection_matrix (cpp/purify/wproj_utilities.cc:132-188) built on the GPU in FP64,
// and the CSR gather used to apply a given sparse gridding matrix.
//
// wprojection_matrix_dev processes the rows of G in batches; per row:
//   1. the chirp image of w (generate_chirp / generate_dde, wproj_utilities.cc:51-109, dde = 1),
//   2. its unnormalised DFT (the unitary FFT times sqrt(N), wproj_utilities.cc:119-126) with the
//      library's own FFT,
//   3. the bisection threshold keeping energy_chirp of the l2 energy (sparsify_row_dense_thres,
//      wproj_utilities.h:146-185) and convert_sparse's keep rule (wproj_utilities.h:89-100),
//   4. the fftshift-centred circular convolution with the row of G (row_wise_sparse_convolution,
//      wproj_utilities.h:222-287),
//   5. the threshold keeping energy_wproj (sparsify_row_thres, wproj_utilities.h:101-145) and the prune.
// Every reduction is a fixed-order block tree, so the result does not depend on scheduling.
#include <algorithm>
#include <cmath>

#include "../../include/purify_amd.h"
#include "pf_internal.h"

namespace pf {

namespace {

constexpr double PI_M = 3.14159265358979323846;
constexpr int WM_THREADS = 256;

__device__ inline int64_t mod_pos(int64_t a, int64_t n) {
  const int64_t r = a % n;
  return r < 0 ? r + n : r;
}

// Chirp images of B rows: pixel (l, m) of the xs x ys image at flat index l + m * ys (the reference
// fills chirp(l, m) of a column-major matrix), exp(-2 pi i w (sqrt(1 - x^2 - y^2) - 1)) / nz.
__global__ void k_wm_chirp(int B, const double* __restrict__ w, double dx, double dy, int xs, int ys, double nz,
                           cpx<double>* __restrict__ out) {
  const int64_t N = (int64_t)xs * ys;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * B) return;
  const int b = (int)(t / N);
  const int64_t p = t - (int64_t)b * N;
  const int m = (int)(p / ys), l = (int)(p - (int64_t)m * ys);
  const double x = (l + 0.5 - xs * 0.5) * dx, y = (m + 0.5 - ys * 0.5) * dy;
  const double ph = -2 * PI_M * w[b] * (sqrt(1 - x * x - y * y) - 1);
  out[t] = mk<double>(cos(ph) / nz, sin(ph) / nz);
}

__global__ void k_wm_abs(int64_t n, const cpx<double>* __restrict__ v, double* __restrict__ a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = hypot(v[t].x, v[t].y);
}

__device__ double block_reduce(double v, double* sh, bool is_max) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = WM_THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = is_max ? fmax(sh[threadIdx.x], sh[threadIdx.x + s])
                                                       : sh[threadIdx.x] + sh[threadIdx.x + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// One block per row: the hard threshold that keeps `energy` of the row's l2 energy, by the reference's
// bisection (at most 200 halvings, stop when the threshold moved < 1e-3 relatively and the kept energy
// is within 1 % above the target; not converged: the lower bound).  Also the row's largest magnitude.
__global__ void __launch_bounds__(WM_THREADS) k_wm_thres(const double* __restrict__ a, int64_t N, double energy,
                                                         double* __restrict__ thr, double* __restrict__ mxo) {
  __shared__ double sh[WM_THREADS];
  const double* r = a + (int64_t)blockIdx.x * N;
  double mx = 0, tot = 0;
  for (int64_t k = threadIdx.x; k < N; k += WM_THREADS) {
    mx = fmax(mx, r[k]);
    tot += r[k] * r[k];
  }
  mx = block_reduce(mx, sh, true);
  tot = block_reduce(tot, sh, false);
  double cur = 0;
  if (energy < 1) {
    const double target = tot * energy;
    double lower = 0, upper = mx, old = upper;
    cur = upper * 0.5;
    bool conv = false;
    for (int it = 0; it < 200; ++it) {
      double s = 0;
      for (int64_t k = threadIdx.x; k < N; k += WM_THREADS)
        if (r[k] > cur) s += r[k] * r[k];
      s = block_reduce(s, sh, false);
      if ((fabs(cur - old) / fabs(old) < 1e-3) && (s >= target) && (((s / target) - 1) < 0.01)) {
        conv = true;
        break;
      }
      if (s < target) upper = cur;
      else lower = cur;
      old = cur;
      cur = (lower + upper) * 0.5;
    }
    if (!conv) cur = lower;
  }
  if (threadIdx.x == 0) {
    thr[blockIdx.x] = cur;
    mxo[blockIdx.x] = mx;
  }
}

// convert_sparse: keep |c| > max(1e-10 max|c|, threshold); count the kept entries of each row.
__global__ void k_wm_keep(int64_t N, int B, cpx<double>* __restrict__ ch, const double* __restrict__ a,
                          const double* __restrict__ thr, const double* __restrict__ mx, int* __restrict__ cnt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * B) return;
  const int b = (int)(t / N);
  const double keep = fmax(mx[b] * 1e-10, thr[b]);
  if (a[t] > keep) atomicAdd(&cnt[b], 1);
  else ch[t] = mk<double>(0.0, 0.0);
}

// row_wise_sparse_convolution for B rows: out[k] = sum over the row's G entries of G * chirp(pos) with
// the fftshift-centred bounds check; hit[b] counts the positions reached through a kept chirp entry
// (the reference's non_zero_W).
__global__ void k_wm_conv(int B, int xs, int ys, const int64_t* __restrict__ gptr, const int32_t* __restrict__ gcol,
                          const cpx<double>* __restrict__ gval, const cpx<double>* __restrict__ ch,
                          cpx<double>* __restrict__ out, int* __restrict__ hit) {
  const int64_t N = (int64_t)xs * ys;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * B) return;
  const int b = (int)(t / N);
  const int64_t k = t - (int64_t)b * N;
  const int xc = xs / 2, yc = ys / 2, xcu = (xs + 1) / 2, ycu = (ys + 1) / 2;
  // utilities::ind2sub(k, y_size, x_size) = (k / y_size, k % y_size)
  const int js = (int)(k / ys), is = (int)(k % ys);
  const int iW = (int)mod_pos(is + xc, xs), jW = (int)mod_pos(js + yc, ys);
  const cpx<double>* c = ch + (int64_t)b * N;
  double sr = 0, si = 0;
  bool any = false;
  for (int64_t e = gptr[b]; e < gptr[b + 1]; ++e) {
    const int g = gcol[e];
    const int jsg = g / ys, isg = g % ys;
    const int iG = (int)mod_pos(isg + xc, xs), jG = (int)mod_pos(jsg + yc, ys);
    const int iC = iW - iG, jC = jW - jG;
    if (-xc <= iC && iC < xcu && -yc <= jC && jC < ycu) {
      const int64_t pos = mod_pos(jC, ys) * xs + mod_pos(iC, xs);
      const cpx<double> gv = gval[e], cv = c[pos];
      sr += gv.x * cv.x - gv.y * cv.y;
      si += gv.x * cv.y + gv.y * cv.x;
      any |= (cv.x != 0.0 || cv.y != 0.0);
    }
  }
  out[t] = mk<double>(sr, si);
  if (any) atomicAdd(&hit[b], 1);
}

template <typename R>
__global__ void __launch_bounds__(256) k_csr_gather(int64_t rows, const int64_t* __restrict__ ptr,
                                                    const int32_t* __restrict__ col, const cpx<R>* __restrict__ val,
                                                    const cpx<R>* __restrict__ x, cpx<R>* __restrict__ y, R scl) {
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);  // one wave per row
  const int lane = threadIdx.x & 63;
  if (r >= rows) return;
  cpx<R> acc = mk<R>(R(0), R(0));
  for (int64_t k = ptr[r] + lane; k < ptr[r + 1]; k += 64) {
    const cpx<R> a = val[k], g = x[col[k]];
    acc.x += a.x * g.x - a.y * g.y;
    acc.y += a.x * g.y + a.y * g.x;
  }
  acc = wave_sum(acc);
  if (lane == 0) y[r] = scale(acc, scl);
}

unsigned nblocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

void wprojection_matrix_dev(int64_t rows, const int64_t* rowptr, const int32_t* col, const double* val, int x_size,
                            int y_size, const double* w, double cell_x, double cell_y, double e_chirp, double e_wproj,
                            HostCsr& out) {
  if (x_size != y_size) throw Error(PF_ERR_UNSUPPORTED, "the chirp w-projection matrix needs a square grid");
  if (x_size < 2 || (x_size & (x_size - 1))) throw Error(PF_ERR_UNSUPPORTED, "grid side must be a power of two");
  const int64_t N = (int64_t)x_size * y_size;
  out.rows = rows;
  out.cols = N;
  out.rowptr.assign((size_t)rows + 1, 0);
  out.col.clear();
  out.val.clear();
  if (rows == 0) return;
  for (int64_t k = 0; k < rowptr[rows]; ++k)
    if (col[k] < 0 || col[k] >= N) throw Error(PF_ERR_ARG, "G column index outside the grid");
  // create_chirp_row(w, cell_x, cell_y, ftsizev = x_size, ftsizeu = y_size) calls
  // generate_chirp(w, cell_x, cell_y, x_size = y_size, y_size = x_size): image sides xs = y_size, ys = x_size
  const int xs = y_size, ys = x_size;
  const double L = 2 * std::sin(PI_M / 180. * (cell_x * xs) / (60. * 60.) * 0.5);  // fov_cosines
  const double Mc = 2 * std::sin(PI_M / 180. * (cell_y * ys) / (60. * 60.) * 0.5);
  const double dx = L / xs, dy = Mc / ys, nz = (double)ys * xs;
  const int64_t B = std::max<int64_t>(1, std::min<int64_t>(rows, (int64_t(1) << 22) / N));
  Fft2dPlan plan;
  fft_plan_init(plan, y_size, x_size, 1);
  cpx<double>* d_ch = dev_alloc_n<cpx<double>>(B * N);
  cpx<double>* d_ws = dev_alloc_n<cpx<double>>(B * N);
  cpx<double>* d_out = dev_alloc_n<cpx<double>>(B * N);
  double* d_abs = dev_alloc_n<double>(B * N);
  double* d_w = dev_alloc_n<double>(B);
  double *d_thr = dev_alloc_n<double>(B), *d_mx = dev_alloc_n<double>(B);
  int *d_cnt = dev_alloc_n<int>(B), *d_hit = dev_alloc_n<int>(B);
  const int64_t maxg = [&] {
    int64_t m = 0;
    for (int64_t r = 0; r < rows; ++r) m = std::max(m, rowptr[r + 1] - rowptr[r]);
    return m;
  }();
  int64_t* d_gptr = dev_alloc_n<int64_t>(B + 1);
  int32_t* d_gcol = dev_alloc_n<int32_t>(B * maxg);
  cpx<double>* d_gval = dev_alloc_n<cpx<double>>(B * maxg);
  std::vector<double> h_out((size_t)(B * N) * 2), h_abs((size_t)(B * N)), h_thr(B);
  std::vector<int> h_cnt(B), h_hit(B);
  std::vector<int64_t> gp(B + 1);
  hipStream_t s = cstream();
  try {
    for (int64_t r0 = 0; r0 < rows; r0 += B) {
      const int nb = (int)std::min(B, rows - r0);
      const int64_t NB = N * nb;
      // the batch's rows of G and their w
      for (int b = 0; b <= nb; ++b) gp[b] = rowptr[r0 + b] - rowptr[r0];
      const int64_t ge = gp[nb];
      h2d(d_gptr, gp.data(), sizeof(int64_t) * (nb + 1));
      if (ge) {
        h2d(d_gcol, col + rowptr[r0], sizeof(int32_t) * ge);
        h2d(d_gval, val + 2 * rowptr[r0], sizeof(cpx<double>) * ge);
      }
      h2d(d_w, w + r0, sizeof(double) * nb);
      PF_HIP(hipMemsetAsync(d_cnt, 0, sizeof(int) * nb, s));
      PF_HIP(hipMemsetAsync(d_hit, 0, sizeof(int) * nb, s));
      // 1-2: chirp images and their unnormalised DFT (row-major y_size x x_size, as the FFT operator sees them)
      hipLaunchKernelGGL(k_wm_chirp, dim3(nblocks(NB, 256)), dim3(256), 0, s, nb, d_w, dx, dy, xs, ys, nz, d_ch);
      SZ<double> sz;
      fft_rows<double>(plan, ROW_PLAIN, false, d_ch, d_ws, nb * y_size, sz, s);
      for (int b = 0; b < nb; ++b)
        fft_cols<double>(plan, false, d_ws + (int64_t)b * N, d_out + (int64_t)b * N, RowWindow{0, 0}, false, false, s);
      // 3: chirp threshold and keep rule (d_out holds the chirp rows now)
      hipLaunchKernelGGL(k_wm_abs, dim3(nblocks(NB, 256)), dim3(256), 0, s, NB, d_out, d_abs);
      hipLaunchKernelGGL(k_wm_thres, dim3(nb), dim3(WM_THREADS), 0, s, d_abs, N, e_chirp, d_thr, d_mx);
      hipLaunchKernelGGL(k_wm_keep, dim3(nblocks(NB, 256)), dim3(256), 0, s, N, nb, d_out, d_abs, d_thr, d_mx, d_cnt);
      // 4: convolution with the rows of G (into d_ch)
      hipLaunchKernelGGL(k_wm_conv, dim3(nblocks(NB, 256)), dim3(256), 0, s, nb, x_size, y_size, d_gptr, d_gcol, d_gval,
                         d_out, d_ch, d_hit);
      // 5: row threshold
      hipLaunchKernelGGL(k_wm_abs, dim3(nblocks(NB, 256)), dim3(256), 0, s, NB, d_ch, d_abs);
      hipLaunchKernelGGL(k_wm_thres, dim3(nb), dim3(WM_THREADS), 0, s, d_abs, N, e_wproj, d_thr, d_mx);
      PF_HIP(hipGetLastError());
      d2h(h_out.data(), d_ch, sizeof(cpx<double>) * NB);
      d2h(h_abs.data(), d_abs, sizeof(double) * NB);
      d2h(h_thr.data(), d_thr, sizeof(double) * nb);
      d2h(h_cnt.data(), d_cnt, sizeof(int) * nb);
      d2h(h_hit.data(), d_hit, sizeof(int) * nb);
      for (int b = 0; b < nb; ++b) {
        // wproj_utilities.h:246-247
        if ((int64_t)h_hit[b] < std::min<int64_t>(gp[b + 1] - gp[b], h_cnt[b]))
          throw Error(PF_ERR_ARG, "Gridding kernel or chirp kernel is zero.");
        const double* o = h_out.data() + (size_t)b * N * 2;
        const double* a = h_abs.data() + (size_t)b * N;
        for (int64_t k = 0; k < N; ++k) {
          if ((o[2 * k] != 0.0 || o[2 * k + 1] != 0.0) && a[k] > h_thr[b]) {
            out.col.push_back((int32_t)k);
            out.val.push_back(o[2 * k]);
            out.val.push_back(o[2 * k + 1]);
          }
        }
        out.rowptr[r0 + b + 1] = (int64_t)out.col.size();
      }
    }
  } catch (...) {
    cstream_sync();
    for (void* p : {(void*)d_ch, (void*)d_ws, (void*)d_out, (void*)d_abs, (void*)d_w, (void*)d_thr, (void*)d_mx,
                    (void*)d_cnt, (void*)d_hit, (void*)d_gptr, (void*)d_gcol, (void*)d_gval})
      dev_free(p);
    fft_plan_free(plan);
    throw;
  }
  for (void* p : {(void*)d_ch, (void*)d_ws, (void*)d_out, (void*)d_abs, (void*)d_w, (void*)d_thr, (void*)d_mx,
                  (void*)d_cnt, (void*)d_hit, (void*)d_gptr, (void*)d_gcol, (void*)d_gval})
    dev_free(p);
  fft_plan_free(plan);
}

template <typename R>
void csr_gather(int64_t rows, const int64_t* rowptr, const int32_t* col, const cpx<R>* val, const cpx<R>* x, cpx<R>* y,
                R scale, hipStream_t s, const char* name) {
  if (rows == 0) return;
  ProfScope prof(name, s);
  hipLaunchKernelGGL((k_csr_gather<R>), dim3(nblocks(rows, 4)), dim3(256), 0, s, rows, rowptr, col, val, x, y, scale);
  PF_HIP(hipGetLastError());
}
template void csr_gather<float>(int64_t, const int64_t*, const int32_t*, const cpx<float>*, const cpx<float>*,
                                cpx<float>*, float, hipStream_t, const char*);
template void csr_gather<double>(int64_t, const int64_t*, const int32_t*, const cpx<double>*, const cpx<double>*,
                                 cpx<double>*, double, hipStream_t, const char*);

}  // namespace pf
