// pf_chirp.hip -- chirp-truncated w-projection rows for a whole operator, built on the device.
//
// wproj_utilities::wprojection_matrix (cpp/purify/wproj_utilities.cc:132-188) convolves every row of the
// plain Kaiser-Bessel gridding matrix G (operators.cc:8-46) with the hard-thresholded DFT of its w chirp
// (create_chirp_row, wproj_utilities.cc:110-130) and hard-thresholds the result.  Done literally that is
// one N_g-point FFT per visibility.  Here the chirp image is c_w = exp(-2 pi i w psi) / N_g with the fixed
// phase screen psi = sqrt(1 - x^2 - y^2) - 1 on the chirp grid (generate_chirp, wproj_utilities.cc:76-109),
// so its DFT is the power series
//     C_w = sum_j (-2 pi i w)^j / j! * Psi_j,      Psi_j = DFT(psi^j / N_g),
// whose spectra Psi_j are computed once per operator (one FP64 FFT each) on a window of +-Rw cells around
// the zero frequency.  The series is summed until its remaining terms are below 1e-17 (|a_j Psi_j| <=
// theta^j / j! with theta = 2 pi max|w| max|psi|).  Each row (one wave per visibility) then takes, in FP64:
//   1. C_w on the window;
//   2. the bisection threshold keeping energy_chirp of the row's l2 energy (sparsify_row_dense_thres,
//      wproj_utilities.h:146-185; the total energy is 1 by Parseval, |c| = 1 / N_g) and convert_sparse's
//      keep rule |C| > max(1e-10 max|C|, threshold) (wproj_utilities.h:89-100);
//   3. the convolution with the row of G in the fftshift-centred coordinates of
//      row_wise_sparse_convolution (wproj_utilities.h:222-287): a G tap at centred position iG and a
//      chirp offset c contribute only if iG + c stays inside [0, N);
//   4. the bisection threshold keeping energy_wproj (sparsify_row_thres, wproj_utilities.h:101-145) and
//      the prune |value| > threshold.
// The window is valid when every bisection trial threshold of step 2 exceeds the window's border values
// (the spectrum decays away from the zero frequency); each row checks it and the builder doubles Rw
// until it holds.  A first pass finds the largest kept chirp radius R over all rows, which fixes the
// stored footprint (J + 2R)^2 of every row; the second pass writes the rows.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/purify_amd.h"
#include "pf_internal.h"

namespace pf {

namespace {

constexpr double PI_C = 3.14159265358979323846;
constexpr int CR_MAXJ = 16;   // kernel support
constexpr int CR_MAXRW = 32;  // window radius
constexpr int CR_MAXT = 64;   // series terms

__device__ inline double wsum(double a) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) a += __shfl_xor(a, m, 64);
  return a;
}
__device__ inline double wmax(double a) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) a = fmax(a, __shfl_xor(a, m, 64));
  return a;
}
__device__ inline int wmaxi(int a) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) a = max(a, __shfl_xor(a, m, 64));
  return a;
}

// psi^j / nz on the chirp grid, row-major [m][l] (the flat order the reference's FFT operator sees:
// chirp(l, m) of a column-major matrix), x = (l + 1/2 - N/2) dx, y = (m + 1/2 - N/2) dy
__global__ void k_psi_pow(int N, double dx, double dy, int j, double nz, cpx<double>* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)N * N) return;
  const int m = (int)(t / N), l = (int)(t - (int64_t)m * N);
  const double x = (l + 0.5 - N * 0.5) * dx, y = (m + 0.5 - N * 0.5) * dy;
  const double psi = sqrt(1 - x * x - y * y) - 1;
  double p = 1.0 / nz;
  for (int k = 0; k < j; ++k) p *= psi;
  out[t] = mk<double>(p, 0.0);
}
// win[(dv + R) W + du + R] = spec[(dv mod N) N + (du mod N)], |du|, |dv| <= R
__global__ void k_window(int N, int R, const cpx<double>* __restrict__ spec, cpx<double>* __restrict__ win) {
  const int W = 2 * R + 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= W * W) return;
  const int dv = t / W - R, du = t % W - R;
  win[t] = spec[(int64_t)((dv + N) % N) * N + (du + N) % N];
}

struct ChirpArgs {
  int64_t M;
  int N, J, Rw, JT;
  int R;  // the stored footprint radius (fill pass)
  double e_chirp, e_wproj;
  const cpx<double>* uv;  // pixels, caller's order
  const double* w;        // wavelengths, caller's order
  const cpx<double>* wt;  // weights, caller's order (null: unit)
  const int32_t* perm;    // fill: sorted position -> caller's index
  KParams<double> ku, kv;
  const cpx<double>* basis;  // JT windows of (2 Rw + 1)^2
  int* rmax;                 // count: largest kept chirp radius
  int* flag;                 // count: a window too small for some row
  unsigned long long* nnz;   // fill: kept entries
  void* rows;                // fill: M * (J + 2R)^2 values in the operator's precision
  int prec;
  uint64_t* mask;            // fill with (J + 2R)^2 <= 64: each row's keep mask (bit ov K + ou), or null
};

// The bisection of sparsify_row_(dense_)thres on the values a[0..n) held by the wave (lane-strided in
// LDS), with total energy `tot` and largest value `mx`: the threshold, and the smallest trial.
__device__ inline double bisect(const double* a, int n, double tot, double mx, double energy, double* tmin) {
  *tmin = mx;
  if (energy >= 1) {
    *tmin = 0;
    return 0;
  }
  const int lane = threadIdx.x;
  const double target = tot * energy;
  double lower = 0, upper = mx, old = upper, cur = upper * 0.5;
  bool conv = false;
  for (int it = 0; it < 200; ++it) {
    double s = 0;
    for (int k = lane; k < n; k += 64)
      if (a[k] > cur) s += a[k] * a[k];
    s = wsum(s);
    *tmin = fmin(*tmin, cur);
    if ((fabs(cur - old) / fabs(old) < 1e-3) && (s >= target) && (((s / target) - 1) < 0.01)) {
      conv = true;
      break;
    }
    if (s < target) upper = cur;
    else lower = cur;
    old = cur;
    cur = (lower + upper) * 0.5;
  }
  return conv ? cur : lower;
}

template <bool FILL>
__global__ void __launch_bounds__(64) k_chirp_rows(ChirpArgs g) {
  extern __shared__ unsigned char smem_raw[];
  const int W = 2 * g.Rw + 1, WW = W * W;
  cpx<double>* C = reinterpret_cast<cpx<double>*>(smem_raw);  // chirp window
  double* A = reinterpret_cast<double*>(C + WW);               // its magnitudes
  const int K = g.J + 2 * g.R, KK = K * K;
  cpx<double>* O = reinterpret_cast<cpx<double>*>(A + WW);     // the row (fill)
  double* OA = reinterpret_cast<double*>(O + KK);
  double* fu = OA + KK;                                        // G taps along u and v (fill)
  double* fv = fu + CR_MAXJ;
  const int lane = threadIdx.x;
  const int N = g.N, J = g.J;
  int my_rmax = 0, my_flag = 0;
  unsigned long long my_nnz = 0;
  for (int64_t i = blockIdx.x; i < g.M; i += gridDim.x) {
    const int64_t m = FILL ? (int64_t)g.perm[i] : i;
    const cpx<double> uvm = g.uv[m];
    const double w = g.w[m];
    // 1. C_w = sum_j a_j Psi_j with a_j = (-2 pi i w)^j / j!  (a_(j+1) = a_j (i z) / (j + 1), z = -2 pi w)
    const double z = -2 * PI_C * w;
    for (int k = lane; k < WW; k += 64) {
      double ar = 1.0, ai = 0.0, sr = 0.0, si = 0.0;
      for (int j = 0; j < g.JT; ++j) {
        const cpx<double> p = g.basis[(int64_t)j * WW + k];
        sr += ar * p.x - ai * p.y;
        si += ar * p.y + ai * p.x;
        const double nr = -ai * z / (j + 1), ni = ar * z / (j + 1);
        ar = nr;
        ai = ni;
      }
      C[k] = mk<double>(sr, si);
      A[k] = hypot(sr, si);
    }
    __syncthreads();
    // 2. threshold and keep rule; the border and the kept radius
    double mx = 0, border = 0;
    for (int k = lane; k < WW; k += 64) {
      mx = fmax(mx, A[k]);
      const int dv = k / W - g.Rw, du = k % W - g.Rw;
      if (abs(du) == g.Rw || abs(dv) == g.Rw) border = fmax(border, A[k]);
    }
    mx = wmax(mx);
    border = wmax(border);
    double tmin;
    const double thr = bisect(A, WW, 1.0, mx, g.e_chirp, &tmin);
    const double keep = fmax(mx * 1e-10, thr);
    int rk = 0;
    for (int k = lane; k < WW; k += 64) {
      if (A[k] > keep) {
        const int dv = k / W - g.Rw, du = k % W - g.Rw;
        rk = max(rk, max(abs(du), abs(dv)));
      } else {
        C[k] = mk<double>(0.0, 0.0);
      }
    }
    rk = wmaxi(rk);
    __syncthreads();
    if (!FILL) {
      my_rmax = max(my_rmax, rk);
      if (!(border < tmin) || rk >= g.Rw) my_flag = 1;
      continue;
    }
    // 3. the convolution with G's row (J x J taps at cells k + 1 + a, k = floor(u - J/2))
    const double kuf = floor(uvm.x - J * 0.5), kvf = floor(uvm.y - J * 0.5);
    const int64_t ku = (int64_t)kuf, kv = (int64_t)kvf;
    if (lane < J) fu[lane] = kern1d<double>(g.ku, uvm.x - (double)(ku + 1 + lane));
    if (lane >= 32 && lane < 32 + J) fv[lane - 32] = kern1d<double>(g.kv, uvm.y - (double)(kv + 1 + lane - 32));
    __syncthreads();
    const cpx<double> wt = g.wt ? g.wt[m] : mk<double>(1.0, 0.0);
    const int hN = N / 2;
    double tot = 0, omx = 0;
    for (int o = lane; o < KK; o += 64) {
      const int ov = o / K, ou = o - ov * K;
      double sr = 0, si = 0;
      for (int b = 0; b < J; ++b) {
        const int cv = ov - g.R - b;
        if (cv < -g.R || cv > g.R) continue;
        const int iGv = (int)(((kv + 1 + b) % N + N + hN) % N);  // centred position of the tap row
        if (iGv + cv < 0 || iGv + cv >= N) continue;
        for (int a = 0; a < J; ++a) {
          const int cu = ou - g.R - a;
          if (cu < -g.R || cu > g.R) continue;
          const int iGu = (int)(((ku + 1 + a) % N + N + hN) % N);
          if (iGu + cu < 0 || iGu + cu >= N) continue;
          const cpx<double> c = C[(cv + g.Rw) * W + cu + g.Rw];
          // G value: (-1)^(k_u + 1 + a + k_v + 1 + b) K_u K_v (operators.cc:38-41), the weight below
          double gv = fu[a] * fv[b];
          if ((ku + kv + a + b) & 1) gv = -gv;
          sr += gv * c.x;
          si += gv * c.y;
        }
      }
      const cpx<double> val = mk<double>(sr * wt.x - si * wt.y, sr * wt.y + si * wt.x);
      O[o] = val;
      OA[o] = hypot(val.x, val.y);
      tot += OA[o] * OA[o];
      omx = fmax(omx, OA[o]);
    }
    tot = wsum(tot);
    omx = wmax(omx);
    __syncthreads();
    // 4. the row's threshold and prune; store in the operator's precision
    double tmin2;
    const double thr2 = bisect(OA, KK, tot, omx, g.e_wproj, &tmin2);
    for (int o = lane; o < KK; o += 64) {
      const bool kept = OA[o] > thr2;
      const cpx<double> val = kept ? O[o] : mk<double>(0.0, 0.0);
      my_nnz += kept ? 1 : 0;
      if (g.prec == PF_C64)
        static_cast<cpx<float>*>(g.rows)[i * KK + o] = mk<float>((float)val.x, (float)val.y);
      else
        static_cast<cpx<double>*>(g.rows)[i * KK + o] = val;
    }
    if (g.mask) {  // KK <= 64: entry o is lane o's
      const unsigned long long bits = __ballot(lane < KK && OA[lane] > thr2);
      if (lane == 0) g.mask[i] = bits;
    }
    __syncthreads();
  }
  if (!FILL) {
    if (lane == 0) {
      atomicMax(g.rmax, my_rmax);
      if (my_flag) atomicOr(g.flag, 1);
    }
  } else {
    // one atomic per wave: the per-lane counts summed first
    unsigned long long s = my_nnz;
#pragma unroll
    for (int mm = 32; mm >= 1; mm >>= 1) s += __shfl_xor(s, mm, 64);
    if (lane == 0) atomicAdd(g.nnz, s);
  }
}

template <typename T>
T* cr_alloc(size_t n) {
  return dev_alloc_n<T>(n);
}

}  // namespace

// The chirp-row builder's device state across its two passes.
struct ChirpBuild {
  ChirpArgs a{};
  int blocks = 0;
  std::vector<void*> bufs;
  ~ChirpBuild() {
    cstream_sync();
    for (void* p : bufs) dev_free(p);
  }
  template <typename T>
  T* alloc(size_t n) {
    T* p = cr_alloc<T>(n);
    bufs.push_back(p);
    return p;
  }
};

namespace {
size_t chirp_lds(int Rw, int J, int R) {
  const int W = 2 * Rw + 1, K = J + 2 * R;
  return (size_t)W * W * 24 + (size_t)K * K * 24 + 2 * CR_MAXJ * sizeof(double);
}

// Psi_j windows for j < JT at radius Rw on the device
void chirp_basis(ChirpBuild& cb, int N, double cellx, double celly, int Rw, int JT) {
  const double Lx = 2 * std::sin(PI_C / 180. * (cellx * N) / 3600. * 0.5);  // fov_cosines
  const double Ly = 2 * std::sin(PI_C / 180. * (celly * N) / 3600. * 0.5);
  const double dx = Lx / N, dy = Ly / N, nz = (double)N * N;
  const int W = 2 * Rw + 1;
  cpx<double>* basis = cb.alloc<cpx<double>>((size_t)JT * W * W);
  cpx<double>* img = cr_alloc<cpx<double>>((size_t)N * N);
  cpx<double>* ws = cr_alloc<cpx<double>>((size_t)N * N);
  Fft2dPlan plan;
  try {
    fft_plan_init(plan, N, N, 1);
    hipStream_t s = cstream();
    const unsigned nb = (unsigned)(((int64_t)N * N + 255) / 256);
    for (int j = 0; j < JT; ++j) {
      hipLaunchKernelGGL(k_psi_pow, dim3(nb), dim3(256), 0, s, N, dx, dy, j, nz, img);
      SZ<double> sz;
      fft_rows<double>(plan, ROW_PLAIN, false, img, ws, N, sz, s);
      fft_cols<double>(plan, false, ws, img, RowWindow{0, 0}, false, false, s);
      hipLaunchKernelGGL(k_window, dim3((W * W + 255) / 256), dim3(256), 0, s, N, Rw, img, basis + (int64_t)j * W * W);
    }
    PF_HIP(hipStreamSynchronize(s));
  } catch (...) {
    cstream_sync();
    dev_free(img);
    dev_free(ws);
    fft_plan_free(plan);
    throw;
  }
  dev_free(img);
  dev_free(ws);
  fft_plan_free(plan);
  cb.a.basis = basis;
  cb.a.Rw = Rw;
  cb.a.JT = JT;
}
}  // namespace

// Pass 1: uploads the caller-order coordinates, finds the series length and a valid window, and returns
// the largest kept chirp radius R over all rows (the footprint is then (J + 2R)^2).
int chirp_rows_radius(ChirpBuild& cb, int64_t M, int N, int J, int kind, int Jeval_u, int Jeval_v, double os,
                      double cellx, double celly, double e_chirp, double e_wproj, const double* u, const double* v,
                      const double* w, const double* weights) {
  if (!(e_chirp > 0 && e_chirp < 1))
    throw Error(PF_ERR_UNSUPPORTED, "chirp w-projection rows need 0 < energy_chirp < 1 (1 keeps every chirp entry)");
  if (!(e_wproj > 0 && e_wproj <= 1)) throw Error(PF_ERR_ARG, "energy_wproj must be in (0, 1]");
  if (J > CR_MAXJ) throw Error(PF_ERR_UNSUPPORTED, "chirp w-projection rows support J <= 16");
  ChirpArgs& a = cb.a;
  a.M = M;
  a.N = N;
  a.J = J;
  a.e_chirp = e_chirp;
  a.e_wproj = e_wproj;
  a.ku = make_kparams<double>(kind, Jeval_u, os);
  a.kv = make_kparams<double>(kind, Jeval_v, os);
  std::vector<cpx<double>> uv((size_t)M);
  double wmax = 0;
  for (int64_t m = 0; m < M; ++m) {
    uv[m] = mk<double>(u[m], v[m]);
    if (!std::isfinite(w[m])) throw Error(PF_ERR_ARG, "non-finite w");
    wmax = std::max(wmax, std::fabs(w[m]));
  }
  cpx<double>* d_uv = cb.alloc<cpx<double>>(M);
  double* d_w = cb.alloc<double>(M);
  if (M) {
    h2d(d_uv, uv.data(), sizeof(cpx<double>) * M);
    h2d(d_w, w, sizeof(double) * M);
  }
  a.uv = d_uv;
  a.w = d_w;
  if (weights) {
    cpx<double>* d_wt = cb.alloc<cpx<double>>(M);
    h2d(d_wt, weights, sizeof(cpx<double>) * M);
    a.wt = d_wt;
  }
  // series length: |a_j Psi_j| <= theta^j / j!, theta = 2 pi max|w| max|psi| (psi is largest at the corner)
  const double Lx = 2 * std::sin(PI_C / 180. * (cellx * N) / 3600. * 0.5);
  const double Ly = 2 * std::sin(PI_C / 180. * (celly * N) / 3600. * 0.5);
  const double xc = (0.5 - N * 0.5) * Lx / N, yc = (0.5 - N * 0.5) * Ly / N;
  const double r2 = xc * xc + yc * yc;
  if (!(r2 < 1)) throw Error(PF_ERR_ARG, "the chirp's field of view reaches the horizon");
  const double psimax = 1 - std::sqrt(1 - r2);
  const double theta = 2 * PI_C * wmax * psimax;
  if (theta > 20) throw Error(PF_ERR_UNSUPPORTED, "w times the field's curvature is too large for the chirp series");
  int JT = 1;
  double term = 1;  // theta^JT / JT!
  while (JT < CR_MAXT && term * std::exp(theta) > 1e-17) {
    term *= theta / JT;
    ++JT;
  }
  a.JT = JT;
  int* d_int = cb.alloc<int>(2);
  a.rmax = d_int;
  a.flag = d_int + 1;
  int dev = 0, ncu = 0;
  PF_HIP(hipGetDevice(&dev));
  PF_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  cb.blocks = (int)std::max<int64_t>(1, std::min<int64_t>(M, (int64_t)std::max(ncu, 1) * 32));
  for (int Rw = 8;; Rw *= 2) {
    if (Rw > CR_MAXRW || 2 * Rw + 1 > N) throw Error(PF_ERR_UNSUPPORTED, "chirp spectrum wider than the window allows");
    if (a.basis) {
      dev_free(const_cast<cpx<double>*>(a.basis));
      cb.bufs.erase(std::find(cb.bufs.begin(), cb.bufs.end(), (void*)a.basis));
      a.basis = nullptr;
    }
    chirp_basis(cb, N, cellx, celly, Rw, JT);
    a.R = 0;
    PF_HIP(hipMemsetAsync(d_int, 0, 2 * sizeof(int), cstream()));
    if (chirp_lds(Rw, J, Rw - 1) > 160 * 1024)
      throw Error(PF_ERR_UNSUPPORTED, "chirp spectrum wider than the window allows");
    if (M) {
      const size_t shm = chirp_lds(Rw, J, 0);
      allow_lds((const void*)k_chirp_rows<false>, shm);
      hipLaunchKernelGGL(k_chirp_rows<false>, dim3(cb.blocks), dim3(64), shm, cstream(), a);
      PF_HIP(hipGetLastError());
    }
    int h[2] = {0, 0};
    d2h(h, d_int, 2 * sizeof(int));
    if (!h[1]) {
      a.R = h[0];
      return h[0];
    }
  }
}

// Pass 2: the rows, in the sorted order perm (device: sorted position -> caller's index), into rows
// (M * (J + 2R)^2 values in precision prec), and with (J + 2R)^2 <= 64 each row's keep mask into mask
// (if not null); returns the kept entries.
int64_t chirp_rows_fill(ChirpBuild& cb, const int32_t* perm_dev, void* rows, int prec, uint64_t* mask) {
  ChirpArgs& a = cb.a;
  if (mask && (a.J + 2 * a.R) * (a.J + 2 * a.R) > 64) throw Error(PF_ERR_ARG, "keep masks need (J + 2R)^2 <= 64");
  a.prec = prec;
  a.rows = rows;
  a.mask = mask;
  unsigned long long* d_nnz = cb.alloc<unsigned long long>(1);
  PF_HIP(hipMemsetAsync(d_nnz, 0, sizeof(unsigned long long), cstream()));
  a.perm = perm_dev;
  a.nnz = d_nnz;
  if (a.M) {
    const size_t shm = chirp_lds(a.Rw, a.J, a.R);
    allow_lds((const void*)k_chirp_rows<true>, shm);
    hipLaunchKernelGGL(k_chirp_rows<true>, dim3(cb.blocks), dim3(64), shm, cstream(), a);
    PF_HIP(hipGetLastError());
  }
  unsigned long long h = 0;
  d2h(&h, d_nnz, sizeof(h));
  return (int64_t)h;
}

ChirpBuild* chirp_build_new() { return new ChirpBuild(); }
void chirp_build_free(ChirpBuild* cb) { delete cb; }

}  // namespace pf
