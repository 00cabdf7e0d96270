// pf_solver.hip -- the imaging PADMM of PURIFY's pipeline on the device: SOPT's
// sopt::algorithm::ImagingProximalADMM with the parameters padmm_factory sets
// (cpp/purify/algorithm_factory.h:36-81): the closed-form l1 proximal for a tight frame (one
// orthonormal wavelet: main.cc:121-126 sets tight_frame = sara.size() < 2), or the iterative,
// constrained one (SARA; algorithm_factory.h:60-66: l1_proximal_tolerance 1e-2, nu 1, itermax 50,
// positivity and real constraints).
//
// SOPT is not in the reference tree, so the iteration is restated from the published proximal ADMM
// (Carrillo, McEwen & Wiaux 2014, "PURIFY: a new approach to radio-interferometric imaging",
// section 3; Onose et al. 2016, algorithm 1) in SOPT's variable names:
//   z      = P_eps(-lambda - r)                         l2-ball projection, radius eps
//   x~     = x - Phi^dagger(r + lambda + z) / nu
//   x      = x~ + Psi(soft(Psi^dagger x~, gamma nu) - Psi^dagger x~) / nu   tight-frame l1 proximal
//   r      = Phi x - y
//   lambda = lambda + rho (r + z)                        rho = lagrange_update_scale
// from x = Phi^dagger y / nu, r = Phi x - y, lambda = 0.  Convergence: ||x - x_prev|| / ||x|| <
// relative_variation and ||r|| < residual_tolerance.  soft(c, t) = c max(0, 1 - t / |c|).
//
// The iterative proximal, x = argmin 1/2 ||x - x~||^2 + gamma ||Psi^dagger x||_1 over the constraint set
// C (Re x >= 0 and Im x >= 0 with positivity, Im x = 0 with the real constraint), is the dual
// forward-backward iteration (Combettes & Pesquet 2011, "Proximal splitting methods in signal
// processing", the dual of the constrained l1 problem; Fadili & Starck 2009) with step 1 / l1_nu:
//   u_0 = 0, x_0 = P_C(x~);  u_k = P_gamma(u_(k-1) + Psi^dagger x_(k-1) / l1_nu),  x_k = P_C(x~ - Psi u_k)
// P_gamma clamps each coefficient's modulus to gamma.  It stops after l1_itermax steps or once the
// objective 1/2 ||x_k - x~||^2 + gamma ||Psi^dagger x_k||_1 changes by less than l1_tolerance relative:
// that test reads two FP64 sums on the host, so the iterative mode runs eagerly (no graph).
//
// The same handle type also runs SOPT's ImagingForwardBackward as cpp/example/fb_random_coverage.cc:
// 89-106 configures it (PhiTPhi the compact operator, the target the dirty image Phi^dagger y), restated
// from the published forward-backward splitting (Combettes & Wajs 2005):
//   x~ = x - (beta / sigma^2) (Phi^dagger Phi x - Phi^dagger y),   x = prox_(mu beta ||Psi^dagger .||_1)(x~)
// from x = 0, with either l1 proximal above; convergence on ||x - x_prev|| / ||x|| alone (the example's
// residual tolerance is 0), the image-domain residual Phi^dagger Phi x - Phi^dagger y recorded beside it.
//
// Vectors stay in the operator's precision; every norm is reduced in FP64 on the device in a fixed
// order, and the projection reads its norm from device memory, so an iteration has no host
// synchronisation and is captured once into a hipGraph that is replayed.  Per-iteration FP64
// statistics (||x - x_prev||^2, ||x||^2, ||r||^2) are appended to a device array through a device
// counter, so a replayed graph records every iteration.
#include <algorithm>
#include <cstring>
#include <cmath>
#include <memory>
#include <vector>

#include "../../include/purify_amd.h"
#include "pf_internal.h"

namespace pf {
namespace {

constexpr int SR_BLOCKS = 512, SR_THREADS = 256;

template <typename R>
__device__ inline double norm2d(cpx<R> a) {
  return (double)a.x * (double)a.x + (double)a.y * (double)a.y;
}

// partials[b] = sum |a - b|^2 (B given) or |a|^2 (ABS: sum |a|) over a grid-stride slice; fixed order
template <typename R, bool DIFF, bool ABS = false>
__global__ void __launch_bounds__(SR_THREADS) k_sq_partial(const cpx<R>* __restrict__ a, const cpx<R>* __restrict__ b,
                                                           int64_t n, double* __restrict__ part) {
  __shared__ double sh[SR_THREADS];
  double acc = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * SR_THREADS + threadIdx.x; k < n; k += (int64_t)SR_BLOCKS * SR_THREADS) {
    cpx<R> v = a[k];
    if (DIFF) {
      const cpx<R> w = b[k];
      v.x -= w.x;
      v.y -= w.y;
    }
    acc += ABS ? sqrt(norm2d(v)) : norm2d(v);
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = SR_THREADS / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ void __launch_bounds__(SR_THREADS) k_sq_final(const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double sh[SR_THREADS];
  double acc = 0.0;
  for (int k = threadIdx.x; k < SR_BLOCKS; k += SR_THREADS) acc += part[k];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = SR_THREADS / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}
template <typename R>
void sq_norm(const cpx<R>* a, const cpx<R>* b, int64_t n, double* part, double* out, hipStream_t s) {
  if (b) hipLaunchKernelGGL((k_sq_partial<R, true>), dim3(SR_BLOCKS), dim3(SR_THREADS), 0, s, a, b, n, part);
  else hipLaunchKernelGGL((k_sq_partial<R, false>), dim3(SR_BLOCKS), dim3(SR_THREADS), 0, s, a, b, n, part);
  hipLaunchKernelGGL(k_sq_final, dim3(1), dim3(SR_THREADS), 0, s, part, out);
}

// *out = sum |a| (FP64, fixed order)
template <typename R>
void abs_sum(const cpx<R>* a, int64_t n, double* part, double* out, hipStream_t s) {
  hipLaunchKernelGGL((k_sq_partial<R, false, true>), dim3(SR_BLOCKS), dim3(SR_THREADS), 0, s, a, nullptr, n, part);
  hipLaunchKernelGGL(k_sq_final, dim3(1), dim3(SR_THREADS), 0, s, part, out);
}

inline dim3 vgrid(int64_t n) { return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096))); }
#define GRID_STRIDE(k, n) for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < (n); k += (int64_t)gridDim.x * blockDim.x)

// ---- fused PADMM vector passes: each also writes its FP64 partial sums (one per block, fixed order);
// k_fr_final adds them in block order.
constexpr int FR_BLOCKS = 2048;

template <int NV>
__device__ inline void reduce_partials(double (&acc)[NV], double* __restrict__ part) {
  __shared__ double sh[NV][SR_THREADS];
  const int t = threadIdx.x;
#pragma unroll
  for (int v = 0; v < NV; ++v) sh[v][t] = acc[v];
  __syncthreads();
  for (int w = SR_THREADS / 2; w > 0; w >>= 1) {
    if (t < w)
#pragma unroll
      for (int v = 0; v < NV; ++v) sh[v][t] += sh[v][t + w];
    __syncthreads();
  }
  if (t == 0)
#pragma unroll
    for (int v = 0; v < NV; ++v) part[v * FR_BLOCKS + blockIdx.x] = sh[v][0];
}
// The kernels of the PADMM iteration take their arguments as one struct; BT marks a batch launch, whose
// blockIdx.z selects the arguments from tab (see Recorder).
struct FrFinalArgs {
  const double* part;
  double* out;
};
// out[v] = sum of the FR_BLOCKS partials of quantity v (block v of the launch)
template <bool BT = false>
__global__ void __launch_bounds__(SR_THREADS) k_fr_final(FrFinalArgs a0, const FrFinalArgs* __restrict__ tab) {
  const FrFinalArgs q = BT ? tab[blockIdx.z] : a0;
  const double* __restrict__ part = q.part;
  double* __restrict__ out = q.out;
  __shared__ double sh[SR_THREADS];
  const double* p = part + blockIdx.x * FR_BLOCKS;
  double acc = 0.0;
  for (int k = threadIdx.x; k < FR_BLOCKS; k += SR_THREADS) acc += p[k];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = SR_THREADS / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}
#define FR_STRIDE(k, n) for (int64_t k = (int64_t)blockIdx.x * SR_THREADS + threadIdx.x; k < (n); k += (int64_t)FR_BLOCKS * SR_THREADS)

// out = ||lambda + r||^2 (= ||v||^2 for v = -lambda - r)
template <typename R>
__global__ void __launch_bounds__(SR_THREADS) k_ball_norm(const cpx<R>* __restrict__ lam, const cpx<R>* __restrict__ r,
                                                          int64_t n, double* part) {
  double acc[1] = {0.0};
  FR_STRIDE(k, n) {
    const cpx<R> a = lam[k], b = r[k];
    acc[0] += norm2d(mk<R>(-a.x - b.x, -a.y - b.y));
  }
  reduce_partials<1>(acc, part);
}
// v = -lambda - r; z = P_eps(v) (v scaled by min(1, eps / ||v||), ||v||^2 read from *nrm2); w = r + lambda + z
template <typename R>
struct BallOutArgs {
  const double* nrm2;
  double eps;
  const cpx<R>* r;
  const cpx<R>* lam;
  cpx<R>* z;
  cpx<R>* w;
  int64_t n;
};
template <typename R, bool BT = false>
__global__ void k_ball_out(BallOutArgs<R> a0, const BallOutArgs<R>* __restrict__ tab) {
  const BallOutArgs<R> q = BT ? tab[blockIdx.z] : a0;
  const double* __restrict__ nrm2 = q.nrm2;
  const double eps = q.eps;
  const cpx<R>* __restrict__ r = q.r;
  const cpx<R>* __restrict__ lam = q.lam;
  cpx<R>* __restrict__ z = q.z;
  cpx<R>* __restrict__ w = q.w;
  const int64_t n = q.n;
  const double nv = sqrt(*nrm2);
  const R f = (R)(nv > eps ? eps / nv : 1.0);
  GRID_STRIDE(k, n) {
    const cpx<R> b = r[k], c = lam[k];
    const cpx<R> zz = mk<R>((-c.x - b.x) * f, (-c.y - b.y) * f);
    z[k] = zz;
    w[k] = mk<R>(b.x + c.x + zz.x, b.y + c.y + zz.y);
  }
}
// x = xu + d / nu (d null: x as it is); out[0] = ||x - xprev||^2, out[1] = ||x||^2.  xprev may be x itself
// (its old value is read before the new one is written), so x and xprev are not restrict.
template <typename R>
struct ProxOutArgs {
  const cpx<R>* xu;
  const cpx<R>* d;
  cpx<R>* x;
  const cpx<R>* xprev;
  R inv_nu;
  int64_t n;
  double* part;
};
template <typename R, bool BT = false>
__global__ void __launch_bounds__(SR_THREADS) k_prox_out_norm(ProxOutArgs<R> a0, const ProxOutArgs<R>* __restrict__ tab) {
  const ProxOutArgs<R> q = BT ? tab[blockIdx.z] : a0;
  const cpx<R>* __restrict__ xu = q.xu;
  const cpx<R>* __restrict__ d = q.d;
  cpx<R>* x = q.x;
  const cpx<R>* xprev = q.xprev;
  const R inv_nu = q.inv_nu;
  const int64_t n = q.n;
  double* part = q.part;
  double acc[2] = {0.0, 0.0};
  FR_STRIDE(k, n) {
    const cpx<R> p = xprev[k];
    cpx<R> a;
    if (d) {
      const cpx<R> u = xu[k], b = d[k];
      a = mk<R>(u.x + b.x * inv_nu, u.y + b.y * inv_nu);
      x[k] = a;
    } else {
      a = x[k];
    }
    acc[0] += norm2d(mk<R>(a.x - p.x, a.y - p.y));
    acc[1] += norm2d(a);
  }
  reduce_partials<2>(acc, part);
}
// r = fx - y; lambda += rho (r + z); out[0] = ||r||^2, out[1] = ||lambda + r||^2 (the next iteration's
// ball norm, ||v||^2 for v = -lambda - r)
template <typename R>
struct ResidualArgs {
  const cpx<R>* fx;
  const cpx<R>* y;
  const cpx<R>* z;
  cpx<R>* r;
  cpx<R>* lam;
  R rho;
  int64_t n;
  double* part;
};
template <typename R, bool BT = false>
__global__ void __launch_bounds__(SR_THREADS) k_residual_dual_norm(ResidualArgs<R> a0,
                                                                   const ResidualArgs<R>* __restrict__ tab) {
  const ResidualArgs<R> q = BT ? tab[blockIdx.z] : a0;
  const cpx<R>* __restrict__ fx = q.fx;
  const cpx<R>* __restrict__ y = q.y;
  const cpx<R>* __restrict__ z = q.z;
  cpx<R>* __restrict__ r = q.r;
  cpx<R>* __restrict__ lam = q.lam;
  const R rho = q.rho;
  const int64_t n = q.n;
  double* part = q.part;
  double acc[2] = {0.0, 0.0};
  FR_STRIDE(k, n) {
    const cpx<R> a = fx[k], b = y[k], c = z[k];
    const cpx<R> rr = mk<R>(a.x - b.x, a.y - b.y);
    r[k] = rr;
    const cpx<R> l = lam[k];
    const cpx<R> ln = mk<R>(l.x + rho * (rr.x + c.x), l.y + rho * (rr.y + c.y));
    lam[k] = ln;
    acc[0] += norm2d(rr);
    acc[1] += norm2d(mk<R>(-ln.x - rr.x, -ln.y - rr.y));
  }
  reduce_partials<2>(acc, part);
}
// xu = x - g / nu (COPY: and xprev = x)
template <typename R>
struct GradArgs {
  cpx<R>* x;
  const cpx<R>* g;
  cpx<R>* xprev;
  cpx<R>* xu;
  R inv_nu;
  int64_t n;
};
template <typename R, bool COPY, bool BT = false>
__global__ void k_grad_step(GradArgs<R> a0, const GradArgs<R>* __restrict__ tab) {
  const GradArgs<R> q = BT ? tab[blockIdx.z] : a0;
  cpx<R>* __restrict__ x = q.x;
  const cpx<R>* __restrict__ g = q.g;
  cpx<R>* __restrict__ xprev = q.xprev;
  cpx<R>* __restrict__ xu = q.xu;
  const R inv_nu = q.inv_nu;
  const int64_t n = q.n;
  GRID_STRIDE(k, n) {
    const cpx<R> a = x[k], b = g[k];
    if (COPY) xprev[k] = a;
    xu[k] = mk<R>(a.x - b.x * inv_nu, a.y - b.y * inv_nu);
  }
}
// c <- soft(c, t) - c
template <typename R>
struct SoftArgs {
  cpx<R>* c;
  R t;
  int64_t n;
};
template <typename R, bool BT = false>
__global__ void k_soft_diff(SoftArgs<R> a0, const SoftArgs<R>* __restrict__ tab) {
  const SoftArgs<R> q = BT ? tab[blockIdx.z] : a0;
  cpx<R>* __restrict__ c = q.c;
  const R t = q.t;
  const int64_t n = q.n;
  GRID_STRIDE(k, n) {
    const cpx<R> a = c[k];
    const R m = sqrt(a.x * a.x + a.y * a.y);
    const R f = m > t ? (R(1) - t / m) : R(0);
    c[k] = mk<R>(a.x * f - a.x, a.y * f - a.y);
  }
}
// x = xu + d / nu
template <typename R>
__global__ void k_prox_out(const cpx<R>* __restrict__ xu, const cpx<R>* __restrict__ d, cpx<R>* __restrict__ x, R inv_nu,
                           int64_t n) {
  GRID_STRIDE(k, n) {
    const cpx<R> a = xu[k], b = d[k];
    x[k] = mk<R>(a.x + b.x * inv_nu, a.y + b.y * inv_nu);
  }
}
// the constraint set of the iterative proximal: positivity keeps both parts >= 0, real zeroes Im
template <typename R>
__device__ inline cpx<R> project_c(cpx<R> a, bool pos, bool real) {
  if (real) a.y = R(0);
  if (pos) {
    a.x = a.x > R(0) ? a.x : R(0);
    a.y = a.y > R(0) ? a.y : R(0);
  }
  return a;
}
// x = P_C(x0 - d) (d null: P_C(x0))
template <typename R>
__global__ void k_l1_primal(const cpx<R>* __restrict__ x0, const cpx<R>* __restrict__ d, cpx<R>* __restrict__ x, int pos,
                            int real, int64_t n) {
  GRID_STRIDE(k, n) {
    cpx<R> a = x0[k];
    if (d) {
      const cpx<R> b = d[k];
      a = mk<R>(a.x - b.x, a.y - b.y);
    }
    x[k] = project_c(a, pos != 0, real != 0);
  }
}
// u = P_gamma(u + c / nu): each coefficient's modulus clamped to gamma
template <typename R>
__global__ void k_l1_dual(cpx<R>* __restrict__ u, const cpx<R>* __restrict__ c, R inv_nu, R gamma, int64_t n) {
  GRID_STRIDE(k, n) {
    const cpx<R> a = u[k], b = c[k];
    const cpx<R> w = mk<R>(a.x + b.x * inv_nu, a.y + b.y * inv_nu);
    const R m = sqrt(w.x * w.x + w.y * w.y);
    const R f = m > gamma ? gamma / m : R(1);
    u[k] = mk<R>(w.x * f, w.y * f);
  }
}
// r = fx - y
template <typename R>
__global__ void k_sub(const cpx<R>* __restrict__ a, const cpx<R>* __restrict__ b, cpx<R>* __restrict__ r, int64_t n) {
  GRID_STRIDE(k, n) {
    const cpx<R> u = a[k], v = b[k];
    r[k] = mk<R>(u.x - v.x, u.y - v.y);
  }
}
template <typename R>
__global__ void k_scale_copy(const cpx<R>* __restrict__ a, cpx<R>* __restrict__ b, R f, int64_t n) {
  GRID_STRIDE(k, n) {
    const cpx<R> u = a[k];
    b[k] = mk<R>(u.x * f, u.y * f);
  }
}
template <typename R>
__global__ void k_zero(cpx<R>* __restrict__ a, int64_t n) {
  GRID_STRIDE(k, n) a[k] = mk<R>(R(0), R(0));
}
// stats[*counter] = cur (3 doubles); ++*counter (capacity checked: the run never exceeds it)
struct RecordArgs {
  const double* cur;
  double* stats;
  int* counter;
  int cap;
};
template <bool BT = false>
__global__ void k_record(RecordArgs a0, const RecordArgs* __restrict__ tab) {
  const RecordArgs q = BT ? tab[blockIdx.z] : a0;
  const double* __restrict__ cur = q.cur;
  double* __restrict__ stats = q.stats;
  int* __restrict__ counter = q.counter;
  const int cap = q.cap;
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const int c = *counter;
    if (c < cap) {
      stats[3 * c] = cur[0];
      stats[3 * c + 1] = cur[1];
      stats[3 * c + 2] = cur[2];
    }
    *counter = c + 1;
  }
}
__global__ void k_set_int(int* p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// launch (or record) helpers of the iteration's kernels
inline void fr_final(const double* part, double* out, int nq, hipStream_t s) {
  launch_or_record<FrFinalArgs, k_fr_final<>, k_fr_final<true>>(dim3(nq), dim3(SR_THREADS), 0, false,
                                                                 FrFinalArgs{part, out}, s);
}
template <typename R>
void ball_out(const double* nrm2, double eps, const cpx<R>* r, const cpx<R>* lam, cpx<R>* z, cpx<R>* w, int64_t n,
              hipStream_t s) {
  launch_or_record<BallOutArgs<R>, k_ball_out<R>, k_ball_out<R, true>>(vgrid(n), dim3(256), 0, true,
                                                                       BallOutArgs<R>{nrm2, eps, r, lam, z, w, n}, s);
}
template <typename R>
void prox_out_norm(const cpx<R>* xu, const cpx<R>* d, cpx<R>* x, const cpx<R>* xprev, R inv_nu, int64_t n, double* part,
                   hipStream_t s) {
  launch_or_record<ProxOutArgs<R>, k_prox_out_norm<R>, k_prox_out_norm<R, true>>(
      dim3(FR_BLOCKS), dim3(SR_THREADS), 0, false, ProxOutArgs<R>{xu, d, x, xprev, inv_nu, n, part}, s);
}
template <typename R>
void residual_dual_norm(const cpx<R>* fx, const cpx<R>* y, const cpx<R>* z, cpx<R>* r, cpx<R>* lam, R rho, int64_t n,
                        double* part, hipStream_t s) {
  launch_or_record<ResidualArgs<R>, k_residual_dual_norm<R>, k_residual_dual_norm<R, true>>(
      dim3(FR_BLOCKS), dim3(SR_THREADS), 0, false, ResidualArgs<R>{fx, y, z, r, lam, rho, n, part}, s);
}
template <typename R, bool COPY>
void grad_step(cpx<R>* x, const cpx<R>* g, cpx<R>* xprev, cpx<R>* xu, R inv_nu, int64_t n, hipStream_t s) {
  launch_or_record<GradArgs<R>, k_grad_step<R, COPY>, k_grad_step<R, COPY, true>>(
      vgrid(n), dim3(256), 0, true, GradArgs<R>{x, g, xprev, xu, inv_nu, n}, s);
}
template <typename R>
void soft_diff(cpx<R>* c, R t, int64_t n, hipStream_t s) {
  launch_or_record<SoftArgs<R>, k_soft_diff<R>, k_soft_diff<R, true>>(vgrid(n), dim3(256), 0, true,
                                                                      SoftArgs<R>{c, t, n}, s);
}
inline void record_stats(const double* cur, double* stats, int* counter, int cap, hipStream_t s) {
  launch_or_record<RecordArgs, k_record<>, k_record<true>>(dim3(1), dim3(64), 0, false,
                                                           RecordArgs{cur, stats, counter, cap}, s);
}

}  // namespace
}  // namespace pf

using namespace pf;

struct pf_padmm {
  pf_op* phi = nullptr;       // not owned
  pf_wavelet* psi = nullptr;  // not owned
  int prec = PF_C64;
  int64_t N = 0, M = 0, NC = 0;
  pf_padmm_params par{};
  double gamma = 0;
  // device state (operator precision)
  void *x = nullptr, *xprev = nullptr, *xu = nullptr, *g = nullptr, *c = nullptr, *d = nullptr;
  void *y = nullptr, *r = nullptr, *lam = nullptr, *z = nullptr, *v = nullptr, *fx = nullptr;
  double* part = nullptr;  // reduction partials
  double* fpart = nullptr;  // fused reductions' partials (2 x FR_BLOCKS)
  double* cur = nullptr;   // [0] ||x - xprev||^2, [1] ||x||^2, [2] ||r||^2, [3] ||v||^2, [4] ||x - x~||^2, [5] ||Psi^dagger x||_1
  bool tight = true;       // closed-form tight-frame proximal, else the iterative one
  bool fb = false;         // forward-backward instead of PADMM
  double fb_step = 0;      // beta / sigma^2
  void* dimage = nullptr;  // forward-backward target Phi^dagger y (N)
  void* u = nullptr;       // the iterative proximal's dual variable (NC coefficients)
  int l1_last = 0;         // inner iterations of the last iterative proximal
  int64_t l1_total = 0;
  double* stats = nullptr;
  int* counter = nullptr;
  int cap = 0;
  int iters_done = 0;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t captured_gen = 0;  // the operator's generation when the graph was captured
  double captured_gamma = 0;  // and the l1 weight its threshold kernel was given (reset recomputes it)
  std::vector<void*> allocs;
  void* dalloc(size_t bytes) {
    void* p = dev_alloc(std::max<size_t>(bytes, 16));
    allocs.push_back(p);
    return p;
  }
  ~pf_padmm() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (stream) (void)hipStreamSynchronize(stream);
    for (void* p : allocs) dev_free(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace pf {
namespace {

template <typename R>
cpx<R>* V(void* p) {
  return static_cast<cpx<R>*>(p);
}

// x = prox(xu) by the iterative constrained proximal (eager: the stopping test reads the objective)
template <typename R>
void l1_prox_iterative(pf_padmm* h, double gamma) {
  hipStream_t s = h->stream;
  const int64_t N = h->N, NC = h->NC;
  const bool pos = h->par.positivity != 0, real = h->par.real_constraint != 0;
  const R inv_nu = (R)(1.0 / h->par.l1_nu), gam = (R)gamma;
  auto objective = [&]() {
    sq_norm<R>(V<R>(h->x), V<R>(h->xu), N, h->part, h->cur + 4, s);
    abs_sum<R>(V<R>(h->c), NC, h->part, h->cur + 5, s);
    double o[2];
    PF_HIP(hipMemcpyAsync(o, h->cur + 4, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    PF_HIP(hipStreamSynchronize(s));
    return 0.5 * o[0] + gamma * o[1];
  };
  hipLaunchKernelGGL((k_zero<R>), vgrid(NC), dim3(256), 0, s, V<R>(h->u), NC);
  hipLaunchKernelGGL((k_l1_primal<R>), vgrid(N), dim3(256), 0, s, V<R>(h->xu), (const cpx<R>*)nullptr, V<R>(h->x),
                     (int)pos, (int)real, N);
  wavelet_analysis<R>(h->psi, V<R>(h->x), V<R>(h->c), s);
  double obj = objective();
  int k = 0;
  while (k < h->par.l1_itermax) {
    ++k;
    hipLaunchKernelGGL((k_l1_dual<R>), vgrid(NC), dim3(256), 0, s, V<R>(h->u), V<R>(h->c), inv_nu, gam, NC);
    wavelet_synthesis<R>(h->psi, V<R>(h->u), V<R>(h->d), s);
    hipLaunchKernelGGL((k_l1_primal<R>), vgrid(N), dim3(256), 0, s, V<R>(h->xu), V<R>(h->d), V<R>(h->x), (int)pos,
                       (int)real, N);
    wavelet_analysis<R>(h->psi, V<R>(h->x), V<R>(h->c), s);
    const double o = objective();
    const bool stop = std::abs(o - obj) < h->par.l1_tolerance * std::abs(o);
    obj = o;
    if (stop) break;
  }
  h->l1_last = k;
  h->l1_total += k;
  PF_HIP(hipGetLastError());
}

template <typename R>
void fb_iteration(pf_padmm* h);
template <typename R>
void padmm_step(pf_padmm* h);
// one iteration of the handle's solver on h->stream (no host synchronisation with the tight-frame
// proximal)
template <typename R>
void padmm_iteration(pf_padmm* h) {
  if (h->fb) fb_iteration<R>(h);
  else padmm_step<R>(h);
}
template <typename R>
void padmm_step(pf_padmm* h) {
  hipStream_t s = h->stream;
  const int64_t N = h->N, M = h->M, NC = h->NC;
  const R inv_nu = (R)(1.0 / h->par.nu);
  // cur[3] = ||lambda + r||^2 from the previous iteration's residual pass (or the initialisation)
  // fx holds r + lambda + z until the forward apply below overwrites it
  ball_out<R>(h->cur + 3, h->par.epsilon, V<R>(h->r), V<R>(h->lam), V<R>(h->z), V<R>(h->fx), M, s);
  op_adjoint_sorted<R>(h->phi, V<R>(h->fx), V<R>(h->g), s);
  if (h->tight) {
    // x keeps its old value until the proximal output replaces it, so no copy of x_prev is made
    grad_step<R, false>(V<R>(h->x), V<R>(h->g), V<R>(h->xprev), V<R>(h->xu), inv_nu, N, s);
    wavelet_analysis<R>(h->psi, V<R>(h->xu), V<R>(h->c), s);
    // the l1 proximal's own frame bound l1_nu (SOPT's L1TightFrame nu), not the ADMM step nu
    soft_diff<R>(V<R>(h->c), (R)(h->gamma * h->par.l1_nu), NC, s);
    wavelet_synthesis<R>(h->psi, V<R>(h->c), V<R>(h->d), s);
    prox_out_norm<R>(V<R>(h->xu), V<R>(h->d), V<R>(h->x), V<R>(h->x), (R)(1.0 / h->par.l1_nu), N, h->fpart, s);
  } else {
    grad_step<R, true>(V<R>(h->x), V<R>(h->g), V<R>(h->xprev), V<R>(h->xu), inv_nu, N, s);
    l1_prox_iterative<R>(h, h->gamma);
    prox_out_norm<R>(V<R>(h->xu), (const cpx<R>*)nullptr, V<R>(h->x), V<R>(h->xprev), inv_nu, N, h->fpart, s);
  }
  fr_final(h->fpart, h->cur + 0, 2, s);
  op_forward_sorted<R>(h->phi, V<R>(h->x), V<R>(h->fx), s);
  residual_dual_norm<R>(V<R>(h->fx), V<R>(h->y), V<R>(h->z), V<R>(h->r), V<R>(h->lam), (R)h->par.lagrange_update_scale,
                        M, h->fpart, s);
  fr_final(h->fpart, h->cur + 2, 2, s);
  record_stats(h->cur, h->stats, h->counter, h->cap, s);
  PF_HIP(hipGetLastError());
}

// x~ = x - step (g - dimage) (g = Phi^dagger Phi x); xprev = x
template <typename R>
__global__ void k_fb_grad(const cpx<R>* __restrict__ x, const cpx<R>* __restrict__ g, const cpx<R>* __restrict__ dimg,
                          cpx<R>* __restrict__ xprev, cpx<R>* __restrict__ xu, cpx<R>* __restrict__ res, R step, int64_t n) {
  GRID_STRIDE(k, n) {
    const cpx<R> a = x[k], b = g[k], c = dimg[k];
    const cpx<R> r = mk<R>(b.x - c.x, b.y - c.y);
    res[k] = r;
    xprev[k] = a;
    xu[k] = mk<R>(a.x - step * r.x, a.y - step * r.y);
  }
}

// one forward-backward iteration; stats: ||x - x_prev||^2, ||x||^2 and the squared norm of the image-domain
// residual Phi^dagger Phi x_prev - Phi^dagger y the step used
template <typename R>
void fb_iteration(pf_padmm* h) {
  hipStream_t s = h->stream;
  const int64_t N = h->N, NC = h->NC;
  const double gam = h->gamma;  // mu beta
  op_normal<R>(h->phi, V<R>(h->x), V<R>(h->g), s);
  hipLaunchKernelGGL((k_fb_grad<R>), vgrid(N), dim3(256), 0, s, V<R>(h->x), V<R>(h->g), V<R>(h->dimage), V<R>(h->xprev),
                     V<R>(h->xu), V<R>(h->v), (R)h->fb_step, N);
  if (h->tight) {
    wavelet_analysis<R>(h->psi, V<R>(h->xu), V<R>(h->c), s);
    soft_diff<R>(V<R>(h->c), (R)(gam * h->par.l1_nu), NC, s);
    wavelet_synthesis<R>(h->psi, V<R>(h->c), V<R>(h->d), s);
    hipLaunchKernelGGL((k_prox_out<R>), vgrid(N), dim3(256), 0, s, V<R>(h->xu), V<R>(h->d), V<R>(h->x),
                       (R)(1.0 / h->par.l1_nu), N);
  } else {
    l1_prox_iterative<R>(h, gam);
  }
  sq_norm<R>(V<R>(h->x), V<R>(h->xprev), N, h->part, h->cur + 0, s);
  sq_norm<R>(V<R>(h->x), nullptr, N, h->part, h->cur + 1, s);
  sq_norm<R>(V<R>(h->v), nullptr, N, h->part, h->cur + 2, s);
  record_stats(h->cur, h->stats, h->counter, h->cap, s);
  PF_HIP(hipGetLastError());
}

template <typename R>
void fb_init(pf_padmm* h) {
  hipStream_t s = h->stream;
  hipLaunchKernelGGL((k_zero<R>), vgrid(h->N), dim3(256), 0, s, V<R>(h->x), h->N);
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, h->counter, 0);
  PF_HIP(hipStreamSynchronize(s));
  h->iters_done = 0;
  h->l1_last = 0;
  h->l1_total = 0;
}

// x = Phi^dagger y / nu, r = Phi x - y, lambda = 0; gamma from |Psi^dagger Phi^dagger y| if not given
template <typename R>
void padmm_init(pf_padmm* h) {
  hipStream_t s = h->stream;
  const int64_t N = h->N, M = h->M;
  op_adjoint_sorted<R>(h->phi, V<R>(h->y), V<R>(h->g), s);
  if (!(h->par.gamma > 0)) {
    // algorithm_factory.h:70-73: gamma = max |Psi^dagger Phi^dagger y| * 1e-3
    wavelet_analysis<R>(h->psi, V<R>(h->g), V<R>(h->c), s);
    std::vector<cpx<R>> hc((size_t)h->NC);
    PF_HIP(hipMemcpyAsync(hc.data(), h->c, sizeof(cpx<R>) * h->NC, hipMemcpyDeviceToHost, s));
    PF_HIP(hipStreamSynchronize(s));
    double mx = 0;
    for (const auto& a : hc) mx = std::max(mx, std::sqrt((double)a.x * a.x + (double)a.y * a.y));
    h->gamma = mx * 1e-3;
  } else {
    h->gamma = h->par.gamma;
  }
  hipLaunchKernelGGL((k_scale_copy<R>), vgrid(N), dim3(256), 0, s, V<R>(h->g), V<R>(h->x), (R)(1.0 / h->par.nu), N);
  op_forward_sorted<R>(h->phi, V<R>(h->x), V<R>(h->fx), s);
  hipLaunchKernelGGL((k_sub<R>), vgrid(M), dim3(256), 0, s, V<R>(h->fx), V<R>(h->y), V<R>(h->r), M);
  hipLaunchKernelGGL((k_zero<R>), vgrid(M), dim3(256), 0, s, V<R>(h->lam), M);
  hipLaunchKernelGGL((k_ball_norm<R>), dim3(FR_BLOCKS), dim3(SR_THREADS), 0, s, V<R>(h->lam), V<R>(h->r), M, h->fpart);
  fr_final(h->fpart, h->cur + 3, 1, s);
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, h->counter, 0);
  PF_HIP(hipStreamSynchronize(s));
  h->iters_done = 0;
  h->l1_last = 0;
  h->l1_total = 0;
}

// A captured graph holds the operator's launch arguments (its norm in the correction scale, the
// gridding variant, the FFT plan's twiddle table).  pf_op_set_norm, pf_op_set_tuning and
// pf_op_autotune change them, so the graph is captured again when the operator's generation moved.
template <typename R>
void padmm_capture(pf_padmm* h) {
  hipStream_t s = h->stream;
  const uint64_t gen = op_generation(h->phi);
  if (h->exec && (h->captured_gen != gen || h->captured_gamma != h->gamma)) {
    PF_HIP(hipStreamSynchronize(s));
    (void)hipGraphExecDestroy(h->exec);
    (void)hipGraphDestroy(h->graph);
    h->exec = nullptr;
    h->graph = nullptr;
    // run the operator's (possibly new) kernels once eagerly, so their attributes are set before capture;
    // fx and g are scratch between iterations (each iteration writes them before reading them)
    if (h->fb) {
      op_normal<R>(h->phi, V<R>(h->x), V<R>(h->g), s);
    } else {
      op_forward_sorted<R>(h->phi, V<R>(h->x), V<R>(h->fx), s);
      op_adjoint_sorted<R>(h->phi, V<R>(h->fx), V<R>(h->g), s);
    }
    PF_HIP(hipStreamSynchronize(s));
  }
  if (!h->exec) {
    if (h->iters_done == 0) {
      // one eager iteration first, so every kernel attribute is set before capture; then
      // re-initialise so the captured run starts from the initial state
      padmm_iteration<R>(h);
      PF_HIP(hipStreamSynchronize(s));
      if (h->fb) fb_init<R>(h);
      else padmm_init<R>(h);
    }
    // capture records the iteration without running it: the state is unchanged
    PF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    padmm_iteration<R>(h);
    PF_HIP(hipStreamEndCapture(s, &h->graph));
    PF_HIP(hipGraphInstantiate(&h->exec, h->graph, nullptr, nullptr, 0));
    h->captured_gen = gen;
    h->captured_gamma = h->gamma;
  }
}

template <typename R>
void padmm_enqueue(pf_padmm* h, int iters, bool use_graph) {
  if (h->iters_done + iters > h->cap) throw Error(PF_ERR_ARG, "more iterations than the solver was created for");
  use_graph = use_graph && h->tight;
  if (use_graph) padmm_capture<R>(h);
  for (int i = 0; i < iters; ++i) {
    if (use_graph) PF_HIP(hipGraphLaunch(h->exec, h->stream));
    else padmm_iteration<R>(h);
  }
  h->iters_done += iters;
}

void padmm_result(pf_padmm* h, bool converged, pf_padmm_result* res) {
  hipStream_t s = h->stream;
  PF_HIP(hipStreamSynchronize(s));
  if (res) {
    double st[3] = {0, 0, 0};
    // stream-ordered copies only: a legacy-stream copy is refused while another thread captures
    if (h->iters_done > 0) {
      PF_HIP(hipMemcpyAsync(st, h->stats + 3 * (h->iters_done - 1), 3 * sizeof(double), hipMemcpyDeviceToHost, s));
      PF_HIP(hipStreamSynchronize(s));
    }
    res->niters = h->iters_done;
    res->converged = converged ? 1 : 0;
    res->relative_variation = st[1] > 0 ? std::sqrt(st[0] / st[1]) : 0.0;
    res->residual_norm = std::sqrt(st[2]);
    res->gamma = h->gamma;
    res->l1_iters_last = h->l1_last;
    res->l1_iters_total = h->l1_total;
  }
}

template <typename R>
void padmm_run(pf_padmm* h, int iters, bool use_graph, int check_every, pf_padmm_result* res) {
  hipStream_t s = h->stream;
  if (h->iters_done + iters > h->cap) throw Error(PF_ERR_ARG, "more iterations than the solver was created for");
  use_graph = use_graph && h->tight;
  if (use_graph) padmm_capture<R>(h);
  const int chunk = check_every > 0 ? check_every : iters;
  int done = 0;
  bool converged = false;
  std::vector<double> st;
  while (done < iters && !converged) {
    const int k = std::min(chunk, iters - done);
    for (int i = 0; i < k; ++i) {
      if (use_graph) PF_HIP(hipGraphLaunch(h->exec, s));
      else padmm_iteration<R>(h);
    }
    done += k;
    h->iters_done += k;
    if (check_every > 0) {
      // convergence of the chunk's last iteration (SOPT checks after every iteration: chunk 1)
      st.resize(3);
      PF_HIP(hipMemcpyAsync(st.data(), h->stats + 3 * (h->iters_done - 1), 3 * sizeof(double), hipMemcpyDeviceToHost, s));
      PF_HIP(hipStreamSynchronize(s));
      const double relvar = st[1] > 0 ? std::sqrt(st[0] / st[1]) : 0.0;
      converged = relvar < h->par.relative_variation && std::sqrt(st[2]) < h->par.residual_tolerance;
    }
  }
  padmm_result(h, converged, res);
}

}  // namespace
}  // namespace pf

extern "C" {

int pf_padmm_params_init(pf_padmm_params* p) {
  if (!p) return fail(PF_ERR_ARG, "null argument");
  *p = pf_padmm_params{};
  p->gamma = 0;  // computed from the data (algorithm_factory.h:70-73)
  p->epsilon = 0;
  p->nu = 1.0;
  p->lagrange_update_scale = 0.9;
  p->relative_variation = 1e-3;
  p->residual_tolerance = 0;  // 0: epsilon (algorithm_factory.h:74)
  p->itermax = 500;
  // algorithm_factory.h:60-66
  p->tight_frame = -1;
  p->l1_itermax = 50;
  p->l1_tolerance = 1e-2;
  p->l1_nu = 1.0;
  p->positivity = 1;
  p->real_constraint = 1;
  return PF_OK;
}

int pf_padmm_create(pf_op* phi, pf_wavelet* psi, const void* y, int y_loc, const pf_padmm_params* params,
                    pf_padmm** out) {
  if (!phi || !psi || !y || !params || !out) return fail(PF_ERR_ARG, "null argument");
  *out = nullptr;
  std::unique_ptr<pf_padmm> h(new pf_padmm());
  int rc = guarded([&] {
    const OpShape os = op_shape(phi);
    const WaveletShape ws = wavelet_shape(psi);
    if (os.ny != ws.ny || os.nx != ws.nx) throw Error(PF_ERR_ARG, "operator and dictionary image sizes differ");
    if (os.prec != ws.prec) throw Error(PF_ERR_ARG, "operator and dictionary precisions differ");
    if (params->tight_frame > 0 && ws.nwav > 1)
      throw Error(PF_ERR_ARG,
                  "l1 proximal not consistent: a tight frame was requested, but the dictionary has more than one basis");
    if (params->tight_frame == 0 || (params->tight_frame < 0 && ws.nwav > 1)) {
      if (!(params->l1_nu > 0)) throw Error(PF_ERR_ARG, "l1_nu must be positive");
      if (params->l1_itermax < 1) throw Error(PF_ERR_ARG, "l1_itermax must be positive");
      if (!(params->l1_tolerance >= 0)) throw Error(PF_ERR_ARG, "l1_tolerance must be >= 0");
      h->tight = false;
    }
    if (!(params->epsilon > 0)) throw Error(PF_ERR_ARG, "epsilon (the l2-ball radius) must be positive");
    if (!(params->nu > 0)) throw Error(PF_ERR_ARG, "nu must be positive");
    if (params->itermax < 1) throw Error(PF_ERR_ARG, "itermax must be positive");
    h->phi = phi;
    h->psi = psi;
    h->prec = os.prec;
    h->N = (int64_t)os.ny * os.nx;
    h->M = os.M;
    h->NC = h->N * ws.nwav;
    h->par = *params;
    if (!(h->par.residual_tolerance > 0)) h->par.residual_tolerance = h->par.epsilon;
    const size_t e = os.prec == PF_C64 ? 8 : 16;
    for (void** p : {&h->x, &h->xprev, &h->xu, &h->g, &h->d}) *p = h->dalloc(e * h->N);
    h->c = h->dalloc(e * h->NC);
    if (!h->tight) h->u = h->dalloc(e * h->NC);
    for (void** p : {&h->y, &h->r, &h->lam, &h->z, &h->fx}) *p = h->dalloc(e * std::max<int64_t>(h->M, 1));
    h->part = (double*)h->dalloc(sizeof(double) * SR_BLOCKS);
    h->fpart = (double*)h->dalloc(sizeof(double) * 2 * FR_BLOCKS);
    h->cur = (double*)h->dalloc(sizeof(double) * 6);
    h->cap = params->itermax;
    h->stats = (double*)h->dalloc(sizeof(double) * 3 * h->cap);
    h->counter = (int*)h->dalloc(sizeof(int));
    PF_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    // the visibilities: host complex128 or device in the operator's precision
    if (y_loc == PF_HOST) {
      if (h->prec == PF_C64) {
        void* tmp = dev_alloc(16 * std::max<int64_t>(h->M, 1));
        PF_HIP(hipMemcpyAsync(tmp, y, 16 * h->M, hipMemcpyHostToDevice, h->stream));
        c128_to_c64((const cpx<double>*)tmp, (cpx<float>*)h->y, h->M, h->stream);
        PF_HIP(hipStreamSynchronize(h->stream));
        dev_free(tmp);
      } else {
        PF_HIP(hipMemcpyAsync(h->y, y, 16 * h->M, hipMemcpyHostToDevice, h->stream));
      }
    } else if (y_loc == PF_DEVICE) {
      PF_HIP(hipMemcpyAsync(h->y, y, e * h->M, hipMemcpyDeviceToDevice, h->stream));
    } else {
      throw Error(PF_ERR_ARG, "location must be PF_HOST or PF_DEVICE");
    }
    // every visibility-domain vector of the iteration (y, r, lambda, z) lives in the operator's sorted
    // order, so the iterations apply Phi and Phi^dagger without permutations
    if (h->prec == PF_C64) op_to_sorted<float>(phi, V<float>(h->y), V<float>(h->fx), h->stream);
    else op_to_sorted<double>(phi, V<double>(h->y), V<double>(h->fx), h->stream);
    std::swap(h->y, h->fx);
    if (h->prec == PF_C64) padmm_init<float>(h.get());
    else padmm_init<double>(h.get());
  });
  if (rc == PF_OK) *out = h.release();
  return rc;
}

int pf_fb_params_init(pf_fb_params* p) {
  if (!p) return fail(PF_ERR_ARG, "null argument");
  *p = pf_fb_params{};
  p->mu = 0;     // 1e-5 N (fb_random_coverage.cc:92)
  p->beta = 0;   // 0.5 / N (fb_random_coverage.cc:93)
  p->sigma = 0;  // required
  p->relative_variation = 1e-3;
  p->itermax = 500;
  p->tight_frame = -1;  // fb_random_coverage.cc:98 passes false with SARA
  p->l1_itermax = 50;
  p->l1_tolerance = 1e-2;
  p->l1_nu = 1.0;
  p->positivity = 1;
  p->real_constraint = 1;
  return PF_OK;
}

int pf_fb_create(pf_op* phi, pf_wavelet* psi, const void* y, int y_loc, const pf_fb_params* params, pf_padmm** out) {
  if (!phi || !psi || !y || !params || !out) return fail(PF_ERR_ARG, "null argument");
  *out = nullptr;
  std::unique_ptr<pf_padmm> h(new pf_padmm());
  int rc = guarded([&] {
    const OpShape os = op_shape(phi);
    const WaveletShape ws = wavelet_shape(psi);
    if (os.ny != ws.ny || os.nx != ws.nx) throw Error(PF_ERR_ARG, "operator and dictionary image sizes differ");
    if (os.prec != ws.prec) throw Error(PF_ERR_ARG, "operator and dictionary precisions differ");
    if (!(params->sigma > 0)) throw Error(PF_ERR_ARG, "sigma must be positive");
    if (params->itermax < 1) throw Error(PF_ERR_ARG, "itermax must be positive");
    if (params->tight_frame > 0 && ws.nwav > 1)
      throw Error(PF_ERR_ARG,
                  "l1 proximal not consistent: a tight frame was requested, but the dictionary has more than one basis");
    h->fb = true;
    h->phi = phi;
    h->psi = psi;
    h->prec = os.prec;
    h->N = (int64_t)os.ny * os.nx;
    h->M = os.M;
    h->NC = h->N * ws.nwav;
    h->tight = !(params->tight_frame == 0 || (params->tight_frame < 0 && ws.nwav > 1));
    if (!h->tight) {
      if (!(params->l1_nu > 0)) throw Error(PF_ERR_ARG, "l1_nu must be positive");
      if (params->l1_itermax < 1) throw Error(PF_ERR_ARG, "l1_itermax must be positive");
      if (!(params->l1_tolerance >= 0)) throw Error(PF_ERR_ARG, "l1_tolerance must be >= 0");
    }
    const double N = (double)h->N;
    const double mu = params->mu > 0 ? params->mu : 1e-5 * N;
    const double beta = params->beta > 0 ? params->beta : 0.5 / N;
    h->fb_step = beta / (params->sigma * params->sigma);
    h->gamma = mu * beta;
    // the shared parameter block: l1 settings, convergence on the relative variation alone
    h->par = pf_padmm_params{};
    h->par.nu = 1.0;
    h->par.relative_variation = params->relative_variation;
    h->par.residual_tolerance = HUGE_VAL;
    h->par.itermax = params->itermax;
    h->par.tight_frame = params->tight_frame;
    h->par.l1_itermax = params->l1_itermax;
    h->par.l1_tolerance = params->l1_tolerance;
    h->par.l1_nu = params->l1_nu > 0 ? params->l1_nu : 1.0;
    h->par.positivity = params->positivity;
    h->par.real_constraint = params->real_constraint;
    const size_t e = os.prec == PF_C64 ? 8 : 16;
    for (void** p : {&h->x, &h->xprev, &h->xu, &h->g, &h->d, &h->v, &h->dimage}) *p = h->dalloc(e * h->N);
    h->r = h->v;  // PF_PADMM_RESIDUAL: the image-domain residual
    h->c = h->dalloc(e * h->NC);
    if (!h->tight) h->u = h->dalloc(e * h->NC);
    h->part = (double*)h->dalloc(sizeof(double) * SR_BLOCKS);
    h->cur = (double*)h->dalloc(sizeof(double) * 6);
    h->cap = params->itermax;
    h->stats = (double*)h->dalloc(sizeof(double) * 3 * h->cap);
    h->counter = (int*)h->dalloc(sizeof(int));
    PF_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    // the target Phi^dagger y
    void* yd = dev_alloc(16 * std::max<int64_t>(h->M, 1));
    if (y_loc == PF_HOST) {
      if (h->prec == PF_C64) {
        void* tmp = dev_alloc(16 * std::max<int64_t>(h->M, 1));
        PF_HIP(hipMemcpyAsync(tmp, y, 16 * h->M, hipMemcpyHostToDevice, h->stream));
        c128_to_c64((const cpx<double>*)tmp, (cpx<float>*)yd, h->M, h->stream);
        PF_HIP(hipStreamSynchronize(h->stream));
        dev_free(tmp);
      } else {
        PF_HIP(hipMemcpyAsync(yd, y, 16 * h->M, hipMemcpyHostToDevice, h->stream));
      }
    } else if (y_loc == PF_DEVICE) {
      PF_HIP(hipMemcpyAsync(yd, y, e * h->M, hipMemcpyDeviceToDevice, h->stream));
    } else {
      dev_free(yd);
      throw Error(PF_ERR_ARG, "location must be PF_HOST or PF_DEVICE");
    }
    if (h->prec == PF_C64) {
      op_adjoint<float>(phi, (const cpx<float>*)yd, V<float>(h->dimage), h->stream);
      op_normal<float>(phi, V<float>(h->dimage), V<float>(h->g), h->stream);  // sets up the operator's buffers
    } else {
      op_adjoint<double>(phi, (const cpx<double>*)yd, V<double>(h->dimage), h->stream);
      op_normal<double>(phi, V<double>(h->dimage), V<double>(h->g), h->stream);
    }
    PF_HIP(hipStreamSynchronize(h->stream));
    dev_free(yd);
    if (h->prec == PF_C64) fb_init<float>(h.get());
    else fb_init<double>(h.get());
  });
  if (rc == PF_OK) *out = h.release();
  return rc;
}

int pf_padmm_run(pf_padmm* h, int32_t iters, int32_t use_graph, int32_t check_every, pf_padmm_result* res) {
  return guarded([&] {
    if (!h) throw Error(PF_ERR_ARG, "null handle");
    if (iters < 0) throw Error(PF_ERR_ARG, "iterations must be >= 0");
    if (iters == 0) return;
    if (h->prec == PF_C64) padmm_run<float>(h, iters, use_graph != 0, check_every, res);
    else padmm_run<double>(h, iters, use_graph != 0, check_every, res);
  });
}

int pf_padmm_enqueue(pf_padmm* h, int32_t iters, int32_t use_graph) {
  return guarded([&] {
    if (!h) throw Error(PF_ERR_ARG, "null handle");
    if (iters < 0) throw Error(PF_ERR_ARG, "iterations must be >= 0");
    if (h->prec == PF_C64) padmm_enqueue<float>(h, iters, use_graph != 0);
    else padmm_enqueue<double>(h, iters, use_graph != 0);
  });
}

int pf_padmm_sync(pf_padmm* h, pf_padmm_result* res) {
  return guarded([&] {
    if (!h) throw Error(PF_ERR_ARG, "null handle");
    padmm_result(h, false, res);
  });
}

int pf_padmm_reset(pf_padmm* h) {
  return guarded([&] {
    if (!h) throw Error(PF_ERR_ARG, "null handle");
    if (h->fb) {
      if (h->prec == PF_C64) fb_init<float>(h);
      else fb_init<double>(h);
    } else if (h->prec == PF_C64) {
      padmm_init<float>(h);
    } else {
      padmm_init<double>(h);
    }
  });
}

int pf_padmm_get(pf_padmm* h, int32_t what, void* out, int out_loc) {
  return guarded([&] {
    if (!h || !out) throw Error(PF_ERR_ARG, "null argument");
    PF_HIP(hipStreamSynchronize(h->stream));
    const void* src = what == PF_PADMM_IMAGE ? h->x : what == PF_PADMM_RESIDUAL ? h->r : nullptr;
    if (!src) throw Error(PF_ERR_ARG, "what must be PF_PADMM_IMAGE or PF_PADMM_RESIDUAL");
    const int64_t n = what == PF_PADMM_IMAGE || h->fb ? h->N : h->M;
    if (what == PF_PADMM_RESIDUAL && !h->fb) {
      // r is in the operator's sorted order; fx is scratch between iterations
      if (h->prec == PF_C64) op_to_caller<float>(h->phi, V<float>(h->r), V<float>(h->fx), h->stream);
      else op_to_caller<double>(h->phi, V<double>(h->r), V<double>(h->fx), h->stream);
      src = h->fx;
    }
    const size_t e = h->prec == PF_C64 ? 8 : 16;
    if (out_loc == PF_DEVICE) {
      PF_HIP(hipMemcpyAsync(out, src, e * n, hipMemcpyDeviceToDevice, h->stream));
      PF_HIP(hipStreamSynchronize(h->stream));
    } else if (out_loc == PF_HOST) {
      if (h->prec == PF_C64) {
        void* tmp = dev_alloc(16 * std::max<int64_t>(n, 1));
        c64_to_c128((const cpx<float>*)src, (cpx<double>*)tmp, n, h->stream);
        PF_HIP(hipMemcpyAsync(out, tmp, 16 * n, hipMemcpyDeviceToHost, h->stream));
        PF_HIP(hipStreamSynchronize(h->stream));
        dev_free(tmp);
      } else {
        PF_HIP(hipMemcpyAsync(out, src, 16 * n, hipMemcpyDeviceToHost, h->stream));
        PF_HIP(hipStreamSynchronize(h->stream));
      }
    } else {
      throw Error(PF_ERR_ARG, "location must be PF_HOST or PF_DEVICE");
    }
  });
}

int pf_padmm_stats(pf_padmm* h, double* out, int32_t cap, int32_t* count) {
  return guarded([&] {
    if (!h) throw Error(PF_ERR_ARG, "null handle");
    PF_HIP(hipStreamSynchronize(h->stream));
    const int n = std::min(h->iters_done, cap);
    if (n > 0 && out) {
      PF_HIP(hipMemcpyAsync(out, h->stats, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, h->stream));
      PF_HIP(hipStreamSynchronize(h->stream));
    }
    if (count) *count = h->iters_done;
  });
}

int pf_padmm_destroy(pf_padmm* h) {
  delete h;
  return PF_OK;
}

}  // extern "C"

// ---- channel batches: one launch per kernel of the iteration for every group of alike channels ----
// Each member solver's tight-frame PADMM iteration is recorded (Recorder: the launch sites record their
// arguments and shapes instead of launching, inside a stream capture that must stay empty, so that a
// launch without a batch twin is caught).  Channels whose recordings agree position by position (the
// same kernels, block shapes, LDS and grids, or grids that differ only where the kernel strides over
// gridDim.x) form a group; each group runs on its own stream, every position launched once with the
// group's argument table, blockIdx.z selecting the channel.
struct pf_padmm_batch {
  struct Pos {
    void (*launch)(dim3, dim3, size_t, hipStream_t, const void*);
    dim3 grid, block;
    size_t shm, off;
  };
  struct Group {
    std::vector<int> members;
    std::vector<Pos> pos;
    hipStream_t s = nullptr;
    hipEvent_t done = nullptr;
  };
  std::vector<pf_padmm*> h;  // not owned
  std::vector<Group> groups;
  void* d_tab = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uint64_t> rec_gen;  // each member's operator generation and l1 weight when recorded
  std::vector<double> rec_gamma;
  bool ran = false;
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  void drop_groups() {
    for (Group& g : groups) {
      if (g.s) (void)hipStreamDestroy(g.s);
      if (g.done) (void)hipEventDestroy(g.done);
    }
    groups.clear();
    dev_free(d_tab);
    d_tab = nullptr;
  }
  ~pf_padmm_batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graph();
    drop_groups();
    if (fork) (void)hipEventDestroy(fork);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace pf {
namespace {

bool same_shape(const LaunchRec& a, const LaunchRec& b) {
  return a.key == b.key && a.block.x == b.block.x && a.block.y == b.block.y && a.block.z == b.block.z &&
         a.shm == b.shm && a.grid.y == b.grid.y && a.grid.z == b.grid.z && a.args.size() == b.args.size() &&
         a.grid_x_max == b.grid_x_max && (a.grid_x_max || a.grid.x == b.grid.x);
}

void batch_record(pf_padmm_batch* b) {
  const int n = (int)b->h.size();
  std::vector<std::vector<LaunchRec>> recs((size_t)n);
  hipStream_t cap = nullptr;
  PF_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
  PF_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeRelaxed));
  size_t stray = 0;
  try {
    for (int c = 0; c < n; ++c) {
      pf_padmm* h = b->h[(size_t)c];
      Recorder r;
      recorder() = &r;
      const hipStream_t own = h->stream;
      h->stream = cap;
      try {
        if (h->prec == PF_C64) padmm_step<float>(h);
        else padmm_step<double>(h);
      } catch (...) {
        h->stream = own;
        recorder() = nullptr;
        throw;
      }
      h->stream = own;
      recorder() = nullptr;
      recs[(size_t)c] = std::move(r.recs);
    }
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(cap, &g);
    if (g) (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(cap);
    throw;
  }
  hipGraph_t g = nullptr;
  PF_HIP(hipStreamEndCapture(cap, &g));
  if (g) {
    (void)hipGraphGetNodes(g, nullptr, &stray);
    (void)hipGraphDestroy(g);
  }
  (void)hipStreamDestroy(cap);
  if (stray)
    throw Error(PF_ERR_UNSUPPORTED, "the iteration queues work that a channel batch cannot take (" +
                                        std::to_string(stray) + " operations; is operator profiling on?)");
  // groups of channels whose recordings agree position by position
  b->drop_graph();
  b->drop_groups();
  for (int c = 0; c < n; ++c) {
    int gi = -1;
    for (size_t k = 0; k < b->groups.size() && gi < 0; ++k) {
      const auto& rep = recs[(size_t)b->groups[k].members[0]];
      const auto& mine = recs[(size_t)c];
      bool ok = rep.size() == mine.size();
      for (size_t p = 0; ok && p < rep.size(); ++p) ok = same_shape(rep[p], mine[p]);
      if (ok) gi = (int)k;
    }
    if (gi < 0) {
      b->groups.emplace_back();
      gi = (int)b->groups.size() - 1;
    }
    b->groups[(size_t)gi].members.push_back(c);
  }
  if (b->groups.size() > 65535) throw Error(PF_ERR_UNSUPPORTED, "too many channel groups");
  // the argument tables: per group and position, the members' argument structs side by side
  std::vector<unsigned char> host;
  for (auto& gr : b->groups) {
    const auto& rep = recs[(size_t)gr.members[0]];
    for (size_t p = 0; p < rep.size(); ++p) {
      pf_padmm_batch::Pos pos;
      pos.launch = rep[p].launch;
      pos.grid = rep[p].grid;
      pos.block = rep[p].block;
      pos.shm = rep[p].shm;
      pos.off = (host.size() + 15) & ~size_t(15);
      host.resize(pos.off);
      const size_t sz = rep[p].args.size();  // the kernel's stride: sizeof of its argument struct
      for (int c : gr.members) {
        const LaunchRec& lr = recs[(size_t)c][p];
        if (lr.grid_x_max) pos.grid.x = std::max(pos.grid.x, lr.grid.x);
        const size_t at = host.size();
        host.resize(at + sz, 0);
        std::memcpy(host.data() + at, lr.args.data(), lr.args.size());
      }
      pos.grid.z = (unsigned)gr.members.size();
      if (pos.shm > 65536)
        PF_HIP(hipFuncSetAttribute(rep[p].key, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pos.shm));
      gr.pos.push_back(pos);
    }
    if (b->groups.size() > 1) {
      PF_HIP(hipStreamCreateWithFlags(&gr.s, hipStreamNonBlocking));
      PF_HIP(hipEventCreateWithFlags(&gr.done, hipEventDisableTiming));
    }
  }
  b->d_tab = dev_alloc(std::max<size_t>(host.size(), 16));
  h2d(b->d_tab, host.data(), host.size());
  b->rec_gen.resize((size_t)n);
  b->rec_gamma.resize((size_t)n);
  for (int c = 0; c < n; ++c) {
    b->rec_gen[(size_t)c] = op_generation(b->h[(size_t)c]->phi);
    b->rec_gamma[(size_t)c] = b->h[(size_t)c]->gamma;
  }
}

// one iteration of every member, queued on the batch's stream (forked over the groups' streams)
void batch_iteration(pf_padmm_batch* b) {
  const bool fork = b->groups.size() > 1;
  if (fork) PF_HIP(hipEventRecord(b->fork, b->stream));
  for (auto& gr : b->groups) {
    hipStream_t s = fork ? gr.s : b->stream;
    if (fork) PF_HIP(hipStreamWaitEvent(s, b->fork, 0));
    for (const auto& p : gr.pos)
      p.launch(p.grid, p.block, p.shm, s, static_cast<const unsigned char*>(b->d_tab) + p.off);
    if (fork) {
      PF_HIP(hipEventRecord(gr.done, s));
      PF_HIP(hipStreamWaitEvent(b->stream, gr.done, 0));
    }
  }
  PF_HIP(hipGetLastError());
}

void batch_run(pf_padmm_batch* b, int iters, bool use_graph) {
  const int n = (int)b->h.size();
  bool stale = b->groups.empty();
  for (int c = 0; c < n; ++c) {
    pf_padmm* h = b->h[(size_t)c];
    if (h->iters_done + iters > h->cap) throw Error(PF_ERR_ARG, "more iterations than a member solver was created for");
    if (!stale && (op_generation(h->phi) != b->rec_gen[(size_t)c] || h->gamma != b->rec_gamma[(size_t)c])) stale = true;
    PF_HIP(hipStreamSynchronize(h->stream));  // the member's own queued work first
  }
  if (stale) batch_record(b);
  int done = 0;
  if (use_graph && !b->exec && iters > 0 && !b->ran) {
    batch_iteration(b);  // the first launches of the batch twins run eagerly, before any capture
    done = 1;
  }
  if (use_graph && !b->exec && iters > done) {
    PF_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
    try {
      batch_iteration(b);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(b->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    PF_HIP(hipStreamEndCapture(b->stream, &b->graph));
    PF_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  for (; done < iters; ++done) {
    if (use_graph && b->exec) PF_HIP(hipGraphLaunch(b->exec, b->stream));
    else batch_iteration(b);
  }
  if (iters > 0) b->ran = true;
  PF_HIP(hipStreamSynchronize(b->stream));
  for (pf_padmm* h : b->h) h->iters_done += iters;
}

}  // namespace
}  // namespace pf

extern "C" {

int pf_padmm_batch_create(pf_padmm* const* solvers, int32_t n, pf_padmm_batch** out) {
  if (!solvers || !out || n <= 0) return fail(PF_ERR_ARG, "null argument or empty batch");
  *out = nullptr;
  std::unique_ptr<pf_padmm_batch> b(new pf_padmm_batch());
  int rc = guarded([&] {
    for (int c = 0; c < n; ++c) {
      pf_padmm* h = solvers[c];
      if (!h) throw Error(PF_ERR_ARG, "null solver in the batch");
      if (h->fb || !h->tight) throw Error(PF_ERR_UNSUPPORTED, "channel batches take tight-frame PADMM solvers");
      for (int k = 0; k < c; ++k)
        if (solvers[k] == h) throw Error(PF_ERR_ARG, "a solver appears twice in the batch");
      b->h.push_back(h);
    }
    PF_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    PF_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    for (pf_padmm* h : b->h) PF_HIP(hipStreamSynchronize(h->stream));
    batch_record(b.get());
  });
  if (rc == PF_OK) *out = b.release();
  return rc;
}

int pf_padmm_batch_run(pf_padmm_batch* b, int32_t iters, int32_t use_graph) {
  return guarded([&] {
    if (!b) throw Error(PF_ERR_ARG, "null batch");
    if (iters < 0) throw Error(PF_ERR_ARG, "iterations must be >= 0");
    batch_run(b, iters, use_graph != 0);
  });
}

int pf_padmm_batch_info(const pf_padmm_batch* b, int32_t* groups, int32_t* launches_per_iteration) {
  if (!b) return fail(PF_ERR_ARG, "null batch");
  int64_t l = 0;
  for (const auto& g : b->groups) l += (int64_t)g.pos.size();
  if (groups) *groups = (int32_t)b->groups.size();
  if (launches_per_iteration) *launches_per_iteration = (int32_t)l;
  return PF_OK;
}

int pf_padmm_batch_destroy(pf_padmm_batch* b) {
  delete b;
  return PF_OK;
}

}  // extern "C"
