// pf_wproj.hip -- the w-projection gridding matrix (w_term operators): rows of (Jv + Jw - 1) x
// (Ju + Jw - 1) taps per visibility, built once on the host in FP64 and stored in HBM in the operator's
// precision (replaces details::init_gridding_matrix_2d with w_term, cpp/purify/operators.cc:49-99, and
// projection_kernels::projection, cpp/purify/projection_kernels.cc:59-85).
//
//  * degrid_wrows: one 256-thread block per owner tile stages the tile and its (K - 1)-wide halo in LDS;
//    each wave then takes one visibility at a time, its lanes striding over the row's taps (coalesced
//    reads of the stored row), and sums the products with a fixed butterfly: deterministic.
//  * grid_wrows: one wave per tile accumulates in LDS; for each visibility of the tile's list (in
//    sorted order) the lanes add the taps that fall inside the tile.  The taps of one visibility are
//    distinct grid cells (K <= N), so the lanes' read-modify-writes never collide and every cell's sum
//    follows the list order: reproducible bit for bit.
#include <algorithm>
#include <cmath>
#include <complex>

#include "../../include/purify_amd.h"
#include "pf_internal.h"

namespace pf {

// ---------------------------------------------------------------------------------------------
// Host: one projection row in FP64
// ---------------------------------------------------------------------------------------------
namespace {
const double PI_W = 3.14159265358979323846;

// w_projection_kernel_approx (projection_kernels.cc:31-47) with w_proj_approx (16-22): the chirp
// exp(-i pi (a^2 + b^2) / w) / (i w) at a = u du, b = v dv inside |a/w| < 1, |b/w| < 1, else 1.
std::complex<double> wkernel_approx(double u, double v, double w, double du, double dv) {
  if (std::abs(u * du / w) < 1 && std::abs(v * dv / w) < 1) {
    const double a = u * du, b = v * dv;
    const std::complex<double> I(0.0, 1.0);
    return std::exp(-PI_W * I * (a * a + b * b) / w) / w / I;
  }
  return 1.0;
}
}  // namespace

void wrow_host(const WRowSpec& sp, double u, double v, double w, double wre, double wim, double* out) {
  const int Ju = sp.Ju, Jv = sp.Jv, Jw = sp.Jw;
  const int Ku = Jw + Ju - 1, Kv = Jw + Jv - 1;
  // 1-D kernel samples; both axes use the u-slot kernel (projection_kernels.cc:66-74)
  double fu[16], fv[16];
  const double du = u - std::floor(u - Ju * 0.5);
  for (int a = 0; a < Ju; ++a) fu[a] = host_kernel(sp.kind, du - a - 1, sp.Jeval, sp.os);
  const double dv = v - std::floor(v - Jv * 0.5);
  for (int b = 0; b < Jv; ++b) fv[b] = host_kernel(sp.kind, dv - b - 1, sp.Jeval, sp.os);
  // Jw x Jw w-kernel samples g[jv][ju] (projection_kernels.cc:75-80)
  std::complex<double> g[16][16];
  const double dwu = u - std::floor(u - Jw * 0.5), dwv = v - std::floor(v - Jw * 0.5);
  for (int jv = 0; jv < Jw; ++jv)
    for (int ju = 0; ju < Jw; ++ju)
      g[jv][ju] = wkernel_approx(dwu - ju - 1, dwv - jv - 1, w, sp.du_lambda, sp.dv_lambda);
  // linear_convol_2d (convolution.h:54-67): columns with fv, then rows with fu.  A 1-D pass gives
  // out[j] = sum_t f[t] * x[Jx - 1 - j + t] over the valid indices of x (convolution.h:43-52).
  std::complex<double> buf[32][16];
  for (int i = 0; i < Jw; ++i)
    for (int r = 0; r < Kv; ++r) {
      std::complex<double> acc = 0.0;
      for (int t = 0; t < Jv; ++t) {
        const int k = Jw - 1 - r + t;
        if (k >= 0 && k < Jw) acc += fv[t] * g[k][i];
      }
      buf[r][i] = acc;
    }
  const std::complex<double> wt(wre, wim);
  const double ku = std::floor(u - Ku * 0.5), kv = std::floor(v - Kv * 0.5);
  for (int r = 0; r < Kv; ++r)
    for (int c = 0; c < Ku; ++c) {
      std::complex<double> acc = 0.0;
      for (int t = 0; t < Ju; ++t) {
        const int k = Jw - 1 - c + t;
        if (k >= 0 && k < Jw) acc += fu[t] * buf[r][k];
      }
      // (-1)^(kw_u + c + 1 + kw_v + r + 1): the exact value of the reference's
      // exp(-2 pi i ((kw_u + ju) / 2 + (kw_v + jv) / 2)) (operators.cc:93)
      const long long par = (long long)ku + (long long)kv + c + r + 2;
      const std::complex<double> val = ((par & 1) ? -acc : acc) * wt;
      out[2 * (r * Ku + c)] = val.real();
      out[2 * (r * Ku + c) + 1] = val.imag();
    }
}

// ---------------------------------------------------------------------------------------------
// Device
// ---------------------------------------------------------------------------------------------
__device__ inline int wrapw(int a, int n) { return a & (n - 1); }

constexpr int WD_THREADS = 256;
constexpr int WD_ROWS = 4;  // staged rows per wave per round

template <typename R>
__global__ void __launch_bounds__(WD_THREADS) k_degrid_wrows(int Nu, int Nv, WRows<R> w, const int32_t* __restrict__ perm,
                                                             TileLists tl, const cpx<R>* __restrict__ grid,
                                                             cpx<R>* __restrict__ y, R scl) {
  extern __shared__ unsigned char smem_raw[];
  cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
  const uint4 sl = tl.slot_own[blockIdx.x];
  const int tile = (int)sl.x;
  const int tv = tile / tl.ntu, tu = tile - tv * tl.ntu;
  const int tx0 = tu * tl.TU, ty0 = tv * tl.TV;
  const int W = tl.TU + w.Ku - 1, H = tl.TV + w.Kv - 1;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = WD_THREADS / 64;
  // stage the tile and its halo (wrapping at the grid edges), loads before stores
  for (int r0 = wave * WD_ROWS; r0 < H; r0 += nw * WD_ROWS) {
    for (int lx = lane; lx < W; lx += 64) {
      const int gx = wrapw(tx0 + lx, Nu);
      cpx<R> tmp[WD_ROWS];
#pragma unroll
      for (int k = 0; k < WD_ROWS; ++k) {
        const int gy = wrapw(ty0 + min(r0 + k, H - 1), Nv);
        tmp[k] = grid[(int64_t)gy * Nu + gx];
      }
#pragma unroll
      for (int k = 0; k < WD_ROWS; ++k)
        if (r0 + k < H) lds[(r0 + k) * W + lx] = tmp[k];
    }
  }
  __syncthreads();
  const int KK = w.Ku * w.Kv;
  for (uint32_t i = sl.y + wave; i < sl.z; i += nw) {
    const int2 k0 = w.kw[i];
    const int base = (wrapw(k0.y + 1, Nv) - ty0) * W + (wrapw(k0.x + 1, Nu) - tx0);
    const cpx<R>* row = w.vals + (int64_t)i * KK;
    cpx<R> acc = mk<R>(R(0), R(0));
    for (int k = lane; k < KK; k += 64) {
      const int jv = k / w.Ku, ju = k - jv * w.Ku;
      const cpx<R> g = lds[base + jv * W + ju];
      const cpx<R> a = row[k];
      acc.x += a.x * g.x - a.y * g.y;
      acc.y += a.x * g.y + a.y * g.x;
    }
    acc = wave_sum(acc);
    if (lane == 0) y[perm ? perm[i] : i] = scale(acc, scl);
  }
}

template <typename R>
__global__ void __launch_bounds__(64) k_grid_wrows(int Nu, int Nv, WRows<R> w, TileLists tl,
                                                   const cpx<R>* __restrict__ ys, cpx<R>* __restrict__ grid, R scl) {
  extern __shared__ unsigned char smem_raw[];
  cpx<R>* acc = reinterpret_cast<cpx<R>*>(smem_raw);
  const uint4 sl = tl.slot[blockIdx.x];
  const int tile = (int)sl.x;
  const int TU = tl.TU, TV = tl.TV;
  const int tv = tile / tl.ntu, tu = tile - tv * tl.ntu;
  const int tx0 = tu * TU, ty0 = tv * TV;
  const int lane = threadIdx.x;
  for (int e = lane; e < TU * TV; e += 64) acc[e] = mk<R>(R(0), R(0));
  __syncthreads();
  const int KK = w.Ku * w.Kv;
  for (uint32_t e = sl.y; e < sl.z; ++e) {
    const uint32_t i = tl.list[e];
    const int2 k0 = w.kw[i];
    const cpx<R> s = scale(ys[i], scl);
    const cpx<R>* row = w.vals + (int64_t)i * KK;
    for (int k = lane; k < KK; k += 64) {
      const int jv = k / w.Ku, ju = k - jv * w.Ku;
      const int x = wrapw(k0.x + 1 + ju - tx0, Nu), yy = wrapw(k0.y + 1 + jv - ty0, Nv);
      if (x < TU && yy < TV) {
        const cpx<R> a = row[k];
        // conj(G) * y
        cpx<R>& c = acc[yy * TU + x];
        c.x += a.x * s.x + a.y * s.y;
        c.y += a.x * s.y - a.y * s.x;
      }
    }
  }
  __syncthreads();
  for (int e = lane; e < TU * TV; e += 64) {
    const int ly = e / TU, lx = e - ly * TU;
    grid[(int64_t)(ty0 + ly) * Nu + tx0 + lx] = acc[e];
  }
}

template <typename R>
void degrid_wrows(int Nu, int Nv, const WRows<R>& w, const int32_t* perm, const TileLists& t, const cpx<R>* grid,
                  cpx<R>* y, R scale, hipStream_t s) {
  if (t.n_own == 0) return;
  const size_t shm = sizeof(cpx<R>) * (size_t)(t.TU + w.Ku - 1) * (t.TV + w.Kv - 1);
  if (shm > 65536) throw Error(PF_ERR_UNSUPPORTED, "w-projection degridding tile and halo exceed 64 KiB of LDS");
  ProfScope prof("degrid_wrows", s);
  hipLaunchKernelGGL((k_degrid_wrows<R>), dim3((unsigned)t.n_own), dim3(WD_THREADS), shm, s, Nu, Nv, w, perm, t, grid,
                     y, scale);
  PF_HIP(hipGetLastError());
}

template <typename R>
void grid_wrows(int Nu, int Nv, const WRows<R>& w, const TileLists& t, const cpx<R>* ys, cpx<R>* grid, R scale,
                hipStream_t s) {
  const size_t shm = sizeof(cpx<R>) * (size_t)t.TU * t.TV;
  ProfScope prof("grid_wrows", s);
  hipLaunchKernelGGL((k_grid_wrows<R>), dim3((unsigned)(t.ntu * t.ntv)), dim3(64), shm, s, Nu, Nv, w, t, ys, grid,
                     scale);
  PF_HIP(hipGetLastError());
}

template void degrid_wrows<float>(int, int, const WRows<float>&, const int32_t*, const TileLists&, const cpx<float>*,
                                  cpx<float>*, float, hipStream_t);
template void degrid_wrows<double>(int, int, const WRows<double>&, const int32_t*, const TileLists&,
                                   const cpx<double>*, cpx<double>*, double, hipStream_t);
template void grid_wrows<float>(int, int, const WRows<float>&, const TileLists&, const cpx<float>*, cpx<float>*, float,
                                hipStream_t);
template void grid_wrows<double>(int, int, const WRows<double>&, const TileLists&, const cpx<double>*, cpx<double>*,
                                 double, hipStream_t);

}  // namespace pf
