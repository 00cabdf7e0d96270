// pf_wproj.hip -- the w-projection gridding matrix (w_term operators): rows of (Jv + Jw - 1) x
// (Ju + Jw - 1) taps per visibility, built once on the host in FP64 and stored in HBM in the operator's
// precision (replaces details::init_gridding_matrix_2d with w_term, cpp/purify/operators.cc:49-99, and
// projection_kernels::projection, cpp/purify/projection_kernels.cc:59-85); built on the device in FP64.
//
//  * degrid_wrows: one 256-thread block per owner tile stages the tile and its (K - 1)-wide halo in LDS;
//    each wave then takes eight visibilities at a time, eight lanes each striding over a row's taps, and
//    each group sums its products with a fixed butterfly: deterministic.
//  * grid_wrows: one wave per tile accumulates in LDS; the tile's list (in sorted order) is walked in
//    batches of 64 entries loaded together, one entry at a time with the next entry's row prefetched,
//    the lanes adding the taps that fall inside the tile.  The taps of one visibility are distinct grid
//    cells (K <= N), so the lanes' read-modify-writes never collide and every cell's sum follows the
//    list order: reproducible bit for bit.
#include <algorithm>
#include <cmath>
#include <complex>

#include "../../include/purify_amd.h"
#include "pf_internal.h"

namespace pf {

// ---------------------------------------------------------------------------------------------
// Construction: the projection rows in FP64 on the device, one thread per (visibility, row of taps)
// ---------------------------------------------------------------------------------------------
namespace {
constexpr double PI_W = 3.14159265358979323846;

// w_projection_kernel_approx (projection_kernels.cc:31-47) with w_proj_approx (16-22): the chirp
// exp(-i pi (a^2 + b^2) / w) / (i w) at a = u du, b = v dv inside |a/w| < 1, |b/w| < 1, else 1.
__device__ inline cpx<double> wkernel_approx(double u, double v, double w, double du, double dv) {
  if (fabs(u * du / w) < 1 && fabs(v * dv / w) < 1) {
    const double a = u * du, b = v * dv;
    double sn, cs;
    sincos(-PI_W * (a * a + b * b) / w, &sn, &cs);
    // (cs + i sn) / w / i
    return mk<double>(sn / w, -cs / w);
  }
  return mk<double>(1.0, 0.0);
}

// Row r (of Kv) of the projection row of sorted visibility i, as projection_kernels.cc:59-85 composes
// it: 1-D kernel samples on both axes from the u-slot kernel (projection_kernels.cc:66-74), the Jw x Jw
// w-kernel samples g[jv][ju] (75-80), linear_convol_2d (convolution.h:54-67: columns with fv, then rows
// with fu; a 1-D pass gives out[j] = sum_t f[t] x[Jx - 1 - j + t] over the valid indices of x,
// convolution.h:43-52), the exact (-1)^(kw_u + c + 1 + kw_v + r + 1) phase of operators.cc:93 and the
// weight; w enters as w - mean(w) (operators.h:323).
template <typename R>
__global__ void k_wrows_build(WRowArgs g, const cpx<double>* __restrict__ uv, const double* __restrict__ w,
                              const cpx<double>* __restrict__ wt, const int32_t* __restrict__ perm, int64_t M,
                              cpx<R>* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = t / g.Kv;
  const int r = (int)(t - i * g.Kv);
  if (i >= M) return;
  const int64_t m = perm[i];
  const cpx<double> c = uv[m];
  const double wv = w[m] - g.wmean;
  double fu[16], fv[16];
  const double du = c.x - floor(c.x - g.Ju * 0.5);
  for (int a = 0; a < g.Ju; ++a) fu[a] = kern1d<double>(g.kp, du - a - 1);
  const double dv = c.y - floor(c.y - g.Jv * 0.5);
  for (int b = 0; b < g.Jv; ++b) fv[b] = kern1d<double>(g.kp, dv - b - 1);
  const double dwu = c.x - floor(c.x - g.Jw * 0.5), dwv = c.y - floor(c.y - g.Jw * 0.5);
  // buf[q] = row r of the column pass: sum_t fv[t] g[Jw - 1 - r + t][q]
  cpx<double> buf[16];
  for (int q = 0; q < g.Jw; ++q) {
    double sr = 0, si = 0;
    for (int tt = 0; tt < g.Jv; ++tt) {
      const int k = g.Jw - 1 - r + tt;
      if (k >= 0 && k < g.Jw) {
        const cpx<double> gk = wkernel_approx(dwu - q - 1, dwv - k - 1, wv, g.du_l, g.dv_l);
        sr += fv[tt] * gk.x;
        si += fv[tt] * gk.y;
      }
    }
    buf[q] = mk<double>(sr, si);
  }
  const cpx<double> wgt = wt ? wt[m] : mk<double>(1.0, 0.0);
  const double kwu = floor(c.x - g.Ku * 0.5), kwv = floor(c.y - g.Kv * 0.5);
  cpx<R>* o = out + i * (int64_t)g.Ku * g.Kv + (int64_t)r * g.Ku;
  for (int col = 0; col < g.Ku; ++col) {
    double sr = 0, si = 0;
    for (int tt = 0; tt < g.Ju; ++tt) {
      const int k = g.Jw - 1 - col + tt;
      if (k >= 0 && k < g.Jw) {
        sr += fu[tt] * buf[k].x;
        si += fu[tt] * buf[k].y;
      }
    }
    const long long par = (long long)kwu + (long long)kwv + col + r + 2;
    if (par & 1) {
      sr = -sr;
      si = -si;
    }
    o[col] = mk<R>((R)(sr * wgt.x - si * wgt.y), (R)(sr * wgt.y + si * wgt.x));
  }
}
}  // namespace

template <typename R>
void wrows_build(const WRowArgs& g, const cpx<double>* uv, const double* w, const cpx<double>* wt, const int32_t* perm,
                 int64_t M, cpx<R>* out) {
  if (M == 0) return;
  const int64_t n = M * g.Kv;
  hipLaunchKernelGGL((k_wrows_build<R>), dim3((unsigned)((n + 127) / 128)), dim3(128), 0, cstream(), g, uv, w, wt, perm, M,
                     out);
  PF_HIP(hipGetLastError());
}
template void wrows_build<float>(const WRowArgs&, const cpx<double>*, const double*, const cpx<double>*, const int32_t*,
                                 int64_t, cpx<float>*);
template void wrows_build<double>(const WRowArgs&, const cpx<double>*, const double*, const cpx<double>*,
                                  const int32_t*, int64_t, cpx<double>*);

// ---------------------------------------------------------------------------------------------
// Device
// ---------------------------------------------------------------------------------------------
__device__ inline int wrapw(int a, int n) { return a & (n - 1); }

constexpr int WD_THREADS = 256;
constexpr int WD_ROWS = 4;  // staged rows per wave per round

constexpr int WG_LANES = 8;  // lanes per visibility in degridding: 8 visibilities per wave step

// Degridding: the block stages its owner tile and the (K - 1)-wide halo; each wave then takes 8
// visibilities at a time, 8 lanes each: lane r of a group sums the taps k = r, r + 8, ... of its
// visibility's stored row (8 consecutive taps per step: one 64-byte segment of the row for complex64),
// and the group adds its 8 partial sums with a fixed butterfly.
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
  const int grp = lane / WG_LANES, r = lane % WG_LANES;
  constexpr int GPW = 64 / WG_LANES;  // visibilities per wave step
  for (uint32_t i0 = sl.y + (uint32_t)wave * GPW; i0 < sl.z; i0 += (uint32_t)nw * GPW) {
    const uint32_t i = i0 + (uint32_t)grp;
    cpx<R> acc = mk<R>(R(0), R(0));
    if (i < sl.z) {
      const int2 k0 = w.kw[i];
      const int base = (wrapw(k0.y + 1, Nv) - ty0) * W + (wrapw(k0.x + 1, Nu) - tx0);
      const cpx<R>* row = w.vals + (int64_t)i * KK;
      for (int k = r; k < KK; k += WG_LANES) {
        const int jv = k / w.Ku, ju = k - jv * w.Ku;
        const cpx<R> g = lds[base + jv * W + ju];
        const cpx<R> a = row[k];
        acc.x += a.x * g.x - a.y * g.y;
        acc.y += a.x * g.y + a.y * g.x;
      }
    }
#pragma unroll
    for (int m = WG_LANES / 2; m >= 1; m >>= 1) {
      acc.x += __shfl_xor(acc.x, m, 64);
      acc.y += __shfl_xor(acc.y, m, 64);
    }
    if (r == 0 && i < sl.z) y[perm ? perm[i] : i] = scale(acc, scl);
  }
}

// Gridding: one wave per tile accumulates in LDS.  The tile's list is taken in batches of 64 entries:
// each lane loads one entry's index, tap origin and value, then the wave walks the batch entry by entry
// (broadcasting the entry from its lane), every lane adding its taps of the entry's stored row that
// fall inside the tile, with the next D entries' row segments already requested.  The taps of one
// visibility are distinct cells (K <= N), so the lanes' read-modify-writes never collide and each
// cell's sum follows the list order: reproducible bit for bit.  TPL: taps per lane (KK <= 64 TPL); D:
// the prefetch depth in entries (D * TPL row segments in registers).
template <typename R, int TPL, int D>
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
  // this lane's taps: k = lane + 64 t, their row and column offsets (lanes past the row have none: they
  // must not touch any cell, or they would race the lane that holds the same cell's tap)
  int dv[TPL], du[TPL];
  bool has[TPL];
#pragma unroll
  for (int t = 0; t < TPL; ++t) {
    const int k = lane + 64 * t;
    has[t] = k < KK;
    dv[t] = has[t] ? k / w.Ku : 0;
    du[t] = has[t] ? k - (k / w.Ku) * w.Ku : 0;
  }
  for (uint32_t b0 = sl.y; b0 < sl.z; b0 += 64) {
    const int n = (int)min(64u, sl.z - b0);
    uint32_t my_i = 0;
    int2 my_k = make_int2(0, 0);
    cpx<R> my_s = mk<R>(R(0), R(0));
    if (lane < n) {
      my_i = tl.list[b0 + lane];
      my_k = w.kw[my_i];
      my_s = scale(ys[my_i], scl);
    }
    // a ring of the next D entries' row segments, indexed statically (the loop is unrolled by it)
    cpx<R> ring[D][TPL];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const uint32_t i = __shfl(my_i, d, 64);
      const cpx<R>* row = w.vals + (int64_t)i * KK;
#pragma unroll
      for (int t = 0; t < TPL; ++t)
        ring[d][t] = (d < n && has[t]) ? row[lane + 64 * t] : mk<R>(R(0), R(0));
    }
    for (int j0 = 0; j0 < n; j0 += D) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const int j = j0 + d;
        if (j >= n) break;
        cpx<R> a[TPL];
#pragma unroll
        for (int t = 0; t < TPL; ++t) a[t] = ring[d][t];
        if (j + D < n) {  // entry j + D's row, requested before this entry's updates
          const uint32_t i1 = __shfl(my_i, j + D, 64);
          const cpx<R>* row = w.vals + (int64_t)i1 * KK;
#pragma unroll
          for (int t = 0; t < TPL; ++t)
            if (has[t]) ring[d][t] = row[lane + 64 * t];
        }
        const int kx = __shfl(my_k.x, j, 64), ky = __shfl(my_k.y, j, 64);
        cpx<R> s;
        s.x = __shfl(my_s.x, j, 64);
        s.y = __shfl(my_s.y, j, 64);
#pragma unroll
        for (int t = 0; t < TPL; ++t) {
          const int x = wrapw(kx + 1 + du[t] - tx0, Nu), yy = wrapw(ky + 1 + dv[t] - ty0, Nv);
          if (has[t] && x < TU && yy < TV) {
            // conj(G) * y
            cpx<R>& c = acc[yy * TU + x];
            c.x += a[t].x * s.x + a[t].y * s.y;
            c.y += a[t].x * s.y - a[t].y * s.x;
          }
        }
      }
    }
  }
  __syncthreads();
  if (sl.w == PART_NONE) {
    for (int e = lane; e < TU * TV; e += 64) {
      const int ly = e / TU, lx = e - ly * TU;
      grid[(int64_t)(ty0 + ly) * Nu + tx0 + lx] = acc[e];
    }
  } else {  // one part of a split tile: its partial sums, combined below
    cpx<R>* part = static_cast<cpx<R>*>(tl.part) + (int64_t)sl.w * TU * TV;
    for (int e = lane; e < TU * TV; e += 64) part[e] = acc[e];
  }
}

// The split tiles: each cell sums its parts' partial sums in part order (the list order of the parts'
// ranges), so the result is the same on every run.
template <typename R>
__global__ void __launch_bounds__(256) k_combine_parts(int Nu, TileLists tl, cpx<R>* __restrict__ grid) {
  const uint4 c = tl.comb[blockIdx.x];
  const int TU = tl.TU, TT = tl.TU * tl.TV;
  const int tv = (int)c.x / tl.ntu, tu = (int)c.x - tv * tl.ntu;
  const cpx<R>* part = static_cast<const cpx<R>*>(tl.part) + (int64_t)c.y * TT;
  for (int e = threadIdx.x; e < TT; e += 256) {
    cpx<R> sum = part[e];
    for (uint32_t p = 1; p < c.z; ++p) {
      const cpx<R> q = part[(int64_t)p * TT + e];
      sum.x += q.x;
      sum.y += q.y;
    }
    const int ly = e / TU, lx = e - ly * TU;
    grid[(int64_t)(tv * tl.TV + ly) * Nu + tu * TU + lx] = sum;
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
  const int KK = w.Ku * w.Kv;
  const dim3 blocks((unsigned)t.n_slot);
  ProfScope prof("grid_wrows", s);
  if (KK <= 64)
    hipLaunchKernelGGL((k_grid_wrows<R, 1, 8>), blocks, dim3(64), shm, s, Nu, Nv, w, t, ys, grid, scale);
  else if (KK <= 128)
    hipLaunchKernelGGL((k_grid_wrows<R, 2, 8>), blocks, dim3(64), shm, s, Nu, Nv, w, t, ys, grid, scale);
  else if (KK <= 256)
    hipLaunchKernelGGL((k_grid_wrows<R, 4, 4>), blocks, dim3(64), shm, s, Nu, Nv, w, t, ys, grid, scale);
  else if (KK <= 576)
    hipLaunchKernelGGL((k_grid_wrows<R, 9, 2>), blocks, dim3(64), shm, s, Nu, Nv, w, t, ys, grid, scale);
  else if (KK <= 1024)
    hipLaunchKernelGGL((k_grid_wrows<R, 16, 1>), blocks, dim3(64), shm, s, Nu, Nv, w, t, ys, grid, scale);
  else
    throw Error(PF_ERR_UNSUPPORTED, "w-projection rows above 1024 taps");
  PF_HIP(hipGetLastError());
  if (t.n_comb > 0) {
    hipLaunchKernelGGL((k_combine_parts<R>), dim3((unsigned)t.n_comb), dim3(256), 0, s, Nu, t, grid);
    PF_HIP(hipGetLastError());
  }
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
