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
//  * kept-entry rows: the chirp-truncated rows (K^2 <= 64) keep a small share of their K^2 entries (the
//    reference stores them as a sparse matrix, wproj_utilities.h:89-100), so they are stored as a 64-bit
//    keep mask and an offset per row, the kept values and each one's tap.  Gridding walks the same taps
//    in the same order as with dense rows and skips only zero entries (bit-identical); degridding takes
//    the kept values in stored order (equal up to the order of its additions).
#include <rocprim/device/device_scan.hpp>

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
// Kept-entry rows
// ---------------------------------------------------------------------------------------------
namespace {
// kept counts per row, and their maximum (one atomic per block after a block-wide max)
__global__ void __launch_bounds__(256) k_pack_count(const uint64_t* __restrict__ mask, int64_t M,
                                                    uint64_t* __restrict__ cnt, unsigned* __restrict__ maxc) {
  __shared__ unsigned bm;
  if (threadIdx.x == 0) bm = 0;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned c = 0;
  if (i < M) {
    c = (unsigned)__popcll(mask[i]);
    cnt[i] = c;
  }
  atomicMax(&bm, c);
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(maxc, bm);
}
__global__ void k_pack_rec(const uint64_t* __restrict__ mask, const uint64_t* __restrict__ off, int64_t M,
                           uint4* __restrict__ rec) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const uint64_t m = mask[i], o = off[i];
  rec[i] = make_uint4((uint32_t)m, (uint32_t)(m >> 32), (uint32_t)o, (uint32_t)(o >> 32));
}
__device__ inline uint64_t rec_mask(uint4 r) { return ((uint64_t)r.y << 32) | r.x; }
__device__ inline int64_t rec_off(uint4 r) { return (int64_t)(((uint64_t)r.w << 32) | r.z); }
// one thread per dense entry (row i, entry o): its position among the row's kept entries is the count of
// kept entries before it
template <typename R>
__global__ void k_pack_vals(const cpx<R>* __restrict__ dense, const uint4* __restrict__ rec, int64_t M, int KK,
                            int Ku, cpx<R>* __restrict__ packed, uint8_t* __restrict__ pos) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M * KK) return;
  const int64_t i = t / KK;
  const int o = (int)(t - i * KK);
  const uint4 r = rec[i];
  const uint64_t m = rec_mask(r);
  if ((m >> o) & 1) {
    const int64_t d = rec_off(r) + __popcll(m & ((1ull << o) - 1));
    packed[d] = dense[t];
    pos[d] = (uint8_t)(((o / Ku) << 4) | (o % Ku));
  }
}
template <typename R>
__global__ void k_unpack(const cpx<R>* __restrict__ packed, const uint4* __restrict__ rec, int64_t M, int KK,
                         cpx<R>* __restrict__ dense) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M * KK) return;
  const int64_t i = t / KK;
  const int o = (int)(t - i * KK);
  const uint4 r = rec[i];
  const uint64_t m = rec_mask(r);
  dense[t] = ((m >> o) & 1) ? packed[rec_off(r) + __popcll(m & ((1ull << o) - 1))] : mk<R>(R(0), R(0));
}
unsigned pack_blocks(int64_t n) { return (unsigned)((n + 255) / 256); }
}  // namespace

uint64_t rows_pack_records(const uint64_t* mask, int64_t M, uint4* rec, int* max_kept, hipStream_t s) {
  *max_kept = 0;
  if (M == 0) return 0;
  uint64_t* cnt = dev_alloc_n<uint64_t>((size_t)M);
  uint64_t* off = nullptr;
  unsigned* maxc = nullptr;
  void* tmp = nullptr;
  uint64_t last[2] = {0, 0};
  unsigned mx = 0;
  try {
    off = dev_alloc_n<uint64_t>((size_t)M);
    maxc = dev_alloc_n<unsigned>(1);
    PF_HIP(hipMemsetAsync(maxc, 0, sizeof(unsigned), s));
    hipLaunchKernelGGL(k_pack_count, dim3(pack_blocks(M)), dim3(256), 0, s, mask, M, cnt, maxc);
    PF_HIP(hipGetLastError());
    size_t tb = 0;
    PF_HIP(rocprim::exclusive_scan(nullptr, tb, cnt, off, (uint64_t)0, (size_t)M, rocprim::plus<uint64_t>(), s));
    tmp = dev_alloc(tb ? tb : 1);
    PF_HIP(rocprim::exclusive_scan(tmp, tb, cnt, off, (uint64_t)0, (size_t)M, rocprim::plus<uint64_t>(), s));
    hipLaunchKernelGGL(k_pack_rec, dim3(pack_blocks(M)), dim3(256), 0, s, mask, off, M, rec);
    PF_HIP(hipGetLastError());
    PF_HIP(hipMemcpyAsync(&last[0], off + (M - 1), sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    PF_HIP(hipMemcpyAsync(&last[1], cnt + (M - 1), sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    PF_HIP(hipMemcpyAsync(&mx, maxc, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    PF_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipStreamSynchronize(s);
    dev_free(cnt);
    if (off) dev_free(off);
    if (maxc) dev_free(maxc);
    if (tmp) dev_free(tmp);
    throw;
  }
  dev_free(cnt);
  dev_free(off);
  dev_free(maxc);
  dev_free(tmp);
  *max_kept = (int)mx;
  return last[0] + last[1];
}

template <typename R>
void rows_pack_values(const cpx<R>* dense, const uint4* rec, int64_t M, int KK, int Ku, cpx<R>* packed,
                      uint8_t* pos, hipStream_t s) {
  if (Ku > 15 || KK > 64) throw Error(PF_ERR_ARG, "kept-entry rows need Ku Kv <= 64");
  if (M == 0) return;
  hipLaunchKernelGGL((k_pack_vals<R>), dim3(pack_blocks(M * KK)), dim3(256), 0, s, dense, rec, M, KK, Ku, packed, pos);
  PF_HIP(hipGetLastError());
}
template <typename R>
void rows_unpack(const cpx<R>* packed, const uint4* rec, int64_t M, int KK, cpx<R>* dense, hipStream_t s) {
  if (M == 0) return;
  hipLaunchKernelGGL((k_unpack<R>), dim3(pack_blocks(M * KK)), dim3(256), 0, s, packed, rec, M, KK, dense);
  PF_HIP(hipGetLastError());
}
template void rows_pack_values<float>(const cpx<float>*, const uint4*, int64_t, int, int, cpx<float>*, uint8_t*,
                                      hipStream_t);
template void rows_pack_values<double>(const cpx<double>*, const uint4*, int64_t, int, int, cpx<double>*, uint8_t*,
                                       hipStream_t);
template void rows_unpack<float>(const cpx<float>*, const uint4*, int64_t, int, cpx<float>*, hipStream_t);
template void rows_unpack<double>(const cpx<double>*, const uint4*, int64_t, int, cpx<double>*, hipStream_t);

// ---------------------------------------------------------------------------------------------
// Device
// ---------------------------------------------------------------------------------------------
__device__ inline int wrapw(int a, int n) { return a & (n - 1); }
// c + a g with four fused multiply-adds (both degridding kernels, so that they round alike)
template <typename R>
__device__ inline cpx<R> mac(cpx<R> c, cpx<R> a, cpx<R> g) {
  return mk<R>(fmar(a.x, g.x, fmar(-a.y, g.y, c.x)), fmar(a.x, g.y, fmar(a.y, g.x, c.y)));
}
// c + conj(a) s with four fused multiply-adds (both gridding kernels, so that they round alike)
template <typename R>
__device__ inline cpx<R> mac_conj(cpx<R> c, cpx<R> a, cpx<R> s) {
  return mk<R>(fmar(s.x, a.x, fmar(s.y, a.y, c.x)), fmar(s.y, a.x, fmar(-s.x, a.y, c.y)));
}

constexpr int WD_THREADS = 256;
constexpr int WD_ROWS = 4;  // staged rows per wave per round

constexpr int WG_LANES = 8;  // lanes per visibility in degridding: 8 visibilities per wave step

// Degridding: the block stages its owner tile and the (K - 1)-wide halo; each wave then takes 8
// visibilities at a time, 8 lanes each: lane r of a group sums the taps k = r, r + 8, ... of its
// visibility's stored row (8 consecutive taps per step: one 64-byte segment of the row for complex64),
// and the group adds its 8 partial sums with a fixed butterfly.
// The block's tile and its halo (wrapping at the grid edges) into LDS, loads before stores.
template <typename R>
__device__ inline void stage_halo(cpx<R>* lds, const cpx<R>* __restrict__ grid, int Nu, int Nv, int tx0, int ty0, int W,
                                  int H) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = WD_THREADS / 64;
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
}

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
  stage_halo<R>(lds, grid, Nu, Nv, tx0, ty0, W, H);
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
        acc = mac(acc, row[k], g);
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

// Degridding with dense rows of up to 8 NT taps: k_degrid_wrows's groups and sums (lane r the taps r,
// r + 8, ..., the same multiply-adds in the same order: bit-identical), with each lane's NT taps a
// compile-time count, so that all of a step's row loads are in flight together; the steps of 8
// visibilities run as a pipeline, the next step's row segments, origins and output positions loaded
// while this step sums, in two register sets taken in turn.  Every load is unconditional (taps past the
// row read its last tap, and are not used; visibilities past the slot read the slot's last one).
template <typename R, bool PERM, int NT>
__global__ void __launch_bounds__(WD_THREADS) k_degrid_wrows_dense(int Nu, int Nv, WRows<R> w,
                                                                   const int32_t* __restrict__ perm, TileLists tl,
                                                                   const cpx<R>* __restrict__ grid,
                                                                   cpx<R>* __restrict__ y, R scl) {
  extern __shared__ unsigned char smem_raw[];
  cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
  const uint4 sl = tl.slot_own[blockIdx.x];
  const int tile = (int)sl.x;
  const int tv = tile / tl.ntu, tu = tile - tv * tl.ntu;
  const int tx0 = tu * tl.TU, ty0 = tv * tl.TV;
  const int W = tl.TU + w.Ku - 1, H = tl.TV + w.Kv - 1;
  stage_halo<R>(lds, grid, Nu, Nv, tx0, ty0, W, H);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  constexpr uint32_t STEP = (WD_THREADS / 64) * (64 / WG_LANES);
  const int grp = lane / WG_LANES, r = lane % WG_LANES;
  const uint32_t end = sl.z, last = sl.z - 1;
  uint32_t iw = sl.y + (uint32_t)wave * (64 / WG_LANES);
  if (iw >= end) return;
  const int KK = w.Ku * w.Kv;
  int ofs[NT];  // LDS offsets of this lane's taps from the footprint origin (the last tap's past the row)
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int k = min(r + WG_LANES * t, KK - 1);
    ofs[t] = (k / w.Ku) * W + (k - (k / w.Ku) * w.Ku);
  }
  struct Step {
    cpx<R> a[NT];
    int2 kw;
    uint32_t pos;
  };
  auto load = [&](uint32_t i0, Step& st) {
    const uint32_t i = min(i0 + (uint32_t)grp, last);
    const cpx<R>* row = w.vals + (int64_t)i * KK;
#pragma unroll
    for (int t = 0; t < NT; ++t) st.a[t] = row[min(r + WG_LANES * t, KK - 1)];
    st.kw = w.kw[i];
    if constexpr (PERM) st.pos = (uint32_t)perm[i];
    else st.pos = i;
  };
  auto sum = [&](uint32_t i0, const Step& st) {
    const int base = (wrapw(st.kw.y + 1, Nv) - ty0) * W + (wrapw(st.kw.x + 1, Nu) - tx0);
    cpx<R> g[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) g[t] = lds[base + ofs[t]];
    cpx<R> acc = mk<R>(R(0), R(0));
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const cpx<R> m = mac(acc, st.a[t], g[t]);
      const bool on = r + WG_LANES * t < KK;
      acc.x = on ? m.x : acc.x;  // per component: a select of the pair would go through memory
      acc.y = on ? m.y : acc.y;
    }
#pragma unroll
    for (int m = WG_LANES / 2; m >= 1; m >>= 1) {
      acc.x += __shfl_xor(acc.x, m, 64);
      acc.y += __shfl_xor(acc.y, m, 64);
    }
    if (r == 0 && i0 + (uint32_t)grp < end) y[st.pos] = scale(acc, scl);
  };
  Step sA, sB;
  load(iw, sA);
  for (;;) {
    load(iw + STEP, sB);
    sum(iw, sA);
    iw += STEP;
    if (iw >= end) break;
    load(iw + STEP, sA);
    sum(iw, sB);
    iw += STEP;
    if (iw >= end) break;
  }
}

// Degridding with kept-entry rows (K^2 <= 64): 8 lanes per visibility as in k_degrid_wrows, but the
// lanes take the row's kept values in their stored order: lane r the SL consecutive slots from SL r (SL
// per lane: 2, 4 or 8, the fewest that hold every row's kept taps), each with its tap's position byte, so
// that each load instruction reads 8 short contiguous runs instead of one tap of 8 rows.  The group's 8
// partial sums meet in a fixed butterfly: deterministic, and equal to the dense rows' sums up to the
// order of the additions.  A visibility's slots wait on two dependent loads (its record, then its
// values), so the wave walks its steps of 8 visibilities as a pipeline: the records two steps ahead, the
// values one step ahead, in two register sets taken in turn (every load unconditional: the record
// clamped to the slot's last visibility, and a slot past the row reads the next row's value or a spare
// entry, and is not used).  PERM: the output position in the caller's order is loaded with the record (a
// load issued at the store would wait for every load in flight).
template <typename R, bool PERM, int SL>
__global__ void __launch_bounds__(WD_THREADS) k_degrid_wrows_packed(int Nu, int Nv, WRows<R> w,
                                                                    const int32_t* __restrict__ perm, TileLists tl,
                                                                    const cpx<R>* __restrict__ grid,
                                                                    cpx<R>* __restrict__ y, R scl) {
  extern __shared__ unsigned char smem_raw[];
  cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
  const uint4 sl = tl.slot_own[blockIdx.x];
  const int tile = (int)sl.x;
  const int tv = tile / tl.ntu, tu = tile - tv * tl.ntu;
  const int tx0 = tu * tl.TU, ty0 = tv * tl.TV;
  const int W = tl.TU + w.Ku - 1, H = tl.TV + w.Kv - 1;
  stage_halo<R>(lds, grid, Nu, Nv, tx0, ty0, W, H);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  constexpr uint32_t STEP = (WD_THREADS / 64) * (64 / WG_LANES);  // visibilities per block step
  const int grp = lane / WG_LANES, r = lane % WG_LANES;
  const uint32_t end = sl.z, last = sl.z - 1;
  uint32_t iw = sl.y + (uint32_t)wave * (64 / WG_LANES);  // the wave's first visibility of the step
  if (iw >= end) return;
  struct Slots {
    cpx<R> a[SL];
    uint8_t p[SL];
  };
  auto records = [&](uint32_t i0, uint4& rc, int2& kw, uint32_t& pos) {
    const uint32_t i = min(i0 + (uint32_t)grp, last);
    kw = w.kw[i];
    rc = w.rec[i];
    if constexpr (PERM) pos = (uint32_t)perm[i];
    else pos = i;
  };
  auto values = [&](const uint4& rc, Slots& v) {
    const int64_t j = rec_off(rc) + SL * r;
#pragma unroll
    for (int k = 0; k < SL; ++k) v.a[k] = w.vals[j + k];
#pragma unroll
    for (int k = 0; k < SL; ++k) v.p[k] = w.pos[j + k];
  };
  auto sum = [&](uint32_t i0, const uint4& rc, int base, uint32_t pos, const Slots& v) {
    const int nnz = __popc(rc.x) + __popc(rc.y);
    cpx<R> g[SL];
#pragma unroll
    for (int k = 0; k < SL; ++k) g[k] = lds[base + (int)(v.p[k] >> 4) * W + (int)(v.p[k] & 15u)];
    cpx<R> acc = mk<R>(R(0), R(0));
#pragma unroll
    for (int k = 0; k < SL; ++k) {
      const cpx<R> t = mac(acc, v.a[k], g[k]);
      const bool on = SL * r + k < nnz;
      acc.x = on ? t.x : acc.x;  // per component: a select of the pair would go through memory
      acc.y = on ? t.y : acc.y;
    }
#pragma unroll
    for (int m = WG_LANES / 2; m >= 1; m >>= 1) {
      acc.x += __shfl_xor(acc.x, m, 64);
      acc.y += __shfl_xor(acc.y, m, 64);
    }
    if (r == 0 && i0 + (uint32_t)grp < end) y[pos] = scale(acc, scl);
  };
  auto base_of = [&](int2 k0) { return (wrapw(k0.y + 1, Nv) - ty0) * W + (wrapw(k0.x + 1, Nu) - tx0); };
  uint4 rA, rB;
  int2 kA, kB;
  uint32_t pA, pB;
  Slots aA, aB;
  records(iw, rA, kA, pA);
  records(iw + STEP, rB, kB, pB);
  values(rA, aA);
  for (;;) {
    {  // step in set A: its record, origin and position taken before the set is reloaded two steps ahead
      const uint4 rc = rA;
      const uint32_t pos = pA;
      const int base = base_of(kA);
      records(iw + 2 * STEP, rA, kA, pA);
      values(rB, aB);
      sum(iw, rc, base, pos, aA);
    }
    iw += STEP;
    if (iw >= end) break;
    {  // the same in set B
      const uint4 rc = rB;
      const uint32_t pos = pB;
      const int base = base_of(kB);
      records(iw + 2 * STEP, rB, kB, pB);
      values(rA, aA);
      sum(iw, rc, base, pos, aB);
    }
    iw += STEP;
    if (iw >= end) break;
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
            cpx<R>& c = acc[yy * TU + x];
            c = mac_conj(c, a[t], s);
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

// Gridding with kept-entry rows (K^2 <= 64): as k_grid_wrows with one tap per lane, the lane's tap of
// an entry held only where the entry's mask keeps it; its value is the kept entry numbered by the mask
// bits below the lane (mbcnt).  The kept taps are added in the dense kernel's order with the same
// multiply-adds and the zeros are skipped, so the grid equals the dense rows' bit for bit.
// The kernel is bound by vector instruction issue, so the per-entry work is kept small: the mask itself
// is the lane predicate (inverse ballot: no vector compare), the value is read at a uniform base plus the
// lane's mbcnt (a non-kept lane reads at most one past the row's last kept value, which exists: the kept
// values carry one spare entry), the entry's footprint origin relative to the tile (wrapped, clamped to
// 16 bits) is computed once per batch, and the LDS cell is a uniform offset plus the lane's own.
// The list is walked in batches of 64 entries (one per lane: index, origin, value and record), each
// batch fully unrolled so that every load is unconditional (clamped to the tile's last entry) and the
// waits are counted statically: the D-deep ring of tap values runs across the batch boundary, and the
// next batch's records are requested a quarter of the way through the current one.  Entries are
// broadcast from their lane with readlane (the batch position is a compile-time constant).
__device__ inline int rl(int v, int j) { return __builtin_amdgcn_readlane(v, j); }
__device__ inline uint32_t rl(uint32_t v, int j) { return (uint32_t)__builtin_amdgcn_readlane((int)v, j); }
__device__ inline float rl(float v, int j) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j)); }
__device__ inline double rl(double v, int j) {
  const long long b = __double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, j);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), j);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

template <typename R>
struct GEnt {
  uint32_t xy;  // footprint origin relative to the tile: low 16 bits x, high 16 bits y (signed)
  cpx<R> s;     // the entry's visibility, scaled
  uint4 r;      // its record
};

template <typename R, int D, bool OFF64>
__global__ void __launch_bounds__(64) k_grid_wrows_packed(int Nu, int Nv, WRows<R> w, TileLists tl,
                                                          const cpx<R>* __restrict__ ys, cpx<R>* __restrict__ grid,
                                                          R scl) {
  static_assert(64 % D == 0 && D <= 32, "the ring refills from the next batch after 64 - D entries");
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
  const uint32_t beg = sl.y, end = sl.z;
  if (beg < end) {
    // this lane's tap: row and column offsets (lanes past the row have no mask bit) and its LDS offset
    // from the footprint origin's cell
    const int dv = lane / w.Ku, du = lane - (lane / w.Ku) * w.Ku;
    const int lofs = (dv * TU + du) * (int)sizeof(cpx<R>);
    auto rel16 = [](int a, int n) {  // a wrapped into (-n/2, n/2], clamped to 16 bits
      const int r = ((a + n / 2) & (n - 1)) - n / 2;
      return (uint32_t)(uint16_t)(int16_t)max(-32768, min(32767, r));
    };
    auto load_ent = [&](uint32_t i) {
      GEnt<R> e;
      const int2 k = w.kw[i];
      e.xy = rel16(k.x + 1 - tx0, Nu) | (rel16(k.y + 1 - ty0, Nv) << 16);
      e.s = scale(ys[i], scl);
      e.r = w.rec[i];
      return e;
    };
    // batch entry j's tap of this lane: the entry's mask and this lane's value (read for every lane)
    auto fetch = [&](const GEnt<R>& e, int j, uint64_t& m, cpx<R>& val) {
      const uint32_t mlo = rl(e.r.x, j), mhi = rl(e.r.y, j);
      m = ((uint64_t)mhi << 32) | mlo;
      uint64_t off = rl(e.r.z, j);
      if constexpr (OFF64) off |= (uint64_t)rl(e.r.w, j) << 32;
      const uint32_t idx = __builtin_amdgcn_mbcnt_hi(mhi, __builtin_amdgcn_mbcnt_lo(mlo, 0u));
      val = (w.vals + off)[idx];
    };
    GEnt<R> cur = load_ent(tl.list[min(beg + (uint32_t)lane, end - 1)]);
    uint64_t rm[D];
    cpx<R> ring[D];
#pragma unroll
    for (int d = 0; d < D; ++d) fetch(cur, d, rm[d], ring[d]);
    for (uint32_t b0 = beg; b0 < end; b0 += 64) {
      // the next batch (or this one again after the last: loaded, never used)
      const uint32_t nb = b0 + 64 < end ? b0 + 64 : b0;
      const uint32_t nidx = tl.list[min(nb + (uint32_t)lane, end - 1)];
      GEnt<R> nxt;
#pragma unroll
      for (int j = 0; j < 64; ++j) {
        if (j == 16) nxt = load_ent(nidx);
        const cpx<R> a = ring[j % D];
        const uint64_t ma = b0 + (uint32_t)j < end ? rm[j % D] : 0ull;
        if (j + D < 64) fetch(cur, j + D, rm[j % D], ring[j % D]);
        else fetch(nxt, j + D - 64, rm[j % D], ring[j % D]);
        const uint32_t xy = rl(cur.xy, j);
        const int xs = (int)(int16_t)(xy & 0xffffu), ysh = (int)(int16_t)(xy >> 16);
        const cpx<R> sv = mk<R>(rl(cur.s.x, j), rl(cur.s.y, j));
        if (__builtin_amdgcn_inverse_ballot_w64(ma) && (unsigned)(xs + du) < (unsigned)TU &&
            (unsigned)(ysh + dv) < (unsigned)TV) {
          cpx<R>* c = reinterpret_cast<cpx<R>*>(smem_raw + (ysh * TU + xs) * (int)sizeof(cpx<R>) + lofs);
          *c = mac_conj(*c, a, sv);
        }
      }
      cur = nxt;
    }
  }
  __syncthreads();
  if (sl.w == PART_NONE) {
    for (int e = lane; e < TU * TV; e += 64) {
      const int ly = e / TU, lx = e - ly * TU;
      grid[(int64_t)(ty0 + ly) * Nu + tx0 + lx] = acc[e];
    }
  } else {
    cpx<R>* part = static_cast<cpx<R>*>(tl.part) + (int64_t)sl.w * TU * TV;
    for (int e = lane; e < TU * TV; e += 64) part[e] = acc[e];
  }
}

// Gridding with kept-entry rows, QE entries per step (2 or 4, while no row keeps more than 64 / QE taps):
// the lanes take the kept values in stored order, 64 / QE lanes per entry (lane k of entry h the slot k,
// with its tap byte), so that a step spends its lanes on kept taps only.  Entries of one step may share
// cells, so entry h of each step adds into its own accumulator copy h of the tile; the copies are summed
// in order at the end: deterministic, and equal to the one-entry kernel's sums up to the order of the
// additions.  The batch's entries (index, origin, value, record) are loaded one per lane as in
// k_grid_wrows_packed and handed to the step's lanes by ds_bpermute; the slots' values and tap bytes run
// DS steps ahead in a ring across the batch boundary, every load unconditional (a slot past the row reads
// the next rows' values or the spare entries, and is not used).
template <typename R, int QE, bool OFF64>
__global__ void __launch_bounds__(64) k_grid_wrows_kept(int Nu, int Nv, WRows<R> w, TileLists tl,
                                                        const cpx<R>* __restrict__ ys, cpx<R>* __restrict__ grid,
                                                        R scl) {
  constexpr int LPE = 64 / QE;  // lanes (slots) per entry
  constexpr int SPB = 64 / QE;  // steps per 64-entry batch
  constexpr int DS = 16 / QE;   // steps in flight
  static_assert(QE == 2 || QE == 4, "two or four entries per step");
  extern __shared__ unsigned char smem_raw[];
  cpx<R>* acc = reinterpret_cast<cpx<R>*>(smem_raw);  // QE copies of the tile
  const uint4 sl = tl.slot[blockIdx.x];
  const int tile = (int)sl.x;
  const int TU = tl.TU, TV = tl.TV, TT = TU * TV;
  const int tv = tile / tl.ntu, tu = tile - tv * tl.ntu;
  const int tx0 = tu * TU, ty0 = tv * TV;
  const int lane = threadIdx.x;
  for (int e = lane; e < QE * TT; e += 64) acc[e] = mk<R>(R(0), R(0));
  __syncthreads();
  const uint32_t beg = sl.y, end = sl.z;
  if (beg < end) {
    const int h = lane / LPE, k = lane % LPE;  // this lane's entry of a step and its slot
    cpx<R>* my_acc = acc + h * TT;
    auto rel16 = [](int a, int n) {  // a wrapped into (-n/2, n/2], clamped to 16 bits
      const int r = ((a + n / 2) & (n - 1)) - n / 2;
      return (uint32_t)(uint16_t)(int16_t)max(-32768, min(32767, r));
    };
    struct Ent {
      uint32_t xy, nnz, olo, ohi;
      cpx<R> s;
    };
    auto load_ent = [&](uint32_t i) {
      Ent e;
      const int2 kw = w.kw[i];
      e.xy = rel16(kw.x + 1 - tx0, Nu) | (rel16(kw.y + 1 - ty0, Nv) << 16);
      e.s = scale(ys[i], scl);
      const uint4 r = w.rec[i];
      e.nnz = (uint32_t)(__popc(r.x) + __popc(r.y));
      e.olo = r.z;
      e.ohi = r.w;
      return e;
    };
    auto bp = [](uint32_t v, int src) { return (uint32_t)__shfl((int)v, src, 64); };
    struct Slot {
      cpx<R> a;
      uint32_t p;   // tap byte
      uint32_t on;  // the slot holds a kept tap
    };
    // step t of a batch: this lane's slot of entry QE t + h
    auto fetch = [&](const Ent& e, int t, Slot& sv) {
      const int src = QE * t + h;
      uint64_t off = bp(e.olo, src);
      if constexpr (OFF64) off |= (uint64_t)bp(e.ohi, src) << 32;
      sv.on = (uint32_t)k < bp(e.nnz, src);
      sv.a = w.vals[off + (uint32_t)k];
      sv.p = w.pos[off + (uint32_t)k];
    };
    Ent cur = load_ent(tl.list[min(beg + (uint32_t)lane, end - 1)]);
    Slot ring[DS];
#pragma unroll
    for (int d = 0; d < DS; ++d) fetch(cur, d, ring[d]);
    for (uint32_t b0 = beg; b0 < end; b0 += 64) {
      // the next batch (or this one again after the last: loaded, never used)
      const uint32_t nb = b0 + 64 < end ? b0 + 64 : b0;
      const uint32_t nidx = tl.list[min(nb + (uint32_t)lane, end - 1)];
      Ent nxt;
#pragma unroll
      for (int t = 0; t < SPB; ++t) {
        if (t == SPB / 4) nxt = load_ent(nidx);
        const Slot sv = ring[t % DS];
        if (t + DS < SPB) fetch(cur, t + DS, ring[t % DS]);
        else fetch(nxt, t + DS - SPB, ring[t % DS]);
        const int src = QE * t + h;
        const uint32_t xy = bp(cur.xy, src);
        const cpx<R> s = mk<R>(__shfl(cur.s.x, src, 64), __shfl(cur.s.y, src, 64));
        const int x = (int)(int16_t)(xy & 0xffffu) + (int)(sv.p & 15u);
        const int y = (int)(int16_t)(xy >> 16) + (int)(sv.p >> 4);
        if (sv.on && b0 + (uint32_t)src < end && (unsigned)x < (unsigned)TU && (unsigned)y < (unsigned)TV) {
          cpx<R>& c = my_acc[y * TU + x];
          c = mac_conj(c, sv.a, s);
        }
      }
      cur = nxt;
    }
  }
  __syncthreads();
  // the copies summed in order, into the grid or this part's partial sums
  cpx<R>* out = sl.w == PART_NONE ? nullptr : static_cast<cpx<R>*>(tl.part) + (int64_t)sl.w * TT;
  for (int e = lane; e < TT; e += 64) {
    cpx<R> v = acc[e];
#pragma unroll
    for (int q = 1; q < QE; ++q) v = v + acc[q * TT + e];
    if (out) {
      out[e] = v;
    } else {
      const int ly = e / TU, lx = e - ly * TU;
      grid[(int64_t)(ty0 + ly) * Nu + tx0 + lx] = v;
    }
  }
}

// Gridding with dense rows of up to 64 TPL taps: k_grid_wrows's walk (one wave per tile slot, lane l the
// taps l + 64 t, the same multiply-adds in list order: bit-identical) in k_grid_wrows_packed's form: each
// 64-entry batch fully unrolled with every load unconditional (taps past the row read its last tap, and
// are not used), a D-entry ring of row segments running across the batch boundary, the next batch's
// records requested a quarter of the way in, entries broadcast by readlane, and the footprint origin
// relative to the tile computed once per batch.
template <typename R, int D, int TPL>
__global__ void __launch_bounds__(64) k_grid_wrows_dense(int Nu, int Nv, WRows<R> w, TileLists tl,
                                                         const cpx<R>* __restrict__ ys, cpx<R>* __restrict__ grid,
                                                         R scl) {
  static_assert(64 % D == 0 && D <= 32, "the ring refills from the next batch after 64 - D entries");
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
  const uint32_t beg = sl.y, end = sl.z;
  if (beg < end) {
    const int KK = w.Ku * w.Kv;
    // this lane's taps: row and column offsets, LDS offsets from the footprint origin's cell, and the
    // index it reads (the row's last tap past the row)
    int dv[TPL], du[TPL], lofs[TPL], kix[TPL];
    bool has[TPL];
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
      const int k = lane + 64 * t;
      has[t] = k < KK;
      kix[t] = min(k, KK - 1);
      dv[t] = kix[t] / w.Ku;
      du[t] = kix[t] - dv[t] * w.Ku;
      lofs[t] = (dv[t] * TU + du[t]) * (int)sizeof(cpx<R>);
    }
    auto rel16 = [](int a, int n) {  // a wrapped into (-n/2, n/2], clamped to 16 bits
      const int r = ((a + n / 2) & (n - 1)) - n / 2;
      return (uint32_t)(uint16_t)(int16_t)max(-32768, min(32767, r));
    };
    struct Ent {
      uint32_t i, xy;
      cpx<R> s;
    };
    auto load_ent = [&](uint32_t i) {
      Ent e;
      const int2 k = w.kw[i];
      e.i = i;
      e.xy = rel16(k.x + 1 - tx0, Nu) | (rel16(k.y + 1 - ty0, Nv) << 16);
      e.s = scale(ys[i], scl);
      return e;
    };
    auto fetch = [&](const Ent& e, int j, cpx<R> (&val)[TPL]) {
      const cpx<R>* row = w.vals + (int64_t)rl(e.i, j) * KK;
#pragma unroll
      for (int t = 0; t < TPL; ++t) val[t] = row[kix[t]];
    };
    Ent cur = load_ent(tl.list[min(beg + (uint32_t)lane, end - 1)]);
    cpx<R> ring[D][TPL];
#pragma unroll
    for (int d = 0; d < D; ++d) fetch(cur, d, ring[d]);
    for (uint32_t b0 = beg; b0 < end; b0 += 64) {
      const uint32_t nb = b0 + 64 < end ? b0 + 64 : b0;
      const uint32_t nidx = tl.list[min(nb + (uint32_t)lane, end - 1)];
      Ent nxt;
#pragma unroll
      for (int j = 0; j < 64; ++j) {
        if (j == 16) nxt = load_ent(nidx);
        cpx<R> a[TPL];
#pragma unroll
        for (int t = 0; t < TPL; ++t) a[t] = ring[j % D][t];
        if (j + D < 64) fetch(cur, j + D, ring[j % D]);
        else fetch(nxt, j + D - 64, ring[j % D]);
        const bool live = b0 + (uint32_t)j < end;
        const uint32_t xy = rl(cur.xy, j);
        const int xs = (int)(int16_t)(xy & 0xffffu), ysh = (int)(int16_t)(xy >> 16);
        const cpx<R> sv = mk<R>(rl(cur.s.x, j), rl(cur.s.y, j));
#pragma unroll
        for (int t = 0; t < TPL; ++t) {
          if (live && has[t] && (unsigned)(xs + du[t]) < (unsigned)TU && (unsigned)(ysh + dv[t]) < (unsigned)TV) {
            cpx<R>* c = reinterpret_cast<cpx<R>*>(smem_raw + (ysh * TU + xs) * (int)sizeof(cpx<R>) + lofs[t]);
            *c = mac_conj(*c, a[t], sv);
          }
        }
      }
      cur = nxt;
    }
  }
  __syncthreads();
  if (sl.w == PART_NONE) {
    for (int e = lane; e < TU * TV; e += 64) {
      const int ly = e / TU, lx = e - ly * TU;
      grid[(int64_t)(ty0 + ly) * Nu + tx0 + lx] = acc[e];
    }
  } else {
    cpx<R>* part = static_cast<cpx<R>*>(tl.part) + (int64_t)sl.w * TU * TV;
    for (int e = lane; e < TU * TV; e += 64) part[e] = acc[e];
  }
}

// Gridding with dense rows by owner tile: one wave per degridding slot (a chunk of the visibilities whose
// first tap lies in its tile, in their sorted order), each row read once and summed into the tile extended
// by K - 1 columns and rows in LDS, with k_grid_wrows_dense's walk (lane l the taps l + 64 t, the rows of a
// D-entry ring requested ahead, entries broadcast by readlane); the extended tile goes to the slot's
// partial sums for k_combine_own.  A list-based walk reads a row once per tile its footprint touches
// (1.72 times for 11 x 11 rows in 32 x 32 tiles); this reads it once.
template <typename R, int D, int TPL>
__global__ void __launch_bounds__(64) k_grid_wrows_own(int Nu, int Nv, WRows<R> w, TileLists tl,
                                                       const cpx<R>* __restrict__ ys, R scl) {
  static_assert(64 % D == 0 && D <= 32, "the ring refills from the next batch after 64 - D entries");
  extern __shared__ unsigned char smem_raw[];
  cpx<R>* acc = reinterpret_cast<cpx<R>*>(smem_raw);
  const uint4 sl = tl.slot_own[blockIdx.x];
  const int tile = (int)sl.x;
  const int EW = tl.ext_w, EE = tl.ext_w * tl.ext_h;
  const int tv = tile / tl.ntu, tu = tile - tv * tl.ntu;
  const int tx0 = tu * tl.TU, ty0 = tv * tl.TV;
  const int lane = threadIdx.x;
  for (int e = lane; e < EE; e += 64) acc[e] = mk<R>(R(0), R(0));
  __syncthreads();
  const uint32_t beg = sl.y, end = sl.z;
  if (beg < end) {
    const int KK = w.Ku * w.Kv;
    int lofs[TPL], kix[TPL];
    bool has[TPL];
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
      const int k = lane + 64 * t;
      has[t] = k < KK;
      kix[t] = min(k, KK - 1);
      const int dv = kix[t] / w.Ku, du = kix[t] - dv * w.Ku;
      lofs[t] = (dv * EW + du) * (int)sizeof(cpx<R>);
    }
    struct Ent {
      uint32_t i, off;  // sorted index, LDS byte offset of the footprint origin in the extended tile
      cpx<R> s;
    };
    auto load_ent = [&](uint32_t i) {
      Ent e;
      const int2 k = w.kw[i];
      // the first tap lies in the tile: origin offsets in [0, TU) x [0, TV)
      const int dx = (k.x + 1 - tx0) & (Nu - 1), dy = (k.y + 1 - ty0) & (Nv - 1);
      e.i = i;
      e.off = (uint32_t)((dy * EW + dx) * (int)sizeof(cpx<R>));
      e.s = scale(ys[i], scl);
      return e;
    };
    auto fetch = [&](const Ent& e, int j, cpx<R> (&val)[TPL]) {
      const cpx<R>* row = w.vals + (int64_t)rl(e.i, j) * KK;
#pragma unroll
      for (int t = 0; t < TPL; ++t) val[t] = row[kix[t]];
    };
    Ent cur = load_ent(min(beg + (uint32_t)lane, end - 1));
    cpx<R> ring[D][TPL];
#pragma unroll
    for (int d = 0; d < D; ++d) fetch(cur, d, ring[d]);
    for (uint32_t b0 = beg; b0 < end; b0 += 64) {
      const uint32_t nb = b0 + 64 < end ? b0 + 64 : b0;
      const uint32_t nidx = min(nb + (uint32_t)lane, end - 1);
      Ent nxt;
#pragma unroll
      for (int j = 0; j < 64; ++j) {
        if (j == 16) nxt = load_ent(nidx);
        cpx<R> a[TPL];
#pragma unroll
        for (int t = 0; t < TPL; ++t) a[t] = ring[j % D][t];
        if (j + D < 64) fetch(cur, j + D, ring[j % D]);
        else fetch(nxt, j + D - 64, ring[j % D]);
        const bool live = b0 + (uint32_t)j < end;
        const int off = (int)rl(cur.off, j);
        const cpx<R> sv = mk<R>(rl(cur.s.x, j), rl(cur.s.y, j));
#pragma unroll
        for (int t = 0; t < TPL; ++t) {
          if (live && has[t]) {
            cpx<R>* c = reinterpret_cast<cpx<R>*>(smem_raw + off + lofs[t]);
            *c = mac_conj(*c, a[t], sv);
          }
        }
      }
      cur = nxt;
    }
  }
  __syncthreads();
  cpx<R>* out = static_cast<cpx<R>*>(tl.ext) + (int64_t)blockIdx.x * EE;
  for (int e = lane; e < EE; e += 64) out[e] = acc[e];
}

// Each grid cell of the owner-tile gridding: its tile's slots (in the order of their visibilities), then
// the extensions reaching it from the left, upper and upper-left tiles (periodic grid), in that order,
// so the result is the same on every run.  One block per tile, so that the four tiles' slot ranges are
// block-uniform and each cell's partial sums are independent loads.
template <typename R>
__global__ void __launch_bounds__(256) k_combine_own(int Nu, int Ku, int Kv, TileLists tl, cpx<R>* __restrict__ grid) {
  const int t = (int)blockIdx.x;
  const int TU = tl.TU, TV = tl.TV, EW = tl.ext_w, EE = tl.ext_w * tl.ext_h;
  const int tv = t / tl.ntu, tu = t - tv * tl.ntu;
  const int lu = tu > 0 ? tu - 1 : tl.ntu - 1, uv = tv > 0 ? tv - 1 : tl.ntv - 1;
  const int tl_ = tv * tl.ntu + lu, tup = uv * tl.ntu + tu, tul = uv * tl.ntu + lu;
  const cpx<R>* ext = static_cast<const cpx<R>*>(tl.ext);
  auto add = [&](cpx<R>& sum, int tile, int off) {
    for (int k = tl.own_first[tile]; k < tl.own_first[tile + 1]; ++k) {
      const cpx<R> v = ext[(int64_t)tl.own_ids[k] * EE + off];
      sum.x += v.x;
      sum.y += v.y;
    }
  };
  for (int e = threadIdx.x; e < TU * TV; e += 256) {
    const int ly = e / TU, lx = e - ly * TU;
    cpx<R> sum = mk<R>(R(0), R(0));
    add(sum, t, ly * EW + lx);
    if (lx < Ku - 1) add(sum, tl_, ly * EW + lx + TU);
    if (ly < Kv - 1) add(sum, tup, (ly + TV) * EW + lx);
    if (lx < Ku - 1 && ly < Kv - 1) add(sum, tul, (ly + TV) * EW + lx + TU);
    grid[(int64_t)(tv * TV + ly) * Nu + tu * TU + lx] = sum;
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

// A launch with dynamic LDS beyond the default 64 KiB allowed for the kernel first.
template <typename... P, typename... A>
static void launch_lds(void (*kernel)(P...), dim3 g, dim3 b, size_t shm, hipStream_t s, A... args) {
  allow_lds((const void*)kernel, shm);
  hipLaunchKernelGGL(kernel, g, b, shm, s, args...);
}

template <typename R>
void degrid_wrows(int Nu, int Nv, const WRows<R>& w, const int32_t* perm, const TileLists& t, const cpx<R>* grid,
                  cpx<R>* y, R scale, hipStream_t s) {
  if (t.n_own == 0) return;
  // the staged tile and halo: up to the CU's 160 KiB of LDS
  const size_t shm = sizeof(cpx<R>) * (size_t)(t.TU + w.Ku - 1) * (t.TV + w.Kv - 1);
  if (shm > 160 * 1024) throw Error(PF_ERR_UNSUPPORTED, "w-projection degridding tile and halo exceed 160 KiB of LDS");
  ProfScope prof("degrid_wrows", s);
  if (w.rec) {
    if (w.Ku * w.Kv > 64) throw Error(PF_ERR_ARG, "kept-entry rows need Ku Kv <= 64");
#define PF_DWP(PP, SS) \
  launch_lds(k_degrid_wrows_packed<R, PP, SS>, dim3((unsigned)t.n_own), dim3(WD_THREADS), shm, s, Nu, Nv, w, perm, t, \
             grid, y, scale)
    if (w.slots == 2) {
      if (perm) PF_DWP(true, 2);
      else PF_DWP(false, 2);
    } else if (w.slots == 4) {
      if (perm) PF_DWP(true, 4);
      else PF_DWP(false, 4);
    } else if (w.slots == 8) {
      if (perm) PF_DWP(true, 8);
      else PF_DWP(false, 8);
    } else {
      throw Error(PF_ERR_ARG, "kept-entry degridding: 2, 4 or 8 slots per lane");
    }
#undef PF_DWP
  } else {
    const int KK = w.Ku * w.Kv;
#define PF_DWD(PP, NN) \
  launch_lds(k_degrid_wrows_dense<R, PP, NN>, dim3((unsigned)t.n_own), dim3(WD_THREADS), shm, s, Nu, Nv, w, perm, t, \
             grid, y, scale)
    if (KK <= 64) {
      if (perm) PF_DWD(true, 8);
      else PF_DWD(false, 8);
    } else if (KK <= 128) {
      if (perm) PF_DWD(true, 16);
      else PF_DWD(false, 16);
    } else if (KK <= 192) {
      if (perm) PF_DWD(true, 24);
      else PF_DWD(false, 24);
    } else {
      launch_lds(k_degrid_wrows<R>, dim3((unsigned)t.n_own), dim3(WD_THREADS), shm, s, Nu, Nv, w, perm, t, grid, y,
                 scale);
    }
#undef PF_DWD
  }
  PF_HIP(hipGetLastError());
}

template <typename R>
void grid_wrows(int Nu, int Nv, const WRows<R>& w, const TileLists& t, const cpx<R>* ys, cpx<R>* grid, R scale,
                hipStream_t s) {
  // the tile's accumulator: up to the CU's 160 KiB of LDS
  const size_t shm = sizeof(cpx<R>) * (size_t)t.TU * t.TV;
  if (shm > 160 * 1024) throw Error(PF_ERR_UNSUPPORTED, "w-projection gridding tile exceeds 160 KiB of LDS");
  const int KK = w.Ku * w.Kv;
  const dim3 blocks((unsigned)t.n_slot);
  ProfScope prof("grid_wrows", s);
  if (w.rec && w.qe > 1) {
    if (KK > 64) throw Error(PF_ERR_ARG, "kept-entry rows need Ku Kv <= 64");
    const size_t shq = shm * (size_t)w.qe;
    if (shq > 160 * 1024) throw Error(PF_ERR_UNSUPPORTED, "kept-entry gridding copies exceed 160 KiB of LDS");
#define PF_GWK(QQ, O64) launch_lds(k_grid_wrows_kept<R, QQ, O64>, blocks, dim3(64), shq, s, Nu, Nv, w, t, ys, grid, scale)
    if (w.qe == 4) {
      if (w.off64) PF_GWK(4, true);
      else PF_GWK(4, false);
    } else {
      if (w.off64) PF_GWK(2, true);
      else PF_GWK(2, false);
    }
#undef PF_GWK
  } else if (w.rec) {
    if (KK > 64) throw Error(PF_ERR_ARG, "kept-entry rows need Ku Kv <= 64");
#define PF_GWP(DD, O64) launch_lds(k_grid_wrows_packed<R, DD, O64>, blocks, dim3(64), shm, s, Nu, Nv, w, t, ys, grid, scale)
    if (w.off64) {
      if (w.depth == 32) PF_GWP(32, true);
      else if (w.depth == 8) PF_GWP(8, true);
      else PF_GWP(16, true);
    } else {
      if (w.depth == 32) PF_GWP(32, false);
      else if (w.depth == 8) PF_GWP(8, false);
      else PF_GWP(16, false);
    }
#undef PF_GWP
  } else if (w.own && t.ext && KK > 64 && KK <= 576) {
    // dense rows by owner tile, then the combine pass
    const size_t she = sizeof(cpx<R>) * (size_t)t.ext_w * t.ext_h;
    if (she > 160 * 1024) throw Error(PF_ERR_UNSUPPORTED, "owner-tile gridding tile exceeds 160 KiB of LDS");
    const dim3 ob((unsigned)t.n_own);
    constexpr bool F = sizeof(R) == 4;
    if (t.n_own > 0) {
      if (KK <= 128) launch_lds(k_grid_wrows_own<R, F ? 16 : 8, 2>, ob, dim3(64), she, s, Nu, Nv, w, t, ys, scale);
      else if (KK <= 256) launch_lds(k_grid_wrows_own<R, F ? 8 : 4, 4>, ob, dim3(64), she, s, Nu, Nv, w, t, ys, scale);
      else launch_lds(k_grid_wrows_own<R, F ? 4 : 2, 9>, ob, dim3(64), she, s, Nu, Nv, w, t, ys, scale);
    }
    hipLaunchKernelGGL(k_combine_own<R>, dim3((unsigned)(t.ntu * t.ntv)), dim3(256), 0, s, Nu, w.Ku, w.Kv, t, grid);
    PF_HIP(hipGetLastError());
    return;
  } else if (KK <= 64)
    launch_lds(k_grid_wrows_dense<R, 16, 1>, blocks, dim3(64), shm, s, Nu, Nv, w, t, ys, grid, scale);
  else if (KK <= 128)
    launch_lds(k_grid_wrows_dense<R, sizeof(R) == 4 ? 16 : 8, 2>, blocks, dim3(64), shm, s, Nu, Nv, w, t, ys, grid,
               scale);
  else if (KK <= 256)
    launch_lds(k_grid_wrows<R, 4, 4>, blocks, dim3(64), shm, s, Nu, Nv, w, t, ys, grid, scale);
  else if (KK <= 576)
    launch_lds(k_grid_wrows<R, 9, 2>, blocks, dim3(64), shm, s, Nu, Nv, w, t, ys, grid, scale);
  else if (KK <= 1024)
    launch_lds(k_grid_wrows<R, 16, 1>, blocks, dim3(64), shm, s, Nu, Nv, w, t, ys, grid, scale);
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
