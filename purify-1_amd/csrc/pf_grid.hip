// pf_grid.hip -- the gridding matrix G (degridding gather) and G^H (gridding scatter), evaluated on the
// fly from the visibility coordinates instead of from a stored CSR matrix
// (replaces details::init_gridding_matrix_2d, cpp/purify/operators.cc:8-46, and the CSR SpMV
// utilities::sparse_multiply_matrix, cpp/purify/utilities.h:144-159).
//
// Row m of G has Ju x Jv taps at columns (k_u + ju, k_v + jv) mod (Nu, Nv), ju, jv = 1..J,
// k = floor(u - J/2), value (-1)^(k_u+ju+k_v+jv) * K_u(u - k_u - ju) * K_v(v - k_v - jv) * weight_m.
// The 1-D kernel factors are evaluated once per visibility (2J evaluations, not J^2).
//
// Visibilities are sorted at construction by the grid tile that holds their first tap.
//  * degrid: one lane per sorted visibility, so a wave's J x J reads fall in one or two tiles and hit
//    the cache; the result goes to y[perm[i]] in the caller's order.
//  * grid: one wave per grid tile (tile-owner).  The tile's list holds every visibility with a tap
//    inside the tile, in ascending sorted order.  The wave accumulates into its own LDS copy of the
//    tile (no other wave touches it), then writes the tile once.  Every grid cell is written by
//    exactly one wave, so the grid needs no zero fill and there are no HBM atomics; the sum order
//    inside a tile is fixed by the list order and the lane order of the LDS adds, which makes the
//    result run-to-run reproducible.
#include "pf_internal.h"

namespace pf {

template <typename R>
struct VisCoord {
  int ku, kv;
  R fu, fv;  // u - k_u, v - k_v
};

__device__ inline VisCoord<float> vis_coord(const VisData<float>& v, int64_t i, int, int) {
  const int2 k = v.kk[i];
  const cpx<float> f = v.ff[i];
  return VisCoord<float>{k.x, k.y, f.x, f.y};
}
__device__ inline VisCoord<double> vis_coord(const VisData<double>& v, int64_t i, int Ju, int Jv) {
  const cpx<double> uv = v.uv[i];
  // operators.cc:34-35 in double, as the reference computes it
  const double ku = floor(uv.x - Ju * 0.5), kv = floor(uv.y - Jv * 0.5);
  return VisCoord<double>{(int)ku, (int)kv, uv.x - ku, uv.y - kv};
}

__device__ inline int wrap(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

constexpr int JMAX = 16;

// 1-D tap weights with the (-1)^(k+j) phase folded in: w[a] = (-1)^(k + a) K(f - 1 - a).
template <typename R, int J>
__device__ inline void taps(const KParams<R>& p, int k, R f, int Jrt, R (&w)[J == 0 ? JMAX : J]) {
  constexpr int JJ = J == 0 ? JMAX : J;
#pragma unroll
  for (int a = 0; a < JJ; ++a) {
    if (J == 0 && a >= Jrt) break;
    R val = kern1d<R>(p, f - R(1 + a));
    w[a] = ((k + a) & 1) ? -val : val;
  }
}

// ---------------------------------------------------------------------------------------------
// Degridding  y = G grid
// ---------------------------------------------------------------------------------------------
template <typename R, int JU, int JV>
__global__ void __launch_bounds__(256) k_degrid(GridGeom<R> g, VisData<R> v, const cpx<R>* __restrict__ grid,
                                                cpx<R>* __restrict__ y, R scl) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= v.M) return;
  const int Ju = JU ? JU : g.Ju, Jv = JV ? JV : g.Jv;
  const VisCoord<R> c = vis_coord(v, i, Ju, Jv);
  R wu[JU == 0 ? JMAX : JU], wv[JV == 0 ? JMAX : JV];
  taps<R, JU>(g.ku, c.ku, c.fu, Ju, wu);
  taps<R, JV>(g.kv, c.kv, c.fv, Jv, wv);
  const int q0 = wrap(c.ku + 1, g.Nu), p0 = wrap(c.kv + 1, g.Nv);
  cpx<R> acc = mk<R>(R(0), R(0));
  constexpr int JJV = JV == 0 ? JMAX : JV, JJU = JU == 0 ? JMAX : JU;
#pragma unroll
  for (int b = 0; b < JJV; ++b) {
    if (JV == 0 && b >= Jv) break;
    int p = p0 + b;
    if (p >= g.Nv) p -= g.Nv;
    const cpx<R>* row = grid + (int64_t)p * g.Nu;
    cpx<R> racc = mk<R>(R(0), R(0));
#pragma unroll
    for (int a = 0; a < JJU; ++a) {
      if (JU == 0 && a >= Ju) break;
      int q = q0 + a;
      if (q >= g.Nu) q -= g.Nu;
      const cpx<R> x = row[q];
      racc.x += wu[a] * x.x;
      racc.y += wu[a] * x.y;
    }
    acc.x += wv[b] * racc.x;
    acc.y += wv[b] * racc.y;
  }
  if (v.weights) acc = cmul(acc, v.weights[i]);
  y[v.perm[i]] = scale(acc, scl);
}

// ---------------------------------------------------------------------------------------------
// Gridding  grid = G^H y   (one 64-lane wave per TU x TV tile)
// ---------------------------------------------------------------------------------------------
template <typename R, int JU, int JV>
__global__ void __launch_bounds__(64) k_grid_tiles(GridGeom<R> g, VisData<R> v, TileLists tl,
                                                   const cpx<R>* __restrict__ yin, cpx<R>* __restrict__ grid, R scl) {
  extern __shared__ unsigned char smem_raw[];
  const int TU = tl.TU, TV = tl.TV, TT = TU * TV;
  R* acc_re = reinterpret_cast<R*>(smem_raw);
  R* acc_im = acc_re + TT;
  const int tile = blockIdx.x;
  const int tv = tile / tl.ntu, tu = tile - tv * tl.ntu;
  const int tx0 = tu * TU, ty0 = tv * TV;
  for (int e = threadIdx.x; e < TT; e += blockDim.x) {
    acc_re[e] = R(0);
    acc_im[e] = R(0);
  }
  __syncthreads();
  const int Ju = JU ? JU : g.Ju, Jv = JV ? JV : g.Jv;
  constexpr int JJV = JV == 0 ? JMAX : JV, JJU = JU == 0 ? JMAX : JU;
  const uint32_t end = tl.ptr[tile + 1];
  for (uint32_t base = tl.ptr[tile]; base < end; base += blockDim.x) {
    const uint32_t idx = base + threadIdx.x;
    if (idx < end) {
      const uint32_t i = tl.list[idx];
      const VisCoord<R> c = vis_coord(v, i, Ju, Jv);
      cpx<R> s = yin[v.perm[i]];
      if (v.weights) s = cmulc(s, v.weights[i]);  // conj(G) = phase * K_u * K_v * conj(weight)
      s = scale(s, scl);
      R wu[JJU], wv[JJV];
      taps<R, JU>(g.ku, c.ku, c.fu, Ju, wu);
      taps<R, JV>(g.kv, c.kv, c.fv, Jv, wv);
      const int dx = wrap(c.ku + 1 - tx0, g.Nu), dy = wrap(c.kv + 1 - ty0, g.Nv);
#pragma unroll
      for (int b = 0; b < JJV; ++b) {
        if (JV == 0 && b >= Jv) break;
        int ly = dy + b;
        if (ly >= g.Nv) ly -= g.Nv;
        if (ly < TV) {
          const R sr = s.x * wv[b], si = s.y * wv[b];
#pragma unroll
          for (int a = 0; a < JJU; ++a) {
            if (JU == 0 && a >= Ju) break;
            int lx = dx + a;
            if (lx >= g.Nu) lx -= g.Nu;
            if (lx < TU) {
              atomicAdd(&acc_re[ly * TU + lx], sr * wu[a]);
              atomicAdd(&acc_im[ly * TU + lx], si * wu[a]);
            }
          }
        }
      }
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < TT; e += blockDim.x) {
    const int ly = e / TU, lx = e - ly * TU;
    grid[(int64_t)(ty0 + ly) * g.Nu + tx0 + lx] = mk<R>(acc_re[e], acc_im[e]);
  }
}

template <typename R>
void degrid(const GridGeom<R>& g, const VisData<R>& v, const cpx<R>* grid, cpx<R>* y, R scale, hipStream_t s) {
  if (v.M == 0) return;
  const int threads = 256;
  const dim3 blocks((unsigned)((v.M + threads - 1) / threads));
  ProfScope prof("degrid", s);
#define PF_DG(A, B) hipLaunchKernelGGL((k_degrid<R, A, B>), blocks, dim3(threads), 0, s, g, v, grid, y, scale)
  if (g.Ju == g.Jv) {
    switch (g.Ju) {
      case 2: PF_DG(2, 2); break;
      case 3: PF_DG(3, 3); break;
      case 4: PF_DG(4, 4); break;
      case 5: PF_DG(5, 5); break;
      case 6: PF_DG(6, 6); break;
      case 7: PF_DG(7, 7); break;
      case 8: PF_DG(8, 8); break;
      default: PF_DG(0, 0);
    }
  } else {
    PF_DG(0, 0);
  }
#undef PF_DG
  PF_HIP(hipGetLastError());
}

template <typename R>
void grid_tiles(const GridGeom<R>& g, const VisData<R>& v, const TileLists& t, const cpx<R>* y, cpx<R>* grid, R scale,
                hipStream_t s) {
  const dim3 blocks((unsigned)(t.ntu * t.ntv));
  const size_t shm = 2 * sizeof(R) * (size_t)t.TU * t.TV;
  ProfScope prof("grid_tiles", s);
#define PF_GT(A, B) hipLaunchKernelGGL((k_grid_tiles<R, A, B>), blocks, dim3(64), shm, s, g, v, t, y, grid, scale)
  if (g.Ju == g.Jv) {
    switch (g.Ju) {
      case 2: PF_GT(2, 2); break;
      case 3: PF_GT(3, 3); break;
      case 4: PF_GT(4, 4); break;
      case 5: PF_GT(5, 5); break;
      case 6: PF_GT(6, 6); break;
      case 7: PF_GT(7, 7); break;
      case 8: PF_GT(8, 8); break;
      default: PF_GT(0, 0);
    }
  } else {
    PF_GT(0, 0);
  }
#undef PF_GT
  PF_HIP(hipGetLastError());
}

template void degrid<float>(const GridGeom<float>&, const VisData<float>&, const cpx<float>*, cpx<float>*, float,
                            hipStream_t);
template void degrid<double>(const GridGeom<double>&, const VisData<double>&, const cpx<double>*, cpx<double>*,
                             double, hipStream_t);
template void grid_tiles<float>(const GridGeom<float>&, const VisData<float>&, const TileLists&, const cpx<float>*,
                                cpx<float>*, float, hipStream_t);
template void grid_tiles<double>(const GridGeom<double>&, const VisData<double>&, const TileLists&,
                                 const cpx<double>*, cpx<double>*, double, hipStream_t);

// ---------------------------------------------------------------------------------------------
// Vector helpers: FP64 norms (fixed-order two-level reduction), scaling, precision conversion
// ---------------------------------------------------------------------------------------------
constexpr int RED_BLOCKS = 1024, RED_THREADS = 256;

template <typename R>
__global__ void __launch_bounds__(RED_THREADS) k_norm2_partial(const cpx<R>* __restrict__ x, int64_t n,
                                                                double* __restrict__ part) {
  __shared__ double sh[RED_THREADS];
  double acc = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; k < n; k += (int64_t)RED_BLOCKS * RED_THREADS) {
    const cpx<R> a = x[k];
    acc += (double)a.x * (double)a.x + (double)a.y * (double)a.y;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(RED_THREADS) k_sum_partials(const double* __restrict__ part, int n,
                                                               double* __restrict__ out) {
  __shared__ double sh[RED_THREADS];
  double acc = 0.0;
  for (int k = threadIdx.x; k < n; k += RED_THREADS) acc += part[k];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = RED_THREADS / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

template <typename R>
void norm2_fp64(const cpx<R>* x, int64_t n, double* partials, double* out, hipStream_t s) {
  hipLaunchKernelGGL((k_norm2_partial<R>), dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, x, n, partials);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(RED_THREADS), 0, s, partials, RED_BLOCKS, out);
  PF_HIP(hipGetLastError());
}

template <typename R>
__global__ void k_scale(cpx<R>* x, int64_t n, double sc) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    cpx<R> a = x[k];
    a.x = (R)((double)a.x * sc);
    a.y = (R)((double)a.y * sc);
    x[k] = a;
  }
}

template <typename R>
void scale_vec(cpx<R>* x, int64_t n, double sc, hipStream_t s) {
  hipLaunchKernelGGL((k_scale<R>), dim3(1024), dim3(256), 0, s, x, n, sc);
  PF_HIP(hipGetLastError());
}

__global__ void k_c128_to_c64(const cpx<double>* in, cpx<float>* out, int64_t n) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    out[k] = mk<float>((float)in[k].x, (float)in[k].y);
}
__global__ void k_c64_to_c128(const cpx<float>* in, cpx<double>* out, int64_t n) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    out[k] = mk<double>((double)in[k].x, (double)in[k].y);
}
void c128_to_c64(const cpx<double>* in, cpx<float>* out, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_c128_to_c64, dim3(1024), dim3(256), 0, s, in, out, n);
  PF_HIP(hipGetLastError());
}
void c64_to_c128(const cpx<float>* in, cpx<double>* out, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_c64_to_c128, dim3(1024), dim3(256), 0, s, in, out, n);
  PF_HIP(hipGetLastError());
}

template void norm2_fp64<float>(const cpx<float>*, int64_t, double*, double*, hipStream_t);
template void norm2_fp64<double>(const cpx<double>*, int64_t, double*, double*, hipStream_t);
template void scale_vec<float>(cpx<float>*, int64_t, double, hipStream_t);
template void scale_vec<double>(cpx<double>*, int64_t, double, hipStream_t);

}  // namespace pf
