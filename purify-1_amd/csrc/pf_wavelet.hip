// pf_wavelet.hip -- the sparsity dictionary Psi of PURIFY's solvers on the device: SOPT's
// sopt::wavelets::SARA wrapped by sopt::linear_transform(sara, imsizey, imsizex), as
// factory::wavelet_operator_factory builds it (cpp/purify/wavelet_operator_factory.h:20-45).
//
// Each dictionary entry is the Dirac basis or a Daubechies wavelet with a number of levels.  One level
// of the 2-D analysis transform works on the top-left (h, w) block of the coefficient image: every
// row of the block, then every column, goes through a periodic down-convolution with the low-pass
// and the high-pass filter, low-pass results first:
//     lo[i] = sum_k h[k] x[(2i + k) mod w],  hi[i] = sum_k g[k] x[(2i + k) mod w],  i < w / 2,
// and the next level works on the (h/2, w/2) low-low block.  The filters form an orthonormal pair
// (g[k] = (-1)^k h[L-1-k]), so the synthesis transform (the adjoint, coarsest level first, columns
// then rows) is also the inverse.  SARA stacks the entries' analysis coefficients, each divided by
// sqrt(number of entries); its synthesis sums the entries' reconstructions with the same factor.
//
// The Daubechies scaling filters are computed on the host from their spectral factorisation
// (Daubechies, "Ten Lectures on Wavelets", 1992, section 6.1): the roots y of
// P(y) = sum_{k<N} C(N-1+k, k) y^k map to z through y = -(z - 1)^2 / (4 z); the roots with |z| < 1,
// with N roots at z = -1, give the extremal-phase filter, normalised to sum sqrt(2).
#include <algorithm>
#include <cmath>
#include <complex>
#include <memory>
#include <vector>

#include "../../include/purify_amd.h"
#include "pf_internal.h"

namespace pf {

constexpr int WAV_MAXL = 2 * PF_DB_MAX;  // longest filter

// Daubechies scaling filter with N vanishing moments (length 2N), in double, sum sqrt(2).
std::vector<double> daubechies_filter(int N) {
  if (N < 1 || N > PF_DB_MAX) throw Error(PF_ERR_ARG, "Daubechies order must be 1.." + std::to_string(PF_DB_MAX));
  using C = std::complex<long double>;
  if (N == 1) return {1.0 / std::sqrt(2.0), 1.0 / std::sqrt(2.0)};
  // P(y) coefficients, highest degree first
  std::vector<long double> P(N);
  for (int k = 0; k < N; ++k) {
    long double c = 1;  // C(N-1+k, k)
    for (int j = 1; j <= k; ++j) c = c * (N - 1 + j) / j;
    P[N - 1 - k] = c;
  }
  // roots of P (degree N-1) by Durand-Kerner with Newton polishing, in long double
  const int deg = N - 1;
  auto evalP = [&](C y) {
    C r = 0;
    for (int k = 0; k <= deg; ++k) r = r * y + P[k];
    return r;
  };
  auto evaldP = [&](C y) {
    C r = 0;
    for (int k = 0; k < deg; ++k) r = r * y + P[k] * (long double)(deg - k);
    return r;
  };
  std::vector<C> r(deg);
  for (int i = 0; i < deg; ++i) r[i] = std::pow(C(0.4L, 0.9L), i + 1);
  for (int it = 0; it < 2000; ++it) {
    long double delta = 0;
    for (int i = 0; i < deg; ++i) {
      C den = P[0];
      for (int j = 0; j < deg; ++j)
        if (j != i) den *= (r[i] - r[j]);
      const C step = evalP(r[i]) / den;
      r[i] -= step;
      delta = std::max(delta, std::abs(step));
    }
    if (delta < 1e-30L) break;
  }
  for (int i = 0; i < deg; ++i)
    for (int it = 0; it < 5; ++it) {
      const C d = evaldP(r[i]);
      if (std::abs(d) == 0) break;
      r[i] -= evalP(r[i]) / d;
    }
  // y = -(z - 1)^2 / (4 z)  <=>  z^2 + (4y - 2) z + 1 = 0: keep the root inside the unit circle
  std::vector<C> z;
  for (const C& y : r) {
    const C b = 4.0L * y - 2.0L;
    const C s = std::sqrt(b * b - 4.0L);
    C z1 = (-b + s) / 2.0L, z2 = (-b - s) / 2.0L;
    z.push_back(std::abs(z1) < std::abs(z2) ? z1 : z2);
  }
  for (int k = 0; k < N; ++k) z.push_back(C(-1.0L, 0.0L));
  // polynomial with those roots (lowest degree first), real up to rounding
  std::vector<C> poly{C(1.0L)};
  for (const C& root : z) {
    std::vector<C> nxt(poly.size() + 1, C(0));
    for (size_t k = 0; k < poly.size(); ++k) {
      nxt[k + 1] += poly[k];
      nxt[k] -= root * poly[k];
    }
    poly.swap(nxt);
  }
  long double sum = 0;
  for (const C& c : poly) sum += c.real();
  std::vector<double> h(poly.size());
  // highest power first gives the extremal-phase ordering (energy at the start of the filter)
  for (size_t k = 0; k < poly.size(); ++k)
    h[k] = (double)(poly[poly.size() - 1 - k].real() * std::sqrt(2.0L) / sum);
  return h;
}

template <typename R>
struct WavFilt {
  R lo[WAV_MAXL], hi[WAV_MAXL];
  int L;
};

// One analysis level over the (h, w) block in one launch, for a compile-time filter length L: each row
// is filtered into low (lo) and high (hi) halves, out[k] = sum_j f[j] in[(2 k + j) mod n] with j increasing,
// then each column of that result the same way, giving the four bands.  Block (x, y) produces row-pass
// outputs p in [16 x, 16 x + 16) (both halves) and column outputs i in [16 y, 16 y + 16): it stages the
// input rows 2 i .. 2 i + L - 1 and columns 2 p .. 2 p + L - 1 it needs (wrapped periodically) in LDS,
// filters the rows into LDS, then the columns.  The three detail bands go to their places in dst (row
// stride ld) multiplied by sd, the approximation band (low-low, h/2 x w/2) to apx (row stride la), which
// may be dst itself, multiplied by sa.  ls: row stride of src.  32-bit indices (an image holds < 2^31
// pixels).  src must not overlap dst or apx.
constexpr int DW_T = 16;
template <typename R>
struct DwtArgs {
  const cpx<R>* src;
  int ls;
  cpx<R>* dst;
  int ld;
  cpx<R>* apx;
  int la, h, w;
  WavFilt<R> f;
  R sd, sa;
};
// BT: a batch launch, whose blockIdx.z selects the arguments from tab
template <typename R, int L, bool BT = false>
__global__ void __launch_bounds__(256) k_dwt2d(DwtArgs<R> a0, const DwtArgs<R>* __restrict__ tab) {
  const DwtArgs<R> q = BT ? tab[blockIdx.z] : a0;
  const cpx<R>* __restrict__ src = q.src;
  cpx<R>* __restrict__ dst = q.dst;
  cpx<R>* __restrict__ apx = q.apx;
  const int ls = q.ls, ld = q.ld, la = q.la, h = q.h, w = q.w;
  const WavFilt<R>& f = q.f;
  const R sd = q.sd, sa = q.sa;
  constexpr int RI = 2 * DW_T + L - 2;  // staged rows and columns
  __shared__ cpx<R> in_t[RI][RI + 1];
  __shared__ cpx<R> rp_t[RI][2 * DW_T + 1];
  const int hh = h / 2, hw = w / 2;
  const int p0 = (int)blockIdx.x * DW_T, i0 = (int)blockIdx.y * DW_T;
  const int t = (int)threadIdx.x;
  for (int idx = t; idx < RI * RI; idx += 256) {
    const int r = idx / RI, cc = idx - r * RI;
    int row = 2 * i0 + r, col = 2 * p0 + cc;
    while (row >= h) row -= h;
    while (col >= w) col -= w;
    in_t[r][cc] = src[row * ls + col];
  }
  __syncthreads();
  for (int idx = t; idx < RI * DW_T; idx += 256) {
    const int r = idx / DW_T, q = idx - r * DW_T;
    cpx<R> a = mk<R>(R(0), R(0)), b = mk<R>(R(0), R(0));
#pragma unroll
    for (int k = 0; k < L; ++k) {
      const cpx<R> x = in_t[r][2 * q + k];
      a.x += f.lo[k] * x.x;
      a.y += f.lo[k] * x.y;
      b.x += f.hi[k] * x.x;
      b.y += f.hi[k] * x.y;
    }
    rp_t[r][q] = a;
    rp_t[r][DW_T + q] = b;
  }
  __syncthreads();
  for (int idx = t; idx < DW_T * 2 * DW_T; idx += 256) {
    const int ii = idx / (2 * DW_T), jj = idx - ii * (2 * DW_T);
    const int q = jj < DW_T ? jj : jj - DW_T;
    if (i0 + ii >= hh || p0 + q >= hw) continue;
    cpx<R> a = mk<R>(R(0), R(0)), b = mk<R>(R(0), R(0));
#pragma unroll
    for (int k = 0; k < L; ++k) {
      const cpx<R> x = rp_t[2 * ii + k][jj];
      a.x += f.lo[k] * x.x;
      a.y += f.lo[k] * x.y;
      b.x += f.hi[k] * x.x;
      b.y += f.hi[k] * x.y;
    }
    const int col = (jj < DW_T ? 0 : hw) + p0 + q;
    if (jj < DW_T)
      apx[(i0 + ii) * la + col] = scale(a, sa);
    else
      dst[(i0 + ii) * ld + col] = scale(a, sd);
    dst[(hh + i0 + ii) * ld + col] = scale(b, sd);
  }
}

// One synthesis level in one launch, for a compile-time filter length L: first the columns, then the
// rows.  Along a line of n samples whose low half holds lo[0 .. n/2) and high half hi[0 .. n/2),
// outputs 2m and 2m + 1 read the same L/2 pairs i = m, m - 1, ... (mod n/2): output 2m takes the even
// taps, 2m + 1 the odd ones, in increasing tap order (out[2m + e] = sum_k f.lo[2k + e] lo[m - k] +
// f.hi[2k + e] hi[m - k]); the column results pass to the row step through LDS.  Block (x, y) produces output rows
// 2 i .. 2 i + 1 and columns 2 p .. 2 p + 1 for i in [16 y, 16 y + 16), p in [16 x, 16 x + 16): it
// stages the CI = 16 + L/2 - 1 coefficient rows i - L/2 + 1 .. i + 15 (mod h/2) of the low and of the
// high half, over the CI columns of the same window in each half of the columns; the approximation
// quadrant comes from apx (row stride la; the coarser level's output) or, when apx is null, from c
// scaled by cs, and every other quadrant from c (row stride ld) scaled by cs.  Output to out (row stride
// lout), added to its values when ACC.  apx and c must not overlap out.
template <typename R>
struct IdwtArgs {
  const cpx<R>* c;
  int ld;
  const cpx<R>* apx;
  int la;
  cpx<R>* out;
  int lout, h, w;
  WavFilt<R> f;
  R cs;
};
template <typename R, bool ACC, int L, bool BT = false>
__global__ void __launch_bounds__(256) k_idwt2d(IdwtArgs<R> a0, const IdwtArgs<R>* __restrict__ tab) {
  const IdwtArgs<R> q = BT ? tab[blockIdx.z] : a0;
  const cpx<R>* __restrict__ c = q.c;
  const cpx<R>* __restrict__ apx = q.apx;
  cpx<R>* __restrict__ out = q.out;
  const int ld = q.ld, la = q.la, lout = q.lout, h = q.h, w = q.w;
  const WavFilt<R>& f = q.f;
  const R cs = q.cs;
  constexpr int K = L / 2, CI = DW_T + K - 1;
  __shared__ cpx<R> in_t[2 * CI][2 * CI + 1];
  __shared__ cpx<R> cp_t[2 * DW_T][2 * CI + 1];
  const int hh = h / 2, hw = w / 2;
  const int p0 = (int)blockIdx.x * DW_T, i0 = (int)blockIdx.y * DW_T;
  const int t = (int)threadIdx.x;
  for (int idx = t; idx < 4 * CI * CI; idx += 256) {
    const int rr = idx / (2 * CI), cc = idx - rr * (2 * CI);
    const bool hi_r = rr >= CI, hi_c = cc >= CI;
    int i = i0 - (K - 1) + (hi_r ? rr - CI : rr), j = p0 - (K - 1) + (hi_c ? cc - CI : cc);
    while (i < 0) i += hh;
    while (i >= hh) i -= hh;
    while (j < 0) j += hw;
    while (j >= hw) j -= hw;
    const int row = (hi_r ? hh : 0) + i, col = (hi_c ? hw : 0) + j;
    in_t[rr][cc] = apx && !hi_r && !hi_c ? apx[i * la + j] : scale(c[row * ld + col], cs);
  }
  __syncthreads();
  for (int idx = t; idx < DW_T * 2 * CI; idx += 256) {
    const int mi = idx / (2 * CI), cc = idx - mi * (2 * CI);
    cpx<R> ae = mk<R>(R(0), R(0)), ao = mk<R>(R(0), R(0));
#pragma unroll
    for (int kk = 0; kk < K; ++kk) {
      const cpx<R> cl = in_t[K - 1 + mi - kk][cc], ch = in_t[CI + K - 1 + mi - kk][cc];
      ae.x += f.lo[2 * kk] * cl.x + f.hi[2 * kk] * ch.x;
      ae.y += f.lo[2 * kk] * cl.y + f.hi[2 * kk] * ch.y;
      ao.x += f.lo[2 * kk + 1] * cl.x + f.hi[2 * kk + 1] * ch.x;
      ao.y += f.lo[2 * kk + 1] * cl.y + f.hi[2 * kk + 1] * ch.y;
    }
    cp_t[2 * mi][cc] = ae;
    cp_t[2 * mi + 1][cc] = ao;
  }
  __syncthreads();
  for (int idx = t; idx < 2 * DW_T * DW_T; idx += 256) {
    const int rl = idx / DW_T, q = idx - rl * DW_T;
    if (i0 + rl / 2 >= hh || p0 + q >= hw) continue;
    cpx<R> ae = mk<R>(R(0), R(0)), ao = mk<R>(R(0), R(0));
#pragma unroll
    for (int kk = 0; kk < K; ++kk) {
      const cpx<R> cl = cp_t[rl][K - 1 + q - kk], ch = cp_t[rl][CI + K - 1 + q - kk];
      ae.x += f.lo[2 * kk] * cl.x + f.hi[2 * kk] * ch.x;
      ae.y += f.lo[2 * kk] * cl.y + f.hi[2 * kk] * ch.y;
      ao.x += f.lo[2 * kk + 1] * cl.x + f.hi[2 * kk + 1] * ch.x;
      ao.y += f.lo[2 * kk + 1] * cl.y + f.hi[2 * kk + 1] * ch.y;
    }
    cpx<R>* o = out + (2 * i0 + rl) * lout + 2 * (p0 + q);
    if (ACC) {
      const cpx<R> oe = o[0], oo = o[1];
      ae.x += oe.x;
      ae.y += oe.y;
      ao.x += oo.x;
      ao.y += oo.y;
    }
    o[0] = ae;
    o[1] = ao;
  }
}

// Calls F.template operator()<L>() for the runtime filter length L = 2, 4, ..., WAV_MAXL.
template <class F>
static void dispatch_len(int L, F&& fn) {
  switch (L) {
    case 2: fn.template operator()<2>(); break;
    case 4: fn.template operator()<4>(); break;
    case 6: fn.template operator()<6>(); break;
    case 8: fn.template operator()<8>(); break;
    case 10: fn.template operator()<10>(); break;
    case 12: fn.template operator()<12>(); break;
    case 14: fn.template operator()<14>(); break;
    case 16: fn.template operator()<16>(); break;
    case 18: fn.template operator()<18>(); break;
    case 20: fn.template operator()<20>(); break;
    default: throw Error(-3, "wavelet filter length must be even and at most " + std::to_string(WAV_MAXL));
  }
}

// out[k] = in[k] * scl (ACC: out[k] += in[k] * scl) over the whole image
template <typename R>
struct CopyArgs {
  const cpx<R>* in;
  cpx<R>* out;
  int64_t n;
  R scl;
};
template <typename R, bool ACC, bool BT = false>
__global__ void __launch_bounds__(256) k_copy_scaled(CopyArgs<R> a0, const CopyArgs<R>* __restrict__ tab) {
  const CopyArgs<R> q = BT ? tab[blockIdx.z] : a0;
  const cpx<R>* __restrict__ in = q.in;
  cpx<R>* __restrict__ out = q.out;
  const int64_t n = q.n;
  const R scl = q.scl;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    cpx<R> v = scale(in[k], scl);
    if (ACC) {
      const cpx<R> o = out[k];
      v.x += o.x;
      v.y += o.y;
    }
    out[k] = v;
  }
}

static dim3 grid_for(int64_t n) { return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192))); }

// out = in * scl (acc: out += in * scl) over n values
template <typename R>
void copy_scaled(const cpx<R>* in, cpx<R>* out, int64_t n, R scl, bool acc, hipStream_t s) {
  const CopyArgs<R> a{in, out, n, scl};
  if (acc) {
    if (record_launch<CopyArgs<R>, k_copy_scaled<R, true, true>>(grid_for(n), dim3(256), 0, true, a)) return;
    hipLaunchKernelGGL((k_copy_scaled<R, true>), grid_for(n), dim3(256), 0, s, a, nullptr);
  } else {
    if (record_launch<CopyArgs<R>, k_copy_scaled<R, false, true>>(grid_for(n), dim3(256), 0, true, a)) return;
    hipLaunchKernelGGL((k_copy_scaled<R, false>), grid_for(n), dim3(256), 0, s, a, nullptr);
  }
}

template <typename R>
WavFilt<R> make_filt(const std::vector<double>& h) {
  WavFilt<R> f{};
  f.L = (int)h.size();
  for (int k = 0; k < f.L; ++k) {
    f.lo[k] = (R)h[k];
    f.hi[k] = (R)(((k & 1) ? -1.0 : 1.0) * h[f.L - 1 - k]);
  }
  return f;
}

inline unsigned blocks_of(int n) { return (unsigned)std::max(1, (n + 255) / 256); }

template <typename R>
struct DwtLevel {
  const cpx<R>* src;
  int ls;
  cpx<R>* c;
  int ld;
  cpx<R>* apx;
  int la, h, w;
  const WavFilt<R>& f;
  R sd, sa;
  hipStream_t s;
  template <int L>
  void operator()() {
    const unsigned bx = (unsigned)((w / 2 + DW_T - 1) / DW_T), by = (unsigned)((h / 2 + DW_T - 1) / DW_T);
    const DwtArgs<R> a{src, ls, c, ld, apx, la, h, w, f, sd, sa};
    if (record_launch<DwtArgs<R>, k_dwt2d<R, L, true>>(dim3(bx, by), dim3(256), 0, false, a)) return;
    hipLaunchKernelGGL((k_dwt2d<R, L>), dim3(bx, by), dim3(256), 0, s, a, nullptr);
  }
};
template <typename R, bool ACC>
struct IdwtLevel {
  const cpx<R>* c;
  int ld;
  const cpx<R>* apx;
  int la;
  cpx<R>* out;
  int lout, h, w;
  const WavFilt<R>& f;
  R cs;
  hipStream_t s;
  template <int L>
  void operator()() {
    const unsigned bx = (unsigned)((w / 2 + DW_T - 1) / DW_T), by = (unsigned)((h / 2 + DW_T - 1) / DW_T);
    const IdwtArgs<R> a{c, ld, apx, la, out, lout, h, w, f, cs};
    if (record_launch<IdwtArgs<R>, k_idwt2d<R, ACC, L, true>>(dim3(bx, by), dim3(256), 0, false, a)) return;
    hipLaunchKernelGGL((k_idwt2d<R, ACC, L>), dim3(bx, by), dim3(256), 0, s, a, nullptr);
  }
};


// Analysis of one entry: image x -> coefficients c (both ny x nx), each coefficient multiplied by scl;
// T: work image.  Each level writes its detail bands into c and its approximation band, compact (row
// stride w / 2), into T or T + ny nx / 4 in turn, where the next level reads it: a level never reads
// what it writes, and the two regions (at most ny nx / 4 and ny nx / 16 elements) are disjoint.  The
// last level writes its approximation into c.  Only what lands in c is scaled.
template <typename R>
void dwt2(const cpx<R>* x, cpx<R>* c, cpx<R>* T, int ny, int nx, int levels, const WavFilt<R>& f, R scl,
          hipStream_t s) {
  const cpx<R>* src = x;
  int ls = nx;
  for (int l = 0; l < levels; ++l) {
    const int h = ny >> l, w = nx >> l;
    const bool last = l == levels - 1;
    cpx<R>* apx = last ? c : T + ((l & 1) ? (size_t)ny * nx / 4 : 0);
    const int la = last ? nx : w / 2;
    dispatch_len(f.L, DwtLevel<R>{src, ls, c, nx, apx, la, h, w, f, scl, last ? scl : R(1), s});
    src = apx;
    ls = la;
  }
}

// Synthesis of one entry: coefficients c -> image x (ACC: x += ...); W, T: work images.
template <typename R>
void idwt2(const cpx<R>* c, cpx<R>* x, cpx<R>* W, cpx<R>* T, int ny, int nx, int levels, const WavFilt<R>& f, R scl,
           bool acc, hipStream_t s) {
  const int64_t N = (int64_t)ny * nx;
  if (levels == 0) {
    copy_scaled<R>(c, x, N, scl, acc, s);
    return;
  }
  // One launch per level.  Each level but the finest writes its reconstructed approximation, compact
  // (row stride w), into W or T in turn, and the next finer level reads it from there: no level reads
  // what it writes.  The coefficients enter with the SARA factor applied.
  const cpx<R>* apx = nullptr;
  int la = 0;
  for (int l = levels - 1; l >= 0; --l) {
    const int h = ny >> l, w = nx >> l;
    if (l == 0) {
      if (acc) dispatch_len(f.L, IdwtLevel<R, true>{c, nx, apx, la, x, nx, h, w, f, scl, s});
      else dispatch_len(f.L, IdwtLevel<R, false>{c, nx, apx, la, x, nx, h, w, f, scl, s});
    } else {
      cpx<R>* o = (l & 1) ? W : T;
      dispatch_len(f.L, IdwtLevel<R, false>{c, nx, apx, la, o, w, h, w, f, scl, s});
      apx = o;
      la = w;
    }
  }
}

}  // namespace pf

using namespace pf;

struct pf_wavelet {
  int ny = 0, nx = 0, prec = 0, nwav = 0;
  std::vector<int> db, levels;
  std::vector<std::vector<double>> filters;
  void* d_T = nullptr;  // work images
  void* d_W = nullptr;
  void* d_stage = nullptr;  // complex128 host staging (imsize * nwav)
  void* d_in = nullptr;     // staged input in the operator's precision
  void* d_out = nullptr;
  std::vector<void*> allocs;
  void* dalloc(size_t bytes) {
    void* p = dev_alloc(std::max<size_t>(bytes, 16));
    allocs.push_back(p);
    return p;
  }
  size_t esz() const { return prec == PF_C64 ? 8 : 16; }
  int64_t n() const { return (int64_t)ny * nx; }
  ~pf_wavelet() {
    cstream_sync();
    for (void* p : allocs) dev_free(p);
  }
};

namespace pf {
// Psi^dagger: image -> nwav stacked coefficient images
template <typename R>
void wavelet_analysis(pf_wavelet* wv, const cpx<R>* x, cpx<R>* c, hipStream_t s) {
  const R scl = (R)(1.0 / std::sqrt((double)wv->nwav));
  const int64_t N = wv->n();
  for (int i = 0; i < wv->nwav; ++i) {
    cpx<R>* ci = c + (int64_t)i * N;
    if (wv->db[i] == 0) {
      copy_scaled<R>(x, ci, N, scl, false, s);
      continue;
    }
    const WavFilt<R> f = make_filt<R>(wv->filters[i]);
    if (wv->levels[i] == 0) {
      copy_scaled<R>(x, ci, N, scl, false, s);
      continue;
    }
    // the SARA factor is applied as the coefficients are written (a product by 1 for one basis)
    dwt2<R>(x, ci, static_cast<cpx<R>*>(wv->d_T), wv->ny, wv->nx, wv->levels[i], f, scl, s);
  }
  PF_HIP(hipGetLastError());
}

// Psi: nwav stacked coefficient images -> image
template <typename R>
void wavelet_synthesis(pf_wavelet* wv, const cpx<R>* c, cpx<R>* x, hipStream_t s) {
  const R scl = (R)(1.0 / std::sqrt((double)wv->nwav));
  const int64_t N = wv->n();
  for (int i = 0; i < wv->nwav; ++i) {
    const cpx<R>* ci = c + (int64_t)i * N;
    const WavFilt<R> f = wv->db[i] ? make_filt<R>(wv->filters[i]) : WavFilt<R>{};
    idwt2<R>(ci, x, static_cast<cpx<R>*>(wv->d_W), static_cast<cpx<R>*>(wv->d_T), wv->ny, wv->nx,
             wv->db[i] ? wv->levels[i] : 0, f, scl, i > 0, s);
  }
  PF_HIP(hipGetLastError());
}
WaveletShape wavelet_shape(const pf_wavelet* wv) { return WaveletShape{wv->ny, wv->nx, wv->prec, wv->nwav}; }
template void wavelet_analysis<float>(pf_wavelet*, const cpx<float>*, cpx<float>*, hipStream_t);
template void wavelet_analysis<double>(pf_wavelet*, const cpx<double>*, cpx<double>*, hipStream_t);
template void wavelet_synthesis<float>(pf_wavelet*, const cpx<float>*, cpx<float>*, hipStream_t);
template void wavelet_synthesis<double>(pf_wavelet*, const cpx<double>*, cpx<double>*, hipStream_t);

}  // namespace pf

extern "C" {

int pf_wavelet_filter(int32_t db, double* lo, int32_t* len) {
  return guarded([&] {
    const std::vector<double> h = daubechies_filter(db);
    if (len) *len = (int32_t)h.size();
    if (lo) std::copy(h.begin(), h.end(), lo);
  });
}

int pf_wavelet_create(int32_t imsizey, int32_t imsizex, int32_t nwav, const int32_t* db, const int32_t* levels,
                      int32_t precision, pf_wavelet** out) {
  if (!out || !db || !levels) return fail(PF_ERR_ARG, "null argument");
  *out = nullptr;
  std::unique_ptr<pf_wavelet> wv(new pf_wavelet());
  int rc = guarded([&] {
    if (imsizey < 1 || imsizex < 1 || nwav < 1) throw Error(PF_ERR_ARG, "image size and dictionary size must be positive");
    if (precision != PF_C64 && precision != PF_C128) throw Error(PF_ERR_ARG, "precision must be PF_C64 or PF_C128");
    wv->ny = imsizey;
    wv->nx = imsizex;
    wv->prec = precision;
    wv->nwav = nwav;
    for (int i = 0; i < nwav; ++i) {
      if (db[i] < 0 || db[i] > PF_DB_MAX) throw Error(PF_ERR_ARG, "wavelet must be Dirac (0) or DB1..DB" + std::to_string(PF_DB_MAX));
      if (levels[i] < 0 || (db[i] && ((imsizey % (1 << levels[i])) || (imsizex % (1 << levels[i])))))
        throw Error(PF_ERR_ARG, "image sides must be divisible by 2^levels");
      wv->db.push_back(db[i]);
      wv->levels.push_back(db[i] ? levels[i] : 0);
      wv->filters.push_back(db[i] ? daubechies_filter(db[i]) : std::vector<double>());
      // no minimum line length: the periodic extension wraps a filter around a short line as often as needed
    }
    const size_t b = wv->esz() * (size_t)wv->n();
    wv->d_T = wv->dalloc(b);
    wv->d_W = wv->dalloc(b);
  });
  if (rc == PF_OK) *out = wv.release();
  return rc;
}

// coefficients (nwav * imsizey * imsizex) -> image: Psi, the LinearTransform's direct operator
int pf_wavelet_forward(pf_wavelet* wv, const void* coeffs, void* image, int c_loc, int x_loc, pf_stream stream) {
  return guarded([&] {
    if (!wv || !coeffs || !image) throw Error(PF_ERR_ARG, "null argument");
    hipStream_t s = io_stream(stream, c_loc == PF_HOST && x_loc == PF_HOST);
    const int64_t N = wv->n(), NC = N * wv->nwav;
    const void* din = coeffs;
    if (c_loc == PF_HOST || x_loc == PF_HOST) {
      if (!wv->d_stage) {
        wv->d_stage = wv->dalloc(16 * (size_t)NC);
        wv->d_in = wv->dalloc(wv->esz() * (size_t)NC);
        wv->d_out = wv->dalloc(wv->esz() * (size_t)NC);
      }
    }
    if (c_loc == PF_HOST) {
      PF_HIP(hipMemcpyAsync(wv->prec == PF_C64 ? wv->d_stage : wv->d_in, coeffs, 16 * NC, hipMemcpyHostToDevice, s));
      if (wv->prec == PF_C64) c128_to_c64((const cpx<double>*)wv->d_stage, (cpx<float>*)wv->d_in, NC, s);
      din = wv->d_in;
    } else if (c_loc != PF_DEVICE) {
      throw Error(PF_ERR_ARG, "location must be PF_HOST or PF_DEVICE");
    }
    void* dout = x_loc == PF_DEVICE ? image : wv->d_out;
    if (wv->prec == PF_C64) wavelet_synthesis<float>(wv, (const cpx<float>*)din, (cpx<float>*)dout, s);
    else wavelet_synthesis<double>(wv, (const cpx<double>*)din, (cpx<double>*)dout, s);
    if (x_loc == PF_HOST) {
      if (wv->prec == PF_C64) {
        c64_to_c128((const cpx<float>*)dout, (cpx<double>*)wv->d_stage, N, s);
        PF_HIP(hipMemcpyAsync(image, wv->d_stage, 16 * N, hipMemcpyDeviceToHost, s));
      } else {
        PF_HIP(hipMemcpyAsync(image, dout, 16 * N, hipMemcpyDeviceToHost, s));
      }
      PF_HIP(hipStreamSynchronize(s));
    } else if (x_loc != PF_DEVICE) {
      throw Error(PF_ERR_ARG, "location must be PF_HOST or PF_DEVICE");
    }
  });
}

// image -> coefficients: Psi^dagger, the LinearTransform's adjoint
int pf_wavelet_adjoint(pf_wavelet* wv, const void* image, void* coeffs, int x_loc, int c_loc, pf_stream stream) {
  return guarded([&] {
    if (!wv || !coeffs || !image) throw Error(PF_ERR_ARG, "null argument");
    hipStream_t s = io_stream(stream, c_loc == PF_HOST && x_loc == PF_HOST);
    const int64_t N = wv->n(), NC = N * wv->nwav;
    const void* din = image;
    if (c_loc == PF_HOST || x_loc == PF_HOST) {
      if (!wv->d_stage) {
        wv->d_stage = wv->dalloc(16 * (size_t)NC);
        wv->d_in = wv->dalloc(wv->esz() * (size_t)NC);
        wv->d_out = wv->dalloc(wv->esz() * (size_t)NC);
      }
    }
    if (x_loc == PF_HOST) {
      PF_HIP(hipMemcpyAsync(wv->prec == PF_C64 ? wv->d_stage : wv->d_in, image, 16 * N, hipMemcpyHostToDevice, s));
      if (wv->prec == PF_C64) c128_to_c64((const cpx<double>*)wv->d_stage, (cpx<float>*)wv->d_in, N, s);
      din = wv->d_in;
    } else if (x_loc != PF_DEVICE) {
      throw Error(PF_ERR_ARG, "location must be PF_HOST or PF_DEVICE");
    }
    void* dout = c_loc == PF_DEVICE ? coeffs : wv->d_out;
    if (wv->prec == PF_C64) wavelet_analysis<float>(wv, (const cpx<float>*)din, (cpx<float>*)dout, s);
    else wavelet_analysis<double>(wv, (const cpx<double>*)din, (cpx<double>*)dout, s);
    if (c_loc == PF_HOST) {
      if (wv->prec == PF_C64) {
        c64_to_c128((const cpx<float>*)dout, (cpx<double>*)wv->d_stage, NC, s);
        PF_HIP(hipMemcpyAsync(coeffs, wv->d_stage, 16 * NC, hipMemcpyDeviceToHost, s));
      } else {
        PF_HIP(hipMemcpyAsync(coeffs, dout, 16 * NC, hipMemcpyDeviceToHost, s));
      }
      PF_HIP(hipStreamSynchronize(s));
    } else if (c_loc != PF_DEVICE) {
      throw Error(PF_ERR_ARG, "location must be PF_HOST or PF_DEVICE");
    }
  });
}

int pf_wavelet_destroy(pf_wavelet* wv) {
  delete wv;
  return PF_OK;
}

}  // extern "C"
