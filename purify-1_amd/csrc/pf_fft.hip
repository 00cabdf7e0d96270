// pf_fft.hip -- 2-D FFT F on the oversampled grid (replaces FFTW in cpp/purify/operators.h:209-261).
//
// Each 1-D line is a Stockham auto-sort transform (Govindaraju et al., "High Performance Discrete
// Fourier Transforms on Graphics Processors", SC 2008): stage s with radix R and sub-transform
// length Ns reads x[j + r N/R], twiddles by W^(r (j mod Ns) N / (Ns R)), does an R-point DFT in
// registers and writes x[(j - j mod Ns) R + j mod Ns + r Ns].  Every thread holds 16 elements (32 in
// the complex64 single-pass column kernel); the first stage loads straight from HBM and the last
// stores straight to HBM, the stages in between exchange through LDS.  The row and single-pass
// column kernels read their twiddles from two small LDS tables, so a stage issues no global loads.  The stage plan is fixed at compile time by log2 N (radix 16, plus one
// 2/4/8 stage first for odd log2 N), so every stage unrolls completely.
//
// Row pass: one or more grid rows per block (contiguous lines).  The forward operator's row pass
// reads an image row, multiplies the separable correction S, zero-pads (Z) in registers and writes
// only the ny non-zero grid rows; the adjoint's row pass reads those rows back and writes the
// cropped, corrected image.
// Column pass: a block owns C adjacent columns and the lines are columns.  Up to 512 rows, 16 or more
// columns per block; up to 8192 rows, still one pass with 2 to 8 columns per block in up to 160 KiB
// of padded LDS (a single read and write of the grid).  The alternative for long columns
// (fft_columns = 1) splits Nv = N1*N2 into an N1-point pass over rows of stride N2 (with the
// W_Nv^(r2 k1) twiddle) and an N2-point pass over contiguous row blocks, 512 B row segments each.
#include <algorithm>

#include "pf_internal.h"

namespace pf {

LinePlan make_line_plan(int N) {
  LinePlan p;
  p.N = N;
  int m = 0;
  while ((1 << m) < N) ++m;
  if ((1 << m) != N || N < 2) throw Error(-3, "FFT length must be a power of two >= 2");
  p.log2N = m;
  if (N <= FFT_E) {
    p.nst = 1;
    p.rad[0] = N;
    return p;
  }
  const int n16 = m / 4, rem = m % 4;
  int k = 0;
  if (rem) p.rad[k++] = 1 << rem;
  for (int i = 0; i < n16; ++i) p.rad[k++] = 16;
  p.nst = k;
  return p;
}

// Compile-time stage plan of a 2^L-point line (the same plan make_line_plan describes) with EE
// elements per thread (FFT_E, or 32 for the persistent column pass: E / radix butterflies a stage).
template <int L, int EE = FFT_E>
struct CPlan {
  static constexpr int N = 1 << L;
  static constexpr int E = N < EE ? N : EE;  // elements per thread
  static constexpr int T = N / E;            // threads per line
  static constexpr int REM = N <= FFT_E ? 0 : L % 4;
  static constexpr int NST = N <= FFT_E ? 1 : L / 4 + (REM ? 1 : 0);
  static constexpr int rad(int s) { return N <= FFT_E ? N : ((REM && s == 0) ? (1 << REM) : 16); }
  static constexpr int ns(int s) {
    int r = 1;
    for (int k = 0; k < s; ++k) r *= rad(k);
    return r;
  }
};

// LDS position of line element i.  es > 0: element stride es (column tiles, lines interleaved, which
// is conflict-free because consecutive lanes take consecutive lines).  es == 0: a contiguous line
// padded by one element every 16, so the strided Stockham writes of the early stages spread over the
// banks instead of hitting the same ones.  es < 0: column tiles of -es lines with the same padding
// (the first stage's writes, 16 elements apart, would otherwise fall on one bank group).
__device__ inline int lidx(int i, int es) { return es > 0 ? i * es : es == 0 ? i + (i >> 4) : (i + (i >> 4)) * -es; }

// Element at a 32-bit byte offset from a uniform base: one VGPR per address (the base stays in SGPRs).
// Every grid and image here is below 4 GiB (fft_plan_init checks).
template <typename T>
__device__ inline T& at_byte(T* base, unsigned off) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(const_cast<std::remove_const_t<T>*>(base)) + off);
}

// Twiddle W_N^m of an N-point line, read from the device table of the longest axis (stride twm).
template <typename R>
struct TwGlobal {
  const cpx<R>* __restrict__ tw;
  int twm;
  __device__ cpx<R> operator()(int m) const { return tw[m * twm]; }
};

// The same from two small LDS tables, W_N^m = W_N^(hi B) * W_N^lo for m = hi B + lo, B = 2^ceil(L/2):
// B + N/B entries (128 at N = 4096) filled once per block, so the stages issue no global loads and a
// line's HBM loads in flight are never waited on by a twiddle read.
template <int L>
constexpr int tw_lds_shift() { return (L + 1) / 2; }
template <int L>
constexpr int tw_lds_entries() { return (1 << tw_lds_shift<L>()) + ((1 << L) >> tw_lds_shift<L>()); }
inline int tw_lds_entries_rt(int L) { return (1 << ((L + 1) / 2)) + ((1 << L) >> ((L + 1) / 2)); }

template <typename R, int L>
struct TwLds {
  const cpx<R>* lo;  // W_N^lo, lo < B
  const cpx<R>* hi;  // W_N^(hi B), hi < N / B
  __device__ cpx<R> operator()(int m) const {
    constexpr int SH = tw_lds_shift<L>();
    return cmul(hi[m >> SH], lo[m & ((1 << SH) - 1)]);
  }
};

// Fills the tables at tab from the device table (stride twm per N-point step); ends with a barrier.
template <typename R, int L>
__device__ inline TwLds<R, L> tw_lds_fill(cpx<R>* tab, const cpx<R>* __restrict__ tw, int twm) {
  constexpr int SH = tw_lds_shift<L>(), B = 1 << SH, H = (1 << L) >> SH;
  for (int i = threadIdx.x; i < B + H; i += blockDim.x) tab[i] = i < B ? tw[i * twm] : tw[((i - B) << SH) * twm];
  __syncthreads();
  return TwLds<R, L>{tab, tab + B};
}

template <typename R, int E, int RAD, int NS, int N, bool INV, class Tw>
__device__ inline void stage_bfly(cpx<R> (&v)[E], int t, const Tw& tw) {
  constexpr int T = N / E;
#pragma unroll
  for (int b = 0; b < E / RAD; ++b) {
    if constexpr (NS > 1) {
      const int j = t + b * T;
      const int k = j & (NS - 1);
      const int step = k * (N / (NS * RAD));
#pragma unroll
      for (int r = 1; r < RAD; ++r) {
        const cpx<R> w = tw(step * r);
        v[b * RAD + r] = INV ? cmulc(v[b * RAD + r], w) : cmul(v[b * RAD + r], w);
      }
    }
    DFT<R, RAD, INV>::run(&v[b * RAD]);
  }
}

// The first stage's inputs of thread t: register b RAD + r holds element t + b T + r N / RAD.
template <int L, int EE = FFT_E>
__device__ inline int line_index(int t, int e) {
  using P = CPlan<L, EE>;
  constexpr int N = P::N, T = P::T, RAD = P::rad(0);
  return t + (e / RAD) * T + (e % RAD) * (N / RAD);
}
// The last stage's output position of register e of thread t.
template <int L, int EE = FFT_E>
__device__ inline int store_index(int t, int e) {
  using P = CPlan<L, EE>;
  constexpr int S = P::NST - 1, RAD = P::rad(S), NS = P::ns(S), T = P::T;
  const int b = e / RAD, r = e % RAD;
  const int j = t + b * T, k = j & (NS - 1);
  return (j - k) * RAD + k + r * NS;
}
template <typename R, int L, int EE, class Load>
__device__ inline void line_load(cpx<R> (&v)[(CPlan<L, EE>::E)], int t, Load& load) {
#pragma unroll
  for (int e = 0; e < CPlan<L, EE>::E; ++e) v[e] = load(line_index<L, EE>(t, e));
}

// One stage of a 2^L-point line: (load from HBM | read LDS) -> butterflies -> (store | write LDS),
// then the next stage.  PRE: the first stage's inputs are already in v (line_load).
template <typename R, int L, int EE, int S, bool INV, bool PRE, class Tw, class Load, class Store>
__device__ inline void stages(cpx<R> (&v)[(CPlan<L, EE>::E)], int t, cpx<R>* lds, int es, const Tw& tw, Load& load,
                              Store& store) {
  using P = CPlan<L, EE>;
  constexpr int N = P::N, E = P::E, T = P::T, RAD = P::rad(S), NS = P::ns(S);
  if constexpr (S == 0) {
    if constexpr (!PRE) line_load<R, L, EE>(v, t, load);
  } else {
    __syncthreads();  // the previous stage's LDS writes are visible
#pragma unroll
    for (int b = 0; b < E / RAD; ++b)
#pragma unroll
      for (int r = 0; r < RAD; ++r) v[b * RAD + r] = lds[lidx(t + b * T + r * (N / RAD), es)];
  }
  stage_bfly<R, E, RAD, NS, N, INV>(v, t, tw);
  if constexpr (S == P::NST - 1) {
#pragma unroll
    for (int b = 0; b < E / RAD; ++b) {
      const int j = t + b * T, k = j & (NS - 1), o = (j - k) * RAD + k;
#pragma unroll
      for (int r = 0; r < RAD; ++r) store(o + r * NS, v[b * RAD + r], b * RAD + r);
    }
  } else {
    __syncthreads();  // every thread has read this stage's inputs before anyone overwrites them
#pragma unroll
    for (int b = 0; b < E / RAD; ++b) {
      const int j = t + b * T, k = j & (NS - 1), o = (j - k) * RAD + k;
#pragma unroll
      for (int r = 0; r < RAD; ++r) lds[lidx(o + r * NS, es)] = v[b * RAD + r];
    }
    stages<R, L, EE, S + 1, INV, PRE>(v, t, lds, es, tw, load, store);
  }
}

template <typename R, int L, int EE, bool INV, class Tw, class Load, class Store>
__device__ inline void fft_line(int t, cpx<R>* lds, int es, const Tw& tw, Load& load, Store& store) {
  cpx<R> v[(CPlan<L, EE>::E)];
  stages<R, L, EE, 0, INV, false>(v, t, lds, es, tw, load, store);
}
// The same with the first stage's inputs already loaded into v.
template <typename R, int L, int EE, bool INV, class Tw, class Store>
__device__ inline void fft_line_pre(cpx<R> (&v)[(CPlan<L, EE>::E)], int t, cpx<R>* lds, int es, const Tw& tw,
                                    Store& store) {
  auto none = [](int) { return cpx<R>{}; };
  stages<R, L, EE, 0, INV, true>(v, t, lds, es, tw, none, store);
}

// Calls F.template operator()<L>() for the runtime log2 length L in [1, 13].
template <class F>
static void dispatch_log2(int L, F&& f) {
  switch (L) {
    case 1: f.template operator()<1>(); break;
    case 2: f.template operator()<2>(); break;
    case 3: f.template operator()<3>(); break;
    case 4: f.template operator()<4>(); break;
    case 5: f.template operator()<5>(); break;
    case 6: f.template operator()<6>(); break;
    case 7: f.template operator()<7>(); break;
    case 8: f.template operator()<8>(); break;
    case 9: f.template operator()<9>(); break;
    case 10: f.template operator()<10>(); break;
    case 11: f.template operator()<11>(); break;
    case 12: f.template operator()<12>(); break;
    case 13: f.template operator()<13>(); break;
    default: throw Error(-3, "FFT length out of range");
  }
}

// ---------------------------------------------------------------------------------------------
// Row pass
// ---------------------------------------------------------------------------------------------

template <int L>
constexpr int row_threads() {
  return CPlan<L>::T > 256 ? CPlan<L>::T : 256;
}

// Lines are taken with a stride of gridDim.x * lpb: a grid smaller than the line count makes each
// block loop, and the next line's first-stage loads are issued before the current line's stages, so
// the HBM reads of one line overlap the LDS work of the previous one.
// Waves per SIMD the row kernel is compiled for: 3 blocks of 4 waves per CU in complex64 (at 4 the
// current and prefetched lines and S's column factors spill a few registers; 3 measured 1-1.5 us faster
// per 4096 x 4096 pass) and 2 in complex128.
template <typename R>
constexpr int row_waves_per_simd() { return sizeof(R) == 4 ? 3 : 2; }

// BT: a batch launch, whose blockIdx.z selects the arguments from tab (one entry per grid of the batch)
template <typename R, int L, bool INV, int MODE, bool BT = false>
__global__ void __launch_bounds__(row_threads<L>(), row_waves_per_simd<R>())
    k_fft_rows(RowArgs<R> a0, const RowArgs<R>* __restrict__ tab) {
  const RowArgs<R> a = BT ? tab[blockIdx.z] : a0;
  extern __shared__ unsigned char smem_raw[];
  cpx<R>* smem = reinterpret_cast<cpx<R>*>(smem_raw);
  constexpr int N = CPlan<L>::N, T = CPlan<L>::T, E = CPlan<L>::E;
  constexpr bool SZM = MODE == ROW_FWD_SZ || MODE == ROW_ADJ_CROP || MODE == ROW_ADJ_X;  // S's column factors
  constexpr bool SYM = MODE == ROW_FWD_SZ || MODE == ROW_ADJ_CROP;  // S's row factors applied here
  constexpr bool CROP = MODE == ROW_ADJ_CROP || MODE == ROW_ADJ_X;  // output cropped to the image columns
  constexpr bool PAD = MODE == ROW_FWD_SZ || MODE == ROW_FWD_X;     // input zero-padded from the image width
  const int l = threadIdx.x / T, t = threadIdx.x - l * T;
  cpx<R>* lds = smem + l * (N + N / 16);
  const TwLds<R, L> tw = tw_lds_fill<R, L>(smem + a.lpb * (N + N / 16), a.tw, a.twm);
  const SZ<R>& sz = a.sz;
  // S's column factors of this lane's elements, the same on every line: at the first stage's load
  // positions (forward) or the last stage's store positions (adjoint); 0 outside the image columns
  R sxr[SZM ? E : 1];
  if constexpr (SZM) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = (MODE == ROW_FWD_SZ ? line_index<L>(t, e) : store_index<L>(t, e)) - sz.x0;
      sxr[e] = (unsigned)i < (unsigned)sz.nx ? sz.sx[i] : R(0);
    }
  }
  // row base pointers, then 32-bit unsigned element offsets (one VGPR per address)
  auto loader = [&](int line) {
    const bool valid = line < a.nlines;
    const cpx<R>* row = PAD                    ? sz.img_in + (int64_t)line * sz.nx
                        : MODE == ROW_ADJ_CROP ? a.in + (int64_t)(sz.y0 + line) * N
                                               : a.in + (int64_t)line * N;  // ROW_PLAIN, ROW_ADJ_X
    return [=, &sz](int idx) -> cpx<R> {
      if (!valid) return mk<R>(R(0), R(0));
      if constexpr (PAD) {  // the image-width row, zero-padded
        const unsigned i = (unsigned)(idx - sz.x0);
        if (i >= (unsigned)sz.nx) return mk<R>(R(0), R(0));
        return row[i];
      } else {
        return row[(unsigned)idx];
      }
    };
  };
  const int stride = gridDim.x * a.lpb;
  const int first = blockIdx.x * a.lpb;  // the block's first line: the loop bound is block-uniform
  // the next line's elements and S row factor are loaded together, before the current line's stages;
  // the stages read only LDS (lines and twiddles), so nothing waits on them until the next iteration
  cpx<R> nxt[E];
  R nsy = R(0);
  {
    auto ld = loader(first + l);
    line_load<R, L, FFT_E>(nxt, t, ld);
    if constexpr (SYM)
      if (first + l < a.nlines) nsy = sz.sy[first + l];
  }
  for (int base = first; base < a.nlines; base += stride) {
    // an opaque copy of the lane index keeps the address arithmetic inside the loop (hoisted, every
    // stage's addresses would stay live in registers across it)
    int tl = t;
    asm volatile("" : "+v"(tl));
    const int line = base + l;
    const bool valid = line < a.nlines;
    cpx<R> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = nxt[e];
    const R sy = nsy;
    if (base + stride < a.nlines) {
      auto ld = loader(line + stride);
      line_load<R, L, FFT_E>(nxt, tl, ld);
      if constexpr (SYM)
        if (line + stride < a.nlines) nsy = sz.sy[line + stride];
    }
    if constexpr (MODE == ROW_FWD_SZ) {  // S: sy[line] sx[i] and the folded scale
      const R sl = sy * sz.scale;
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = scale(v[e], sl * sxr[e]);
    }
    cpx<R>* orow = MODE == ROW_FWD_SZ     ? a.out + (int64_t)(sz.y0 + line) * N
                   : MODE == ROW_ADJ_CROP ? sz.img_out + (int64_t)line * sz.nx
                   : MODE == ROW_ADJ_X    ? a.out + (int64_t)line * sz.nx
                                          : a.out + (int64_t)line * N;
    auto store = [&](int idx, cpx<R> val, int e) {
      if (!valid) return;
      if constexpr (CROP) {
        const unsigned i = (unsigned)(idx - sz.x0);
        if (i >= (unsigned)sz.nx) return;
        const R s = MODE == ROW_ADJ_X ? sxr[e] * sz.scale : sy * sxr[e] * sz.scale;
        orow[i] = scale(val, s);
      } else {
        orow[(unsigned)idx] = val;
      }
    };
    fft_line_pre<R, L, FFT_E, INV>(v, tl, lds, 0, tw, store);
  }
}

template <typename R>
struct RowLaunch {
  const RowArgs<R>& a;
  int mode;
  bool inv;
  int threads, blocks;
  size_t shm;
  hipStream_t s;
  template <int L, bool I, int MODE>
  void go() {
    if (record_launch<RowArgs<R>, k_fft_rows<R, L, I, MODE, true>>(dim3(blocks), dim3(threads), shm, false, a)) return;
    void (*k)(RowArgs<R>, const RowArgs<R>*) = k_fft_rows<R, L, I, MODE>;
    allow_lds((const void*)k, shm);
    hipLaunchKernelGGL(k, dim3(blocks), dim3(threads), shm, s, a, nullptr);
  }
  template <int L>
  void operator()() {
    if (mode == ROW_FWD_SZ) go<L, false, ROW_FWD_SZ>();
    else if (mode == ROW_ADJ_CROP) go<L, true, ROW_ADJ_CROP>();
    else if (mode == ROW_ADJ_X) go<L, true, ROW_ADJ_X>();
    else if (mode == ROW_FWD_X) go<L, false, ROW_FWD_X>();
    else if (inv) go<L, true, ROW_PLAIN>();
    else go<L, false, ROW_PLAIN>();
  }
};

template <typename R>
void fft_rows(const Fft2dPlan& p, int mode, bool inv, const cpx<R>* in, cpx<R>* out, int nlines, const SZ<R>& sz,
              hipStream_t s) {
  RowArgs<R> a;
  a.in = in;
  a.out = out;
  a.tw = static_cast<const cpx<R>*>(p.tw);
  a.twm = p.Lmax / p.Nu;
  a.Nu = p.Nu;
  a.nlines = nlines;
  a.sz = sz;
  const int N = p.Nu;
  const int E = N < FFT_E ? N : FFT_E;
  const int T = N / E;
  int lpb = 256 / T;
  if (lpb < 1) lpb = 1;
  if (lpb > nlines) lpb = nlines;
  a.lpb = lpb;
  const int threads = lpb * T;
  const size_t shm = ((size_t)lpb * (N + N / 16) + tw_lds_entries_rt(p.row.log2N)) * sizeof(cpx<R>);
  int blocks = (nlines + lpb - 1) / lpb;
  if (p.row_bpc > 0) blocks = std::min(blocks, p.row_bpc * p.n_cu);
  if (blocks == 0) return;
  ProfScope prof(mode == ROW_FWD_SZ     ? "fft_rows_fwd_sz"
                 : mode == ROW_ADJ_CROP ? "fft_rows_adj_crop"
                 : mode == ROW_ADJ_X    ? "fft_rows_adj_x"
                 : mode == ROW_FWD_X    ? "fft_rows_fwd_x"
                                        : "fft_rows",
                 s);
  if (((mode == ROW_FWD_SZ || mode == ROW_FWD_X) && inv) || ((mode == ROW_ADJ_CROP || mode == ROW_ADJ_X) && !inv))
    throw Error(-1, "fused row pass direction mismatch");
  RowLaunch<R> launch{a, mode, inv, threads, blocks, shm, s};
  dispatch_log2(p.row.log2N, launch);
  PF_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------
// Column passes
// ---------------------------------------------------------------------------------------------
enum ColKind { COL_SINGLE = 0, COL_STEP_A = 1, COL_STEP_B = 2 };
// Column-pass input and output forms (compile-time, so that no branch sits among the loads): plain;
// PAD: window row y read from row y - y0 of an image, times isy[y - y0] isx[col] iscale (rin); CROP:
// window row y written at row y - y0, times osy[y - y0] (rout)
enum ColIo { COL_IO_PLAIN = 0, COL_IO_PAD = 1, COL_IO_CROP = 2 };

template <typename R, int L, bool INV, int KIND, bool BT = false, int IO = COL_IO_PLAIN>
__global__ void __launch_bounds__(512) k_fft_cols(ColArgs<R> a0, const ColArgs<R>* __restrict__ tab) {
  const ColArgs<R> a = BT ? tab[blockIdx.z] : a0;
  extern __shared__ unsigned char smem_raw[];
  cpx<R>* smem = reinterpret_cast<cpx<R>*>(smem_raw);
  const int C = a.C;
  const int c = threadIdx.x % C, t = threadIdx.x / C;
  const int col = blockIdx.x * C + c;
  const int by = blockIdx.y;  // r2 for step A, k1 for step B
  const int64_t Nu = a.Nu;
  auto in_window = [&](int row) { return (unsigned)(row - a.y0) < (unsigned)a.ny; };
  auto load = [&](int idx) -> cpx<R> {
    int row;
    if constexpr (KIND == COL_SINGLE) row = idx;
    else if constexpr (KIND == COL_STEP_A) row = a.N2 * idx + by;
    else row = a.N2 * by + idx;
    if (a.rin && !in_window(row)) return mk<R>(R(0), R(0));
    if constexpr (IO == COL_IO_PAD) {
      const int r = row - a.y0;
      return scale(a.in[r * Nu + col], a.isy[r] * (a.isx[col] * a.iscale));
    }
    return a.in[row * Nu + col];
  };
  auto store = [&](int idx, cpx<R> val, int) {
    int row;
    if constexpr (KIND == COL_SINGLE) {
      row = idx;
    } else if constexpr (KIND == COL_STEP_A) {
      row = a.N2 * idx + by;
      const cpx<R> w = a.tw[(by * idx) * a.twm_full];
      val = INV ? cmulc(val, w) : cmul(val, w);
    } else {
      row = by + a.N1 * idx;
    }
    if (a.rout && !in_window(row)) return;
    if constexpr (IO == COL_IO_CROP) {
      row -= a.y0;
      val = scale(val, a.osy[row]);
    }
    a.out[row * Nu + col] = val;
  };
  const TwGlobal<R> tw{a.tw, a.twm};
  fft_line<R, L, FFT_E, INV>(t, smem + c, C, tw, load, store);
}

// Single-pass columns of 1024..8192 rows: a block owns C adjacent columns (C fixed by the element size
// and L: padded LDS tiles and the twiddle tables in 160 KiB, at most 1024 threads, 512 in FP64 for
// room for 256 VGPRs).  The kernel walks column groups with a grid stride, the next group's
// first-stage loads issued before the current group's stages; it is launched with one block per
// group, which measured faster at 4096 x 4096 than a persistent grid (74.6 against 80.6 us in
// complex64: with one 140 KiB block per CU, the CU's next block overlaps better than the prefetch).
// With the group count and the grid both multiples of 8, the blocks of one XCD (workgroups go
// round-robin over the 8 XCDs) take one contiguous band of groups, so that its concurrently running
// blocks share the rows' 128-byte lines in its L2.
// Elements per thread of that pass: 32 in complex64 (512 threads hold four 4096-point columns, with
// registers for the prefetched group beside the current one), 16 in complex128.
constexpr int big_col_e(int esz) { return esz == 16 ? FFT_E : 2 * FFT_E; }
constexpr int big_col_c(int esz, int L) {
  const int N = 1 << L, T = N / big_col_e(esz);
  const int max_threads = esz == 16 ? 512 : 1024;
  const int tab = (1 << ((L + 1) / 2)) + (N >> ((L + 1) / 2));
  int C = 8;
  while (C > 1 && ((N + N / 16) * C * esz + tab * esz > 160 * 1024 || C * T > max_threads)) C /= 2;
  return C;
}

template <typename R, int L>
constexpr int big_col_threads() { return big_col_c(sizeof(cpx<R>), L) * ((1 << L) / big_col_e(sizeof(cpx<R>))); }

template <typename R, int L, bool INV, bool BT = false, int IO = COL_IO_PLAIN>
__global__ void __launch_bounds__((big_col_threads<R, L>()))
    k_fft_cols_big(ColArgs<R> a0, const ColArgs<R>* __restrict__ tab) {
  const ColArgs<R> a = BT ? tab[blockIdx.z] : a0;
  extern __shared__ unsigned char smem_raw[];
  cpx<R>* smem = reinterpret_cast<cpx<R>*>(smem_raw);
  constexpr int EE = big_col_e(sizeof(cpx<R>));
  constexpr int C = big_col_c(sizeof(cpx<R>), L), N = CPlan<L, EE>::N, E = CPlan<L, EE>::E;
  const int c = threadIdx.x % C, t = threadIdx.x / C;
  // PAD / CROP: S's row factors in LDS behind the twiddles (filled before tw_lds_fill's barrier), so
  // that scaling a loaded line or a stored value reads LDS rather than waiting on further global loads
  R* lsy = reinterpret_cast<R*>(smem + C * (N + N / 16) + tw_lds_entries<L>());
  if constexpr (IO != COL_IO_PLAIN) {
    const R* fac = IO == COL_IO_PAD ? a.isy : a.osy;
    for (int i = threadIdx.x; i < a.ny; i += blockDim.x) lsy[i] = fac[i];
  }
  const TwLds<R, L> tw = tw_lds_fill<R, L>(smem + C * (N + N / 16), a.tw, a.twm);
  const unsigned Nu = (unsigned)a.Nu;
  const int G = a.Nu / C, P = gridDim.x;
  const bool band = !(G & 7) && !(P & 7);
  const int q = band ? (int)(blockIdx.x >> 3) : (int)blockIdx.x;
  const int qstep = band ? P >> 3 : P, qend = band ? G >> 3 : G;
  const int gbase = band ? (int)(blockIdx.x & 7) * (G >> 3) : 0;
  auto in_window = [&](int row) { return (unsigned)(row - a.y0) < (unsigned)a.ny; };
  // 32-bit unsigned element offsets from the column's base (the grid holds < 2^32 cells)
  auto load_group = [&](cpx<R> (&v)[E], int g, int tl) {
    const unsigned col = (unsigned)(g * C + c);
    auto ld = [&](int row) -> cpx<R> {
      if (a.rin && !in_window(row)) return mk<R>(R(0), R(0));
      // PAD: window row y is the image's row y - y0
      const unsigned r = IO == COL_IO_PAD ? (unsigned)(row - a.y0) : (unsigned)row;
      return at_byte(a.in, (r * Nu + col) * (unsigned)sizeof(cpx<R>));
    };
    line_load<R, L, EE>(v, tl, ld);
  };
  // PAD: the group's column factor isx[col] iscale, loaded with the group
  auto col_factor = [&](int g) -> R {
    if constexpr (IO == COL_IO_PAD) return a.isx[g * C + c] * a.iscale;
    else return R(1);
  };
  cpx<R> nxt[E];
  R ncs = R(1);
  if (q < qend) {
    load_group(nxt, gbase + q, t);
    ncs = col_factor(gbase + q);
  }
  for (int k = q; k < qend; k += qstep) {
    // an opaque copy of the lane index keeps the address arithmetic inside the loop
    int tl = t;
    asm volatile("" : "+v"(tl));
    cpx<R> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = nxt[e];
    const R cs = ncs;
    if (k + qstep < qend) {
      load_group(nxt, gbase + k + qstep, tl);
      ncs = col_factor(gbase + k + qstep);
    }
    if constexpr (IO == COL_IO_PAD) {  // S on the loaded line (rows outside the window are zero already)
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int r = min(max(line_index<L, EE>(tl, e) - a.y0, 0), a.ny - 1);
        v[e] = scale(v[e], lsy[r] * cs);
      }
    }
    const unsigned col = (unsigned)((gbase + k) * C + c);
    auto store = [&](int row, cpx<R> val, int) {
      if (a.rout && !in_window(row)) return;
      if constexpr (IO == COL_IO_CROP) {
        row -= a.y0;
        val = scale(val, lsy[row]);
      }
      at_byte(a.out, ((unsigned)row * Nu + col) * (unsigned)sizeof(cpx<R>)) = val;
    };
    fft_line_pre<R, L, EE, INV>(v, tl, smem + c, -C, tw, store);
  }
}

template <typename R>
struct BigColLaunch {
  const ColArgs<R>& a;
  bool inv;
  hipStream_t s;
  int max_blocks = 0;  // > 0: at most this many blocks, each walking its groups with the next one prefetched
  template <int L, bool I, int IO>
  void go() {
    constexpr int C = big_col_c(sizeof(cpx<R>), L), N = 1 << L;
    const size_t shm = ((size_t)(N + N / 16) * C + tw_lds_entries<L>()) * sizeof(cpx<R>) +
                       (IO != COL_IO_PLAIN ? (size_t)a.ny * sizeof(R) : 0);
    const int groups = a.Nu / C;
    const dim3 grid((unsigned)(max_blocks > 0 ? std::min(groups, max_blocks) : groups)), block(big_col_threads<R, L>());
    if (record_launch<ColArgs<R>, k_fft_cols_big<R, L, I, true, IO>>(grid, block, shm, false, a)) return;
    void (*k)(ColArgs<R>, const ColArgs<R>*) = k_fft_cols_big<R, L, I, false, IO>;
    allow_lds((const void*)k, shm);
    hipLaunchKernelGGL(k, grid, block, shm, s, a, nullptr);
  }
  template <int L>
  void operator()() {
    if constexpr (L >= 10 && L <= 13) {
      // the padded input form is a forward pass's, the cropped output form an inverse pass's
      if (inv) {
        if (a.osy) go<L, true, COL_IO_CROP>();
        else go<L, true, COL_IO_PLAIN>();
      } else {
        if (a.isy) go<L, false, COL_IO_PAD>();
        else go<L, false, COL_IO_PLAIN>();
      }
    } else {
      throw Error(-3, "single-pass column length must be 1024..8192");
    }
  }
};

template <typename R>
struct ColLaunch {
  const ColArgs<R>& a;
  int kind;
  bool inv;
  dim3 grid;
  int threads;
  size_t shm;
  hipStream_t s;
  template <int L, bool I, int KIND, int IO>
  void go_kind() {
    if (record_launch<ColArgs<R>, k_fft_cols<R, L, I, KIND, true, IO>>(grid, dim3(threads), shm, false, a)) return;
    void (*k)(ColArgs<R>, const ColArgs<R>*) = k_fft_cols<R, L, I, KIND, false, IO>;
    allow_lds((const void*)k, shm);
    hipLaunchKernelGGL(k, grid, dim3(threads), shm, s, a, nullptr);
  }
  template <int L, bool I>
  void go() {
    // the padded input form only in a forward single pass, the cropped output form only in an inverse
    // single pass or second step
    if constexpr (!I) {
      if (a.isy && kind == COL_SINGLE) return go_kind<L, I, COL_SINGLE, COL_IO_PAD>();
    } else {
      if (a.osy && kind == COL_SINGLE) return go_kind<L, I, COL_SINGLE, COL_IO_CROP>();
      if (a.osy && kind == COL_STEP_B) return go_kind<L, I, COL_STEP_B, COL_IO_CROP>();
    }
    if (a.isy || a.osy) throw Error(-1, "column pass form not instantiated");
    if (kind == COL_STEP_A) go_kind<L, I, COL_STEP_A, COL_IO_PLAIN>();
    else if (kind == COL_STEP_B) go_kind<L, I, COL_STEP_B, COL_IO_PLAIN>();
    else go_kind<L, I, COL_SINGLE, COL_IO_PLAIN>();
  }
  template <int L>
  void operator()() {
    if constexpr (L <= 9) {
      if (inv) go<L, true>();
      else go<L, false>();
    } else {
      throw Error(-3, "column line length above 512");
    }
  }
};

template <typename R>
static void launch_cols(int kind, bool inv, int log2n, const ColArgs<R>& a, dim3 grid, hipStream_t s,
                        const char* name = nullptr) {
  const int N = 1 << log2n;
  const int E = N < FFT_E ? N : FFT_E;
  const int threads = a.C * (N / E);
  const size_t shm = (size_t)N * a.C * sizeof(cpx<R>);
  ProfScope prof(name                  ? name
                 : kind == COL_STEP_A ? "fft_cols_stepA"
                 : kind == COL_STEP_B ? "fft_cols_stepB"
                 : inv                 ? "fft_cols_adj"
                                       : "fft_cols_fwd",
                 s);
  ColLaunch<R> launch{a, kind, inv, grid, threads, shm, s};
  dispatch_log2(log2n, launch);
  PF_HIP(hipGetLastError());
}

template <typename R>
void fft_cols(const Fft2dPlan& p, bool inv, cpx<R>* in, cpx<R>* out, RowWindow win, bool restrict_in,
              bool restrict_out, hipStream_t s) {
  ColArgs<R> a;
  a.in = in;
  a.out = out;
  a.tw = static_cast<const cpx<R>*>(p.tw);
  a.Nu = p.Nu;
  a.Nv = p.Nv;
  a.twm_full = p.Lmax / p.Nv;
  a.y0 = win.y0;
  a.ny = win.ny ? win.ny : p.Nv;
  a.N1 = p.N1;
  a.N2 = p.N2;
  a.osy = nullptr;
  a.isy = nullptr;
  a.isx = nullptr;
  a.iscale = R(1);
  if (p.col_big) {
    a.twm = p.Lmax / p.Nv;
    a.C = p.big_C;
    a.rin = restrict_in;
    a.rout = restrict_out;
    ProfScope prof(inv ? "fft_cols_adj" : "fft_cols_fwd", s);
    BigColLaunch<R> launch{a, inv, s};
    dispatch_log2(p.col.log2N, launch);
    PF_HIP(hipGetLastError());
    return;
  }
  if (!p.col_two_step) {
    a.twm = p.Lmax / p.Nv;
    a.C = p.col_C;
    a.rin = restrict_in;
    a.rout = restrict_out;
    launch_cols<R>(COL_SINGLE, inv, p.col.log2N, a, dim3(p.Nu / a.C), s);
    return;
  }
  a.C = p.two_C;
  // step A in place on `in`: N1-point transforms over rows of stride N2
  a.twm = p.Lmax / p.N1;
  a.out = in;
  a.rin = restrict_in;
  a.rout = 0;
  launch_cols<R>(COL_STEP_A, inv, p.colA.log2N, a, dim3(p.Nu / a.C, p.N2), s);
  // step B `in` -> `out`: N2-point transforms over contiguous row blocks
  a.twm = p.Lmax / p.N2;
  a.in = in;
  a.out = out;
  a.rin = 0;
  a.rout = restrict_out;
  launch_cols<R>(COL_STEP_B, inv, p.colB.log2N, a, dim3(p.Nu / a.C, p.N1), s);
}

// The image-column passes with at most two column groups per compute unit: one block per compute unit,
// each walking its groups with the next group's loads issued before the current group's stages (at
// the 2048-column image, complex64: 43.1 -> 39.2 us forward, 42.9 -> 39.4 us adjoint).  With more
// groups (complex128's 2-column groups: 1024) one block per group is faster (72.8 against 82.5 us).
static int q_cols_blocks(const Fft2dPlan& p, int width) {
  return p.col_big && width / p.big_C <= 2 * p.n_cu ? p.n_cu : 0;
}

bool fft_cols_width_ok(const Fft2dPlan& p, int width) {
  const int C = p.col_big ? p.big_C : p.col_two_step ? p.two_C : p.col_C;
  return width > 0 && width <= p.Nu && width % C == 0;
}

template <typename R>
void fft_cols_crop(const Fft2dPlan& p, bool inv, cpx<R>* in, cpx<R>* out, int width, RowWindow win, const R* osy,
                   hipStream_t s) {
  if (!fft_cols_crop_ok(p, width, win.ny) || !osy || win.ny <= 0)
    throw Error(-1, "cropped column pass does not fit the plan");
  Fft2dPlan q = p;  // the same transforms over `width` columns
  q.Nu = width;
  ColArgs<R> a;
  a.in = in;
  a.out = out;
  a.tw = static_cast<const cpx<R>*>(q.tw);
  a.Nu = q.Nu;
  a.Nv = q.Nv;
  a.twm_full = q.Lmax / q.Nv;
  a.y0 = win.y0;
  a.ny = win.ny;
  a.N1 = q.N1;
  a.N2 = q.N2;
  a.osy = nullptr;
  a.isy = nullptr;
  a.isx = nullptr;
  a.iscale = R(1);
  if (q.col_big) {
    a.twm = q.Lmax / q.Nv;
    a.C = q.big_C;
    a.rin = 0;
    a.rout = 1;
    a.osy = osy;
    ProfScope prof(inv ? "fft_cols_adj_x" : "fft_cols_fwd_x", s);
    BigColLaunch<R> launch{a, inv, s, q_cols_blocks(p, width)};
    dispatch_log2(q.col.log2N, launch);
    PF_HIP(hipGetLastError());
    return;
  }
  if (!q.col_two_step) {
    a.twm = q.Lmax / q.Nv;
    a.C = q.col_C;
    a.rin = 0;
    a.rout = 1;
    a.osy = osy;
    launch_cols<R>(COL_SINGLE, inv, q.col.log2N, a, dim3(q.Nu / a.C), s, "fft_cols_adj_x");
    return;
  }
  a.C = q.two_C;
  a.twm = q.Lmax / q.N1;
  a.out = in;
  a.rin = 0;
  a.rout = 0;
  launch_cols<R>(COL_STEP_A, inv, q.colA.log2N, a, dim3(q.Nu / a.C, q.N2), s, "fft_cols_stepA_x");
  a.twm = q.Lmax / q.N2;
  a.in = in;
  a.out = out;
  a.rout = 1;
  a.osy = osy;
  launch_cols<R>(COL_STEP_B, inv, q.colB.log2N, a, dim3(q.Nu / a.C, q.N1), s, "fft_cols_stepB_x");
}
template void fft_cols_crop<float>(const Fft2dPlan&, bool, cpx<float>*, cpx<float>*, int, RowWindow, const float*,
                                   hipStream_t);
template void fft_cols_crop<double>(const Fft2dPlan&, bool, cpx<double>*, cpx<double>*, int, RowWindow,
                                    const double*, hipStream_t);

// the single-pass kernel's LDS with the ny row factors: padded columns, twiddles and the factors
static bool big_cols_lds_ok(const Fft2dPlan& p, int ny) {
  if (!p.col_big) return true;
  const size_t esz = p.prec ? 16 : 8, N = (size_t)1 << p.col.log2N;
  return ((N + N / 16) * p.big_C + tw_lds_entries_rt(p.col.log2N)) * esz + (size_t)ny * (esz / 2) <= 160 * 1024;
}
bool fft_cols_pad_ok(const Fft2dPlan& p, int width, int ny) {
  return !p.col_two_step && fft_cols_width_ok(p, width) && big_cols_lds_ok(p, ny);
}
bool fft_cols_crop_ok(const Fft2dPlan& p, int width, int ny) {
  return fft_cols_width_ok(p, width) && big_cols_lds_ok(p, ny);
}

template <typename R>
void fft_cols_pad(const Fft2dPlan& p, bool inv, const cpx<R>* in, cpx<R>* out, int width, RowWindow win,
                  const R* isy, const R* isx, R iscale, hipStream_t s) {
  if (!fft_cols_pad_ok(p, width, win.ny) || !isy || !isx || win.ny <= 0)
    throw Error(-1, "padded column pass does not fit the plan");
  ColArgs<R> a;
  a.in = const_cast<cpx<R>*>(in);  // read only: a single-pass plan writes `out` alone
  a.out = out;
  a.tw = static_cast<const cpx<R>*>(p.tw);
  a.Nu = width;
  a.Nv = p.Nv;
  a.twm_full = p.Lmax / p.Nv;
  a.twm = p.Lmax / p.Nv;
  a.y0 = win.y0;
  a.ny = win.ny;
  a.N1 = p.N1;
  a.N2 = p.N2;
  a.rin = 1;
  a.rout = 0;
  a.osy = nullptr;
  a.isy = isy;
  a.isx = isx;
  a.iscale = iscale;
  if (p.col_big) {
    a.C = p.big_C;
    ProfScope prof(inv ? "fft_cols_adj_x" : "fft_cols_fwd_x", s);
    BigColLaunch<R> launch{a, inv, s, q_cols_blocks(p, width)};
    dispatch_log2(p.col.log2N, launch);
    PF_HIP(hipGetLastError());
    return;
  }
  a.C = p.col_C;
  launch_cols<R>(COL_SINGLE, inv, p.col.log2N, a, dim3(width / a.C), s, "fft_cols_fwd_x");
}
template void fft_cols_pad<float>(const Fft2dPlan&, bool, const cpx<float>*, cpx<float>*, int, RowWindow,
                                  const float*, const float*, float, hipStream_t);
template void fft_cols_pad<double>(const Fft2dPlan&, bool, const cpx<double>*, cpx<double>*, int, RowWindow,
                                   const double*, const double*, double, hipStream_t);

// ---------------------------------------------------------------------------------------------
// Plans and twiddles
// ---------------------------------------------------------------------------------------------
static void twiddle_table(int L, std::vector<double>& re, std::vector<double>& im) {
  // exp(-2 pi i m / L) from the first octant by symmetry, so W^m and its mirror images agree
  // bit for bit; sin/cos of arguments <= pi/4 are evaluated in long double.
  re.resize(L);
  im.resize(L);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int m = 0; m < L; ++m) {
    const int64_t num = 8LL * m;  // angle in units of pi/4 is num / L
    const int o = (int)(num / L);
    const int64_t rmd = num - (int64_t)o * L;  // in [0, L)
    long double a;
    bool swap = false;
    if (o % 2 == 0) {
      a = (pi / 4) * (long double)rmd / (long double)L;
    } else {
      a = (pi / 4) * (long double)(L - rmd) / (long double)L;
      swap = true;
    }
    long double c = cosl(a), sn = sinl(a);
    if (swap) {
      long double tmp = c;
      c = sn;
      sn = tmp;
    }
    long double cr, sr;
    switch ((o / 2) % 4) {
      case 0: cr = c; sr = sn; break;
      case 1: cr = -sn; sr = c; break;
      case 2: cr = -c; sr = -sn; break;
      default: cr = sn; sr = -c; break;
    }
    re[m] = (double)cr;
    im[m] = (double)(-sr);
  }
}

void fft_plan_setup(Fft2dPlan& p, int Nv, int Nu, int prec, int col_mode) {
  if (Nv > MAX_AXIS || Nu > MAX_AXIS)
    throw Error(-3, "grid axis larger than " + std::to_string(MAX_AXIS) + " is not supported");
  if ((uint64_t)Nv * Nu * (prec ? 16 : 8) >= ((uint64_t)1 << 32))
    throw Error(-3, "grid of 4 GiB or more is not supported (32-bit byte offsets)");
  p.Nv = Nv;
  p.Nu = Nu;
  p.prec = prec;
  p.Lmax = Nv > Nu ? Nv : Nu;
  p.row = make_line_plan(Nu);
  {
    int dev = 0, ncu = 0;
    PF_HIP(hipGetDevice(&dev));
    PF_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    p.n_cu = ncu > 0 ? ncu : 256;
  }
  // row pass: a grid of 3 (complex64) or 2 (complex128) resident blocks per CU (row_waves_per_simd), each
  // looping over lines with the next line prefetched
  p.row_bpc = prec ? 2 : 3;
  const size_t esz = prec ? 16 : 8;
  // single pass while N * C * esz <= 64 KiB with C >= 16 (fp32) / 8 (fp64) columns
  const int Cmin = prec ? 8 : 16;
  if ((size_t)Nv * Cmin * esz <= 65536 || Nv <= 32) {
    p.col_two_step = false;
    p.col = make_line_plan(Nv);
    int C = (int)(65536 / ((size_t)Nv * esz));
    const int E = Nv < FFT_E ? Nv : FFT_E;
    while ((size_t)C * (Nv / E) > 512) C /= 2;  // threads per block
    if (C > 64) C = 64;
    if (C > Nu) C = Nu;
    if (C < 1) C = 1;
    p.col_C = C;
  } else if (col_mode == 0 && Nv <= 8192 && Nu >= 8) {
    // one pass: C columns of Nv rows in <= 160 KiB of padded LDS and <= 1024 threads
    p.col_two_step = false;
    p.col_big = true;
    p.col = make_line_plan(Nv);
    p.big_C = big_col_c((int)esz, p.col.log2N);  // the kernel's own compile-time choice
  } else {
    p.col_two_step = true;
    int m = 0;
    while ((1 << m) < Nv) ++m;
    p.N2 = 1 << (m / 2);
    p.N1 = Nv / p.N2;
    p.colA = make_line_plan(p.N1);
    p.colB = make_line_plan(p.N2);
    int C = prec ? 32 : 64;
    while ((size_t)p.N1 * C * esz > 65536 && C > 8) C /= 2;
    if (C > Nu) C = Nu;
    p.two_C = C;
  }
}

void fft_plan_init(Fft2dPlan& p, int Nv, int Nu, int prec, int col_mode) {
  fft_plan_setup(p, Nv, Nu, prec, col_mode);
  std::vector<double> re, im;
  twiddle_table(p.Lmax, re, im);
  if (prec) {
    std::vector<cpx<double>> h(p.Lmax);
    for (int m = 0; m < p.Lmax; ++m) h[m] = mk<double>(re[m], im[m]);
    p.tw = dev_alloc(sizeof(cpx<double>) * p.Lmax);
    h2d(p.tw, h.data(), sizeof(cpx<double>) * p.Lmax);
  } else {
    std::vector<cpx<float>> h(p.Lmax);
    for (int m = 0; m < p.Lmax; ++m) h[m] = mk<float>((float)re[m], (float)im[m]);
    p.tw = dev_alloc(sizeof(cpx<float>) * p.Lmax);
    h2d(p.tw, h.data(), sizeof(cpx<float>) * p.Lmax);
  }
}

void fft_plan_free(Fft2dPlan& p) {
  dev_free(p.tw);
  p.tw = nullptr;
}

template void fft_rows<float>(const Fft2dPlan&, int, bool, const cpx<float>*, cpx<float>*, int, const SZ<float>&,
                              hipStream_t);
template void fft_rows<double>(const Fft2dPlan&, int, bool, const cpx<double>*, cpx<double>*, int,
                               const SZ<double>&, hipStream_t);
template void fft_cols<float>(const Fft2dPlan&, bool, cpx<float>*, cpx<float>*, RowWindow, bool, bool, hipStream_t);
template void fft_cols<double>(const Fft2dPlan&, bool, cpx<double>*, cpx<double>*, RowWindow, bool, bool,
                               hipStream_t);

}  // namespace pf
