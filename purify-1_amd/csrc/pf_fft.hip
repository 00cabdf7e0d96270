// pf_fft.hip -- 2-D FFT F on the oversampled grid (replaces FFTW in cpp/purify/operators.h:209-261).
//
// Each 1-D line is a Stockham auto-sort transform (Govindaraju et al., "High Performance Discrete
// Fourier Transforms on Graphics Processors", SC 2008): stage s with radix R and sub-transform
// length Ns reads x[j + r N/R], twiddles by W^(r (j mod Ns) N / (Ns R)), does an R-point DFT in
// registers and writes x[(j - j mod Ns) R + j mod Ns + r Ns].  Every thread holds 16 elements; the
// first stage loads straight from HBM and the last stores straight to HBM, the stages in between
// exchange through LDS.  The stage plan is fixed at compile time by log2 N (radix 16, plus one
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

// Compile-time stage plan of a 2^L-point line (the same plan make_line_plan describes).
template <int L>
struct CPlan {
  static constexpr int N = 1 << L;
  static constexpr int E = N < FFT_E ? N : FFT_E;  // elements per thread
  static constexpr int T = N / E;                   // threads per line
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

template <typename R, int E, int RAD, int NS, int N, bool INV>
__device__ inline void stage_bfly(cpx<R> (&v)[E], int t, const cpx<R>* __restrict__ tw, int twm) {
  constexpr int T = N / E;
#pragma unroll
  for (int b = 0; b < E / RAD; ++b) {
    if constexpr (NS > 1) {
      const int j = t + b * T;
      const int k = j & (NS - 1);
      const int step = k * (N / (NS * RAD)) * twm;
#pragma unroll
      for (int r = 1; r < RAD; ++r) {
        const cpx<R> w = tw[step * r];
        v[b * RAD + r] = INV ? cmulc(v[b * RAD + r], w) : cmul(v[b * RAD + r], w);
      }
    }
    DFT<R, RAD, INV>::run(&v[b * RAD]);
  }
}

// The first stage's inputs of thread t: register b RAD + r holds element t + b T + r N / RAD.
template <int L>
__device__ inline int line_index(int t, int e) {
  using P = CPlan<L>;
  constexpr int N = P::N, T = P::T, RAD = P::rad(0);
  return t + (e / RAD) * T + (e % RAD) * (N / RAD);
}
template <typename R, int L, class Load>
__device__ inline void line_load(cpx<R> (&v)[CPlan<L>::E], int t, Load& load) {
#pragma unroll
  for (int e = 0; e < CPlan<L>::E; ++e) v[e] = load(line_index<L>(t, e));
}

// One stage of a 2^L-point line: (load from HBM | read LDS) -> butterflies -> (store | write LDS),
// then the next stage.  PRE: the first stage's inputs are already in v (line_load).
template <typename R, int L, int S, bool INV, bool PRE, class Load, class Store>
__device__ inline void stages(cpx<R> (&v)[CPlan<L>::E], int t, cpx<R>* lds, int es, const cpx<R>* tw, int twm,
                              Load& load, Store& store) {
  using P = CPlan<L>;
  constexpr int N = P::N, E = P::E, T = P::T, RAD = P::rad(S), NS = P::ns(S);
  if constexpr (S == 0) {
    if constexpr (!PRE) line_load<R, L>(v, t, load);
  } else {
    __syncthreads();  // the previous stage's LDS writes are visible
#pragma unroll
    for (int b = 0; b < E / RAD; ++b)
#pragma unroll
      for (int r = 0; r < RAD; ++r) v[b * RAD + r] = lds[lidx(t + b * T + r * (N / RAD), es)];
  }
  stage_bfly<R, E, RAD, NS, N, INV>(v, t, tw, twm);
  if constexpr (S == P::NST - 1) {
#pragma unroll
    for (int b = 0; b < E / RAD; ++b) {
      const int j = t + b * T, k = j & (NS - 1), o = (j - k) * RAD + k;
#pragma unroll
      for (int r = 0; r < RAD; ++r) store(o + r * NS, v[b * RAD + r]);
    }
  } else {
    __syncthreads();  // every thread has read this stage's inputs before anyone overwrites them
#pragma unroll
    for (int b = 0; b < E / RAD; ++b) {
      const int j = t + b * T, k = j & (NS - 1), o = (j - k) * RAD + k;
#pragma unroll
      for (int r = 0; r < RAD; ++r) lds[lidx(o + r * NS, es)] = v[b * RAD + r];
    }
    stages<R, L, S + 1, INV, PRE>(v, t, lds, es, tw, twm, load, store);
  }
}

template <typename R, int L, bool INV, class Load, class Store>
__device__ inline void fft_line(int t, cpx<R>* lds, int es, const cpx<R>* tw, int twm, Load& load, Store& store) {
  cpx<R> v[CPlan<L>::E];
  stages<R, L, 0, INV, false>(v, t, lds, es, tw, twm, load, store);
}
// The same with the first stage's inputs already loaded into v.
template <typename R, int L, bool INV, class Store>
__device__ inline void fft_line_pre(cpx<R> (&v)[CPlan<L>::E], int t, cpx<R>* lds, int es, const cpx<R>* tw, int twm,
                                    Store& store) {
  auto none = [](int) { return cpx<R>{}; };
  stages<R, L, 0, INV, true>(v, t, lds, es, tw, twm, none, store);
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
template <typename R>
struct RowArgs {
  const cpx<R>* in;
  cpx<R>* out;
  const cpx<R>* tw;
  int twm, Nu, nlines, lpb;
  SZ<R> sz;
};

template <int L>
constexpr int row_threads() {
  return CPlan<L>::T > 256 ? CPlan<L>::T : 256;
}

// Lines are taken with a stride of gridDim.x * lpb: a grid smaller than the line count makes each
// block loop, and the next line's first-stage loads are issued before the current line's stages, so
// the HBM reads of one line overlap the LDS work of the previous one.
// Waves per SIMD the row kernel is compiled for: the LDS of a 4096-point line allows 4 blocks of 4
// waves per CU in complex64 and 2 in complex128, so registers are budgeted for that occupancy (the
// current and the prefetched line in registers).
template <typename R>
constexpr int row_waves_per_simd() { return sizeof(R) == 4 ? 4 : 2; }

template <typename R, int L, bool INV, int MODE>
__global__ void __launch_bounds__(row_threads<L>(), row_waves_per_simd<R>()) k_fft_rows(RowArgs<R> a) {
  extern __shared__ unsigned char smem_raw[];
  cpx<R>* smem = reinterpret_cast<cpx<R>*>(smem_raw);
  constexpr int N = CPlan<L>::N, T = CPlan<L>::T, E = CPlan<L>::E;
  const int l = threadIdx.x / T, t = threadIdx.x - l * T;
  cpx<R>* lds = smem + l * (N + N / 16);
  const SZ<R>& sz = a.sz;
  auto loader = [&](int line) {
    const bool valid = line < a.nlines;
    return [&, line, valid](int idx) -> cpx<R> {
      if (!valid) return mk<R>(R(0), R(0));
      if constexpr (MODE == ROW_FWD_SZ) {  // the image row, zero-padded; S is applied after the prefetch
        const int i = idx - sz.x0;
        if ((unsigned)i >= (unsigned)sz.nx) return mk<R>(R(0), R(0));
        return sz.img_in[(int64_t)line * sz.nx + i];
      } else if constexpr (MODE == ROW_ADJ_CROP) {
        return a.in[(int64_t)(sz.y0 + line) * N + idx];
      } else {
        return a.in[(int64_t)line * N + idx];
      }
    };
  };
  const int stride = gridDim.x * a.lpb;
  const int first = blockIdx.x * a.lpb;  // the block's first line: the loop bound is block-uniform
  cpx<R> nxt[E];
  {
    auto ld = loader(first + l);
    line_load<R, L>(nxt, t, ld);
  }
  for (int base = first; base < a.nlines; base += stride) {
    // opaque copies of the lane index and the twiddle pointer keep the address arithmetic and the
    // twiddle loads inside the loop (hoisted, they would hold every stage's addresses and twiddles in
    // registers across the loop)
    const cpx<R>* tw = a.tw;
    int tl = t;
    asm volatile("" : "+s"(tw), "+v"(tl));
    const int line = base + l;
    const bool valid = line < a.nlines;
    cpx<R> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = nxt[e];
    if (base + stride < a.nlines) {
      auto ld = loader(line + stride);
      line_load<R, L>(nxt, tl, ld);
    }
    if constexpr (MODE == ROW_FWD_SZ) {  // S: sy[line] sx[i] and the folded scale
      if (valid) {
        const R sl = sz.sy[line] * sz.scale;
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int i = line_index<L>(tl, e) - sz.x0;
          if ((unsigned)i < (unsigned)sz.nx) v[e] = scale(v[e], sl * sz.sx[i]);
        }
      }
    }
    auto store = [&](int idx, cpx<R> val) {
      if (!valid) return;
      if constexpr (MODE == ROW_FWD_SZ) {
        a.out[(int64_t)(sz.y0 + line) * N + idx] = val;
      } else if constexpr (MODE == ROW_ADJ_CROP) {
        const int i = idx - sz.x0;
        if ((unsigned)i >= (unsigned)sz.nx) return;
        const R s = sz.sy[line] * sz.sx[i] * sz.scale;
        sz.img_out[(int64_t)line * sz.nx + i] = scale(val, s);
      } else {
        a.out[(int64_t)line * N + idx] = val;
      }
    };
    fft_line_pre<R, L, INV>(v, tl, lds, 0, tw, a.twm, store);
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
  void go(void (*k)(RowArgs<R>)) {
    allow_lds((const void*)k, shm);
    hipLaunchKernelGGL(k, dim3(blocks), dim3(threads), shm, s, a);
  }
  template <int L>
  void operator()() {
    if (mode == ROW_FWD_SZ) go(k_fft_rows<R, L, false, ROW_FWD_SZ>);
    else if (mode == ROW_ADJ_CROP) go(k_fft_rows<R, L, true, ROW_ADJ_CROP>);
    else if (inv) go(k_fft_rows<R, L, true, ROW_PLAIN>);
    else go(k_fft_rows<R, L, false, ROW_PLAIN>);
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
  const size_t shm = (size_t)lpb * (N + N / 16) * sizeof(cpx<R>);
  int blocks = (nlines + lpb - 1) / lpb;
  if (p.row_bpc > 0) blocks = std::min(blocks, p.row_bpc * p.n_cu);
  if (blocks == 0) return;
  ProfScope prof(mode == ROW_FWD_SZ ? "fft_rows_fwd_sz" : mode == ROW_ADJ_CROP ? "fft_rows_adj_crop" : "fft_rows", s);
  if ((mode == ROW_FWD_SZ && inv) || (mode == ROW_ADJ_CROP && !inv))
    throw Error(-1, "fused row pass direction mismatch");
  RowLaunch<R> launch{a, mode, inv, threads, blocks, shm, s};
  dispatch_log2(p.row.log2N, launch);
  PF_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------
// Column passes
// ---------------------------------------------------------------------------------------------
enum ColKind { COL_SINGLE = 0, COL_STEP_A = 1, COL_STEP_B = 2 };

template <typename R>
struct ColArgs {
  const cpx<R>* in;
  cpx<R>* out;
  const cpx<R>* tw;
  int twm;       // Lmax / line length
  int twm_full;  // Lmax / Nv (inter-step twiddle)
  int Nu, Nv, C, N1, N2;
  int y0, ny;    // row window
  int rin, rout; // restrict input / output rows to the window
};

template <typename R, int L, bool INV, int KIND>
__global__ void __launch_bounds__(512) k_fft_cols(ColArgs<R> a) {
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
    return a.in[row * Nu + col];
  };
  auto store = [&](int idx, cpx<R> val) {
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
    a.out[row * Nu + col] = val;
  };
  fft_line<R, L, INV>(t, smem + c, C, a.tw, a.twm, load, store);
}

// Single-pass columns up to 8192 rows: C columns per block, 1024 threads at most, padded LDS tiles
// above 64 KiB.  Adjacent column blocks share 128-byte lines of every row, so the block index is
// remapped to give each XCD (workgroups go round-robin over the 8 XCDs) a contiguous band of
// columns and its L2 the shared lines.
template <typename R>
constexpr int big_col_threads() { return sizeof(R) == 4 ? 1024 : 512; }  // FP64: room for 256 VGPRs

template <typename R, int L, bool INV>
__global__ void __launch_bounds__(big_col_threads<R>()) k_fft_cols_big(ColArgs<R> a) {
  extern __shared__ unsigned char smem_raw[];
  cpx<R>* smem = reinterpret_cast<cpx<R>*>(smem_raw);
  const int C = a.C;
  const int c = threadIdx.x % C, t = threadIdx.x / C;
  const int G = gridDim.x;
  const int blk = (G & 7) ? (int)blockIdx.x : (int)(blockIdx.x & 7) * (G >> 3) + (int)(blockIdx.x >> 3);
  const int col = blk * C + c;
  const int64_t Nu = a.Nu;
  auto in_window = [&](int row) { return (unsigned)(row - a.y0) < (unsigned)a.ny; };
  auto load = [&](int row) -> cpx<R> {
    if (a.rin && !in_window(row)) return mk<R>(R(0), R(0));
    return a.in[row * Nu + col];
  };
  auto store = [&](int row, cpx<R> val) {
    if (a.rout && !in_window(row)) return;
    a.out[row * Nu + col] = val;
  };
  fft_line<R, L, INV>(t, smem + c, -C, a.tw, a.twm, load, store);
}

template <typename R>
struct BigColLaunch {
  const ColArgs<R>& a;
  bool inv;
  dim3 grid;
  int threads;
  size_t shm;
  hipStream_t s;
  template <int L>
  void operator()() {
    if constexpr (L >= 10 && L <= 13) {
      void (*k)(ColArgs<R>) = inv ? k_fft_cols_big<R, L, true> : k_fft_cols_big<R, L, false>;
      allow_lds((const void*)k, shm);
      hipLaunchKernelGGL(k, grid, dim3(threads), shm, s, a);
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
  template <int L, bool I>
  void go() {
    void (*k)(ColArgs<R>) = kind == COL_STEP_A   ? k_fft_cols<R, L, I, COL_STEP_A>
                            : kind == COL_STEP_B ? k_fft_cols<R, L, I, COL_STEP_B>
                                                 : k_fft_cols<R, L, I, COL_SINGLE>;
    allow_lds((const void*)k, shm);
    hipLaunchKernelGGL(k, grid, dim3(threads), shm, s, a);
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
static void launch_cols(int kind, bool inv, int log2n, const ColArgs<R>& a, dim3 grid, hipStream_t s) {
  const int N = 1 << log2n;
  const int E = N < FFT_E ? N : FFT_E;
  const int threads = a.C * (N / E);
  const size_t shm = (size_t)N * a.C * sizeof(cpx<R>);
  ProfScope prof(kind == COL_STEP_A ? "fft_cols_stepA" : kind == COL_STEP_B ? "fft_cols_stepB" : "fft_cols", s);
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
  if (p.col_big) {
    a.twm = p.Lmax / p.Nv;
    a.C = p.big_C;
    a.rin = restrict_in;
    a.rout = restrict_out;
    const int N = p.Nv;
    const size_t shm = (size_t)(N + N / 16) * a.C * sizeof(cpx<R>);
    ProfScope prof("fft_cols", s);
    BigColLaunch<R> launch{a, inv, dim3(p.Nu / a.C), a.C * (N / FFT_E), shm, s};
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

void fft_plan_init(Fft2dPlan& p, int Nv, int Nu, int prec, int col_mode) {
  if (Nv > MAX_AXIS || Nu > MAX_AXIS)
    throw Error(-3, "grid axis larger than " + std::to_string(MAX_AXIS) + " is not supported");
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
  // row pass: a grid of 4 (complex64) or 2 (complex128) resident blocks per CU, each looping over lines
  // with the next line prefetched; at 4096 x 4096 this takes the row passes from 43 to 38 us (c64)
  p.row_bpc = prec ? 2 : 4;
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
    const int max_threads = prec ? 512 : 1024;
    int C = 8;
    while (C > 1 && ((size_t)(Nv + Nv / 16) * C * esz > 160 * 1024 - 1024 || C * (Nv / FFT_E) > max_threads)) C /= 2;
    if (C > Nu) C = Nu;
    p.big_C = C;
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
  std::vector<double> re, im;
  twiddle_table(p.Lmax, re, im);
  if (prec) {
    std::vector<cpx<double>> h(p.Lmax);
    for (int m = 0; m < p.Lmax; ++m) h[m] = mk<double>(re[m], im[m]);
    PF_HIP(hipMalloc(&p.tw, sizeof(cpx<double>) * p.Lmax));
    PF_HIP(hipMemcpy(p.tw, h.data(), sizeof(cpx<double>) * p.Lmax, hipMemcpyHostToDevice));
  } else {
    std::vector<cpx<float>> h(p.Lmax);
    for (int m = 0; m < p.Lmax; ++m) h[m] = mk<float>((float)re[m], (float)im[m]);
    PF_HIP(hipMalloc(&p.tw, sizeof(cpx<float>) * p.Lmax));
    PF_HIP(hipMemcpy(p.tw, h.data(), sizeof(cpx<float>) * p.Lmax, hipMemcpyHostToDevice));
  }
}

void fft_plan_free(Fft2dPlan& p) {
  if (p.tw) (void)hipFree(p.tw);
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
