// pf_internal.h -- host-side structures and kernel launchers shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "pf_device.h"

namespace pf {

// ---- errors -------------------------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
void hip_check(hipError_t e, const char* what);
#define PF_HIP(call) ::pf::hip_check((call), #call)

// ---- device memory and the construction stream ----------------------------------------------
// Device blocks come from a per-device cache of freed blocks (best fit within a quarter of the request,
// else hipMalloc): hipFree waits for the whole device, so a handle released while another handle's work
// runs would stall behind it.  dev_free returns a block to the cache at once: every piece of work that
// uses it must have completed (the library synchronises its own streams before it frees; a caller
// synchronises the streams it passed).  The cache holds at most PF_CACHE_MAX bytes per device; above that,
// and when hipMalloc fails, cached blocks are released with hipFree.
constexpr size_t PF_CACHE_MAX = size_t(16) << 30;
void* dev_alloc(size_t bytes);
void dev_free(void* p);
size_t dev_release_cached();
void dev_cache_stats(int64_t& live_bytes, int64_t& cached_bytes);
template <typename T>
T* dev_alloc_n(size_t n) {
  return static_cast<T*>(dev_alloc(sizeof(T) * (n ? n : 1)));
}
// The calling thread's construction stream (non-blocking, default priority; created on first use, on the
// current device): builds, power methods and other library-internal work run on it rather than on the
// legacy null stream, which every thread shares and which would order one request's construction behind
// another's.  h2d / d2h copy on it and wait for the copy.
hipStream_t cstream();
void cstream_sync() noexcept;  // waits for the thread's construction stream, if it has one
void h2d(void* dst, const void* src, size_t bytes);
void d2h(void* dst, const void* src, size_t bytes);
// The stream of an apply: the caller's, or, when it passes none and both vectors are on the host (the
// call is synchronous), the thread's construction stream instead of the shared legacy null stream.
inline hipStream_t io_stream(void* caller, bool host_io) {
  return (!caller && host_io) ? cstream() : static_cast<hipStream_t>(caller);
}

// ---- launch recording (channel batches) ----------------------------------------------------
// While a Recorder is installed on the calling thread, the launch sites of the solver iteration's kernels
// record their argument struct and launch geometry instead of launching.  A channel batch records one
// iteration per channel, checks that every channel recorded the same sequence of kernels and shapes, and
// then launches each position once for all channels: the kernel's batch twin (template flag BT) reads
// its arguments from a device table at blockIdx.z.  The per-channel arithmetic is the single launch's, so
// a channel's results in the batch equal its results alone, bit for bit.
struct LaunchRec {
  const void* key;  // the batch twin
  dim3 grid, block;
  size_t shm;
  bool grid_x_max;  // the kernel strides over gridDim.x: channels may differ there, the batch takes the max
  std::vector<unsigned char> args;
  void (*launch)(dim3 grid, dim3 block, size_t shm, hipStream_t s, const void* tab);
};
struct Recorder {
  std::vector<LaunchRec> recs;
};
Recorder*& recorder();  // the calling thread's, or null

// (a batch sets the twin's LDS limit when it records, before any launch or capture)
template <class A, void (*K)(A, const A*)>
void launch_batch_twin(dim3 grid, dim3 block, size_t shm, hipStream_t s, const void* tab) {
  hipLaunchKernelGGL(K, grid, block, shm, s, A{}, static_cast<const A*>(tab));
}
// Records the launch of K's batch twin KB with arguments a when a Recorder is installed (and returns true).
template <class A, void (*KB)(A, const A*)>
bool record_launch(dim3 grid, dim3 block, size_t shm, bool grid_x_max, const A& a) {
  Recorder* r = recorder();
  if (!r) return false;
  LaunchRec rec;
  rec.key = reinterpret_cast<const void*>(KB);
  rec.grid = grid;
  rec.block = block;
  rec.shm = shm;
  rec.grid_x_max = grid_x_max;
  rec.args.assign(reinterpret_cast<const unsigned char*>(&a), reinterpret_cast<const unsigned char*>(&a) + sizeof(A));
  rec.launch = &launch_batch_twin<A, KB>;
  r->recs.push_back(std::move(rec));
  return true;
}

// Launches K(a) on s, or records its batch twin KB when a Recorder is installed.
template <class A, void (*K)(A, const A*), void (*KB)(A, const A*)>
void launch_or_record(dim3 grid, dim3 block, size_t shm, bool grid_x_max, const A& a, hipStream_t s) {
  if (record_launch<A, KB>(grid, block, shm, grid_x_max, a)) return;
  hipLaunchKernelGGL(K, grid, block, shm, s, a, nullptr);
}

// The calling thread's last error message (pf_last_error) and the C-ABI wrappers that set it.
std::string& last_error();
inline int fail(int code, const std::string& msg) {
  last_error() = msg;
  return code;
}
template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const Error& e) {
    return fail(e.code, e.what());
  } catch (const std::bad_alloc&) {
    return fail(-4, "host allocation failed");  // PF_ERR_NOMEM
  } catch (const std::exception& e) {
    return fail(-1, e.what());  // PF_ERR_ARG
  }
}

// ---- per-kernel timing --------------------------------------------------------------------
// When an operator has profiling enabled it installs its Profiler as the calling thread's current
// one; every launch site then brackets its kernel with HIP events on the launching stream.
struct Profiler {
  struct Rec {
    const char* name;
    hipEvent_t a, b;
  };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  std::vector<std::string> names;
  std::vector<double> total_ms;
  std::vector<int64_t> count;
  hipEvent_t get_event();
  void flush();  // synchronises on the recorded events and accumulates
  void clear();
  ~Profiler();
};
Profiler*& current_profiler();
struct ProfScope {
  Profiler* p;
  const char* name;
  hipStream_t s;
  hipEvent_t a = nullptr;
  ProfScope(const char* n, hipStream_t st);
  ~ProfScope();
};

// ---- FFT plans ----------------------------------------------------------------------------
constexpr int FFT_E = 16;      // complex elements each thread holds per line
constexpr int MAX_AXIS = 8192; // largest supported grid axis

struct LinePlan {
  int N = 0, log2N = 0, nst = 0;
  int rad[8] = {0};
};
LinePlan make_line_plan(int N);

struct Fft2dPlan {
  int Nv = 0, Nu = 0, prec = 0;
  int Lmax = 0;        // twiddle table length (max(Nv, Nu))
  void* tw = nullptr;  // device table exp(-2 pi i m / Lmax), m < Lmax, in the plan's precision
  LinePlan row;        // N = Nu
  int row_lpb = 1;     // lines per block of the row pass
  int row_bpc = 0;     // row pass grid: row_bpc blocks per CU, each looping over lines (0: one pass)
  int n_cu = 256;      // compute units of the device
  bool col_two_step = false;
  bool col_big = false;  // single pass over columns longer than 512 rows
  int big_C = 4;         // columns per block of that pass
  LinePlan col;        // single-pass column plan, N = Nv
  int col_C = 16;      // columns per block of the single-pass column kernel
  int N1 = 0, N2 = 0;  // Nv = N1 * N2 for the two-step column pass
  LinePlan colA, colB; // N1- and N2-point plans
  int two_C = 64;      // columns per block of the two-step kernels
};
// col_mode: 0 = columns in one pass where the LDS allows (up to 8192 rows), 1 = the two-step split
// beyond 512 rows
void fft_plan_init(Fft2dPlan& p, int Nv, int Nu, int prec, int col_mode = 0);
// the plan without its twiddle table (p.tw stays null; the caller provides the table): no allocation
void fft_plan_setup(Fft2dPlan& p, int Nv, int Nu, int prec, int col_mode = 0);
void fft_plan_free(Fft2dPlan& p);

// Which rows of the grid hold data (forward after the row pass) or are wanted (adjoint crop).
struct RowWindow {
  int y0 = 0, ny = 0;  // rows [y0, y0+ny); ny = 0 means all rows
};

// Fused row-pass variants.
// ROW_ADJ_X: inverse transforms of every line, each cropped to the image columns [x0, x0 + nx) and
// scaled by sx and the folded scale (S's row factors are left to the column pass that follows);
// ROW_FWD_X: forward transforms of lines of nx values (img_in, line stride nx) zero-padded to the grid
// width at x0, unscaled (S applied by the column pass before)
enum RowMode { ROW_PLAIN = 0, ROW_FWD_SZ = 1, ROW_ADJ_CROP = 2, ROW_ADJ_X = 3, ROW_FWD_X = 4 };

template <typename R>
struct SZ {
  const cpx<R>* img_in = nullptr;  // ROW_FWD_SZ input image
  cpx<R>* img_out = nullptr;       // ROW_ADJ_CROP output image
  const R* sx = nullptr;           // length nx
  const R* sy = nullptr;           // length ny
  R scale = R(1);
  int nx = 0, ny = 0, x0 = 0, y0 = 0;
};

template <typename R>
void fft_rows(const Fft2dPlan& p, int mode, bool inv, const cpx<R>* in, cpx<R>* out, int nlines, const SZ<R>& sz,
              hipStream_t s);

// Kernel arguments of the row and column passes (recorded by a batch: see Recorder).
template <typename R>
struct RowArgs {
  const cpx<R>* in;
  cpx<R>* out;
  const cpx<R>* tw;
  int twm, Nu, nlines, lpb;
  SZ<R> sz;
};
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
  const R* osy;  // non-null (with rout): window row y is written at row y - y0, times osy[y - y0]
  const R* isy;  // non-null (with rin): window row y is read at row y - y0, times isy[y - y0] isx[col] iscale
  const R* isx;
  R iscale;
};

// Column pass.  restrict_in: rows outside win are zero and not read; restrict_out: only rows in
// win are written.  Two-step plans overwrite `in` with the intermediate result.
template <typename R>
void fft_cols(const Fft2dPlan& p, bool inv, cpx<R>* in, cpx<R>* out, RowWindow win, bool restrict_in,
              bool restrict_out, hipStream_t s);
// Column pass over the `width` columns of a p.Nv x width array (width < p.Nu), writing only the window's
// rows, row y at y - win.y0 of `out` (win.ny x width) times osy[y - win.y0].  Usable when
// fft_cols_crop_ok(p, width, win.ny).
template <typename R>
void fft_cols_crop(const Fft2dPlan& p, bool inv, cpx<R>* in, cpx<R>* out, int width, RowWindow win, const R* osy,
                   hipStream_t s);
bool fft_cols_width_ok(const Fft2dPlan& p, int width);
bool fft_cols_crop_ok(const Fft2dPlan& p, int width, int ny);
// Column pass over `width` columns of a win.ny x width image, S applied as it is read (row y - win.y0
// times isy[y - win.y0] isx[col] iscale, zero outside the window), into a p.Nv x width array.  Not for
// two-step plans (fft_cols_pad_ok).
template <typename R>
void fft_cols_pad(const Fft2dPlan& p, bool inv, const cpx<R>* in, cpx<R>* out, int width, RowWindow win,
                  const R* isy, const R* isx, R iscale, hipStream_t s);
bool fft_cols_pad_ok(const Fft2dPlan& p, int width, int ny);

// ---- gridding -----------------------------------------------------------------------------
template <typename R>
struct VisData {
  // complex64 path: kk = (k_u, k_v) = floor(u - Ju/2), floor(v - Jv/2); ff = (u - k_u, v - k_v)
  // complex128 path: uv = (u, v) in pixels (k computed on the device exactly as the reference)
  const int2* kk = nullptr;
  const cpx<float>* ff = nullptr;
  const cpx<double>* uv = nullptr;
  const cpx<R>* weights = nullptr;  // nullptr: unit weights
  const int32_t* perm = nullptr;    // degridding writes y[perm[i]] (caller's order); nullptr: y[i] (sorted)
  int64_t M = 0;
};

// The permutation between the caller's order and the sorted order.  A random access costs one memory
// transaction whatever its size up to a cache line, so a one-pass scatter of 8-byte values is bound by
// the transaction count; complex64 therefore goes through two passes that each move whole lines, and
// complex128 (16-byte values, half the transactions per byte) uses the one-pass scatter.
//
// Two-pass plan (two_pass): the caller's order is cut into chunks and the sorted order into buckets
// of S visibilities each (nb of both).  A staging vector holds bucket b in [b*S, b*S + size of b),
// and inside it the visibilities grouped by chunk, each group in the caller's order: so a chunk sends
// one contiguous run to each bucket, and a bucket gathers one run from each chunk.
//   chunk pass:  one block stages its chunk in LDS ordered by bucket, then writes (reads) the runs.
//   bucket pass: one block stages its bucket in LDS, then writes (reads) the sorted range coalesced.
// One-pass plan: ys[inv[m]] = y[m], and degridding writes y[perm[i]] itself.
struct VisPerm {
  bool two_pass = false;
  const uint16_t* l1 = nullptr;  // caller index m -> position inside chunk m / S, ordered by bucket
  const int32_t* g1 = nullptr;   // (chunk c, position j) at c * S + j -> staging index
  const uint16_t* p2 = nullptr;  // staging index t -> position inside bucket t / S of the sorted order
  int S = 0;                     // chunk and bucket size, a power of two <= perm_max_S(precision)
  int nb = 0;                    // ceil(M / S)
  const int32_t* perm = nullptr;  // one-pass: sorted position -> caller's index
  const int32_t* inv = nullptr;   // one-pass: caller's index -> sorted position
  int64_t M = 0;
};
// largest chunk: 128 KiB of LDS
inline int perm_max_S(int esz) { return esz == 8 ? 16384 : 8192; }

// Dynamic LDS above 64 KiB (gfx950 has 160 KiB per CU) must be opted into per kernel.  The attribute is
// set once per kernel and device (and raised when a launch needs more): a host call per launch would
// stall a stream of short launches behind the host.
inline void allow_lds(const void* fn, size_t shm) {
  if (shm <= 65536) return;
  int dev = 0;
  PF_HIP(hipGetDevice(&dev));
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, size_t> set;
  std::lock_guard<std::mutex> lk(mu);
  size_t& cur = set[{fn, dev}];
  if (cur >= shm) return;
  PF_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
  cur = shm;
}

// complex64 tap tables: all J taps of a visibility sit at offsets delta + J/2 - 1 - a of one fractional
// delta = (u - k_u) - J/2 in (0, 1], so row q of the table holds the J kernel values at delta = q / TAB_Q
// (T[0..7]) and their forward differences to row q + 1 (D[0..7]); a visibility reads one 64-byte row
// and interpolates linearly.  Interpolation error about 1e-7 relative; used for J <= 8.
constexpr int TAB_Q = 4096;
constexpr int TAB_ROW = 16;

template <typename R>
struct GridGeom {
  int Nu = 0, Nv = 0, Ju = 0, Jv = 0;
  KParams<R> ku, kv;               // u-axis and v-axis kernels
  const float* tab_u = nullptr;    // complex64 path only: tap tables (nullptr: evaluate the kernels)
  const float* tab_v = nullptr;
  // complex64, Ju = Jv: per-tap polynomials (nullptr: tables or direct evaluation); both axes use them
  const float* poly = nullptr;
};
// Host: builds the TAB_Q x TAB_ROW table of a kernel evaluated with support Jeval over J taps.
std::vector<float> make_tap_table(int kind, int Jeval, int J, double os);
// Per-tap polynomials: tap a of a visibility with fractional offset delta = (u - k_u) - J/2 in [0, 1)
// is K(delta + J/2 - 1 - a) ~ sum_d c[a][d] s^(D - d), s = 2 delta - 1, degree D = TAP_POLY_D
// (Chebyshev interpolation at D + 1 nodes, in long double, converted to powers of s; highest degree
// first for Horner).  Returns an empty vector unless the kernel is smooth on every tap interval (KB,
// KB-min, Gaussian, PSWF with Jeval = J) and the interpolation error, checked on 257 points per tap,
// is below 1e-7 of the kernel's peak.  The device evaluates taps a < J/2 and mirrors them (the kernels
// are even: tap J-1-a at s is tap a at -s), so the upper half of the table is not read.
constexpr int TAP_POLY_D = 9;
std::vector<float> make_tap_poly(int kind, int Jeval, int J, double os, double* max_rel_err = nullptr);

constexpr uint32_t PART_NONE = 0xffffffffu;
struct TileLists {
  int TU = 32, TV = 32, ntu = 0, ntv = 0;
  const uint32_t* list = nullptr;  // sorted vis indices per tile (every visibility with a tap in it)
  // launch slots, heaviest first: (tile, begin, end, part) of a range of the tile's entries in `list`.
  // Without splitting, one slot per tile (all tiles) with part = PART_NONE.  Split lists (stored-row
  // gridding) cut a tile with more than a chunk of entries into parts: each part's slot accumulates a
  // partial tile into `part` (TU * TV values at part * TU * TV), and the combine pass sums a split tile's
  // parts in part order into the grid; comb holds (tile, first part, parts, 0) per split tile.
  const uint4* slot = nullptr;
  int n_slot = 0;
  const uint4* comb = nullptr;
  int n_comb = 0;
  void* part = nullptr;
  // launch slots of the degridding kernel: (tile, begin, end, 0) of a chunk of the sorted visibilities
  // whose first tap lies in the tile (chunks of at most max(512, M / (8 CUs)) visibilities, so that a
  // dense tile does not form the launch's tail), for the tiles that own any, heaviest first
  const uint4* slot_own = nullptr;
  int n_own = 0;
  // owner-tile gridding of dense stored rows (footprints of K <= tile side + 1): each degridding slot
  // sums its visibilities' rows into its tile extended by the footprint's K - 1 columns and rows
  // (ext_w x ext_h cells at ext + slot * ext_w * ext_h), and a combine pass adds, for each grid cell, the
  // owner tile's slots, then the left, upper and upper-left neighbours' extensions, in that order.
  // own_first[t] .. own_first[t + 1] index own_ids, tile t's slots in the order of their visibilities.
  const int* own_first = nullptr;
  const int* own_ids = nullptr;
  void* ext = nullptr;
  int ext_w = 0, ext_h = 0;
  // complex64 matrix-core lists: each list entry as one record (8 * sorted index, the byte offset of
  // its value; first-tap cell (k_u mod Nu) | (k_v mod Nv) << 16; the fractions u - k_u and v - k_v
  // as float bits)
  const int4* ent = nullptr;
};

// y = G grid, written y[v.perm[i]] in the caller's order, or y[i] in the sorted order when v.perm is null
// (permute_to_caller then brings it to the caller's order)
template <typename R>
void degrid(const GridGeom<R>& g, const VisData<R>& v, const TileLists& t, const cpx<R>* grid, cpx<R>* y_sorted,
            R scale, hipStream_t s);
bool grid_rc_ok(int Nu, int Nv, int TU, int TV, int Ju, int Jv);
// side of the matrix-core gridding tiles; usable when N >= MFMA_TILE + J - 1 on both axes
constexpr int MFMA_TILE = 16;
bool grid_mfma_ok(int Nu, int Nv, int Ju, int Jv);
// gridding kernel variants: residue-class lanes (needs J <= 8 and N >= T + 8), LDS atomics, or
// rank-4 updates of 16 x 16 tiles on the matrix cores (lists `tm`)
enum GridVariant { GRID_RESIDUE = 0, GRID_ATOMIC = 1, GRID_MFMA = 2 };
// grid = G^H y with y in the operator's sorted order (see permute_to_sorted).  `t` are the lists of
// the TU x TV tiles, `tm` those of the 16 x 16 tiles (GRID_MFMA), gridded by mfma_waves waves that
// each take every mfma_waves-th tile (0: one wave per tile).
template <typename R>
void grid_tiles(const GridGeom<R>& g, const VisData<R>& v, const TileLists& t, const TileLists& tm,
                const cpx<R>* y_sorted, cpx<R>* grid, R scale, hipStream_t s, int variant, int mfma_waves = 0);
// caller's order -> sorted order (ys) and back; the two-pass plan stages through `tmp` (M values).
// With the one-pass plan permute_to_caller gathers y[m] = ys[inv[m]] (the operator's own applies do not
// use it: their degridding writes the caller's order directly).
template <typename R>
void permute_to_sorted(const VisPerm& p, const cpx<R>* y, cpx<R>* tmp, cpx<R>* ys, hipStream_t s);
template <typename R>
void permute_to_caller(const VisPerm& p, const cpx<R>* ys, cpx<R>* tmp, cpx<R>* y, hipStream_t s);

// ---- projection rows (w_term) -------------------------------------------------------------
// Row m of the w-projection gridding matrix (operators.cc:49-99): Kv x Ku taps, K = Jw + J - 1, at
// columns (kw_u + ju, kw_v + jv) mod (Nu, Nv), ju, jv = 1..K, kw = floor(u - K/2).  The values (the
// projection kernel of projection_kernels.cc:59-85, the (-1)^(kw_u+ju+kw_v+jv) phase and the weight)
// are built once per operator and stored per sorted visibility, row-major [jv][ju].
template <typename R>
struct WRows {
  int Ku = 0, Kv = 0;
  const int2* kw = nullptr;      // (kw_u, kw_v) per sorted visibility
  const cpx<R>* vals = nullptr;  // M * Kv * Ku, or the kept entries only when rec is set
  // kept-entry rows (Ku Kv <= 64): rec[i] = {mask lo, mask hi, offset lo, offset hi}; bit jv Ku + ju of
  // the mask is set for a kept entry, and row i's kept values lie at vals[offset ..] in bit order
  const uint4* rec = nullptr;
  const uint8_t* pos = nullptr;  // kept-entry rows: each kept value's tap, jv << 4 | ju (beside vals)
  int slots = 8;                 // kept-entry degridding: slots per lane (8 lanes per row): 2, 4 or 8
  int qe = 1;
  bool own = true;  // dense rows: owner-tile gridding where the tile lists carry its partial sums                    // kept-entry gridding: entries per step (1: one lane per tap position)
  int depth = 16;      // kept-entry gridding: entries whose taps are requested ahead (8, 16 or 32)
  bool off64 = false;  // kept-entry offsets reach 2^32 (else their high words are zero)
};
// Kept-entry rows from dense rows and their keep masks (Ku Kv <= 64): rows_pack_records writes rec from
// the masks (the offsets an exclusive sum of the kept counts) and returns the total kept entries;
// (and the largest kept count of a row into max_kept); rows_pack_values copies the kept values and writes
// each one's tap (jv << 4 | ju, Ku the row width) into
// pos; both arrays carry ROWS_PACK_PAD spare entries past the last (zero), which the kernels read and do
// not use for a dropped tap or an unused slot; rows_unpack rebuilds the dense rows (zeros elsewhere).
constexpr int ROWS_PACK_PAD = 64;
uint64_t rows_pack_records(const uint64_t* mask, int64_t M, uint4* rec, int* max_kept, hipStream_t s);
template <typename R>
void rows_pack_values(const cpx<R>* dense, const uint4* rec, int64_t M, int KK, int Ku, cpx<R>* packed,
                      uint8_t* pos, hipStream_t s);
template <typename R>
void rows_unpack(const cpx<R>* packed, const uint4* rec, int64_t M, int KK, cpx<R>* dense, hipStream_t s);
// y = G grid: owner tiles staged with a (K - 1)-wide halo, one wave per visibility; y[perm[i]] in the
// caller's order, or y[i] in the sorted order when perm is null
template <typename R>
void degrid_wrows(int Nu, int Nv, const WRows<R>& w, const int32_t* perm, const TileLists& t, const cpx<R>* grid,
                  cpx<R>* y, R scale, hipStream_t s);
// grid = G^H ys (ys in sorted order): one wave per tile, the taps of one visibility at a time
template <typename R>
void grid_wrows(int Nu, int Nv, const WRows<R>& w, const TileLists& t, const cpx<R>* ys, cpx<R>* grid, R scale,
                hipStream_t s);
// Construction of the projection rows on the device, FP64: every row of Kv x Ku values for the sorted
// visibility i (the caller's perm[i]), phase and weight included.  kp: the kernel in the u slot (the
// reference evaluates both axes with it, projection_kernels.cc:72-73); du_l, dv_l: pixel_to_lambda of each
// axis; wmean: the mean of w (operators.h:323).
struct WRowArgs {
  int Ju = 4, Jv = 4, Jw = 6, Ku = 9, Kv = 9;
  double du_l = 1, dv_l = 1, wmean = 0;
  KParams<double> kp;
};
template <typename R>
void wrows_build(const WRowArgs& g, const cpx<double>* uv, const double* w, const cpx<double>* wt, const int32_t* perm,
                 int64_t M, cpx<R>* out);

// ---- chirp-truncated w-projection rows for a whole operator (pf_chirp.hip) ----------------------
struct ChirpBuild;
ChirpBuild* chirp_build_new();
void chirp_build_free(ChirpBuild* cb);
// pass 1 (caller-order inputs, pixels and wavelengths): the largest kept chirp radius R of any row
int chirp_rows_radius(ChirpBuild& cb, int64_t M, int N, int J, int kind, int Jeval_u, int Jeval_v, double os,
                      double cellx, double celly, double e_chirp, double e_wproj, const double* u, const double* v,
                      const double* w, const double* weights);
// pass 2: rows of (J + 2R)^2 values, row i for the caller's visibility perm[i] (device), into rows
// (precision prec)
int64_t chirp_rows_fill(ChirpBuild& cb, const int32_t* perm_dev, void* rows, int prec, uint64_t* mask = nullptr);

// ---- operator construction on the device (pf_build.hip) -------------------------------------------
// the owner-tile sort of the caller's pixel coordinates uv: perm (sorted -> caller) and owner counts
void dev_sort_visibilities(const cpx<double>* d_uv, int64_t M, int Nu, int Nv, int Ku, int Kv, int TU, int TV,
                           int32_t* d_perm, std::vector<uint32_t>& owner_count);
// sorted tap origins floor(u - K/2), complex64 fractions, complex128 pixels, weights (any may be null)
template <typename R>
void dev_gather_sorted(const cpx<double>* d_uv, const cpx<double>* d_wt, const int32_t* d_perm, int64_t M, int Ku,
                       int Kv, int2* kk, cpx<float>* ff, cpx<double>* uvs, cpx<R>* ws);
// tile lists (hipMalloc'd, owned by the caller), their length and per-tile counts
uint32_t* dev_tile_lists(const int2* kk, int64_t M, int Nu, int Nv, int Ku, int Kv, int TU, int TV,
                         std::vector<uint32_t>& count, int64_t& len);
// complex64 matrix-core entry records of a list
void dev_entries(const uint32_t* list, int64_t n, const int2* kk, const cpx<float>* ff, int Nu, int Nv, int4* ent);
void dev_inverse(const int32_t* perm, int64_t M, int32_t* inv);
// the two-pass permutation plan's index arrays
void dev_perm_plan(const int32_t* perm, int64_t M, int S, int nb, uint16_t* l1, int32_t* g1, uint16_t* p2);

// ---- sparse gridding matrices (pf_wmat.hip) -------------------------------------------------
// Host-side CSR (complex128 values), the layout of pf_csr.
struct HostCsr {
  int64_t rows = 0, cols = 0;
  std::vector<int64_t> rowptr;
  std::vector<int32_t> col;
  std::vector<double> val;  // interleaved complex
};
// wproj_utilities::wprojection_matrix on the GPU in FP64 (see pf_wprojection_matrix).
void wprojection_matrix_dev(int64_t rows, const int64_t* rowptr, const int32_t* col, const double* val, int x_size,
                            int y_size, const double* w, double cell_x, double cell_y, double e_chirp, double e_wproj,
                            HostCsr& out);
// y[r] = sum_k val[k] x[col[k]] over the entries of row r (one wave per row, fixed-order sums)
template <typename R>
void csr_gather(int64_t rows, const int64_t* rowptr, const int32_t* col, const cpx<R>* val, const cpx<R>* x, cpx<R>* y,
                R scale, hipStream_t s, const char* name);

// ---- vector helpers -----------------------------------------------------------------------
// sum |x|^2 in FP64 over n complex entries, deterministic two-level reduction into *out (device)
template <typename R>
void norm2_fp64(const cpx<R>* x, int64_t n, double* partials, double* out, hipStream_t s);
template <typename R>
void scale_vec(cpx<R>* x, int64_t n, double s, hipStream_t st);
template <typename R>
void cmul_vec(cpx<R>* x, const cpx<R>* h, int64_t n, double sc, hipStream_t s);
void c128_to_c64(const cpx<double>* in, cpx<float>* out, int64_t n, hipStream_t s);
void c64_to_c128(const cpx<float>* in, cpx<double>* out, int64_t n, hipStream_t s);

// ---- host-side kernel functions (kernels.cc) ------------------------------------------------
double host_bessel_i0(double x);
double host_kernel(int kind, double x, int J, double os);       // gridding kernel value
double host_ft_kernel(int kind, double x, int J, double os);    // correction FT value at x
double kb_alpha(int kind, int J, double os);
template <typename R>
KParams<R> make_kparams(int kind, int J, double os);

}  // namespace pf

// ---- access across translation units (the solver drives the operator and the dictionary) ----
struct pf_op;
struct pf_wavelet;
namespace pf {
// y = Phi x, x = Phi^dagger y on device vectors in the operator's precision, on stream s
template <typename R>
void op_forward(pf_op* op, const cpx<R>* x, cpx<R>* y, hipStream_t s);
template <typename R>
void op_adjoint(pf_op* op, const cpx<R>* y, cpx<R>* x, hipStream_t s);
// out = Phi^dagger Phi x (both normalised), without the two permutations
template <typename R>
void op_normal(pf_op* op, const cpx<R>* x, cpx<R>* out, hipStream_t s);
// y = Phi x and x = Phi^dagger y with the visibility vector in the operator's sorted order (no permutation;
// for an operator on a given matrix the sorted order is the caller's)
template <typename R>
void op_forward_sorted(pf_op* op, const cpx<R>* x, cpx<R>* ys, hipStream_t s);
template <typename R>
void op_adjoint_sorted(pf_op* op, const cpx<R>* ys, cpx<R>* x, hipStream_t s);
// a visibility vector from the caller's order to the sorted order and back (y and ys distinct)
template <typename R>
void op_to_sorted(pf_op* op, const cpx<R>* y, cpx<R>* ys, hipStream_t s);
template <typename R>
void op_to_caller(pf_op* op, const cpx<R>* ys, cpx<R>* y, hipStream_t s);
struct OpShape {
  int ny, nx, prec;
  int64_t M;
};
OpShape op_shape(const pf_op* op);
// changes whenever the operator's launches change (norm, tuning, autotune)
uint64_t op_generation(const pf_op* op);
// Psi^dagger (image -> stacked coefficients) and Psi on device vectors
template <typename R>
void wavelet_analysis(pf_wavelet* wv, const cpx<R>* x, cpx<R>* c, hipStream_t s);
template <typename R>
void wavelet_synthesis(pf_wavelet* wv, const cpx<R>* c, cpx<R>* x, hipStream_t s);
struct WaveletShape {
  int ny, nx, prec, nwav;
};
WaveletShape wavelet_shape(const pf_wavelet* wv);
}  // namespace pf
