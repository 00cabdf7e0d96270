// pf_op.hip -- the measurement operator object and the C-ABI declared in include/purify_amd.h.
//
// Construction (host, once): unit conversion, tap origins, a counting sort of the visibilities by
// the grid tile holding their first tap, the per-tile gridding lists, the separable correction
// vectors and the FFT plan; everything is uploaded once and stays resident in HBM.
// Apply (device, per call):
//   forward  y = G F Z S x / norm : [row FFT fused with S, Z] -> [column FFT] -> [degrid gather]
//   adjoint  x = S Z^T F^H G^H y / norm : [tile gridding] -> [column IFFT] -> [row IFFT fused with crop, S]
// The unitary 1/sqrt(N) and 1/norm scalings are folded into the S multiply, so no pass is spent on them.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <thread>

#include "../../include/purify_amd.h"
#include "pf_internal.h"

namespace pf {

std::string& last_error() {
  static thread_local std::string msg;
  return msg;
}


void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(PF_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// ---- per-kernel timing ----
Profiler*& current_profiler() {
  static thread_local Profiler* p = nullptr;
  return p;
}

hipEvent_t Profiler::get_event() {
  if (!pool.empty()) {
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  hipEvent_t e;
  PF_HIP(hipEventCreate(&e));
  return e;
}

void Profiler::flush() {
  for (Rec& r : pending) {
    PF_HIP(hipEventSynchronize(r.b));
    float ms = 0;
    PF_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    size_t k = 0;
    while (k < names.size() && names[k] != r.name) ++k;
    if (k == names.size()) {
      names.push_back(r.name);
      total_ms.push_back(0);
      count.push_back(0);
    }
    total_ms[k] += ms;
    count[k] += 1;
    pool.push_back(r.a);
    pool.push_back(r.b);
  }
  pending.clear();
}

void Profiler::clear() {
  flush();
  names.clear();
  total_ms.clear();
  count.clear();
}

Profiler::~Profiler() {
  for (Rec& r : pending) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

ProfScope::ProfScope(const char* n, hipStream_t st) : p(current_profiler()), name(n), s(st) {
  if (p) {
    a = p->get_event();
    PF_HIP(hipEventRecord(a, s));
  }
}

ProfScope::~ProfScope() {
  if (!p) return;
  hipEvent_t b = p->get_event();
  if (hipEventRecord(b, s) == hipSuccess) p->pending.push_back(Profiler::Rec{name, a, b});
}

// ---------------------------------------------------------------------------------------------
// Host-side kernel functions (cpp/purify/kernels.cc).  Used for the correction vectors and for the
// Kaiser-Bessel Horner coefficients; the device evaluates kernels through KParams.
// ---------------------------------------------------------------------------------------------
static const double PI = 3.14159265358979323846;

double host_bessel_i0(double x) {
  // power series sum_k (x^2/4)^k / (k!)^2 (all terms positive)
  const double q = 0.25 * x * x;
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 500; ++k) {
    term *= q / ((double)k * (double)k);
    sum += term;
    if (term < 1e-17 * sum) break;
  }
  return sum;
}

double kb_alpha(int kind, int J, double os) {
  if (kind == K_KBMIN) return PI * std::sqrt((double)J * J / (os * os) * (os - 0.5) * (os - 0.5) - 0.8);  // kernels.cc:259
  return 2.34 * J;  // kernels.cc:11
}

static double gauss_sigma(int kind, int J) { return kind == K_GAUSS_ALT ? 1.0 : 0.31 * std::pow(J, 0.52); }

static double pswf_calc(double eta0, int J) {
  // kernels.cc:71-105 (alpha = 1)
  if (J != 6) return 0;
  const double a = std::fabs(eta0);
  double e, num, den;
  if (a <= 0.75) {
    e = eta0 * eta0 - 0.5625;
    num = (((2.312756e-1 * e - 5.335581e-1) * e + 6.278660e-1) * e - 3.644705e-1) * e + 8.203343e-2;
    den = (2.078043e-1 * e + 8.212018e-1) * e + 1.0;
  } else if (a <= 1) {
    e = eta0 * eta0 - 1.0;
    num = (((6.412774e-2 * e - 1.201436e-1) * e + 1.021332e-1) * e - 3.697768e-2) * e + 4.028559e-3;
    den = (2.918724e-1 * e + 9.599102e-1) * e + 1.0;
  } else {
    return 0;
  }
  return num / den;
}

double host_kernel(int kind, double x, int J, double os) {
  switch (kind) {
    case K_KB:
    case K_KBMIN: {
      const double al = kb_alpha(kind, J, os), a = 2 * x / J;
      return host_bessel_i0(al * std::sqrt(1 - a * a)) / host_bessel_i0(al);
    }
    case K_GAUSS:
    case K_GAUSS_ALT: {
      const double a = x / gauss_sigma(kind, J);
      return std::exp(-a * a * 0.5);
    }
    case K_PSWF: {
      const double eta = 2 * x / J;
      return pswf_calc(eta, J) * (1 - eta * eta);
    }
    case K_BOX:
      return 1.0 / J;
  }
  return NAN;
}

double host_ft_kernel(int kind, double x, int J, double os) {
  switch (kind) {
    case K_KB:
    case K_KBMIN: {
      // kernels.cc:24-38: Re[sin(eta)/eta] / 38828.11016883, eta = sqrt((pi x J)^2 - alpha^2)
      const double al = kb_alpha(kind, J, os);
      const double t = (PI * x * J) * (PI * x * J) - al * al;
      const std::complex<double> eta = std::sqrt(std::complex<double>(t, 0.0));
      if (std::abs(eta) == 0.0) return 1.0 / 38828.11016883;
      return std::real(std::sin(eta) / eta) / 38828.11016883;
    }
    case K_GAUSS:
    case K_GAUSS_ALT: {
      const double sg = gauss_sigma(kind, J), a = x * sg * PI;
      return std::sqrt(PI / 2) / sg * std::exp(-a * a * 2);  // kernels.cc:219-230
    }
    case K_PSWF:
      return pswf_calc(2 * x, J);  // kernels.cc:124-142
    case K_BOX: {
      const double t = J * x * PI;  // kernels.cc:197-205
      return t == 0.0 ? 1.0 : std::sin(t) / t;
    }
  }
  return NAN;
}

template <typename R>
KParams<R> make_kparams(int kind, int J, double os) {
  KParams<R> p;
  std::memset(&p, 0, sizeof(p));
  p.kind = kind;
  p.J = J;
  p.half = (R)(0.5 * J);
  p.four_over_J2 = (R)(4.0 / ((double)J * J));
  p.two_over_J = (R)(2.0 / J);
  const double sg = gauss_sigma(kind, J);
  p.inv_two_sigma2 = (R)(1.0 / (2 * sg * sg));
  p.box = (R)(1.0 / J);
  p.ncoef = 1;
  if (kind == K_KB || kind == K_KBMIN) {
    // Chebyshev coefficients of P(t) = I0(alpha sqrt t) / I0(alpha), t = (s + 1) / 2, s in [-1, 1],
    // from the values at 64 Chebyshev nodes (discrete cosine transform), in long double.
    const long double al = kb_alpha(kind, J, os);
    auto i0l = [](long double z2) {  // I0(sqrt(z2)) by its power series in z2 / 4
      const long double q = z2 / 4.0L;
      long double term = 1.0L, sum = 1.0L;
      for (int k = 1; k < 400; ++k) {
        term *= q / ((long double)k * k);
        sum += term;
        if (term < 1e-22L * sum) break;
      }
      return sum;
    };
    const long double pi = 3.141592653589793238462643383279502884L;
    const int NN = 64;
    const long double norm = i0l(al * al);
    long double fv[NN], th[NN];
    for (int j = 0; j < NN; ++j) {
      th[j] = pi * (j + 0.5L) / NN;
      const long double t = (cosl(th[j]) + 1.0L) / 2.0L;
      fv[j] = i0l(al * al * t) / norm;
    }
    long double a[KB_MAX_COEF];
    for (int k = 0; k < KB_MAX_COEF; ++k) {
      long double sacc = 0;
      for (int j = 0; j < NN; ++j) sacc += fv[j] * cosl(k * th[j]);
      a[k] = sacc * 2.0L / NN;
    }
    a[0] /= 2;
    const long double cut = sizeof(R) == 8 ? 1e-18L : 1e-10L;
    int n = KB_MAX_COEF;
    while (n > 1 && fabsl(a[n - 1]) < cut) --n;
    p.ncoef = n;
    for (int k = 0; k < KB_MAX_COEF; ++k) p.coef[k] = (R)(k < n ? a[k] : 0.0L);
  }
  return p;
}
template KParams<float> make_kparams<float>(int, int, double);
template KParams<double> make_kparams<double>(int, int, double);

std::vector<float> make_tap_table(int kind, int Jeval, int J, double os) {
  // row q: K(q/Q + J/2 - 1 - a) for a < J in [0, 8), forward differences to row q + 1 in [8, 16)
  std::vector<double> v((size_t)(TAB_Q + 1) * 8, 0.0);
  for (int q = 0; q <= TAB_Q; ++q)
    for (int a = 0; a < J && a < 8; ++a) v[(size_t)q * 8 + a] = host_kernel(kind, (double)q / TAB_Q + 0.5 * J - 1 - a, Jeval, os);
  std::vector<float> t((size_t)TAB_Q * TAB_ROW, 0.0f);
  for (int q = 0; q < TAB_Q; ++q)
    for (int a = 0; a < 8; ++a) {
      t[(size_t)q * TAB_ROW + a] = (float)v[(size_t)q * 8 + a];
      t[(size_t)q * TAB_ROW + 8 + a] = (float)(v[(size_t)(q + 1) * 8 + a] - v[(size_t)q * 8 + a]);
    }
  return t;
}

std::vector<float> make_tap_poly(int kind, int Jeval, int J, double os, double* max_rel_err) {
  if (max_rel_err) *max_rel_err = -1;
  if (Jeval != J || J < 1 || (kind != K_KB && kind != K_KBMIN && kind != K_GAUSS && kind != K_GAUSS_ALT && kind != K_PSWF))
    return {};
  constexpr int D = TAP_POLY_D, NP = D + 1;
  const long double PI_L = 3.141592653589793238462643383279502884L;
  std::vector<float> out((size_t)J * NP);
  double peak = 0, worst = 0;
  for (int a = 0; a < J; ++a) {
    auto f = [&](long double sv) {
      const double delta = (double)((sv + 1) / 2);
      return (long double)host_kernel(kind, delta + 0.5 * J - 1 - a, Jeval, os);
    };
    // Chebyshev coefficients from the values at the D + 1 Chebyshev nodes
    long double fv[NP], cheb[NP];
    for (int j = 0; j < NP; ++j) fv[j] = f(std::cos(PI_L * (j + 0.5L) / NP));
    for (int k = 0; k < NP; ++k) {
      long double acc = 0;
      for (int j = 0; j < NP; ++j) acc += fv[j] * std::cos(PI_L * k * (j + 0.5L) / NP);
      cheb[k] = acc * 2 / NP;
    }
    cheb[0] /= 2;
    // to powers of s: T_0 = 1, T_1 = s, T_(k+1) = 2 s T_k - T_(k-1)
    long double mono[NP] = {}, tkm1[NP] = {}, tk[NP] = {}, tkp1[NP];
    tkm1[0] = 1;
    tk[1] = 1;
    mono[0] += cheb[0];
    if (D >= 1) mono[1] += cheb[1];
    for (int k = 1; k < D; ++k) {
      for (int d = 0; d < NP; ++d) tkp1[d] = -tkm1[d] + (d > 0 ? 2 * tk[d - 1] : 0);
      for (int d = 0; d < NP; ++d) mono[d] += cheb[k + 1] * tkp1[d];
      std::copy(tk, tk + NP, tkm1);
      std::copy(tkp1, tkp1 + NP, tk);
    }
    for (int d = 0; d < NP; ++d) out[(size_t)a * NP + d] = (float)mono[D - d];  // highest degree first
    // check the float polynomial, evaluated as the device does, against the kernel
    for (int q = 0; q <= 256; ++q) {
      const float sv = 2.0f * ((float)q / 256.0f) - 1.0f;
      double p = out[(size_t)a * NP];
      for (int d = 1; d < NP; ++d) p = p * sv + out[(size_t)a * NP + d];
      const double ex = (double)f(sv);
      if (!std::isfinite(ex)) return {};
      peak = std::max(peak, std::fabs(ex));
      worst = std::max(worst, std::fabs(p - ex));
    }
  }
  if (!(peak > 0)) return {};
  if (max_rel_err) *max_rel_err = worst / peak;
  if (worst > 1e-7 * peak) return {};
  return out;
}

}  // namespace pf

using namespace pf;

// ---------------------------------------------------------------------------------------------
// The operator object
// ---------------------------------------------------------------------------------------------
struct pf_op {
  int prec = PF_C64;
  int ny = 0, nx = 0, Nv = 0, Nu = 0, Ju = 0, Jv = 0, Jw = 0, w_term = 0, kind = 0;
  int x0 = 0, y0 = 0;
  double os = 2, norm = 1, cellx = 1, celly = 1;
  int64_t M = 0;
  bool unit_weights = true;
  Fft2dPlan fft;
  TileLists tl;   // TU x TV tiles: degridding, residue-class and atomic gridding
  TileLists tlm;  // 16 x 16 tiles: matrix-core gridding (empty when the grid is too small)
  int Ku = 0, Kv = 0;         // footprint of one row of G (J, or Jw + J - 1 with w-projection)
  int w_chirp = 0;            // the stored rows are chirp-truncated rows (w_term = 2), footprint J + 2R
  void* d_wrows = nullptr;    // w-projection rows, M * Kv * Ku in the operator's precision
  // chirp rows with Ku Kv <= 64 as their kept entries: per row {mask, offset} and the kept values; the
  // dense rows exist only when asked for ("w_rows_packed" = 0)
  uint4* d_wrec = nullptr;
  void* d_wpack = nullptr;
  uint8_t* d_wpos = nullptr;  // each kept value's tap (jv << 4 | ju)
  int w_slots = 8;            // kept-entry degridding's slots per lane (the largest row's kept taps / 8)
  int w_qe = 1, w_qe_max = 1; // kept-entry gridding's entries per step (tuning knob), and the largest allowed
  bool w_owner = true;         // dense rows: owner-tile gridding where built (tuning knob)
  int w_packed = 0;
  int w_grid_depth = 16;      // tuning knob: kept-entry gridding's prefetch depth in entries
  int64_t nnz_w = 0;
  // a given sparse gridding matrix (pf_op_create_csr): G by rows and conj(G) transposed, by grid cell
  bool g_csr = false;
  int64_t* d_csr_ptr = nullptr;
  int32_t* d_csr_col = nullptr;
  void* d_csr_val = nullptr;
  int64_t* d_csrT_ptr = nullptr;
  int32_t* d_csrT_col = nullptr;
  void* d_csrT_val = nullptr;
  int n_tiles = 0;
  int64_t tile_list_len = 0, tile_list_len_m = 0;
  // device buffers
  int2* d_kk = nullptr;
  cpx<float>* d_ff = nullptr;
  cpx<double>* d_uv = nullptr;
  void* d_weights = nullptr;
  VisPerm vperm;                // caller's order <-> sorted order (arrays below)
  int perm_S_req = 0;           // complex64 permutation bucket requested at build (0: the largest that fits)
  uint16_t* d_perm_l1 = nullptr;
  int32_t* d_perm_g1 = nullptr;
  uint16_t* d_perm_p2 = nullptr;
  int32_t* d_perm = nullptr;   // one-pass plan: sorted -> caller, caller -> sorted
  int32_t* d_inv = nullptr;
  void* d_ysort = nullptr;     // visibilities in sorted order (degridding output, gridding input)
  void* d_ystage = nullptr;    // staging vector of the two-pass permutation
  float* d_tab_u = nullptr;  // complex64 tap tables
  float* d_tab_v = nullptr;
  void* d_sx = nullptr;
  void* d_sy = nullptr;
  void* d_gridA = nullptr;
  void* d_gridB = nullptr;
  void* d_stage128 = nullptr;  // complex128 staging for host-pointer applies
  void* d_vecx = nullptr;      // image-sized scratch (prec) for host applies / power method
  void* d_vecy = nullptr;      // vis-sized scratch (prec)
  double* d_red = nullptr;     // reduction partials + result
  GridGeom<float> gf;
  GridGeom<double> gd;
  std::unique_ptr<Profiler> prof;  // per-kernel timing when enabled
  int grid_variant = GRID_RESIDUE;  // tuning knob: gridding kernel
  int kernel_table = 0;             // tuning knob: 0 no tap tables, 1 in gridding, 2 in gridding and degridding
  int kernel_poly = 0;              // tuning knob: per-tap polynomials (complex64, Ju = Jv), ahead of tables
  float* d_poly = nullptr;
  int fft_col_mode = 0;             // tuning knob: column FFT in one pass (0) or two steps (1)
  int grid_waves = 0;               // tuning knob: matrix-core gridding waves (0: one per 16 x 16 tile)
  int fft_row_bpc = -1;             // tuning knob: row pass blocks per CU (0: one line per block; -1: plan default)
  // bumped by every call that changes what an apply launches (norm, tuning, autotune): a solver that
  // captured this operator's launches into a graph captures again when it differs from its own copy
  uint64_t generation = 0;
  int64_t device_bytes = 0;
  std::vector<void*> allocs;

  void* dalloc(size_t bytes) {
    if (bytes == 0) bytes = 16;
    void* p = dev_alloc(bytes);
    allocs.push_back(p);
    device_bytes += (int64_t)bytes;
    return p;
  }
  ~pf_op() {
    cstream_sync();  // a construction that failed part-way may still have work queued
    for (void* p : allocs) dev_free(p);
    fft_plan_free(fft);
  }
  size_t esz() const { return prec == PF_C128 ? 16 : 8; }
  int64_t nimg() const { return (int64_t)nx * ny; }
  int64_t ngrid() const { return (int64_t)Nu * Nv; }
};

namespace {

bool is_pow2(int n) { return n >= 2 && (n & (n - 1)) == 0; }

int wrapi(int64_t a, int n) {
  int64_t r = a % n;
  return (int)(r < 0 ? r + n : r);
}

// f(begin, end) over [0, n) in contiguous slices on up to 16 host threads (each index in one slice)
template <class F>
void parallel_for(int64_t n, F f) {
  const int nth = (int)std::max<int64_t>(1, std::min<int64_t>({16, (int64_t)std::thread::hardware_concurrency(),
                                                                (n + 65535) / 65536}));
  if (nth == 1) {
    f(int64_t(0), n);
    return;
  }
  std::vector<std::thread> pool;
  for (int k = 0; k < nth; ++k) pool.emplace_back([&, k] { f(n * k / nth, n * (k + 1) / nth); });
  for (auto& th : pool) th.join();
}

// Gridding lists of a TU x TV tiling: for each tile, the sorted index of every visibility with a tap
// in it (ascending), and the tiles in launch order, heaviest first, so that the dense centre of the
// uv-plane does not form the tail of the launch.
// Device scratch of one construction, freed when the scope ends.
struct DevTemp {
  std::vector<void*> bufs;
  ~DevTemp() {
    cstream_sync();
    for (void* p : bufs) dev_free(p);
  }
  template <typename T>
  T* alloc(int64_t n) {
    T* p = dev_alloc_n<T>((size_t)std::max<int64_t>(n, 1));
    bufs.push_back(p);
    return p;
  }
};

// The gridding lists of TU x TV tiles (footprints Ku x Kv at the sorted tap origins kk) into tl: the list
// built on the device, the launch slots (heaviest tile first) from its per-tile counts.  Returns its length.
int device_cus() {
  int dev = 0, ncu = 0;
  PF_HIP(hipGetDevice(&dev));
  PF_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  return std::max(ncu, 1);
}

// split: a tile with more than max(1024, len / (4 CUs)) entries is cut into equal parts of at most that
// many, each a launch slot of its own (see TileLists), so that a dense tile does not form the launch's tail.
int64_t device_lists(pf_op* op, const int2* kk, int64_t M, int Ku, int Kv, int TU, int TV, TileLists& tl,
                     bool split = false) {
  std::vector<uint32_t> count;
  int64_t len = 0;
  uint32_t* list = dev_tile_lists(kk, M, op->Nu, op->Nv, Ku, Kv, TU, TV, count, len);
  op->allocs.push_back(list);
  op->device_bytes += (int64_t)sizeof(uint32_t) * std::max<int64_t>(len, 1);
  const int nt = (int)count.size();
  uint32_t ch = 0xffffffffu;
  if (split) {
    const int64_t per_slot = 4 * (int64_t)device_cus();
    ch = (uint32_t)std::max<int64_t>(1024, (len + per_slot - 1) / per_slot);
  }
  std::vector<uint4> slot, comb;
  slot.reserve((size_t)nt);
  uint32_t run = 0, nparts = 0;
  for (int t = 0; t < nt; ++t) {
    const uint32_t c = count[t];
    if (c <= ch) {
      slot.push_back(make_uint4((uint32_t)t, run, run + c, PART_NONE));
    } else {
      const uint32_t np = (c + ch - 1) / ch, per = (c + np - 1) / np;
      comb.push_back(make_uint4((uint32_t)t, nparts, np, 0u));
      for (uint32_t p = 0; p < np; ++p)
        slot.push_back(make_uint4((uint32_t)t, run + p * per, std::min(run + (p + 1) * per, run + c), nparts++));
    }
    run += c;
  }
  std::stable_sort(slot.begin(), slot.end(), [](const uint4& a, const uint4& b) { return a.z - a.y > b.z - b.y; });
  uint4* d_slot = (uint4*)op->dalloc(sizeof(uint4) * std::max<size_t>(slot.size(), 1));
  h2d(d_slot, slot.data(), sizeof(uint4) * slot.size());
  tl.TU = TU;
  tl.TV = TV;
  tl.ntu = op->Nu / TU;
  tl.ntv = op->Nv / TV;
  tl.list = list;
  tl.slot = d_slot;
  tl.n_slot = (int)slot.size();
  if (!comb.empty()) {
    uint4* d_comb = (uint4*)op->dalloc(sizeof(uint4) * comb.size());
    h2d(d_comb, comb.data(), sizeof(uint4) * comb.size());
    tl.comb = d_comb;
    tl.n_comb = (int)comb.size();
    tl.part = op->dalloc(op->esz() * (size_t)nparts * TU * TV);
  }
  return len;
}

// The permutation plan (VisPerm) for perm (device: sorted position -> caller's index).  complex128: the
// one-pass plan (perm and its inverse; d_perm is the operator's).  complex64: the two-pass plan, whose
// staging layout puts bucket b (sorted positions [b*S, b*S + S)) at staging [b*S, ...), grouped by the
// caller's chunk (chunk c = caller indices [c*S, c*S + S)), each group in the caller's order; chunk c's
// LDS order is by bucket, then by caller index.
void build_perm_dev(pf_op* op, int32_t* d_perm) {
  const int64_t M = op->M;
  VisPerm& p = op->vperm;
  p.M = M;
  p.two_pass = op->prec == PF_C64;
  if (!p.two_pass) {
    op->d_perm = d_perm;
    op->d_inv = (int32_t*)op->dalloc(sizeof(int32_t) * std::max<int64_t>(M, 1));
    dev_inverse(d_perm, M, op->d_inv);
    p.perm = op->d_perm;
    p.inv = op->d_inv;
    return;
  }
  const int Smax = op->perm_S_req > 0 ? std::min(op->perm_S_req, perm_max_S((int)op->esz())) : perm_max_S((int)op->esz());
  p.S = 64;
  while (p.S < M && p.S < Smax) p.S <<= 1;
  p.nb = (int)((M + p.S - 1) / p.S);
  op->d_perm_l1 = (uint16_t*)op->dalloc(sizeof(uint16_t) * std::max<int64_t>(M, 1));
  op->d_perm_g1 = (int32_t*)op->dalloc(sizeof(int32_t) * std::max<int64_t>(M, 1));
  op->d_perm_p2 = (uint16_t*)op->dalloc(sizeof(uint16_t) * std::max<int64_t>(M, 1));
  dev_perm_plan(d_perm, M, p.S, p.nb, op->d_perm_l1, op->d_perm_g1, op->d_perm_p2);
  p.l1 = op->d_perm_l1;
  p.g1 = op->d_perm_g1;
  p.p2 = op->d_perm_p2;
}

void to_pixels(const pf_op_desc& d, int Nu, int Nv, std::vector<double>& u, std::vector<double>& v) {
  const int64_t M = d.M;
  u.assign(d.u, d.u + M);
  v.assign(d.v, d.v + M);
  if (d.units == PF_PIXELS) return;
  double cu = d.cellx, cv = d.celly;
  double su = 1, sv = 1;  // lambda only
  if (d.units == PF_LAMBDA) {
    if (cu == 0 && cv == 0) {
      double mu = 0, mv = 0;
      for (int64_t m = 0; m < M; ++m) {
        mu = std::max(mu, u[m] * u[m]);
        mv = std::max(mv, v[m] * v[m]);
      }
      cu = (180 * 3600) / std::sqrt(mu) / PI / 3;
      cv = (180 * 3600) / std::sqrt(mv) / PI / 3;
    }
    if (cv == 0) cv = cu;
    su = 180 * 3600 / cu / PI;
    sv = 180 * 3600 / cv / PI;
  }
  const bool lambda = d.units == PF_LAMBDA;
  for (int64_t m = 0; m < M; ++m) {
    const double ur = lambda ? u[m] / su * 2 * PI : u[m], vr = lambda ? v[m] / sv * 2 * PI : v[m];
    u[m] = ur / (2 * PI) * Nu;
    v[m] = vr / (2 * PI) * Nv;
  }
}

// w-projection rows (operators.cc:49-99), FP64 on the host threads, in sorted order, streamed to the
// device in chunks in the operator's precision.  G receives w - mean(w) (operators.h:323) and both axes
// of the projection use the u-slot kernel, which is the v-support kernel after the slot swap
// (operators.h:324, projection_kernels.cc:72-73).
// The projection rows of operators.cc:49-99 (w_term = 1), built on the device from the caller's pixel
// coordinates d_uv, w and weights into the sorted order d_perm.
void build_wrows(pf_op* op, const pf_op_desc& d, const cpx<double>* d_uv, const cpx<double>* d_wt,
                 const int32_t* d_perm, DevTemp& tmp) {
  const int64_t M = d.M;
  const int KK = op->Ku * op->Kv;
  op->nnz_w = M * KK;
  double mean = 0;
  for (int64_t m = 0; m < M; ++m) mean += d.w[m];
  if (M > 0) mean /= (double)M;
  WRowArgs g;
  g.Ju = op->Ju;
  g.Jv = op->Jv;
  g.Jw = d.Jw;
  g.Ku = op->Ku;
  g.Kv = op->Kv;
  g.kp = make_kparams<double>(d.kernel, d.Jv, op->os);
  // pixel_to_lambda (projection_kernels.cc:28-30)
  g.du_l = 60. * 60. * 180. / d.cellx / std::floor(op->os * op->nx) / PI;
  g.dv_l = 60. * 60. * 180. / d.celly / std::floor(op->os * op->ny) / PI;
  g.wmean = mean;
  double* d_w = tmp.alloc<double>(M);
  if (M) h2d(d_w, d.w, sizeof(double) * M);
  op->d_wrows = op->dalloc(op->esz() * (size_t)std::max<int64_t>(op->nnz_w, 1));
  if (op->prec == PF_C64) wrows_build<float>(g, d_uv, d_w, d_wt, d_perm, M, static_cast<cpx<float>*>(op->d_wrows));
  else wrows_build<double>(g, d_uv, d_w, d_wt, d_perm, M, static_cast<cpx<double>*>(op->d_wrows));
}

void build_tail(pf_op* op, const pf_op_desc& d);

// Tile side for a requested one (req > 0), else the default: 32, halved while the grid is narrower
// than a tile and its halo.
static int pick_tile(int req, int N) {
  if (req > 0) return std::min(req, N);
  int T = 32;
  while (T > 8 && N < T + 8) T /= 2;
  return std::min(T, N);
}

void build(pf_op* op, const pf_op_desc& d) {
  if (d.imsizex <= 0 || d.imsizey <= 0) throw Error(PF_ERR_ARG, "image size must be positive");
  if (d.M < 0 || (d.M > 0 && (!d.u || !d.v))) throw Error(PF_ERR_ARG, "u and v must be given for M > 0");
  if (d.M > 0x3fffffffLL) throw Error(PF_ERR_UNSUPPORTED, "more than 2^30 visibilities in one operator");
  if (d.kernel < 0 || d.kernel > 5) throw Error(PF_ERR_ARG, "Did not choose valid kernel.");
  if (d.kernel == PF_PSWF && (d.Ju != 6 || d.Jv != 6))
    throw Error(PF_ERR_ARG, "Incorrect input: PSWF requires a support of 6");  // kernels.cc:241-244
  if (d.Ju < 1 || d.Jv < 1 || d.Ju > 16 || d.Jv > 16) throw Error(PF_ERR_UNSUPPORTED, "kernel support must be 1..16");
  if (d.precision != PF_C64 && d.precision != PF_C128) throw Error(PF_ERR_ARG, "precision must be PF_C64 or PF_C128");
  if (d.w_term < 0 || d.w_term > 2) throw Error(PF_ERR_ARG, "w_term must be 0, 1 or 2");
  if (d.w_term) {
    if (d.M > 0 && !d.w) throw Error(PF_ERR_ARG, "w must be given for w_term operators");
    if (d.w_term == 1 && (d.Jw < 1 || d.Jw > 16)) throw Error(PF_ERR_UNSUPPORTED, "w-kernel support must be 1..16");
  }
  op->prec = d.precision;
  op->nx = d.imsizex;
  op->ny = d.imsizey;
  op->os = d.oversample_ratio;
  op->Nu = (int)std::floor(d.imsizex * d.oversample_ratio);
  op->Nv = (int)std::floor(d.imsizey * d.oversample_ratio);
  if (!is_pow2(op->Nu) || !is_pow2(op->Nv))
    throw Error(PF_ERR_UNSUPPORTED, "oversampled grid dimensions must be powers of two (got " + std::to_string(op->Nv) +
                                        " x " + std::to_string(op->Nu) + ")");
  op->x0 = (int)std::floor(op->Nu * 0.5 - op->nx * 0.5);
  op->y0 = (int)std::floor(op->Nv * 0.5 - op->ny * 0.5);
  op->kind = d.kernel;
  op->Ju = std::min(d.Ju, op->Nu);  // operators.cc:26-27
  op->Jv = std::min(d.Jv, op->Nv);
  op->Jw = d.Jw;
  op->w_term = d.w_term ? 1 : 0;
  // footprint of one row of G: J x J taps, or (Jw + J - 1)^2 with w-projection (operators.cc:73-74);
  // chirp rows: J + 2R, R found below
  op->w_chirp = d.w_term == 2 ? 1 : 0;
  op->Ku = op->w_term == 1 && !op->w_chirp ? d.Jw + op->Ju - 1 : op->Ju;
  op->Kv = op->w_term == 1 && !op->w_chirp ? d.Jw + op->Jv - 1 : op->Jv;
  if (op->w_chirp) {
    if (op->Nu != op->Nv) throw Error(PF_ERR_UNSUPPORTED, "chirp w-projection rows need a square grid");
    if (op->Ju != op->Jv) throw Error(PF_ERR_UNSUPPORTED, "chirp w-projection rows need Ju = Jv");
  }
  if (op->w_term && (op->Ku > op->Nu || op->Kv > op->Nv))
    throw Error(PF_ERR_UNSUPPORTED, "w-projection rows wider than the grid");
  op->cellx = d.cellx;
  op->celly = d.celly;
  op->M = d.M;
  op->unit_weights = (d.weights == nullptr);
  if (d.weights) {
    op->unit_weights = true;
    for (int64_t m = 0; m < d.M; ++m)
      if (d.weights[2 * m] != 1.0 || d.weights[2 * m + 1] != 0.0) {
        op->unit_weights = false;
        break;
      }
  }
  fft_plan_init(op->fft, op->Nv, op->Nu, op->prec);

  // Kernel slots: the reference passes (kernelv, kernelu) into the (kernelu, kernelv) slots of the
  // gridding matrix (operators.h:322-324), so the u-axis taps use the v-support kernel and vice versa.
  op->gf.Nu = op->gd.Nu = op->Nu;
  op->gf.Nv = op->gd.Nv = op->Nv;
  op->gf.Ju = op->gd.Ju = op->Ju;
  op->gf.Jv = op->gd.Jv = op->Jv;
  op->gf.ku = make_kparams<float>(d.kernel, d.Jv, op->os);
  op->gf.kv = make_kparams<float>(d.kernel, d.Ju, op->os);
  op->gd.ku = make_kparams<double>(d.kernel, d.Jv, op->os);
  op->gd.kv = make_kparams<double>(d.kernel, d.Ju, op->os);
  if (op->prec == PF_C64 && op->Ju <= 8 && op->Jv <= 8 && !op->w_term) {
    const std::vector<float> tu = make_tap_table(d.kernel, d.Jv, op->Ju, op->os);
    const std::vector<float> tv = make_tap_table(d.kernel, d.Ju, op->Jv, op->os);
    op->d_tab_u = (float*)op->dalloc(sizeof(float) * tu.size());
    op->d_tab_v = (float*)op->dalloc(sizeof(float) * tv.size());
    h2d(op->d_tab_u, tu.data(), sizeof(float) * tu.size());
    h2d(op->d_tab_v, tv.data(), sizeof(float) * tv.size());
    op->gf.tab_u = op->d_tab_u;
    op->gf.tab_v = op->d_tab_v;
    op->kernel_table = 1;
  }
  if (op->prec == PF_C64 && op->Ju == op->Jv && op->Ju <= 8 && !op->w_term) {
    const std::vector<float> pc = make_tap_poly(d.kernel, op->Jv, op->Ju, op->os);
    if (!pc.empty()) {
      op->d_poly = (float*)op->dalloc(sizeof(float) * pc.size());
      h2d(op->d_poly, pc.data(), sizeof(float) * pc.size());
      op->gf.poly = op->d_poly;
      op->kernel_poly = 1;
    }
  }

  // ---- coordinates in pixels, tap origins, sort by owner tile ----
  std::vector<double> u, v;
  to_pixels(d, op->Nu, op->Nv, u, v);
  const int64_t M = d.M;
  std::unique_ptr<ChirpBuild, void (*)(ChirpBuild*)> chirp(nullptr, chirp_build_free);
  if (op->w_chirp) {
    // the kept chirp radius of every row fixes the footprint of the stored rows
    chirp.reset(chirp_build_new());
    const int R = chirp_rows_radius(*chirp, M, op->Nu, op->Ju, d.kernel, d.Jv, d.Ju, op->os, d.cellx, d.celly,
                                    d.energy_chirp, d.energy_wproj, u.data(), v.data(), d.w, d.weights);
    op->Ku = op->Kv = op->Ju + 2 * R;
    if (op->Ku > op->Nu) throw Error(PF_ERR_UNSUPPORTED, "w-projection rows wider than the grid");
  }
  // default tile: 32 x 32, halved on small grids so that N >= T + 8 (the residue-class gridding
  // kernel's condition) whenever the grid allows it
  const int TU = pick_tile(d.tile_u, op->Nu);
  const int TV = pick_tile(d.tile_v, op->Nv);
  if (!is_pow2(TU) && TU != op->Nu) throw Error(PF_ERR_ARG, "tile_u must be a power of two");
  if (!is_pow2(TV) && TV != op->Nv) throw Error(PF_ERR_ARG, "tile_v must be a power of two");
  const int ntu = op->Nu / TU, ntv = op->Nv / TV;
  op->n_tiles = ntu * ntv;
  // ---- on the device: the sort by (owner tile, first-tap cell inside the tile), so that a tile's
  // visibilities form one contiguous range ordered row by row (the gridding lanes stay on one cell), the
  // tile lists, the permutation plan and the sorted records ----
  std::vector<cpx<double>> uvh((size_t)M);
  for (int64_t m = 0; m < M; ++m) uvh[m] = mk<double>(u[m], v[m]);
  DevTemp tmp;
  cpx<double>* d_uvc = tmp.alloc<cpx<double>>(M);
  if (M) h2d(d_uvc, uvh.data(), sizeof(cpx<double>) * M);
  std::vector<cpx<double>>().swap(uvh);
  cpx<double>* d_wtc = nullptr;
  if (!op->unit_weights) {
    d_wtc = tmp.alloc<cpx<double>>(M);
    h2d(d_wtc, d.weights, sizeof(cpx<double>) * M);
  }
  // perm (sorted position -> caller's index): kept as the one-pass plan's for complex128
  int32_t* d_perm = op->prec == PF_C128 ? (int32_t*)op->dalloc(sizeof(int32_t) * std::max<int64_t>(M, 1))
                                        : tmp.alloc<int32_t>(M);
  std::vector<uint32_t> owner_count;
  dev_sort_visibilities(d_uvc, M, op->Nu, op->Nv, op->Ku, op->Kv, TU, TV, d_perm, owner_count);
  // sorted tap origins (footprint origins with w-projection rows)
  int2* d_kk = (int2*)op->dalloc(sizeof(int2) * std::max<int64_t>(M, 1));
  cpx<float>* d_ff = nullptr;
  if (op->prec == PF_C64 && !op->w_term) d_ff = (cpx<float>*)op->dalloc(sizeof(cpx<float>) * std::max<int64_t>(M, 1));
  if (op->prec == PF_C128 && !op->w_term) op->d_uv = (cpx<double>*)op->dalloc(sizeof(cpx<double>) * std::max<int64_t>(M, 1));
  void* d_ws = nullptr;
  if (!op->unit_weights && !op->w_term) d_ws = op->dalloc(op->esz() * std::max<int64_t>(M, 1));
  if (op->prec == PF_C64)
    dev_gather_sorted<float>(d_uvc, d_wtc, d_perm, M, op->Ku, op->Kv, d_kk, d_ff, nullptr, (cpx<float>*)d_ws);
  else
    dev_gather_sorted<double>(d_uvc, d_wtc, d_perm, M, op->Ku, op->Kv, d_kk, nullptr, op->d_uv, (cpx<double>*)d_ws);
  op->d_kk = d_kk;
  op->d_ff = d_ff;
  op->d_weights = d_ws;
  // degridding launch slots: the tiles that own visibilities in chunks, most visibilities first
  {
    std::vector<uint4> so;
    const uint32_t ch = (uint32_t)std::max<int64_t>(512, (M + 8 * device_cus() - 1) / (8 * device_cus()));
    uint32_t run = 0;
    for (int t = 0; t < op->n_tiles; ++t) {
      for (uint32_t b = 0; b < owner_count[t]; b += ch)
        so.push_back(make_uint4((uint32_t)t, run + b, run + std::min(owner_count[t], b + ch), 0u));
      run += owner_count[t];
    }
    std::stable_sort(so.begin(), so.end(), [](const uint4& a, const uint4& b) { return a.z - a.y > b.z - b.y; });
    uint4* dso = (uint4*)op->dalloc(sizeof(uint4) * std::max<size_t>(so.size(), 1));
    if (!so.empty()) h2d(dso, so.data(), sizeof(uint4) * so.size());
    op->tl.slot_own = dso;
    op->tl.n_own = (int)so.size();
    // owner-tile gridding of dense stored rows beyond 8 x 8 taps (up to 24 x 24), when the footprint's
    // extension reaches the neighbouring tiles only: each tile's slots in the order of their
    // visibilities, and the slots' extended partial tiles
    const int EW = TU + op->Ku - 1, EH = TV + op->Kv - 1;
    const size_t ext_bytes = op->esz() * (size_t)EW * EH * so.size();
    if (op->w_term && op->Ku * op->Kv > 64 && op->Ku * op->Kv <= 576 && op->Ku - 1 <= TU && op->Kv - 1 <= TV &&
        op->Nu % TU == 0 && op->Nv % TV == 0 && op->esz() * (size_t)EW * EH <= 160 * 1024 &&
        ext_bytes <= ((size_t)8 << 30)) {
      std::vector<int> first(op->n_tiles + 1, 0), ids(so.size());
      for (const uint4& q : so) ++first[q.x + 1];
      for (int t = 0; t < op->n_tiles; ++t) first[t + 1] += first[t];
      std::vector<int> fill(first.begin(), first.end() - 1);
      std::vector<int> order(so.size());
      for (size_t k = 0; k < so.size(); ++k) order[k] = (int)k;
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return so[a].y < so[b].y; });
      for (int k : order) ids[fill[so[k].x]++] = k;
      int* d_first = (int*)op->dalloc(sizeof(int) * first.size());
      int* d_ids = (int*)op->dalloc(sizeof(int) * std::max<size_t>(ids.size(), 1));
      h2d(d_first, first.data(), sizeof(int) * first.size());
      if (!ids.empty()) h2d(d_ids, ids.data(), sizeof(int) * ids.size());
      op->tl.own_first = d_first;
      op->tl.own_ids = d_ids;
      op->tl.ext = op->dalloc(std::max<size_t>(ext_bytes, 1));
      op->tl.ext_w = EW;
      op->tl.ext_h = EH;
    }
  }
  // gridding lists of the TU x TV tiles (split for the stored-row gridding)
  op->tile_list_len = device_lists(op, d_kk, M, op->Ku, op->Kv, TU, TV, op->tl, op->w_term != 0);
  // the permutation between the caller's order and the sorted order
  build_perm_dev(op, d_perm);
  op->d_ysort = op->dalloc(op->esz() * (size_t)std::max<int64_t>(M, 1));
  if (op->vperm.two_pass) op->d_ystage = op->dalloc(op->esz() * (size_t)std::max<int64_t>(M, 1));
  if (op->w_chirp) {
    // rows of (J + 2R)^2 values at origins floor(u - (J + 2R)/2) = floor(u - J/2) - R; with (J + 2R)^2
    // <= 64 written once with their keep masks and stored as their kept entries
    const int KK = op->Ku * op->Kv;
    if (KK <= 64) {
      void* dense = tmp.alloc<unsigned char>((int64_t)op->esz() * std::max<int64_t>(M * KK, 1));
      uint64_t* mask = tmp.alloc<uint64_t>(M);
      op->nnz_w = chirp_rows_fill(*chirp, d_perm, dense, op->prec, mask);
      op->d_wrec = (uint4*)op->dalloc(sizeof(uint4) * (size_t)std::max<int64_t>(M, 1));
      int max_kept = 0;
      const uint64_t kept = rows_pack_records(mask, M, op->d_wrec, &max_kept, cstream());
      op->w_slots = max_kept <= 16 ? 2 : max_kept <= 32 ? 4 : 8;
      // gridding entries per step: as many as the largest row allows (64 / QE lanes each), while the QE
      // accumulator copies of a tile stay within 16 KiB of LDS (ten waves per CU: with 32 x 32 tiles two
      // entries per step grid complex64 rows 9 % faster than one, and complex128 rows 30 % slower)
      op->w_qe_max = max_kept <= 16 ? 4 : max_kept <= 32 ? 2 : 1;
      op->w_qe = op->w_qe_max;
      while (op->w_qe > 1 && (size_t)op->w_qe * op->tl.TU * op->tl.TV * op->esz() > 16384) op->w_qe /= 2;
      if ((int64_t)kept != op->nnz_w) throw Error(PF_ERR_STATE, "kept-entry count differs from the rows' mask bits");
      // the kept values and their taps, with ROWS_PACK_PAD spare (zero) entries past the last: read, and
      // not used, for the last rows' dropped taps and unused slots
      op->d_wpack = op->dalloc(op->esz() * (size_t)(kept + ROWS_PACK_PAD));
      op->d_wpos = (uint8_t*)op->dalloc((size_t)(kept + ROWS_PACK_PAD));
      PF_HIP(hipMemsetAsync(static_cast<unsigned char*>(op->d_wpack) + op->esz() * kept, 0, op->esz() * ROWS_PACK_PAD,
                            cstream()));
      PF_HIP(hipMemsetAsync(op->d_wpos + kept, 0, ROWS_PACK_PAD, cstream()));
      if (op->prec == PF_C64)
        rows_pack_values<float>(static_cast<const cpx<float>*>(dense), op->d_wrec, M, KK, op->Ku,
                                static_cast<cpx<float>*>(op->d_wpack), op->d_wpos, cstream());
      else
        rows_pack_values<double>(static_cast<const cpx<double>*>(dense), op->d_wrec, M, KK, op->Ku,
                                 static_cast<cpx<double>*>(op->d_wpack), op->d_wpos, cstream());
      op->w_packed = 1;
    } else {
      op->d_wrows = op->dalloc(op->esz() * (size_t)std::max<int64_t>(M * KK, 1));
      op->nnz_w = chirp_rows_fill(*chirp, d_perm, op->d_wrows, op->prec);
    }
    chirp.reset();
  } else if (op->w_term) {
    build_wrows(op, d, d_uvc, d_wtc, d_perm, tmp);
  } else if (grid_mfma_ok(op->Nu, op->Nv, op->Ju, op->Jv) && (op->prec != PF_C64 || M < ((int64_t)1 << 29))) {
    op->tile_list_len_m = device_lists(op, d_kk, M, op->Ju, op->Jv, MFMA_TILE, MFMA_TILE, op->tlm);
    op->grid_variant = GRID_MFMA;
    if (op->prec == PF_C64) {
      // entry records: sorted index, wrapped first-tap cell, fractions (as the kk / ff arrays hold them)
      int4* d_ent = (int4*)op->dalloc(sizeof(int4) * (size_t)std::max<int64_t>(op->tile_list_len_m, 1));
      dev_entries(op->tlm.list, op->tile_list_len_m, d_kk, d_ff, op->Nu, op->Nv, d_ent);
      op->tlm.ent = d_ent;
    }
  }
  PF_HIP(hipStreamSynchronize(cstream()));

  build_tail(op, d);
}

// The separable correction and the grid workspaces, shared by every way of building G.
void build_tail(pf_op* op, const pf_op_desc& d) {
  // ---- correction S(j, i) = sy[j] * sx[i] (operators.cc:101-121) ----
  std::vector<double> sx(op->nx), sy(op->ny);
  for (int i = 0; i < op->nx; ++i)
    sx[i] = 1.0 / host_ft_kernel(d.kernel, (op->x0 + i + 0.5) / op->Nu - 0.5, d.Ju, op->os);
  for (int j = 0; j < op->ny; ++j)
    sy[j] = 1.0 / host_ft_kernel(d.kernel, (op->y0 + j + 0.5) / op->Nv - 0.5, d.Jv, op->os);
  if (op->prec == PF_C64) {
    std::vector<float> fx(sx.begin(), sx.end()), fy(sy.begin(), sy.end());
    op->d_sx = op->dalloc(sizeof(float) * op->nx);
    op->d_sy = op->dalloc(sizeof(float) * op->ny);
    h2d(op->d_sx, fx.data(), sizeof(float) * op->nx);
    h2d(op->d_sy, fy.data(), sizeof(float) * op->ny);
  } else {
    op->d_sx = op->dalloc(sizeof(double) * op->nx);
    op->d_sy = op->dalloc(sizeof(double) * op->ny);
    h2d(op->d_sx, sx.data(), sizeof(double) * op->nx);
    h2d(op->d_sy, sy.data(), sizeof(double) * op->ny);
  }

  // ---- workspaces ----
  op->d_gridA = op->dalloc(op->esz() * op->ngrid());
  op->d_gridB = op->dalloc(op->esz() * op->ngrid());
  op->d_red = (double*)op->dalloc(sizeof(double) * 1040);
}

// An operator on a given sparse gridding matrix G (rows = visibilities in the caller's order, columns
// = grid cells p * Nu + q).  G is stored by rows in the operator's precision; the adjoint uses conj(G)
// transposed, built by a stable counting sort on the column, so every grid cell sums its entries in
// row order: deterministic.
void build_csr(pf_op* op, const pf_op_desc& d, const pf_csr& G) {
  if (d.imsizex <= 0 || d.imsizey <= 0) throw Error(PF_ERR_ARG, "image size must be positive");
  if (d.kernel < 0 || d.kernel > 5) throw Error(PF_ERR_ARG, "Did not choose valid kernel.");
  if (d.kernel == PF_PSWF && (d.Ju != 6 || d.Jv != 6))
    throw Error(PF_ERR_ARG, "Incorrect input: PSWF requires a support of 6");  // kernels.cc:241-244
  if (d.Ju < 1 || d.Jv < 1 || d.Ju > 16 || d.Jv > 16) throw Error(PF_ERR_UNSUPPORTED, "kernel support must be 1..16");
  if (d.precision != PF_C64 && d.precision != PF_C128) throw Error(PF_ERR_ARG, "precision must be PF_C64 or PF_C128");
  op->prec = d.precision;
  op->nx = d.imsizex;
  op->ny = d.imsizey;
  op->os = d.oversample_ratio;
  op->Nu = (int)std::floor(d.imsizex * d.oversample_ratio);
  op->Nv = (int)std::floor(d.imsizey * d.oversample_ratio);
  if (!is_pow2(op->Nu) || !is_pow2(op->Nv))
    throw Error(PF_ERR_UNSUPPORTED, "oversampled grid dimensions must be powers of two (got " + std::to_string(op->Nv) +
                                        " x " + std::to_string(op->Nu) + ")");
  op->x0 = (int)std::floor(op->Nu * 0.5 - op->nx * 0.5);
  op->y0 = (int)std::floor(op->Nv * 0.5 - op->ny * 0.5);
  op->kind = d.kernel;
  op->Ju = d.Ju;
  op->Jv = d.Jv;
  op->Ku = op->Ju;
  op->Kv = op->Jv;
  op->cellx = d.cellx;
  op->celly = d.celly;
  const int64_t Ng = op->ngrid();
  if (G.rows < 0 || G.cols != Ng) throw Error(PF_ERR_ARG, "G must have Nv * Nu columns");
  if (G.rows > 0x3fffffffLL) throw Error(PF_ERR_UNSUPPORTED, "more than 2^30 visibilities in one operator");
  if (!G.rowptr || (G.rows > 0 && G.rowptr[0] != 0)) throw Error(PF_ERR_ARG, "G row pointers must start at 0");
  for (int64_t r = 0; r < G.rows; ++r)
    if (G.rowptr[r + 1] < G.rowptr[r]) throw Error(PF_ERR_ARG, "G row pointers must not decrease");
  const int64_t nnz = G.rows > 0 ? G.rowptr[G.rows] : 0;
  for (int64_t k = 0; k < nnz; ++k)
    if (G.colidx[k] < 0 || G.colidx[k] >= Ng) throw Error(PF_ERR_ARG, "G column index outside the grid");
  op->M = G.rows;
  op->nnz_w = nnz;
  op->g_csr = true;
  fft_plan_init(op->fft, op->Nv, op->Nu, op->prec);
  // transpose: cells in order, each cell's entries in row order
  std::vector<int64_t> tp((size_t)Ng + 1, 0);
  for (int64_t k = 0; k < nnz; ++k) tp[(size_t)G.colidx[k] + 1]++;
  for (int64_t c = 0; c < Ng; ++c) tp[c + 1] += tp[c];
  std::vector<int32_t> trow((size_t)std::max<int64_t>(nnz, 1));
  std::vector<double> tval((size_t)std::max<int64_t>(nnz, 1) * 2);
  {
    std::vector<int64_t> fill(tp.begin(), tp.end() - 1);
    for (int64_t r = 0; r < G.rows; ++r)
      for (int64_t k = G.rowptr[r]; k < G.rowptr[r + 1]; ++k) {
        const int64_t q = fill[G.colidx[k]]++;
        trow[q] = (int32_t)r;
        tval[2 * q] = G.vals[2 * k];
        tval[2 * q + 1] = -G.vals[2 * k + 1];  // conj(G)
      }
  }
  auto upload_vals = [&](const double* v) {
    const size_t n = (size_t)std::max<int64_t>(nnz, 1);
    void* dp = op->dalloc(op->esz() * n);
    if (nnz == 0) return dp;
    if (op->prec == PF_C64) {
      std::vector<float> f((size_t)nnz * 2);
      for (size_t q = 0; q < f.size(); ++q) f[q] = (float)v[q];
      h2d(dp, f.data(), sizeof(float) * f.size());
    } else {
      h2d(dp, v, sizeof(double) * 2 * nnz);
    }
    return dp;
  };
  op->d_csr_ptr = (int64_t*)op->dalloc(sizeof(int64_t) * (G.rows + 1));
  h2d(op->d_csr_ptr, G.rowptr, sizeof(int64_t) * (G.rows + 1));
  op->d_csr_col = (int32_t*)op->dalloc(sizeof(int32_t) * std::max<int64_t>(nnz, 1));
  if (nnz) h2d(op->d_csr_col, G.colidx, sizeof(int32_t) * nnz);
  op->d_csr_val = upload_vals(G.vals);
  op->d_csrT_ptr = (int64_t*)op->dalloc(sizeof(int64_t) * (Ng + 1));
  h2d(op->d_csrT_ptr, tp.data(), sizeof(int64_t) * (Ng + 1));
  op->d_csrT_col = (int32_t*)op->dalloc(sizeof(int32_t) * std::max<int64_t>(nnz, 1));
  if (nnz) h2d(op->d_csrT_col, trow.data(), sizeof(int32_t) * nnz);
  op->d_csrT_val = upload_vals(tval.data());
  build_tail(op, d);
}

template <typename R>
VisData<R> visdata(const pf_op* op) {
  VisData<R> v;
  v.kk = op->d_kk;
  v.ff = op->d_ff;
  v.uv = op->d_uv;
  v.weights = static_cast<const cpx<R>*>(op->d_weights);
  v.perm = op->vperm.two_pass ? nullptr : op->vperm.perm;
  v.M = op->M;
  return v;
}
template <typename R>
const GridGeom<R>& geom(const pf_op* op);
template <>
const GridGeom<float>& geom<float>(const pf_op* op) {
  return op->gf;
}
template <>
const GridGeom<double>& geom<double>(const pf_op* op) {
  return op->gd;
}
// Degridding evaluates the kernel directly unless kernel_table = 2: one dependent table read fewer per
// visibility, which is what bounds the degridding kernel's latency.
template <typename R>
GridGeom<R> degrid_geom(const pf_op* op) {
  GridGeom<R> g = geom<R>(op);
  if (op->kernel_table != 2) g.tab_u = g.tab_v = nullptr;
  return g;
}

template <typename R>
SZ<R> make_sz(const pf_op* op, double scl) {
  SZ<R> sz;
  sz.sx = static_cast<const R*>(op->d_sx);
  sz.sy = static_cast<const R*>(op->d_sy);
  sz.scale = (R)scl;
  sz.nx = op->nx;
  sz.ny = op->ny;
  sz.x0 = op->x0;
  sz.y0 = op->y0;
  return sz;
}

template <typename R>
void fz_forward(pf_op* op, const cpx<R>* x, cpx<R>* grid_out, double scl, hipStream_t s) {
  SZ<R> sz = make_sz<R>(op, scl);
  cpx<R>* A = static_cast<cpx<R>*>(op->d_gridA);
  if (fft_cols_pad_ok(op->fft, op->nx, op->ny)) {
    // columns first, over the image's columns (S applied as they are read, zero-padded to the grid
    // height), into an Nv x nx array; then every row of it, zero-padded to the grid width.  The row
    // pass writes the full grid at its row-contiguous rate; the column pass only the image-wide half
    fft_cols_pad<R>(op->fft, false, x, A, op->nx, RowWindow{op->y0, op->ny}, sz.sy, sz.sx, sz.scale, s);
    sz.img_in = A;
    fft_rows<R>(op->fft, ROW_FWD_X, false, nullptr, grid_out, op->Nv, sz, s);
  } else {
    sz.img_in = x;
    fft_rows<R>(op->fft, ROW_FWD_SZ, false, nullptr, A, op->ny, sz, s);
    fft_cols<R>(op->fft, false, A, grid_out, RowWindow{op->y0, op->ny}, true, false, s);
  }
}

template <typename R>
void fz_adjoint(pf_op* op, cpx<R>* grid_in, cpx<R>* x, double scl, hipStream_t s) {
  SZ<R> sz = make_sz<R>(op, scl);
  sz.img_out = x;
  cpx<R>* B = static_cast<cpx<R>*>(op->d_gridB);
  if (fft_cols_crop_ok(op->fft, op->nx, op->ny)) {
    // rows first, over every row of the grid, each cropped to the image columns (S's column factors and
    // the scale applied); then the columns of that Nv x nx array, cropped to the image rows with S's
    // row factors.  The row pass moves the full grid at its row-contiguous rate, and the column pass,
    // whose 32-byte row segments make it the slower of the two per byte, only the cropped half
    fft_rows<R>(op->fft, ROW_ADJ_X, true, grid_in, B, op->Nv, sz, s);
    fft_cols_crop<R>(op->fft, true, B, x, op->nx, RowWindow{op->y0, op->ny}, sz.sy, s);
  } else {
    fft_cols<R>(op->fft, true, grid_in, B, RowWindow{op->y0, op->ny}, false, true, s);
    fft_rows<R>(op->fft, ROW_ADJ_CROP, true, B, nullptr, op->ny, sz, s);
  }
}

template <typename R>
WRows<R> wrows(const pf_op* op) {
  WRows<R> w;
  w.Ku = op->Ku;
  w.Kv = op->Kv;
  w.kw = op->d_kk;
  w.own = op->w_owner;
  if (op->w_packed) {
    w.vals = static_cast<const cpx<R>*>(op->d_wpack);
    w.rec = op->d_wrec;
    w.pos = op->d_wpos;
    w.slots = op->w_slots;
    w.qe = op->w_qe;
    w.depth = op->w_grid_depth;
    w.off64 = (uint64_t)op->nnz_w >= ((uint64_t)1 << 32);
  } else {
    w.vals = static_cast<const cpx<R>*>(op->d_wrows);
  }
  return w;
}

// y = G grid (the gridding matrix without the norm): Kaiser-Bessel taps or stored projection rows
template <typename R>
void G_dev(pf_op* op, const cpx<R>* grid, cpx<R>* y, hipStream_t s) {
  if (op->g_csr)
    csr_gather<R>(op->M, op->d_csr_ptr, op->d_csr_col, static_cast<const cpx<R>*>(op->d_csr_val), grid, y, R(1), s,
                  "degrid_csr");
  else {
    // complex64: written in the sorted order, then permuted; complex128: written in the caller's order
    const bool two = op->vperm.two_pass;
    cpx<R>* out = two ? static_cast<cpx<R>*>(op->d_ysort) : y;
    if (op->w_term) degrid_wrows<R>(op->Nu, op->Nv, wrows<R>(op), two ? nullptr : op->vperm.perm, op->tl, grid, out, R(1), s);
    else degrid<R>(degrid_geom<R>(op), visdata<R>(op), op->tl, grid, out, R(1), s);
    if (two) permute_to_caller<R>(op->vperm, out, static_cast<cpx<R>*>(op->d_ystage), y, s);
  }
}
// grid = G^H y, y in the caller's order
template <typename R>
void GH_dev(pf_op* op, const cpx<R>* y, cpx<R>* grid, hipStream_t s) {
  if (op->g_csr) {
    csr_gather<R>(op->ngrid(), op->d_csrT_ptr, op->d_csrT_col, static_cast<const cpx<R>*>(op->d_csrT_val), y, grid,
                  R(1), s, "grid_csr");
    return;
  }
  cpx<R>* ys = static_cast<cpx<R>*>(op->d_ysort);
  permute_to_sorted<R>(op->vperm, y, static_cast<cpx<R>*>(op->d_ystage), ys, s);
  if (op->w_term) grid_wrows<R>(op->Nu, op->Nv, wrows<R>(op), op->tl, ys, grid, R(1), s);
  else grid_tiles<R>(geom<R>(op), visdata<R>(op), op->tl, op->tlm, ys, grid, R(1), s, op->grid_variant,
                     op->grid_waves);
}

// y = G grid and grid = G^H y with y in the operator's sorted order (the caller's for a given matrix)
template <typename R>
void G_sorted_dev(pf_op* op, const cpx<R>* grid, cpx<R>* ys, hipStream_t s) {
  if (op->g_csr) {
    G_dev<R>(op, grid, ys, s);
  } else if (op->w_term) {
    degrid_wrows<R>(op->Nu, op->Nv, wrows<R>(op), nullptr, op->tl, grid, ys, R(1), s);
  } else {
    VisData<R> v = visdata<R>(op);
    v.perm = nullptr;  // sorted order out
    degrid<R>(degrid_geom<R>(op), v, op->tl, grid, ys, R(1), s);
  }
}
template <typename R>
void GH_sorted_dev(pf_op* op, const cpx<R>* ys, cpx<R>* grid, hipStream_t s) {
  if (op->g_csr) GH_dev<R>(op, ys, grid, s);
  else if (op->w_term) grid_wrows<R>(op->Nu, op->Nv, wrows<R>(op), op->tl, ys, grid, R(1), s);
  else grid_tiles<R>(geom<R>(op), visdata<R>(op), op->tl, op->tlm, ys, grid, R(1), s, op->grid_variant, op->grid_waves);
}

template <typename R>
void forward_dev(pf_op* op, const cpx<R>* x, cpx<R>* y, hipStream_t s) {
  const double scl = 1.0 / (std::sqrt((double)op->ngrid()) * op->norm);
  cpx<R>* B = static_cast<cpx<R>*>(op->d_gridB);
  fz_forward<R>(op, x, B, scl, s);
  G_dev<R>(op, B, y, s);
}

template <typename R>
void adjoint_dev(pf_op* op, const cpx<R>* y, cpx<R>* x, hipStream_t s) {
  const double scl = 1.0 / (std::sqrt((double)op->ngrid()) * op->norm);
  cpx<R>* A = static_cast<cpx<R>*>(op->d_gridA);
  GH_dev<R>(op, y, A, s);
  fz_adjoint<R>(op, A, x, scl, s);
}

void ensure_stage(pf_op* op);

// Phi^dagger Phi x (compact_operators.h:118-147 chains Z^dagger F^dagger, G^H G and F Z): degridding
// writes the sorted order that gridding reads, so neither permutation runs; the weights enter as
// |w|^2 through the weight in G and its conjugate in G^H.  Exact: no entry of G^H G is pruned.
template <typename R>
void normal_dev(pf_op* op, const cpx<R>* x, cpx<R>* out, hipStream_t s) {
  const double scl = 1.0 / (std::sqrt((double)op->ngrid()) * op->norm);
  cpx<R>* A = static_cast<cpx<R>*>(op->d_gridA);
  cpx<R>* B = static_cast<cpx<R>*>(op->d_gridB);
  fz_forward<R>(op, x, B, scl, s);
  if (op->g_csr) {
    ensure_stage(op);
    cpx<R>* t = static_cast<cpx<R>*>(op->d_vecy);
    G_dev<R>(op, B, t, s);
    GH_dev<R>(op, t, A, s);
  } else {
    cpx<R>* ys = static_cast<cpx<R>*>(op->d_ysort);
    if (op->w_term) {
      degrid_wrows<R>(op->Nu, op->Nv, wrows<R>(op), nullptr, op->tl, B, ys, R(1), s);
      grid_wrows<R>(op->Nu, op->Nv, wrows<R>(op), op->tl, ys, A, R(1), s);
    } else {
      VisData<R> v = visdata<R>(op);
      v.perm = nullptr;  // sorted order out
      degrid<R>(degrid_geom<R>(op), v, op->tl, B, ys, R(1), s);
      grid_tiles<R>(geom<R>(op), v, op->tl, op->tlm, ys, A, R(1), s, op->grid_variant, op->grid_waves);
    }
  }
  fz_adjoint<R>(op, A, out, scl, s);
}

void ensure_stage(pf_op* op) {
  if (op->d_stage128) return;
  const int64_t n = std::max(op->nimg(), op->M);
  op->d_stage128 = op->dalloc(16 * (size_t)std::max<int64_t>(n, 1));
  op->d_vecx = op->dalloc(op->esz() * (size_t)std::max<int64_t>(op->nimg(), 1));
  op->d_vecy = op->dalloc(op->esz() * (size_t)std::max<int64_t>(op->M, 1));
}

// Bring a complex128 host vector (or an op-precision device vector) to an op-precision device vector.
const void* stage_in(pf_op* op, const void* p, int loc, int64_t n, void* dev_scratch, hipStream_t s) {
  if (loc == PF_DEVICE) return p;
  if (loc != PF_HOST) throw Error(PF_ERR_ARG, "location must be PF_HOST or PF_DEVICE");
  ensure_stage(op);
  if (op->prec == PF_C128) {
    PF_HIP(hipMemcpyAsync(dev_scratch, p, 16 * n, hipMemcpyHostToDevice, s));
  } else {
    PF_HIP(hipMemcpyAsync(op->d_stage128, p, 16 * n, hipMemcpyHostToDevice, s));
    c128_to_c64(static_cast<const cpx<double>*>(op->d_stage128), static_cast<cpx<float>*>(dev_scratch), n, s);
  }
  return dev_scratch;
}

void stage_out(pf_op* op, const void* dev, void* p, int64_t n, hipStream_t s) {
  if (op->prec == PF_C128) {
    PF_HIP(hipMemcpyAsync(p, dev, 16 * n, hipMemcpyDeviceToHost, s));
  } else {
    c64_to_c128(static_cast<const cpx<float>*>(dev), static_cast<cpx<double>*>(op->d_stage128), n, s);
    PF_HIP(hipMemcpyAsync(p, op->d_stage128, 16 * n, hipMemcpyDeviceToHost, s));
  }
  PF_HIP(hipStreamSynchronize(s));
}

void apply(pf_op* op, bool adjoint, const void* in, void* out, int in_loc, int out_loc, hipStream_t s) {
  if (!op) throw Error(PF_ERR_ARG, "null operator");
  if ((out_loc != PF_HOST && out_loc != PF_DEVICE) || !in || !out) throw Error(PF_ERR_ARG, "bad buffer arguments");
  const int64_t nin = adjoint ? op->M : op->nimg(), nout = adjoint ? op->nimg() : op->M;
  if (in_loc == PF_HOST || out_loc == PF_HOST) ensure_stage(op);
  const void* din = stage_in(op, in, in_loc, nin, adjoint ? op->d_vecy : op->d_vecx, s);
  void* dout = out_loc == PF_DEVICE ? out : (adjoint ? op->d_vecx : op->d_vecy);
  if (op->prec == PF_C64) {
    if (adjoint) adjoint_dev<float>(op, (const cpx<float>*)din, (cpx<float>*)dout, s);
    else forward_dev<float>(op, (const cpx<float>*)din, (cpx<float>*)dout, s);
  } else {
    if (adjoint) adjoint_dev<double>(op, (const cpx<double>*)din, (cpx<double>*)dout, s);
    else forward_dev<double>(op, (const cpx<double>*)din, (cpx<double>*)dout, s);
  }
  if (out_loc == PF_HOST) stage_out(op, dout, out, nout, s);
}

double read_scalar(pf_op* op, hipStream_t s) {
  double h = 0;
  PF_HIP(hipMemcpyAsync(&h, op->d_red + 1024, sizeof(double), hipMemcpyDeviceToHost, s));
  PF_HIP(hipStreamSynchronize(s));
  return h;
}

template <typename R>
double power_method_t(pf_op* op, int iters, double tol, const double* start) {
  hipStream_t s = cstream();
  ensure_stage(op);
  const int64_t n = op->nimg();
  cpx<R>* x = static_cast<cpx<R>*>(op->d_vecx);
  cpx<R>* y = static_cast<cpx<R>*>(op->d_vecy);
  if (start) {
    stage_in(op, start, PF_HOST, n, x, s);
  } else {
    std::vector<cpx<R>> ones(n, mk<R>(R(1), R(0)));
    PF_HIP(hipMemcpyAsync(x, ones.data(), sizeof(cpx<R>) * n, hipMemcpyHostToDevice, s));
    PF_HIP(hipStreamSynchronize(s));
  }
  norm2_fp64<R>(x, n, op->d_red, op->d_red + 1024, s);
  const double n0 = std::sqrt(read_scalar(op, s));
  if (!(n0 > 0)) throw Error(PF_ERR_ARG, "power method start vector is zero");
  scale_vec<R>(x, n, 1.0 / n0, s);
  double old = 0, val = 1;
  for (int it = 0; it < iters; ++it) {
    forward_dev<R>(op, x, y, s);
    adjoint_dev<R>(op, y, x, s);
    norm2_fp64<R>(x, n, op->d_red, op->d_red + 1024, s);
    val = std::sqrt(read_scalar(op, s));
    if (!(val > 0)) throw Error(PF_ERR_STATE, "Error in operator.");
    if (val != val) throw Error(PF_ERR_STATE, "Error in operator or data corrupted.");
    scale_vec<R>(x, n, 1.0 / val, s);
    if (tol * tol > std::fabs(old - val) / old) {
      old = val;
      break;
    }
    old = val;
  }
  return std::sqrt(old);
}

// Installs the operator's profiler (if any) as the thread's current one for the scope of a call.
struct ProfInstall {
  Profiler* prev;
  explicit ProfInstall(pf_op* op) : prev(current_profiler()) { current_profiler() = op && op->prof ? op->prof.get() : nullptr; }
  ~ProfInstall() { current_profiler() = prev; }
};

template <typename R>
__global__ void k_kernel_eval(KParams<R> p, const double* x, double* out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (double)kern1d<R>(p, (R)x[i]);
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------------------------
namespace pf {
template <typename R>
void op_forward(pf_op* op, const cpx<R>* x, cpx<R>* y, hipStream_t s) {
  forward_dev<R>(op, x, y, s);
}
template <typename R>
void op_adjoint(pf_op* op, const cpx<R>* y, cpx<R>* x, hipStream_t s) {
  adjoint_dev<R>(op, y, x, s);
}
template <typename R>
void op_forward_sorted(pf_op* op, const cpx<R>* x, cpx<R>* ys, hipStream_t s) {
  const double scl = 1.0 / (std::sqrt((double)op->ngrid()) * op->norm);
  cpx<R>* B = static_cast<cpx<R>*>(op->d_gridB);
  fz_forward<R>(op, x, B, scl, s);
  G_sorted_dev<R>(op, B, ys, s);
}
template <typename R>
void op_adjoint_sorted(pf_op* op, const cpx<R>* ys, cpx<R>* x, hipStream_t s) {
  const double scl = 1.0 / (std::sqrt((double)op->ngrid()) * op->norm);
  cpx<R>* A = static_cast<cpx<R>*>(op->d_gridA);
  GH_sorted_dev<R>(op, ys, A, s);
  fz_adjoint<R>(op, A, x, scl, s);
}
template <typename R>
void op_to_sorted(pf_op* op, const cpx<R>* y, cpx<R>* ys, hipStream_t s) {
  if (op->g_csr) PF_HIP(hipMemcpyAsync(ys, y, sizeof(cpx<R>) * op->M, hipMemcpyDeviceToDevice, s));
  else permute_to_sorted<R>(op->vperm, y, static_cast<cpx<R>*>(op->d_ystage), ys, s);
}
template <typename R>
void op_to_caller(pf_op* op, const cpx<R>* ys, cpx<R>* y, hipStream_t s) {
  if (op->g_csr) PF_HIP(hipMemcpyAsync(y, ys, sizeof(cpx<R>) * op->M, hipMemcpyDeviceToDevice, s));
  else permute_to_caller<R>(op->vperm, ys, static_cast<cpx<R>*>(op->d_ystage), y, s);
}
template void op_forward_sorted<float>(pf_op*, const cpx<float>*, cpx<float>*, hipStream_t);
template void op_forward_sorted<double>(pf_op*, const cpx<double>*, cpx<double>*, hipStream_t);
template void op_adjoint_sorted<float>(pf_op*, const cpx<float>*, cpx<float>*, hipStream_t);
template void op_adjoint_sorted<double>(pf_op*, const cpx<double>*, cpx<double>*, hipStream_t);
template void op_to_sorted<float>(pf_op*, const cpx<float>*, cpx<float>*, hipStream_t);
template void op_to_sorted<double>(pf_op*, const cpx<double>*, cpx<double>*, hipStream_t);
template void op_to_caller<float>(pf_op*, const cpx<float>*, cpx<float>*, hipStream_t);
template void op_to_caller<double>(pf_op*, const cpx<double>*, cpx<double>*, hipStream_t);
template <typename R>
void op_normal(pf_op* op, const cpx<R>* x, cpx<R>* out, hipStream_t s) {
  normal_dev<R>(op, x, out, s);
}
template void op_normal<float>(pf_op*, const cpx<float>*, cpx<float>*, hipStream_t);
template void op_normal<double>(pf_op*, const cpx<double>*, cpx<double>*, hipStream_t);
template void op_forward<float>(pf_op*, const cpx<float>*, cpx<float>*, hipStream_t);
template void op_forward<double>(pf_op*, const cpx<double>*, cpx<double>*, hipStream_t);
template void op_adjoint<float>(pf_op*, const cpx<float>*, cpx<float>*, hipStream_t);
template void op_adjoint<double>(pf_op*, const cpx<double>*, cpx<double>*, hipStream_t);
OpShape op_shape(const pf_op* op) { return OpShape{op->ny, op->nx, op->prec, op->M}; }
uint64_t op_generation(const pf_op* op) { return op->generation; }
}  // namespace pf

extern "C" {

void pf_op_desc_init(pf_op_desc* d) {
  std::memset(d, 0, sizeof(*d));
  d->oversample_ratio = 2.0;
  d->kernel = PF_KB;
  d->Ju = d->Jv = 4;
  d->Jw = 6;
  d->cellx = d->celly = 1.0;
  d->units = PF_PIXELS;
  d->precision = PF_C64;
  d->power_iters = 100;
  d->power_tol = 1e-4;
  d->energy_chirp = 0.999;
  d->energy_wproj = 0.999;
}

// The normalisation's power method (operators.h:390-391) runs in FP64 whatever the operator's precision:
// a complex64 operator is normalised by the power iteration of its complex128 twin (the same inputs, built
// in complex128 for the iteration and released), so both precisions get the same norm bit for bit.
extern "C++" {
template <class Build>
static double fp64_norm(const pf_op_desc& desc, Build&& build_fn) {
  pf_op_desc d = desc;
  d.precision = PF_C128;
  d.power_iters = 0;
  std::unique_ptr<pf_op> twin(new pf_op());
  build_fn(twin.get(), d);
  return power_method_t<double>(twin.get(), desc.power_iters, desc.power_tol, nullptr);
}
}

// Seconds per Phi + Phi^dagger pair of a built operator on device vectors (reps pairs after one, HIP
// events on the thread's construction stream).
static double time_pairs(pf_op* op, int reps, int warm = 1) {
  hipStream_t s = cstream();
  DevTemp tmp;
  const int64_t n = (int64_t)op->nx * op->ny;
  struct Events {  // released however the timing ends (a candidate's apply may throw)
    hipEvent_t a = nullptr, b = nullptr;
    ~Events() {
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
    }
  } ev;
  PF_HIP(hipEventCreate(&ev.a));
  PF_HIP(hipEventCreate(&ev.b));
  hipEvent_t a = ev.a, b = ev.b;
  auto run = [&](auto* x, auto* y) {
    using R = std::remove_pointer_t<decltype(x)>;
    using F = decltype(R{}.x);
    PF_HIP(hipMemsetAsync(x, 0, sizeof(R) * (size_t)std::max<int64_t>(n, 1), s));
    PF_HIP(hipMemsetAsync(y, 0, sizeof(R) * (size_t)std::max<int64_t>(op->M, 1), s));
    for (int k = 0; k < std::max(warm, 1); ++k) {
      forward_dev<F>(op, x, y, s);
      adjoint_dev<F>(op, y, x, s);
    }
    PF_HIP(hipEventRecord(a, s));
    for (int k = 0; k < reps; ++k) {
      forward_dev<F>(op, x, y, s);
      adjoint_dev<F>(op, y, x, s);
    }
    PF_HIP(hipEventRecord(b, s));
  };
  if (op->prec == PF_C64) run(tmp.alloc<cpx<float>>(n), tmp.alloc<cpx<float>>(op->M));
  else run(tmp.alloc<cpx<double>>(n), tmp.alloc<cpx<double>>(op->M));
  PF_HIP(hipEventSynchronize(b));
  float ms = 0;
  PF_HIP(hipEventElapsedTime(&ms, a, b));
  return ms * 1e-3 / std::max(reps, 1);
}

// tile_u or tile_v < 0: the tile shape (of the degridding and stored-row gridding tiles, which set the
// visibilities' sort order, so each candidate is a build of its own) is chosen by timing Phi + Phi^dagger
// pairs on a build per candidate: 32 x 32 (the default), 64 x 32, 32 x 64 and 64 x 64 where the grid holds
// them; then, for complex64, the caller-order permutation's bucket (the largest that fits, against half
// of it) on the winning shape.  d gets the winning tile; the winner is returned.
static std::unique_ptr<pf_op> build_tuned(pf_op_desc& d, int reps) {
  const int Nu = (int)std::floor(d.imsizex * d.oversample_ratio), Nv = (int)std::floor(d.imsizey * d.oversample_ratio);
  std::vector<std::pair<int, int>> cand;
  auto add = [&](int tu, int tv) {
    tu = std::min(tu, Nu);
    tv = std::min(tv, Nv);
    for (const auto& c : cand)
      if (c.first == tu && c.second == tv) return;
    cand.emplace_back(tu, tv);
  };
  add(pick_tile(0, Nu), pick_tile(0, Nv));  // the default
  for (const auto& c : {std::make_pair(32, 32), std::make_pair(64, 32), std::make_pair(32, 64), std::make_pair(64, 64)})
    if (Nu >= c.first + 8 && Nv >= c.second + 8) add(c.first, c.second);
  std::unique_ptr<pf_op> best;
  double best_t = 1e30;
  pf_op_desc bd = d;
  for (const auto& c : cand) {
    pf_op_desc dc = d;
    dc.tile_u = c.first;
    dc.tile_v = c.second;
    dc.power_iters = 0;
    std::unique_ptr<pf_op> op(new pf_op());
    double t = 0;
    try {
      build(op.get(), dc);
      t = time_pairs(op.get(), reps, reps);  // warmed up as long as timed: a fresh build idled the GPU
    } catch (const Error& e) {
      // a shape this operator cannot use (a stored-row tile and halo beyond the kernel's LDS, say) is
      // skipped; the default shape comes first, so its failure is the operator's own and propagates
      if (e.code != PF_ERR_UNSUPPORTED || !best) throw;
      continue;
    }
    if (t < best_t) {
      best_t = t;
      best = std::move(op);
      bd = dc;
    }
  }
  if (best->prec == PF_C64 && best->vperm.two_pass && best->vperm.S > 64) {
    pf_op_desc dc = bd;
    std::unique_ptr<pf_op> op(new pf_op());
    op->perm_S_req = best->vperm.S / 2;
    build(op.get(), dc);
    if (time_pairs(op.get(), reps, reps) < best_t) best = std::move(op);
  }
  d.tile_u = bd.tile_u;
  d.tile_v = bd.tile_v;
  return best;
}

int pf_op_create(const pf_op_desc* desc, pf_op** out) {
  if (!desc || !out) return fail(PF_ERR_ARG, "null argument");
  *out = nullptr;
  std::unique_ptr<pf_op> op(new pf_op());
  int rc = guarded([&] {
    pf_op_desc dd = *desc;
    if (dd.tile_u < 0 || dd.tile_v < 0) op = build_tuned(dd, 3);
    else build(op.get(), dd);
    if (desc->power_iters > 0) {
      const double nrm = op->prec == PF_C64 ? fp64_norm(dd, [](pf_op* t, const pf_op_desc& d) { build(t, d); })
                                            : power_method_t<double>(op.get(), desc->power_iters, desc->power_tol, nullptr);
      if (!(nrm > 0)) throw Error(PF_ERR_ARG, "Operator norm is not greater than zero.");  // operators.h:187
      op->norm = nrm;
    }
  });
  if (rc == PF_OK) *out = op.release();
  return rc;
}

int pf_op_create_csr(const pf_op_desc* desc, const pf_csr* G, pf_op** out) {
  if (!desc || !G || !out) return fail(PF_ERR_ARG, "null argument");
  *out = nullptr;
  std::unique_ptr<pf_op> op(new pf_op());
  int rc = guarded([&] {
    build_csr(op.get(), *desc, *G);
    if (desc->power_iters > 0) {
      const double nrm =
          op->prec == PF_C64 ? fp64_norm(*desc, [G](pf_op* t, const pf_op_desc& d) { build_csr(t, d, *G); })
                             : power_method_t<double>(op.get(), desc->power_iters, desc->power_tol, nullptr);
      if (!(nrm > 0)) throw Error(PF_ERR_ARG, "Operator norm is not greater than zero.");  // operators.h:187
      op->norm = nrm;
    }
  });
  if (rc == PF_OK) *out = op.release();
  return rc;
}

int pf_wprojection_matrix(const pf_csr* G, int32_t x_size, int32_t y_size, const double* w, double cell_x,
                          double cell_y, double energy_chirp, double energy_wproj, pf_csr* GW) {
  if (!G || !GW || (G->rows > 0 && (!w || !G->rowptr))) return fail(PF_ERR_ARG, "null argument");
  *GW = pf_csr{0, 0, 0, nullptr, nullptr, nullptr};
  return guarded([&] {
    HostCsr h;
    wprojection_matrix_dev(G->rows, G->rowptr, G->colidx, G->vals, x_size, y_size, w, cell_x, cell_y, energy_chirp,
                           energy_wproj, h);
    const int64_t nnz = h.rowptr.empty() ? 0 : h.rowptr.back();
    pf_csr r{h.rows, h.cols, nnz, nullptr, nullptr, nullptr};
    r.rowptr = (int64_t*)std::malloc(sizeof(int64_t) * (size_t)(h.rows + 1));
    r.colidx = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1));
    r.vals = (double*)std::malloc(sizeof(double) * 2 * (size_t)std::max<int64_t>(nnz, 1));
    if (!r.rowptr || !r.colidx || !r.vals) {
      pf_csr_free(&r);
      throw std::bad_alloc();
    }
    std::memcpy(r.rowptr, h.rowptr.data(), sizeof(int64_t) * (size_t)(h.rows + 1));
    if (nnz) {
      std::memcpy(r.colidx, h.col.data(), sizeof(int32_t) * (size_t)nnz);
      std::memcpy(r.vals, h.val.data(), sizeof(double) * 2 * (size_t)nnz);
    }
    *GW = r;
  });
}

void pf_csr_free(pf_csr* m) {
  if (!m) return;
  std::free(m->rowptr);
  std::free(m->colidx);
  std::free(m->vals);
  m->rowptr = nullptr;
  m->colidx = nullptr;
  m->vals = nullptr;
  m->rows = m->cols = m->nnz = 0;
}

int pf_op_destroy(pf_op* op) {
  delete op;
  return PF_OK;
}

}  // extern "C"

// ---- the super operator over visibility blocks (measurementoperator::init_combine_operators,
// operators.h:513-527): y_b = Phi_b x, x = sum_b Phi_b^dagger y_b.  The blocks share the image, the grid and
// the correction S (checked at creation), so F Z S runs once per apply: forward, one row and column pass
// feed every block's G_b; adjoint, every G_b^H (scaled by 1 / ||Phi_b||) accumulates into one grid that
// one adjoint FFT takes back.  Each block keeps its own norm, as each block's create_operator call
// normalises it in the reference.
struct pf_super {
  std::vector<pf_op*> ops;   // not owned
  std::vector<int64_t> off;  // first visibility of each block in the stacked vector
  int64_t M = 0;
};

namespace pf {
namespace {
template <typename R>
__global__ void k_axpy(const cpx<R>* __restrict__ x, cpx<R>* __restrict__ y, R a, int64_t n) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    const cpx<R> u = x[k], v = y[k];
    y[k] = mk<R>(v.x + a * u.x, v.y + a * u.y);
  }
}
template <typename R>
void axpy(const cpx<R>* x, cpx<R>* y, double a, int64_t n, hipStream_t s) {
  const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192));
  hipLaunchKernelGGL((k_axpy<R>), dim3(blocks), dim3(256), 0, s, x, y, (R)a, n);
}
template <typename R>
void super_forward(pf_super* sp, const cpx<R>* x, cpx<R>* y, hipStream_t s) {
  pf_op* o0 = sp->ops[0];
  cpx<R>* B = static_cast<cpx<R>*>(o0->d_gridB);
  fz_forward<R>(o0, x, B, 1.0 / std::sqrt((double)o0->ngrid()), s);
  for (size_t b = 0; b < sp->ops.size(); ++b) {
    pf_op* ob = sp->ops[b];
    if (ob->M == 0) continue;
    G_dev<R>(ob, B, y + sp->off[b], s);
    scale_vec<R>(y + sp->off[b], ob->M, 1.0 / ob->norm, s);
  }
}
template <typename R>
void super_adjoint(pf_super* sp, const cpx<R>* y, cpx<R>* x, hipStream_t s) {
  pf_op* o0 = sp->ops[0];
  cpx<R>* A = static_cast<cpx<R>*>(o0->d_gridA);
  GH_dev<R>(o0, y, A, s);
  scale_vec<R>(A, o0->ngrid(), 1.0 / o0->norm, s);
  for (size_t b = 1; b < sp->ops.size(); ++b) {
    pf_op* ob = sp->ops[b];
    if (ob->M == 0) continue;
    cpx<R>* Ab = static_cast<cpx<R>*>(ob->d_gridA);
    GH_dev<R>(ob, y + sp->off[b], Ab, s);
    axpy<R>(Ab, A, 1.0 / ob->norm, ob->ngrid(), s);
  }
  fz_adjoint<R>(o0, A, x, 1.0 / std::sqrt((double)o0->ngrid()), s);
}
}  // namespace
}  // namespace pf

extern "C" {

int pf_super_create(pf_op* const* ops, int32_t nblocks, pf_super** out) {
  if (!ops || !out || nblocks < 1) return fail(PF_ERR_ARG, "need at least one operator");
  *out = nullptr;
  return guarded([&] {
    std::unique_ptr<pf_super> sp(new pf_super());
    const pf_op* a = ops[0];
    for (int i = 0; i < nblocks; ++i) {
      const pf_op* b = ops[i];
      if (!b) throw Error(PF_ERR_ARG, "null operator");
      if (b->prec != a->prec || b->ny != a->ny || b->nx != a->nx || b->Nu != a->Nu || b->Nv != a->Nv ||
          b->kind != a->kind || b->Ju != a->Ju || b->Jv != a->Jv || b->os != a->os)
        throw Error(PF_ERR_ARG, "the blocks' operators must share precision, image and grid sizes, kernel and supports");
      sp->ops.push_back(ops[i]);
      sp->off.push_back(sp->M);
      sp->M += b->M;
    }
    *out = sp.release();
  });
}

int pf_super_size(const pf_super* sp, int64_t* M) {
  if (!sp || !M) return fail(PF_ERR_ARG, "null argument");
  *M = sp->M;
  return PF_OK;
}

int pf_super_forward(pf_super* sp, const void* x_dev, void* y_dev, pf_stream stream) {
  return guarded([&] {
    if (!sp || !x_dev || (!y_dev && sp->M)) throw Error(PF_ERR_ARG, "null argument");
    if (sp->ops[0]->prec == PF_C64)
      super_forward<float>(sp, (const cpx<float>*)x_dev, (cpx<float>*)y_dev, (hipStream_t)stream);
    else
      super_forward<double>(sp, (const cpx<double>*)x_dev, (cpx<double>*)y_dev, (hipStream_t)stream);
    PF_HIP(hipGetLastError());
  });
}

int pf_super_adjoint(pf_super* sp, const void* y_dev, void* x_dev, pf_stream stream) {
  return guarded([&] {
    if (!sp || !x_dev || (!y_dev && sp->M)) throw Error(PF_ERR_ARG, "null argument");
    if (sp->ops[0]->prec == PF_C64)
      super_adjoint<float>(sp, (const cpx<float>*)y_dev, (cpx<float>*)x_dev, (hipStream_t)stream);
    else
      super_adjoint<double>(sp, (const cpx<double>*)y_dev, (cpx<double>*)x_dev, (hipStream_t)stream);
    PF_HIP(hipGetLastError());
  });
}

int pf_super_destroy(pf_super* sp) {
  delete sp;
  return PF_OK;
}

int pf_op_forward(pf_op* op, const void* x, void* y, int x_loc, int y_loc, pf_stream stream) {
  return guarded([&] {
    ProfInstall pi(op);
    apply(op, false, x, y, x_loc, y_loc, io_stream(stream, x_loc == PF_HOST && y_loc == PF_HOST));
  });
}

int pf_op_adjoint(pf_op* op, const void* y, void* x, int y_loc, int x_loc, pf_stream stream) {
  return guarded([&] {
    ProfInstall pi(op);
    apply(op, true, y, x, y_loc, x_loc, io_stream(stream, x_loc == PF_HOST && y_loc == PF_HOST));
  });
}

int pf_op_apply_normal(pf_op* op, const void* x, void* out, int x_loc, int out_loc, pf_stream stream) {
  return guarded([&] {
    if (!op) throw Error(PF_ERR_ARG, "null operator");
    if ((out_loc != PF_HOST && out_loc != PF_DEVICE) || !x || !out) throw Error(PF_ERR_ARG, "bad buffer arguments");
    ProfInstall pi(op);
    hipStream_t s = io_stream(stream, x_loc == PF_HOST && out_loc == PF_HOST);
    const int64_t n = op->nimg();
    if (x_loc == PF_HOST || out_loc == PF_HOST) ensure_stage(op);
    const void* din = stage_in(op, x, x_loc, n, op->d_vecx, s);
    // in place is fine: x is consumed into the grid before the output is written
    void* dout = out_loc == PF_DEVICE ? out : op->d_vecx;
    if (op->prec == PF_C64) normal_dev<float>(op, (const cpx<float>*)din, (cpx<float>*)dout, s);
    else normal_dev<double>(op, (const cpx<double>*)din, (cpx<double>*)dout, s);
    if (out_loc == PF_HOST) stage_out(op, dout, out, n, s);
  });
}

// ---------------------------------------------------------------------------------------------
// PSF convolution (init_psf_convolve_2d, compact_operators.h:279-306): psf = Phi^dagger Phi delta
// with delta at (floor(ny/2) - 1, floor(nx/2) - 1); x -> IFFT(FFT(psf) . FFT(x)) on the n x n image
// (unitary FFTs, no padding: a circular convolution), divided by the square of its own power-method
// norm.
// ---------------------------------------------------------------------------------------------
struct pf_psf {
  pf_op* op = nullptr;  // not owned: the measurement operator must outlive this handle
  int ny = 0, nx = 0, prec = 0;
  Fft2dPlan fft;
  void* d_ft = nullptr;  // FFT(psf), n x n in the operator's precision
  void* d_w1 = nullptr;  // work images
  void* d_w2 = nullptr;
  void* d_x = nullptr;   // staged input
  double* d_red = nullptr;
  double norm = 1.0;
  std::vector<void*> allocs;
  void* dalloc(size_t bytes) {
    void* p = dev_alloc(std::max<size_t>(bytes, 16));
    allocs.push_back(p);
    return p;
  }
  size_t esz() const { return prec == PF_C64 ? 8 : 16; }
  int64_t n() const { return (int64_t)ny * nx; }
  ~pf_psf() {
    cstream_sync();
    for (void* p : allocs) dev_free(p);
    fft_plan_free(fft);
  }
};

extern "C++" {
namespace {
// out = IFFT(ft . FFT(x)) * sc; x is not modified (x and out may alias)
template <typename R>
void psf_apply_dev(pf_psf* h, const cpx<R>* x, cpx<R>* out, double sc, hipStream_t s) {
  SZ<R> sz;
  cpx<R>* w1 = static_cast<cpx<R>*>(h->d_w1);
  cpx<R>* w2 = static_cast<cpx<R>*>(h->d_w2);
  fft_rows<R>(h->fft, ROW_PLAIN, false, x, w1, h->ny, sz, s);
  fft_cols<R>(h->fft, false, w1, w2, RowWindow{0, 0}, false, false, s);
  // both unitary scalings (1/sqrt(N) each way) and the normalisation in the pointwise product
  cmul_vec<R>(w2, static_cast<const cpx<R>*>(h->d_ft), h->n(), sc / (double)h->n(), s);
  fft_rows<R>(h->fft, ROW_PLAIN, true, w2, w1, h->ny, sz, s);
  fft_cols<R>(h->fft, true, w1, out, RowWindow{0, 0}, false, false, s);
}

template <typename R>
void psf_build(pf_psf* h, int iters, double tol, const double* start) {
  pf_op* op = h->op;
  hipStream_t s = cstream();
  const int64_t n = h->n();
  // psf = Phi^dagger Phi delta with the measurement operator (its norm included)
  std::vector<cpx<R>> delta((size_t)n, mk<R>(R(0), R(0)));
  const int64_t idx = (int64_t)((int)std::floor(h->ny * 0.5) - 1) * h->nx + ((int)std::floor(h->nx * 0.5) - 1);
  if (idx < 0 || idx >= n) throw Error(PF_ERR_ARG, "image too small for the PSF's delta");
  delta[(size_t)idx] = mk<R>(R(1), R(0));
  cpx<R>* x = static_cast<cpx<R>*>(h->d_x);
  cpx<R>* ft = static_cast<cpx<R>*>(h->d_ft);
  PF_HIP(hipMemcpyAsync(x, delta.data(), sizeof(cpx<R>) * n, hipMemcpyHostToDevice, s));
  normal_dev<R>(op, x, x, s);
  // FFT(psf), unitary (the reference's fftop divides by sqrt(N))
  SZ<R> sz;
  fft_rows<R>(h->fft, ROW_PLAIN, false, x, static_cast<cpx<R>*>(h->d_w1), h->ny, sz, s);
  fft_cols<R>(h->fft, false, static_cast<cpx<R>*>(h->d_w1), ft, RowWindow{0, 0}, false, false, s);
  scale_vec<R>(ft, n, 1.0 / std::sqrt((double)n), s);
  PF_HIP(hipStreamSynchronize(s));
  h->norm = 1.0;
  if (iters <= 0) return;
  // power method on x -> psf_convolve(x) (adjoint: the identity), the iteration of SOPT's
  // power_method; the start vector is the caller's or uniform in [-1, 1] (re, im) from
  // mt19937_64(0x5eed): a constant start is an eigenvector of every circular convolution
  std::vector<cpx<double>> st((size_t)n);
  if (start) {
    for (int64_t k = 0; k < n; ++k) st[(size_t)k] = mk<double>(start[2 * k], start[2 * k + 1]);
  } else {
    std::mt19937_64 rng(0x5eed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    for (auto& c : st) {
      const double re = U(rng);
      c = mk<double>(re, U(rng));
    }
  }
  std::vector<cpx<R>> stR((size_t)n);
  for (int64_t k = 0; k < n; ++k) stR[(size_t)k] = mk<R>((R)st[(size_t)k].x, (R)st[(size_t)k].y);
  PF_HIP(hipMemcpyAsync(x, stR.data(), sizeof(cpx<R>) * n, hipMemcpyHostToDevice, s));
  auto nrm = [&]() {
    norm2_fp64<R>(x, n, h->d_red, h->d_red + 1024, s);
    double v = 0;
    PF_HIP(hipMemcpyAsync(&v, h->d_red + 1024, sizeof(double), hipMemcpyDeviceToHost, s));
    PF_HIP(hipStreamSynchronize(s));
    return std::sqrt(v);
  };
  const double n0 = nrm();
  if (!(n0 > 0)) throw Error(PF_ERR_ARG, "power method start vector is zero");
  scale_vec<R>(x, n, 1.0 / n0, s);
  double old = 0, val = 1;
  for (int it = 0; it < iters; ++it) {
    psf_apply_dev<R>(h, x, x, 1.0, s);
    val = nrm();
    if (!(val > 0)) throw Error(PF_ERR_STATE, "Error in operator.");
    scale_vec<R>(x, n, 1.0 / val, s);
    if (tol * tol > std::fabs(old - val) / old) {
      old = val;
      break;
    }
    old = val;
  }
  // init_normalise(op_norm^2) with op_norm = sqrt(old)
  h->norm = std::sqrt(old);
}
}  // namespace
}  // extern "C++"

int pf_psf_create(pf_op* op, int32_t power_iters, double power_tol, const double* start_host, pf_psf** out) {
  if (!op || !out) return fail(PF_ERR_ARG, "null argument");
  *out = nullptr;
  std::unique_ptr<pf_psf> h(new pf_psf());
  int rc = guarded([&] {
    if (!is_pow2(op->ny) || !is_pow2(op->nx))
      throw Error(PF_ERR_UNSUPPORTED, "PSF convolution needs power-of-two image sides");
    h->op = op;
    h->ny = op->ny;
    h->nx = op->nx;
    h->prec = op->prec;
    fft_plan_init(h->fft, h->ny, h->nx, h->prec);
    const size_t b = h->esz() * (size_t)h->n();
    h->d_ft = h->dalloc(b);
    h->d_w1 = h->dalloc(b);
    h->d_w2 = h->dalloc(b);
    h->d_x = h->dalloc(b);
    h->d_red = (double*)h->dalloc(sizeof(double) * 1025);
    if (h->prec == PF_C64) psf_build<float>(h.get(), power_iters, power_tol, start_host);
    else psf_build<double>(h.get(), power_iters, power_tol, start_host);
  });
  if (rc == PF_OK) *out = h.release();
  return rc;
}

int pf_psf_apply(pf_psf* h, const void* x, void* out, int x_loc, int out_loc, pf_stream stream) {
  return guarded([&] {
    if (!h) throw Error(PF_ERR_ARG, "null handle");
    if ((out_loc != PF_HOST && out_loc != PF_DEVICE) || !x || !out) throw Error(PF_ERR_ARG, "bad buffer arguments");
    hipStream_t s = io_stream(stream, x_loc == PF_HOST && out_loc == PF_HOST);
    const int64_t n = h->n();
    pf_op* op = h->op;
    if (x_loc == PF_HOST || out_loc == PF_HOST) ensure_stage(op);
    const void* din = stage_in(op, x, x_loc, n, h->d_x, s);
    void* dout = out_loc == PF_DEVICE ? out : h->d_x;
    const double sc = 1.0 / (h->norm * h->norm);
    if (h->prec == PF_C64) psf_apply_dev<float>(h, (const cpx<float>*)din, (cpx<float>*)dout, sc, s);
    else psf_apply_dev<double>(h, (const cpx<double>*)din, (cpx<double>*)dout, sc, s);
    if (out_loc == PF_HOST) stage_out(op, dout, out, n, s);
  });
}

int pf_psf_norm(const pf_psf* h, double* norm) {
  if (!h || !norm) return fail(PF_ERR_ARG, "null argument");
  *norm = h->norm;
  return PF_OK;
}

int pf_psf_destroy(pf_psf* h) {
  delete h;
  return PF_OK;
}

static int set_tuning_impl(pf_op* op, const char* key, int64_t value);
int pf_op_set_tuning(pf_op* op, const char* key, int64_t value) {
  const int st = set_tuning_impl(op, key, value);
  if (st == PF_OK) ++op->generation;
  return st;
}
static int set_tuning_impl(pf_op* op, const char* key, int64_t value) {
  if (!op || !key) return fail(PF_ERR_ARG, "null argument");
  const std::string k(key);
  if (k == "grid_kernel") {
    if (value != GRID_RESIDUE && value != GRID_ATOMIC && value != GRID_MFMA)
      return fail(PF_ERR_ARG, "grid_kernel must be 0, 1 or 2");
    if (value == GRID_MFMA && !op->tlm.slot)
      return fail(PF_ERR_UNSUPPORTED, "matrix-core gridding needs grid sides >= 16 + J - 1");
    op->grid_variant = (int)value;
    return PF_OK;
  }
  if (k == "fft_columns") {
    if (value != 0 && value != 1) return fail(PF_ERR_ARG, "fft_columns must be 0 or 1");
    return guarded([&] {
      Fft2dPlan np;  // built first: a mode the grid cannot use leaves the current plan in place
      fft_plan_init(np, op->Nv, op->Nu, op->prec, (int)value);
      fft_plan_free(op->fft);
      op->fft = np;
      if (op->fft_row_bpc >= 0) op->fft.row_bpc = op->fft_row_bpc;
      op->fft_col_mode = (int)value;
    });
  }
  if (k == "grid_waves") {
    if (value < 0 || value > (int64_t)1 << 24) return fail(PF_ERR_ARG, "grid_waves must be 0..2^24");
    op->grid_waves = (int)value;
    return PF_OK;
  }
  if (k == "fft_row_blocks") {
    if (value < 0 || value > 64) return fail(PF_ERR_ARG, "fft_row_blocks must be 0..64");
    op->fft_row_bpc = (int)value;
    op->fft.row_bpc = (int)value;
    return PF_OK;
  }
  if (k == "kernel_poly") {
    if (value != 0 && value != 1) return fail(PF_ERR_ARG, "kernel_poly must be 0 or 1");
    if (value && !op->d_poly)
      return fail(PF_ERR_UNSUPPORTED, "tap polynomials exist only for complex64, Ju = Jv <= 8 and smooth kernels");
    op->kernel_poly = (int)value;
    op->gf.poly = value ? op->d_poly : nullptr;
    return PF_OK;
  }
  if (k == "w_grid_owner") {
    if (value != 0 && value != 1) return fail(PF_ERR_ARG, "w_grid_owner must be 0 or 1");
    if (value && !op->tl.ext)
      return fail(PF_ERR_UNSUPPORTED, "owner-tile gridding needs dense stored rows of 65 to 576 taps");
    op->w_owner = value != 0;
    return PF_OK;
  }
  if (k == "w_grid_entries") {
    if (value != 1 && value != 2 && value != 4) return fail(PF_ERR_ARG, "w_grid_entries must be 1, 2 or 4");
    if (value > 1 && (!op->d_wpack || value > op->w_qe_max))
      return fail(PF_ERR_UNSUPPORTED, "w_grid_entries above 1 needs kept-entry rows of at most 64 / w_grid_entries taps");
    if ((size_t)value * op->tl.TU * op->tl.TV * op->esz() > 160 * 1024)
      return fail(PF_ERR_UNSUPPORTED, "w_grid_entries copies of the tile exceed 160 KiB of LDS");
    op->w_qe = (int)value;
    return PF_OK;
  }
  if (k == "w_grid_depth") {
    if (value != 8 && value != 16 && value != 32) return fail(PF_ERR_ARG, "w_grid_depth must be 8, 16 or 32");
    op->w_grid_depth = (int)value;
    return PF_OK;
  }
  if (k == "w_rows_packed") {
    if (value != 0 && value != 1) return fail(PF_ERR_ARG, "w_rows_packed must be 0 or 1");
    if (!op->d_wpack) {
      if (value) return fail(PF_ERR_UNSUPPORTED, "kept-entry rows exist only for chirp rows of at most 8 x 8 taps");
      return PF_OK;
    }
    return guarded([&] {
      if (!value && !op->d_wrows) {  // the dense rows, rebuilt once from the kept entries
        const int64_t KK = (int64_t)op->Ku * op->Kv;
        void* dense = op->dalloc(op->esz() * (size_t)std::max<int64_t>(op->M * KK, 1));
        if (op->prec == PF_C64)
          rows_unpack<float>(static_cast<const cpx<float>*>(op->d_wpack), op->d_wrec, op->M, (int)KK,
                             static_cast<cpx<float>*>(dense), cstream());
        else
          rows_unpack<double>(static_cast<const cpx<double>*>(op->d_wpack), op->d_wrec, op->M, (int)KK,
                              static_cast<cpx<double>*>(dense), cstream());
        PF_HIP(hipStreamSynchronize(cstream()));
        op->d_wrows = dense;
      }
      op->w_packed = (int)value;
    });
  }
  if (k == "kernel_table") {
    if (value < 0 || value > 2) return fail(PF_ERR_ARG, "kernel_table must be 0, 1 or 2");
    if (value && !op->d_tab_u) return fail(PF_ERR_UNSUPPORTED, "tap tables exist only for complex64 and J <= 8");
    op->kernel_table = (int)value;
    op->gf.tab_u = value ? op->d_tab_u : nullptr;
    op->gf.tab_v = value ? op->d_tab_v : nullptr;
    return PF_OK;
  }
  return fail(PF_ERR_ARG, "unknown tuning key " + k);
}

int pf_op_get_tuning(const pf_op* op, const char* key, int64_t* value) {
  if (!op || !key || !value) return fail(PF_ERR_ARG, "null argument");
  const std::string k(key);
  if (k == "grid_kernel") {
    *value = op->grid_variant;
    return PF_OK;
  }
  if (k == "grid_waves") {
    *value = op->grid_waves;
    return PF_OK;
  }
  if (k == "fft_row_blocks") {
    *value = op->fft.row_bpc;
    return PF_OK;
  }
  if (k == "fft_columns") {
    *value = op->fft_col_mode;
    return PF_OK;
  }
  if (k == "kernel_poly") {
    *value = op->kernel_poly;
    return PF_OK;
  }
  if (k == "kernel_table") {
    *value = op->kernel_table;
    return PF_OK;
  }
  if (k == "tile_u") {
    *value = op->tl.TU;
    return PF_OK;
  }
  if (k == "w_footprint") {
    *value = op->w_term ? op->Ku : 0;
    return PF_OK;
  }
  if (k == "split_tiles") {
    *value = op->tl.n_comb;
    return PF_OK;
  }
  if (k == "w_rows_packed") {
    *value = op->w_packed;
    return PF_OK;
  }
  if (k == "w_grid_depth") {
    *value = op->w_grid_depth;
    return PF_OK;
  }
  if (k == "w_grid_owner") {
    *value = op->tl.ext && op->w_owner ? 1 : 0;
    return PF_OK;
  }
  if (k == "w_grid_entries") {
    *value = op->d_wpack ? op->w_qe : 0;
    return PF_OK;
  }
  if (k == "w_degrid_slots") {
    *value = op->d_wpack ? op->w_slots : 0;
    return PF_OK;
  }
  if (k == "perm_bucket") {
    *value = op->vperm.two_pass ? op->vperm.S : 0;
    return PF_OK;
  }
  if (k == "tile_v") {
    *value = op->tl.TV;
    return PF_OK;
  }
  return fail(PF_ERR_ARG, "unknown tuning key " + k);
}

int pf_op_autotune(pf_op* op, int32_t reps, int64_t* chosen) {
  if (!op) return fail(PF_ERR_ARG, "null operator");
  return guarded([&] {
    // projection rows and a given matrix have one gridding kernel each (kept-entry rows: its entries per
    // step are chosen), then the column FFT
    const bool one_grid_kernel = op->w_term || op->g_csr;
    ensure_stage(op);
    hipStream_t s = cstream();
    PF_HIP(hipMemsetAsync(op->d_vecy, 0, op->esz() * (size_t)std::max<int64_t>(op->M, 1), s));
    hipEvent_t a, b;
    PF_HIP(hipEventCreate(&a));
    PF_HIP(hipEventCreate(&b));
    const int n = std::max(1, (int)reps);
    float best = 1e30f;
    int best_v = op->grid_variant;
    // the LDS-atomic variant's in-wave summation order is not architecturally fixed, so it is never
    // chosen here: every candidate gives run-to-run bit-identical grids
    for (int variant : {GRID_RESIDUE, GRID_MFMA}) {
      if (one_grid_kernel) break;
      if (variant == GRID_RESIDUE && !grid_rc_ok(op->Nu, op->Nv, op->tl.TU, op->tl.TV, op->Ju, op->Jv)) continue;
      if (variant == GRID_MFMA && !op->tlm.slot) continue;
      op->grid_variant = variant;
      auto run = [&] {
        if (op->prec == PF_C64)
          grid_tiles<float>(op->gf, visdata<float>(op), op->tl, op->tlm, (const cpx<float>*)op->d_vecy,
                            (cpx<float>*)op->d_gridA, 1.0f, s, variant, op->grid_waves);
        else
          grid_tiles<double>(op->gd, visdata<double>(op), op->tl, op->tlm, (const cpx<double>*)op->d_vecy,
                             (cpx<double>*)op->d_gridA, 1.0, s, variant, op->grid_waves);
      };
      run();  // warm-up
      PF_HIP(hipEventRecord(a, s));
      for (int k = 0; k < n; ++k) run();
      PF_HIP(hipEventRecord(b, s));
      PF_HIP(hipEventSynchronize(b));
      float ms = 0;
      PF_HIP(hipEventElapsedTime(&ms, a, b));
      if (ms < best) {
        best = ms;
        best_v = variant;
      }
    }
    op->grid_variant = best_v;
    // kept-entry rows: the gridding's entries per step, from one up to what the largest row and the LDS
    // allow (every setting is fixed-order); the settings alternate over three trials, each keeping its best
    if (op->w_term && op->w_packed && op->w_qe_max > 1) {
      float best_q[3] = {1e30f, 1e30f, 1e30f};  // entries 1, 2, 4
      for (int trial = 0; trial < 3; ++trial) {
        for (int qi = 0, q = 1; q <= op->w_qe_max; ++qi, q *= 2) {
          if ((size_t)q * op->tl.TU * op->tl.TV * op->esz() > 160 * 1024) break;
          op->w_qe = q;
          auto run = [&] {
            if (op->prec == PF_C64)
              grid_wrows<float>(op->Nu, op->Nv, wrows<float>(op), op->tl, (const cpx<float>*)op->d_vecy,
                                (cpx<float>*)op->d_gridA, 1.0f, s);
            else
              grid_wrows<double>(op->Nu, op->Nv, wrows<double>(op), op->tl, (const cpx<double>*)op->d_vecy,
                                 (cpx<double>*)op->d_gridA, 1.0, s);
          };
          for (int k = 0; k < n; ++k) run();  // warm-up
          PF_HIP(hipEventRecord(a, s));
          for (int k = 0; k < n; ++k) run();
          PF_HIP(hipEventRecord(b, s));
          PF_HIP(hipEventSynchronize(b));
          float ms = 0;
          PF_HIP(hipEventElapsedTime(&ms, a, b));
          best_q[qi] = std::min(best_q[qi], ms);
        }
      }
      int bi = 0;
      for (int qi = 1; qi < 3; ++qi)
        if (best_q[qi] < best_q[bi]) bi = qi;
      op->w_qe = 1 << bi;
    }
    // column FFT: one pass or two steps, timed over whole Phi + Phi^dagger pairs (a timing of the column
    // passes alone is at the mercy of the host's launch rate).  Both plans are built first and swapped in
    // (a rebuild between trials idles the GPU, and the pairs after it run slow); each trial warms up with
    // as many pairs as it times, the modes alternate over three trials, and each keeps its best.
    Fft2dPlan plans[2];
    double best_t[2] = {1e30, 1e30};
    const Fft2dPlan cur = op->fft;
    try {
      for (int mode : {0, 1}) {
        fft_plan_init(plans[mode], op->Nv, op->Nu, op->prec, mode);
        if (op->fft_row_bpc >= 0) plans[mode].row_bpc = op->fft_row_bpc;
      }
      for (int trial = 0; trial < 3; ++trial)
        for (int mode : {0, 1}) {
          op->fft = plans[mode];
          best_t[mode] = std::min(best_t[mode], time_pairs(op, n, n));
        }
    } catch (...) {
      op->fft = cur;
      for (Fft2dPlan& pl : plans)
        if (pl.tw) fft_plan_free(pl);
      throw;
    }
    const int best_m = best_t[0] <= best_t[1] ? 0 : 1;
    Fft2dPlan old = cur;
    fft_plan_free(old);
    op->fft = plans[best_m];
    fft_plan_free(plans[1 - best_m]);
    op->fft_col_mode = best_m;
    // the row passes' grid (every line computed alike, so results do not depend on it): a few resident
    // blocks per CU looping over lines with the next line prefetched, or one block per group of lines;
    // timed over whole pairs like the column mode (at the 4096^2 grid, complex64, one block per group
    // of lines beat three resident blocks per CU by 4 us per pass once the passes covered all 4096 rows)
    {
      const int cand[4] = {0, 2, 3, 4};
      double bt[4] = {1e30, 1e30, 1e30, 1e30};
      for (int trial = 0; trial < 3; ++trial)
        for (int i = 0; i < 4; ++i) {
          op->fft.row_bpc = cand[i];
          bt[i] = std::min(bt[i], time_pairs(op, n, n));
        }
      int bi = 0;
      for (int i = 1; i < 4; ++i)
        if (bt[i] < bt[bi]) bi = i;
      op->fft.row_bpc = op->fft_row_bpc = cand[bi];
    }
    ++op->generation;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (chosen) *chosen = best_v;
  });
}

int pf_op_profile_enable(pf_op* op, int on) {
  if (!op) return fail(PF_ERR_ARG, "null operator");
  return guarded([&] {
    if (on) {
      if (!op->prof) op->prof.reset(new Profiler());
      op->prof->clear();
    } else {
      op->prof.reset();
    }
  });
}

int pf_op_profile_get(pf_op* op, int32_t idx, char* name, int32_t name_len, double* total_ms, int64_t* count) {
  if (!op || !op->prof) return fail(PF_ERR_STATE, "profiling is not enabled on this operator");
  return guarded([&] {
    op->prof->flush();
    if (idx < 0 || idx >= (int32_t)op->prof->names.size()) throw Error(PF_ERR_ARG, "profile index out of range");
    if (name && name_len > 0) {
      std::strncpy(name, op->prof->names[idx].c_str(), name_len - 1);
      name[name_len - 1] = 0;
    }
    if (total_ms) *total_ms = op->prof->total_ms[idx];
    if (count) *count = op->prof->count[idx];
  });
}

int pf_op_power_method(pf_op* op, int32_t iters, double tol, const double* start, double* norm_out) {
  if (!op || !norm_out) return fail(PF_ERR_ARG, "null argument");
  return guarded([&] {
    if (iters <= 0) {
      *norm_out = 1.0;
      return;
    }
    *norm_out = op->prec == PF_C64 ? power_method_t<float>(op, iters, tol, start)
                                   : power_method_t<double>(op, iters, tol, start);
  });
}

int pf_op_set_norm(pf_op* op, double norm) {
  if (!op) return fail(PF_ERR_ARG, "null operator");
  if (!(norm > 0)) return fail(PF_ERR_ARG, "Operator norm is not greater than zero.");
  op->norm = norm;
  ++op->generation;
  return PF_OK;
}

int pf_op_get_norm(const pf_op* op, double* norm) {
  if (!op || !norm) return fail(PF_ERR_ARG, "null argument");
  *norm = op->norm;
  return PF_OK;
}

int pf_op_get_info(const pf_op* op, pf_op_info* info) {
  if (!op || !info) return fail(PF_ERR_ARG, "null argument");
  std::memset(info, 0, sizeof(*info));
  info->imsizey = op->ny;
  info->imsizex = op->nx;
  info->Nv = op->Nv;
  info->Nu = op->Nu;
  info->Ju = op->Ju;
  info->Jv = op->Jv;
  info->precision = op->prec;
  info->w_term = op->w_term;
  info->M = op->M;
  info->nnz_w = op->nnz_w;
  info->tile_u = op->tl.TU;
  info->tile_v = op->tl.TV;
  info->n_tiles = op->n_tiles;
  info->tile_list_len = op->tile_list_len;
  info->norm = op->norm;
  info->device_bytes = op->device_bytes;
  return PF_OK;
}

int pf_op_apply_FZ(pf_op* op, const void* x_dev, void* grid_dev, pf_stream stream) {
  return guarded([&] {
    ProfInstall pi(op);
    const double scl = 1.0 / std::sqrt((double)op->ngrid());
    if (op->prec == PF_C64) fz_forward<float>(op, (const cpx<float>*)x_dev, (cpx<float>*)grid_dev, scl, (hipStream_t)stream);
    else fz_forward<double>(op, (const cpx<double>*)x_dev, (cpx<double>*)grid_dev, scl, (hipStream_t)stream);
  });
}

int pf_op_apply_FZ_adjoint(pf_op* op, const void* grid_dev, void* x_dev, pf_stream stream) {
  return guarded([&] {
    ProfInstall pi(op);
    const double scl = 1.0 / std::sqrt((double)op->ngrid());
    // the column pass works in place on its input, so copy the caller's grid first
    const size_t bytes = op->esz() * op->ngrid();
    PF_HIP(hipMemcpyAsync(op->d_gridA, grid_dev, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    if (op->prec == PF_C64) fz_adjoint<float>(op, (cpx<float>*)op->d_gridA, (cpx<float>*)x_dev, scl, (hipStream_t)stream);
    else fz_adjoint<double>(op, (cpx<double>*)op->d_gridA, (cpx<double>*)x_dev, scl, (hipStream_t)stream);
  });
}

int pf_op_apply_G(pf_op* op, const void* grid_dev, void* y_dev, pf_stream stream) {
  return guarded([&] {
    ProfInstall pi(op);
    if (op->prec == PF_C64) G_dev<float>(op, (const cpx<float>*)grid_dev, (cpx<float>*)y_dev, (hipStream_t)stream);
    else G_dev<double>(op, (const cpx<double>*)grid_dev, (cpx<double>*)y_dev, (hipStream_t)stream);
  });
}

int pf_op_apply_GH(pf_op* op, const void* y_dev, void* grid_dev, pf_stream stream) {
  return guarded([&] {
    ProfInstall pi(op);
    if (op->prec == PF_C64) GH_dev<float>(op, (const cpx<float>*)y_dev, (cpx<float>*)grid_dev, (hipStream_t)stream);
    else GH_dev<double>(op, (const cpx<double>*)y_dev, (cpx<double>*)grid_dev, (hipStream_t)stream);
  });
}

int pf_fft2d(int32_t Nv, int32_t Nu, int32_t precision, int32_t sign, const void* in_dev, void* out_dev,
             void* workspace_dev, pf_stream stream) {
  return guarded([&] {
    if (sign != -1 && sign != 1) throw Error(PF_ERR_ARG, "sign must be -1 or +1");
    if (!is_pow2(Nv) || !is_pow2(Nu)) throw Error(PF_ERR_UNSUPPORTED, "FFT axes must be powers of two >= 2");
    Fft2dPlan p;
    fft_plan_init(p, Nv, Nu, precision);
    hipStream_t s = (hipStream_t)stream;
    const size_t esz = precision ? 16 : 8;
    void* ws = workspace_dev;
    bool own = false;
    if (!ws) {
      ws = dev_alloc(esz * (size_t)Nv * Nu);
      own = true;
    }
    const bool inv = sign > 0;
    const double scl = 1.0 / std::sqrt((double)Nv * Nu);
    auto run = [&](auto tag) {
      using R = decltype(tag);
      SZ<R> sz;
      fft_rows<R>(p, ROW_PLAIN, inv, (const cpx<R>*)in_dev, (cpx<R>*)ws, Nv, sz, s);
      fft_cols<R>(p, inv, (cpx<R>*)ws, (cpx<R>*)out_dev, RowWindow{0, 0}, false, false, s);
      scale_vec<R>((cpx<R>*)out_dev, (int64_t)Nv * Nu, scl, s);
    };
    if (precision) run(double());
    else run(float());
    PF_HIP(hipStreamSynchronize(s));
    if (own) dev_free(ws);
    fft_plan_free(p);
  });
}

int pf_kernel_eval(int32_t kind, int32_t which, int32_t J, double os, int32_t ftsize, int32_t precision,
                   const double* x_host, double* out_host, int64_t n) {
  return guarded([&] {
    if (kind < 0 || kind > 5) throw Error(PF_ERR_ARG, "Did not choose valid kernel.");
    if (which == 1) {
      // correction functions are evaluated on the host at construction (operators.cc:101-121)
      for (int64_t i = 0; i < n; ++i) out_host[i] = host_ft_kernel(kind, x_host[i] / ftsize - 0.5, J, os);
      return;
    }
    DevTemp tmp;
    double* dx = tmp.alloc<double>(n);
    double* dy = tmp.alloc<double>(n);
    hipStream_t s = cstream();
    h2d(dx, x_host, sizeof(double) * n);
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (blocks) {
      if (precision == PF_C128)
        hipLaunchKernelGGL((k_kernel_eval<double>), dim3(blocks), dim3(256), 0, s, make_kparams<double>(kind, J, os), dx, dy, n);
      else
        hipLaunchKernelGGL((k_kernel_eval<float>), dim3(blocks), dim3(256), 0, s, make_kparams<float>(kind, J, os), dx, dy, n);
      PF_HIP(hipGetLastError());
    }
    d2h(out_host, dy, sizeof(double) * n);
  });
}

// ---- synthetic inputs ------------------------------------------------------------------------
static double uni53(std::mt19937_64& g) { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); }
static double normal(std::mt19937_64& g) {
  // Box-Muller with u1 in (0, 1]
  const double u1 = 1.0 - uni53(g), u2 = uni53(g);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * PI * u2);
}

int pf_synth_uv(uint64_t seed, int64_t M, double sigma, double* u, double* v) {
  if ((M > 0 && (!u || !v)) || M < 0) return fail(PF_ERR_ARG, "bad arguments");
  std::mt19937_64 g(seed);
  for (int64_t m = 0; m < M; ++m) {
    u[m] = std::min(PI, std::max(-PI, sigma * normal(g)));
    v[m] = std::min(PI, std::max(-PI, sigma * normal(g)));
  }
  return PF_OK;
}

int pf_synth_image(uint64_t seed, int32_t ny, int32_t nx, int32_t nsrc, double* img) {
  if (ny <= 0 || nx <= 0 || nsrc < 0 || !img) return fail(PF_ERR_ARG, "bad arguments");
  std::mt19937_64 g(seed);
  std::vector<double> cx(nsrc), cy(nsrc), amp(nsrc), sg(nsrc);
  for (int k = 0; k < nsrc; ++k) {
    cx[k] = uni53(g) * nx;
    cy[k] = uni53(g) * ny;
    amp[k] = 0.1 + 0.9 * uni53(g);
    sg[k] = 1.0 + 7.0 * uni53(g);
  }
  // each source is separable: amp exp(-dx^2 / 2 sg^2) exp(-dy^2 / 2 sg^2), factors tabulated per axis
  std::vector<double> ex((size_t)nsrc * nx), ey((size_t)nsrc * ny);
  for (int k = 0; k < nsrc; ++k) {
    const double c = 1.0 / (2 * sg[k] * sg[k]);
    for (int i = 0; i < nx; ++i) ex[(size_t)k * nx + i] = std::exp(-(i - cx[k]) * (i - cx[k]) * c);
    for (int j = 0; j < ny; ++j) ey[(size_t)k * ny + j] = amp[k] * std::exp(-(j - cy[k]) * (j - cy[k]) * c);
  }
  parallel_for((int64_t)ny * nx, [&](int64_t p0, int64_t p1) {
    for (int64_t p = p0; p < p1; ++p) {
      const int64_t j = p / nx, i = p - j * nx;
      double s = 0;
      for (int k = 0; k < nsrc; ++k) s += ey[(size_t)k * ny + j] * ex[(size_t)k * nx + i];
      img[2 * p] = s;
      img[2 * p + 1] = 0.0;
    }
  });
  return PF_OK;
}

int pf_synth_cnormal(uint64_t seed, int64_t n, double scl, double* out) {
  if (n < 0 || (n > 0 && !out)) return fail(PF_ERR_ARG, "bad arguments");
  std::mt19937_64 g(seed);
  const double s = scl / std::sqrt(2.0);
  for (int64_t k = 0; k < n; ++k) {
    out[2 * k] = s * normal(g);
    out[2 * k + 1] = s * normal(g);
  }
  return PF_OK;
}

const char* pf_last_error(void) { return last_error().c_str(); }
const char* pf_version(void) { return "purify-amd 0.1 (gfx950)"; }

}  // extern "C"
