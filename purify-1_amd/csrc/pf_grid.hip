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
//  * degrid: one block per tile stages the tile and its halo in LDS; one lane per sorted visibility
//    gathers its J x J taps there and writes y in the sorted order (coalesced; complex64) or straight
//    to the caller's order (complex128).
//  * permute_to_sorted / permute_to_caller: the permutation between the caller's order and the sorted
//    order: two LDS-staged passes that move whole cache lines (complex64), or one scatter (complex128).
//  * grid: one wave per grid tile (tile-owner).  The tile's list holds every visibility with a tap
//    inside the tile, in ascending sorted order.  The wave accumulates into its own LDS copy of the
//    tile (no other wave touches it), then writes the tile once.  Every grid cell is written by
//    exactly one wave, so the grid needs no zero fill and there are no HBM atomics; the sum order
//    inside a tile is fixed by the list order and the lane order of the LDS adds, which makes the
//    result run-to-run reproducible.
#include <algorithm>

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

// a mod n for the power-of-two grid sizes the operator accepts (two's complement makes it exact for
// negative a as well)
__device__ inline int wrap(int a, int n) { return a & (n - 1); }

constexpr int JMAX = 16;

// 1-D tap weights with the (-1)^(k+j) phase folded in: w[a] = (-1)^(k + a) K(f - 1 - a).
template <typename R, int J>
__device__ inline void taps(const KParams<R>& p, int k, R f, int Jrt, R (&w)[J == 0 ? JMAX : J]) {
  constexpr int JJ = J == 0 ? JMAX : J;
#pragma unroll
  for (int a = 0; a < JJ; ++a) {
    if (J == 0 && a >= Jrt) break;
    R val = kern1d<R, J == 0 ? 0 : 1>(p, f - R(1 + a));
    w[a] = ((k + a) & 1) ? -val : val;
  }
}

// Row of a tap table for fraction f: four 16-byte loads, interpolated by tab_interp.
struct TabRow {
  float4 t0, t1, d0, d1;
  float r;
};
template <int J>
__device__ inline TabRow tab_row(const float* __restrict__ tab, float f) {
  const float qf = (f - 0.5f * (float)J) * (float)TAB_Q;  // delta in (0, 1] times TAB_Q
  int q = (int)qf;
  q = q < 0 ? 0 : (q > TAB_Q - 1 ? TAB_Q - 1 : q);
  const float4* row = reinterpret_cast<const float4*>(tab + (size_t)q * TAB_ROW);
  return TabRow{row[0], row[1], row[2], row[3], qf - (float)q};
}
template <int J>
__device__ inline void tab_interp(const TabRow& a, int k, float (&w)[J]) {
  const float t[8] = {a.t0.x, a.t0.y, a.t0.z, a.t0.w, a.t1.x, a.t1.y, a.t1.z, a.t1.w};
  const float dd[8] = {a.d0.x, a.d0.y, a.d0.z, a.d0.w, a.d1.x, a.d1.y, a.d1.z, a.d1.w};
#pragma unroll
  for (int e = 0; e < J; ++e) {
    const float val = t[e] + a.r * dd[e];
    w[e] = ((k + e) & 1) ? -val : val;
  }
}

// How a kernel instance computes its taps: taps_both (tables if present, else direct evaluation),
// tables requested ahead by the kernel's own pipeline, or the per-tap polynomials.
enum TapMode { TAPS_EVAL = 0, TAPS_TABLE = 1, TAPS_POLY = 2 };

// Taps from the per-tap polynomials (make_tap_poly) in s = 2 (f - J/2) - 1; the coefficients are
// uniform across the wave (scalar loads).  The kernels are even, so tap J-1-a at s is tap a at -s
// (K(delta + J/2 - 1 - (J-1-a)) = K((1 - delta) + J/2 - 1 - a)): with p_a(s) = E_a(s^2) + s O_a(s^2), the
// pair is E_a +- s O_a, two Horner chains in s^2 for two taps (11 operations instead of 18 at D = 9).
// The taps carry (-1)^a only; the rest of the phase, (-1)^(k_u + k_v), is one sign flip of the
// visibility's value (flip_phase) instead of one per tap.
// The coefficients are read through the constant address space: uniform scalar loads of memory the
// kernel never writes, which the compiler may repeat instead of holding all J (D + 1) values live.
typedef const __attribute__((address_space(4))) float const_float;
template <int J>
__device__ inline void poly_taps(const float* __restrict__ c_, float f, float (&w)[J]) {
  static_assert(TAP_POLY_D % 2 == 1, "the even / odd split assumes an odd degree");
  const const_float* c = (const const_float*)c_;
  const float sv = 2.0f * (f - 0.5f * (float)J) - 1.0f;
  const float t = sv * sv;
#pragma unroll
  for (int a = 0; a < (J + 1) / 2; ++a) {
    // highest degree first: ca[0] s^D (odd), ca[1] s^(D-1) (even), ...
    const const_float* ca = c + a * (TAP_POLY_D + 1);
    float o = ca[0], e = ca[1];
#pragma unroll
    for (int d = 2; d <= TAP_POLY_D; d += 2) {
      o = fmaf(o, t, ca[d]);
      e = fmaf(e, t, ca[d + 1]);
    }
    const float so = sv * o;
    const float pa = e + so;
    w[a] = (a & 1) ? -pa : pa;  // compile-time sign
    if (J - 1 - a != a) {
      const float pm = e - so;
      w[J - 1 - a] = ((J - 1 - a) & 1) ? -pm : pm;
    }
  }
}
// z (-1)^(k_u + k_v): the sign bits flipped when the parity is odd
__device__ inline cpx<float> flip_phase(cpx<float> z, int ku, int kv) {
  const unsigned m = (unsigned)((ku + kv) & 1) << 31;
  return mk<float>(__uint_as_float(__float_as_uint(z.x) ^ m), __uint_as_float(__float_as_uint(z.y) ^ m));
}

// Tap weights of both axes: tables where present (both rows are requested before either is used, so
// a visibility waits for one memory round trip, not two), else direct evaluation.  The kernels that
// use the polynomials call poly_taps instead (a compile-time tap mode, TapMode).
template <typename R, int JU, int JV>
__device__ inline void taps_both(const GridGeom<R>& g, const KParams<R>& kv, const float* tab_v, int ku_, R fu,
                                 int kv_, R fv, int Ju, int Jv, R (&wu)[JU == 0 ? JMAX : JU],
                                 R (&wv)[JV == 0 ? JMAX : JV]) {
  if constexpr (sizeof(R) == 4 && JU > 0 && JU <= 8 && JV > 0 && JV <= 8) {
    if (g.tab_u && tab_v) {
      const TabRow a = tab_row<JU>(g.tab_u, fu);
      const TabRow b = tab_row<JV>(tab_v, fv);
      tab_interp<JU>(a, ku_, wu);
      tab_interp<JV>(b, kv_, wv);
      return;
    }
  }
  taps<R, JU>(g.ku, ku_, fu, Ju, wu);
  taps<R, JV>(kv, kv_, fv, Jv, wv);
}

// With equal compile-time supports the two axis kernels are identical (the reference's kernel-slot swap
// exchanges equal kernels), so both axes read the u-axis parameters and one coefficient set stays in
// registers.
template <int JU, int JV, typename R>
__device__ inline const KParams<R>& kparams_v(const GridGeom<R>& g) {
  if constexpr (JU == JV && JU > 0) return g.ku;
  else return g.kv;
}
template <int JU, int JV, typename R>
__device__ inline const float* table_v(const GridGeom<R>& g) {
  if constexpr (JU == JV && JU > 0) return g.tab_u;
  else return g.tab_v;
}

// Host: the kernels instantiated for a compile-time support sum the Kaiser-Bessel expansion with its
// fixed length, which must cover the host's coefficient count.
template <typename R>
static bool fixed_len_ok(const GridGeom<R>& g) {
  auto ok = [](const KParams<R>& p) { return (p.kind != K_KB && p.kind != K_KBMIN) || p.ncoef <= KBLen<R>::N; };
  return ok(g.ku) && ok(g.kv);
}

// Tile kernels keep the hardware's round-robin spread of workgroups over the 8 XCDs: the visibility
// density is concentrated in the centre of the uv-plane, and giving each XCD a contiguous band of
// tiles (to share halos in one L2) would put all heavy tiles on two XCDs.

// ---------------------------------------------------------------------------------------------
// Degridding  y = G grid: one block per tile stages the tile plus its (J-1)-wide halo in LDS, then
// every visibility whose first tap lies in the tile reads its J x J taps from LDS.
// ---------------------------------------------------------------------------------------------
constexpr int DG_THREADS = 256;
// halo rows one wave stages per round: 4 waves cover 40 rows (32 + J - 1, J <= 9) in one round for
// complex64, in two for complex128 (half the staging registers)
template <typename R>
constexpr int dg_rows() { return sizeof(R) == 4 ? 10 : 5; }

// TAB (complex64, J <= 8, tap tables present): the taps come from the tables, whose rows (addressed by
// the coordinates) are requested one visibility ahead and the coordinates two ahead, so neither round
// trip is waited for in the loop; otherwise the kernel is evaluated directly.
template <typename R>
struct DegridArgs {
  GridGeom<R> g;
  VisData<R> v;
  TileLists tl;
  const cpx<R>* grid;
  cpx<R>* y;
  R scl;
};
// BT: a batch launch, whose blockIdx.z selects the arguments from tab; its grid covers the largest slot
// count of the batch, so blocks past a grid's own slots return
template <typename R, int JU, int JV, int TM, bool BT = false>
__global__ void __launch_bounds__(DG_THREADS) k_degrid_tiles(DegridArgs<R> a0, const DegridArgs<R>* __restrict__ tab) {
  const DegridArgs<R> q = BT ? tab[blockIdx.z] : a0;
  const GridGeom<R>& g = q.g;
  const VisData<R>& v = q.v;
  const TileLists& tl = q.tl;
  const cpx<R>* __restrict__ grid = q.grid;
  cpx<R>* __restrict__ y = q.y;
  const R scl = q.scl;
  if (BT && blockIdx.x >= (unsigned)tl.n_own) return;
  constexpr bool TAB = TM == TAPS_TABLE && sizeof(R) == 4 && JU > 0 && JU <= 8 && JV > 0 && JV <= 8;
  constexpr bool POLY = TM == TAPS_POLY && sizeof(R) == 4 && JU > 0 && JU == JV && JU <= 8;
  extern __shared__ unsigned char smem_raw[];
  cpx<R>* lds = reinterpret_cast<cpx<R>*>(smem_raw);
  const uint4 sl = tl.slot_own[blockIdx.x];
  const int tile = (int)sl.x;
  const uint32_t beg = sl.y, end = sl.z;
  const int Ju = JU ? JU : g.Ju, Jv = JV ? JV : g.Jv;
  constexpr int JJV = JV == 0 ? JMAX : JV, JJU = JU == 0 ? JMAX : JU;
  const int tv = tile / tl.ntu, tu = tile - tv * tl.ntu;
  const int tx0 = tu * tl.TU, ty0 = tv * tl.TV;
  const int W = tl.TU + Ju - 1, H = tl.TV + Jv - 1;
  // this thread's first visibility: its loads go out before the tile staging so the latencies overlap
  uint32_t i = beg + threadIdx.x;
  VisCoord<R> c{}, c2{};  // this visibility's and (TAB) the one after next's coordinates
  TabRow ta{}, tb{};      // TAB: table rows of the visibility being computed
  uint32_t dst = i;
  if (i < end) {
    c = vis_coord(v, i, Ju, Jv);
    if (v.perm) dst = (uint32_t)v.perm[i];
    if constexpr (TAB) {
      ta = tab_row<JU>(g.tab_u, c.fu);
      tb = tab_row<JV>(table_v<JU, JV>(g), c.fv);
    }
  }
  if constexpr (TAB)
    if (i + DG_THREADS < end) c2 = vis_coord(v, i + DG_THREADS, Ju, Jv);
  // stage the tile and its halo: each wave takes DG_ROWS rows per round, every load before any store
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = DG_THREADS / 64;
  constexpr int DG_ROWS = dg_rows<R>();
  for (int r0 = wave * DG_ROWS; r0 < H; r0 += nw * DG_ROWS) {
    for (int lx = lane; lx < W; lx += 64) {
      int gx = tx0 + lx;
      if (gx >= g.Nu) gx -= g.Nu;
      // rows past the halo re-read its last row (always in bounds, never stored): unconditional loads
      // keep tmp in registers
      cpx<R> tmp[DG_ROWS];
#pragma unroll
      for (int k = 0; k < DG_ROWS; ++k) {
        int gy = ty0 + min(r0 + k, H - 1);
        if (gy >= g.Nv) gy -= g.Nv;
        tmp[k] = grid[(int64_t)gy * g.Nu + gx];
      }
#pragma unroll
      for (int k = 0; k < DG_ROWS; ++k)
        if (r0 + k < H) lds[(r0 + k) * W + lx] = tmp[k];
    }
  }
  __syncthreads();
  for (; i < end; i += DG_THREADS) {
    // the next visibility's loads go out before this one's arithmetic
    const uint32_t nxt = i + DG_THREADS;
    VisCoord<R> cn = c;
    uint32_t dn = nxt;
    R wu[JJU], wv[JJV];
    if constexpr (TAB) {
      tab_interp<JU>(ta, c.ku, wu);
      tab_interp<JV>(tb, c.kv, wv);
      if (nxt < end) {
        cn = c2;
        ta = tab_row<JU>(g.tab_u, cn.fu);
        tb = tab_row<JV>(table_v<JU, JV>(g), cn.fv);
        if (v.perm) dn = (uint32_t)v.perm[nxt];
      }
      if (nxt + DG_THREADS < end) c2 = vis_coord(v, nxt + DG_THREADS, Ju, Jv);
    } else {
      if (nxt < end) {
        cn = vis_coord(v, nxt, Ju, Jv);
        if (v.perm) dn = (uint32_t)v.perm[nxt];
      }
      if constexpr (POLY) {
        poly_taps<JU>(g.poly, c.fu, wu);
        poly_taps<JV>(g.poly, c.fv, wv);
      } else {
        taps_both<R, JU, JV>(g, kparams_v<JU, JV>(g), table_v<JU, JV>(g), c.ku, c.fu, c.kv, c.fv, Ju, Jv, wu, wv);
      }
    }
    const int lx0 = wrap(c.ku + 1, g.Nu) - tx0, ly0 = wrap(c.kv + 1, g.Nv) - ty0;
    const cpx<R>* base = lds + ly0 * W + lx0;
    cpx<R> acc = mk<R>(R(0), R(0));
#pragma unroll
    for (int b = 0; b < JJV; ++b) {
      if (JV == 0 && b >= Jv) break;
      cpx<R> racc = mk<R>(R(0), R(0));
#pragma unroll
      for (int a = 0; a < JJU; ++a) {
        if (JU == 0 && a >= Ju) break;
        const cpx<R> x = base[b * W + a];
        racc.x += wu[a] * x.x;
        racc.y += wu[a] * x.y;
      }
      acc.x += wv[b] * racc.x;
      acc.y += wv[b] * racc.y;
    }
    if constexpr (POLY) acc = flip_phase(acc, c.ku, c.kv);
    if (v.weights) acc = cmul(acc, v.weights[i]);
    y[dst] = scale(acc, scl);
    c = cn;
    dst = dn;
  }
}

// ---------------------------------------------------------------------------------------------
// Gridding  grid = G^H y, one 64-lane wave per TU x TV tile, no atomics.
// Lane (a, b) = (lane & 7, lane >> 3) owns every tile cell (x, y) with x = a, y = b (mod 8).  For a
// footprint [x0, x0+Ju) x [y0, y0+Jv) with Ju, Jv <= 8 each tap therefore belongs to a different
// lane, and a cell is only ever touched by its owning lane.  Each lane sums into a register while
// consecutive visibilities hit its current cell and folds the sum into the LDS tile (a plain
// read-modify-write) when its cell changes; visibilities are sorted by first-tap cell inside a tile,
// so that is rare.  The order of every cell's sum is the list order: reproducible bit for bit.
// Requires Nu >= TU + 8 and Nv >= TV + 8 (a footprint meets a tile from one side only).
// ---------------------------------------------------------------------------------------------
// One (weight, coordinate) entry of a visibility's residue-rotated tap table.  Inactive taps carry
// weight 0 and coordinate RC_OFF, so the sum of the u and v coordinates is negative.
template <typename R>
struct alignas(2 * sizeof(R)) RCTap {
  R w;
  int c;
};
constexpr int RC_OFF = -(1 << 20);

template <typename R, int JU, int JV>
__global__ void __launch_bounds__(64) k_grid_rc(GridGeom<R> g, VisData<R> v, TileLists tl,
                                                const cpx<R>* __restrict__ yin, cpx<R>* __restrict__ grid, R scl) {
  extern __shared__ unsigned char smem_raw[];
  const int TU = tl.TU, TV = tl.TV, P = TU + 1;  // padded row stride of the LDS tile
  cpx<R>* acc = reinterpret_cast<cpx<R>*>(smem_raw);
  cpx<R>* sv = acc + TV * P;                                    // 64 scaled visibility values
  RCTap<R>* tu_tab = reinterpret_cast<RCTap<R>*>(sv + 64);     // [64][8]: entry a = tap of residue a
  RCTap<R>* tv_tab = tu_tab + 64 * 8;                           // [64][8]: coordinates pre-multiplied by P
  const uint4 sl = tl.slot[blockIdx.x];
  const int tile = (int)sl.x;
  const int tv = tile / tl.ntu, tu = tile - tv * tl.ntu;
  const int tx0 = tu * TU, ty0 = tv * TV;
  const int lane = threadIdx.x, la = lane & 7, lb = lane >> 3;
  for (int e = lane; e < TV * P; e += 64) acc[e] = mk<R>(R(0), R(0));
  int cur = -1;
  cpx<R> accv = mk<R>(R(0), R(0));
  const uint32_t end = sl.z;
  for (uint32_t base = sl.y; base < end; base += 64) {
    const int nb = (int)min((uint32_t)64, end - base);
    if (lane < nb) {
      // visibility `lane` of the batch: value, tap weights and their residue-rotated placement
      const uint32_t i = tl.list[base + lane];
      const VisCoord<R> c = vis_coord(v, i, JU, JV);
      cpx<R> s = yin[i];  // visibilities in sorted order
      if (v.weights) s = cmulc(s, v.weights[i]);  // conj(G) = phase * K_u * K_v * conj(weight)
      sv[lane] = scale(s, scl);
      R wu[JU], wv[JV];
      taps_both<R, JU, JV>(g, kparams_v<JU, JV>(g), table_v<JU, JV>(g), c.ku, c.fu, c.kv, c.fv, JU, JV, wu, wv);
      int dx = wrap(c.ku + 1 - tx0, g.Nu), dy = wrap(c.kv + 1 - ty0, g.Nv);
      if (dx >= g.Nu - 8) dx -= g.Nu;
      if (dy >= g.Nv - 8) dy -= g.Nv;
#pragma unroll
      for (int d = 0; d < 8; ++d) {
        const int x = dx + d;
        const bool ok = d < JU && (unsigned)x < (unsigned)TU;
        RCTap<R> e;
        e.w = ok ? wu[d < JU ? d : 0] : R(0);
        e.c = ok ? x : RC_OFF;
        tu_tab[lane * 8 + (x & 7)] = e;
      }
#pragma unroll
      for (int d = 0; d < 8; ++d) {
        const int y = dy + d;
        const bool ok = d < JV && (unsigned)y < (unsigned)TV;
        RCTap<R> e;
        e.w = ok ? wv[d < JV ? d : 0] : R(0);
        e.c = ok ? y * P : RC_OFF;
        tv_tab[lane * 8 + (y & 7)] = e;
      }
    }
    __syncthreads();
    // every lane reads its own residue's entries: three independent LDS reads per visibility
    RCTap<R> eu = tu_tab[la], ev = tv_tab[lb];
    cpx<R> s = sv[0];
    for (int t = 0; t < nb; ++t) {
      RCTap<R> eu_n = eu, ev_n = ev;
      cpx<R> s_n = s;
      if (t + 1 < nb) {
        eu_n = tu_tab[(t + 1) * 8 + la];
        ev_n = tv_tab[(t + 1) * 8 + lb];
        s_n = sv[t + 1];
      }
      const int cell = eu.c + ev.c;
      if (cell >= 0) {
        if (cell != cur) {
          if (cur >= 0) {
            const cpx<R> a0 = acc[cur];
            acc[cur] = mk<R>(a0.x + accv.x, a0.y + accv.y);
          }
          cur = cell;
          accv = mk<R>(R(0), R(0));
        }
        const R w = eu.w * ev.w;
        accv.x += s.x * w;
        accv.y += s.y * w;
      }
      eu = eu_n;
      ev = ev_n;
      s = s_n;
    }
    __syncthreads();
  }
  if (cur >= 0) {
    const cpx<R> a0 = acc[cur];
    acc[cur] = mk<R>(a0.x + accv.x, a0.y + accv.y);
  }
  __syncthreads();
  for (int e = lane; e < TU * TV; e += 64) {
    const int ly = e / TU, lx = e - ly * TU;
    grid[(int64_t)(ty0 + ly) * g.Nu + tx0 + lx] = acc[ly * P + lx];
  }
}

// ---------------------------------------------------------------------------------------------
// Gridding fallback for tiny grids or supports above 8: the tile wave adds with LDS atomics.  Only
// this wave touches its tile, so inter-wave order never matters.
// ---------------------------------------------------------------------------------------------
template <typename R, int JU, int JV>
__global__ void __launch_bounds__(64) k_grid_tiles(GridGeom<R> g, VisData<R> v, TileLists tl,
                                                   const cpx<R>* __restrict__ yin, cpx<R>* __restrict__ grid, R scl) {
  extern __shared__ unsigned char smem_raw[];
  const int TU = tl.TU, TV = tl.TV, TT = TU * TV;
  R* acc_re = reinterpret_cast<R*>(smem_raw);
  R* acc_im = acc_re + TT;
  const uint4 sl = tl.slot[blockIdx.x];
  const int tile = (int)sl.x;
  const int tv = tile / tl.ntu, tu = tile - tv * tl.ntu;
  const int tx0 = tu * TU, ty0 = tv * TV;
  for (int e = threadIdx.x; e < TT; e += blockDim.x) {
    acc_re[e] = R(0);
    acc_im[e] = R(0);
  }
  __syncthreads();
  const int Ju = JU ? JU : g.Ju, Jv = JV ? JV : g.Jv;
  constexpr int JJV = JV == 0 ? JMAX : JV, JJU = JU == 0 ? JMAX : JU;
  const uint32_t end = sl.z;
  for (uint32_t base = sl.y; base < end; base += blockDim.x) {
    const uint32_t idx = base + threadIdx.x;
    if (idx < end) {
      const uint32_t i = tl.list[idx];
      const VisCoord<R> c = vis_coord(v, i, Ju, Jv);
      cpx<R> s = yin[i];  // visibilities in sorted order
      if (v.weights) s = cmulc(s, v.weights[i]);
      s = scale(s, scl);
      R wu[JJU], wv[JJV];
      taps_both<R, JU, JV>(g, kparams_v<JU, JV>(g), table_v<JU, JV>(g), c.ku, c.fu, c.kv, c.fv, Ju, Jv, wu, wv);
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

// ---------------------------------------------------------------------------------------------
// Gridding on the matrix cores.  Inside a 16 x 16 tile
//     tile[i][j] = sum_t (s_t wv_t[i - dy_t]) * wu_t[j - dx_t]
// over the tile's list: the product A B of a 16 x K complex matrix (one column per visibility: its
// v-taps times its value, zero outside the footprint) with a K x 16 real one (its u-taps).  One wave
// per tile.  In batches of 64 list entries, lane t evaluates entry t's taps and stores a compact record
// in LDS: the footprint offset (dx, dy), the J_v products s wv and the J_u taps wu.  Each MFMA 16x16x4
// step then adds four entries: lane l gathers A[l % 16][l / 16] and B[l / 16][l % 16] from the records
// of entry 4k + l / 16 (zero outside the footprint) and holds four result rows of column l % 16:
// 4 (l / 16) + r for the FP32 instruction, l / 16 + 4 r for the FP64 one (r < 4).  complex64 sums are
// compensated per tile (the accumulators fold into a TwoSum running sum every MFMA_COMP batches).  The summation
// order is fixed by the list and the instruction: reproducible bit for bit.  Needs N >= 16 + J - 1 on
// both axes, so that a footprint meets a tile in one piece.
// ---------------------------------------------------------------------------------------------
template <typename R>
struct Mfma;
template <>
struct Mfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  __device__ static acc_t step(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int lane, int r) { return (lane >> 4) * 4 + r; }  // result row of register r
};
template <>
struct Mfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  __device__ static acc_t step(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

// hi += b by Knuth's TwoSum (no ordering of |hi| and |b| assumed); returns the rounding error of the
// addition exactly, hi + b = (new hi) + error
__device__ inline float two_sum(float& hi, float b) {
  const float s = hi + b;
  const float bb = s - hi;
  const float e = (hi - (s - bb)) + (b - bb);
  hi = s;
  return e;
}

constexpr int MB = 64;           // list entries per batch (a multiple of 16)
// complex64 compensated tile sums: the MFMA accumulators sum up to MFMA_COMP batches (256 entries: at
// most 64 rounded accumulations per cell); then their partial sum is added to the tile's running sum
// hi by TwoSum, and the accumulators restart from that addition's rounding error, which the next
// period's MFMAs carry (Kahan's compensated summation with the accumulator as the compensation).  A
// cell's error is then bounded by one period's roundings, not by the length of its tile's list.
constexpr int MFMA_COMP = 4;
constexpr int MQ = MB / 4;       // entries per MFMA row group of a batch
constexpr int MFMA_NOTAP = 1024; // footprint offset of an unused entry slot: no row or column matches

// One list entry of the gridding pipeline: coordinates, value and weight.
template <typename R>
struct GridEntry {
  VisCoord<R> c{};
  cpx<R> s{}, w{};
};
template <typename R>
__device__ inline void load_entry(GridEntry<R>& e, const VisData<R>& v, const cpx<R>* __restrict__ yin, uint32_t i,
                                  int Ju, int Jv) {
  e.c = vis_coord(v, i, Ju, Jv);
  e.s = yin[i];  // visibilities in sorted order
  e.w = v.weights ? v.weights[i] : mk<R>(R(1), R(0));
}
// complex64: from an entry record (TileLists::ent); the wrapped first-tap cell keeps the tap phase
// (-1)^(k + a), as the grid sides are even.  The record holds the byte offset of the visibility's
// value (8 * sorted index, below 2^32): a 32-bit offset register, so the loop-carried record needs no
// 64-bit address pair (whose register copy would make the wave wait for the record's load).
__device__ inline void load_entry(GridEntry<float>& e, const VisData<float>& v, const cpx<float>* __restrict__ yin,
                                  const int4 r) {
  const uint32_t off = (uint32_t)r.x;
  e.c = VisCoord<float>{r.y & 0xffff, (int)((unsigned)r.y >> 16), __int_as_float(r.z), __int_as_float(r.w)};
  e.s = *reinterpret_cast<const cpx<float>*>(reinterpret_cast<const char*>(yin) + off);
  e.w = v.weights ? *reinterpret_cast<const cpx<float>*>(reinterpret_cast<const char*>(v.weights) + off)
                  : mk<float>(1.0f, 0.0f);
}

// Position of a batch in a wave's stream: the tile's launch slot and the batch's first list entry.
// Wave w takes the launch slots w, w + W, w + 2W, ... (W waves; heaviest tiles first), each as
// ceil(n / MB) batches of its list, or one empty batch for a tile without entries (its cells are
// written as zeros).  tile < 0: past the end of the stream.
struct BatchPos {
  int tile;
  uint32_t base, end;
};
__device__ inline BatchPos slot_pos(const TileLists& tl, int j, int nslots) {
  if (j >= nslots) return BatchPos{-1, 0u, 0u};
  const uint4 sl = tl.slot[j];
  return BatchPos{(int)sl.x, sl.y, sl.z};
}

// The wave's stream is software-pipelined across batches and tiles: the list entries of a batch are
// requested three batches ahead (complex64: one 16-byte record each), their values two ahead, and
// (TAB: complex64, J <= 8, tap tables present) the tap-table rows one ahead, while the previous
// batch's MFMAs run; the rows' addresses depend on the coordinates, so this hides a dependent round
// trip per batch.  A tile's 16 x 16 sums stay in the MFMA accumulators until its last batch.
template <typename R>
struct GridMfmaArgs {
  GridGeom<R> g;
  VisData<R> v;
  TileLists tl;
  int nslots;
  const cpx<R>* yin;
  cpx<R>* grid;
  R scl;
};
// BT: a batch launch, whose blockIdx.z selects the arguments from tab; its grid (the wave count W) covers
// the largest slot count of the batch, and each tile is still summed by one wave in list order
template <typename R, int JU, int JV, int TM, bool BT = false>
__global__ void __launch_bounds__(64) k_grid_mfma(GridMfmaArgs<R> a0, const GridMfmaArgs<R>* __restrict__ tab) {
  const GridMfmaArgs<R> q = BT ? tab[blockIdx.z] : a0;
  const GridGeom<R>& g = q.g;
  const VisData<R>& v = q.v;
  const TileLists& tl = q.tl;
  const int nslots = q.nslots;
  const cpx<R>* __restrict__ yin = q.yin;
  cpx<R>* __restrict__ grid = q.grid;
  const R scl = q.scl;
  constexpr bool TAB = TM == TAPS_TABLE && sizeof(R) == 4 && JU > 0 && JU <= 8 && JV > 0 && JV <= 8;
  constexpr bool POLY = TM == TAPS_POLY && sizeof(R) == 4 && JU > 0 && JU == JV && JU <= 8;
  constexpr bool ENT = sizeof(R) == 4;
  constexpr int T = MFMA_TILE;
  constexpr int JJV = JV == 0 ? JMAX : JV, JJU = JU == 0 ? JMAX : JU;
  // Record windows [tap 0 .. tap J-1, 0 ...]: a row or column index taken as unsigned and capped at J
  // selects a tap or the zero after the taps without a branch (a negative index wraps to a large value).
  // Odd strides keep the 64 lanes' stores conflict-free; one zero slot (not one on each side) keeps the
  // wave's LDS at 5.5 KiB for complex64 and J = 6: 29 resident waves per CU.
  constexpr int SV = (JJV + 1) | 1, SU = (JJU + 1) | 1;
  // footprint origin relative to the tile, packed as (dx + 512) | (dy + 512) << 16, stored so that the
  // offsets lane l needs in steps 4q .. 4q + 3 (entries 4k + l / 16) are one 16-byte word:
  // entry t at rd[(t & 3) * MQ + (t >> 2)]
  __shared__ __attribute__((aligned(16))) int rd[MB];
  __shared__ cpx<R> ra[MB * SV];  // [s * wv[0..Jv-1], 0 ...]
  __shared__ R rb[MB * SU];       // [wu[0..Ju-1], 0 ...]
  using acc_t = typename Mfma<R>::acc_t;
  const int lane = threadIdx.x, li = lane & 15, lt = lane >> 4;
  for (int q = lane; q < MB * SV; q += 64) ra[q] = mk<R>(R(0), R(0));  // slots past the taps stay zero
  for (int q = lane; q < MB * SU; q += 64) rb[q] = R(0);
  const int Ju = JU ? JU : g.Ju, Jv = JV ? JV : g.Jv;
  const int W = (int)gridDim.x;
  // batch positions: p0 being prepared and multiplied, p1 / p2 / p3 the next three; j3 is p3's slot
  int j3 = (int)blockIdx.x;
  auto advance = [&](BatchPos p) -> BatchPos {
    if (p.tile < 0) return p;
    if (p.base + MB < p.end) return BatchPos{p.tile, p.base + MB, p.end};
    j3 += W;
    return slot_pos(tl, j3, nslots);
  };
  const BatchPos p0i = slot_pos(tl, j3, nslots);
  BatchPos p0 = p0i;
  BatchPos p1 = advance(p0);
  BatchPos p2 = advance(p1);
  BatchPos p3 = advance(p2);
  auto has = [&](const BatchPos& p) { return p.tile >= 0 && p.base + lane < p.end; };
  // lane's entry of batch p0 (cur) and of p1 (nxt), and the list entry of p2 (complex64: its record;
  // complex128: its index)
  GridEntry<R> cur, nxt;
  TabRow ta{}, tb{};  // TAB: table rows of cur
  uint32_t inext = 0;
  int4 enext{};
  if constexpr (ENT) {
    if (has(p0)) load_entry(cur, v, yin, tl.ent[p0.base + lane]);
    if (has(p1)) load_entry(nxt, v, yin, tl.ent[p1.base + lane]);
    if (has(p2)) enext = tl.ent[p2.base + lane];
  } else {
    if (has(p0)) load_entry(cur, v, yin, tl.list[p0.base + lane], Ju, Jv);
    if (has(p1)) load_entry(nxt, v, yin, tl.list[p1.base + lane], Ju, Jv);
    if (has(p2)) inext = tl.list[p2.base + lane];
  }
  if constexpr (TAB) {
    if (has(p0)) {
      ta = tab_row<JU>(g.tab_u, cur.c.fu);
      tb = tab_row<JV>(table_v<JU, JV>(g), cur.c.fv);
    }
  }
  acc_t cre = {R(0), R(0), R(0), R(0)}, cim = {R(0), R(0), R(0), R(0)};
  // complex64: the tile's compensated running sums (see MFMA_COMP)
  constexpr bool COMP = sizeof(R) == 4;
  float hre[4] = {0.f, 0.f, 0.f, 0.f}, him[4] = {0.f, 0.f, 0.f, 0.f};
  int nacc = 0;  // batches summed in cre / cim since the last fold
  while (p0.tile >= 0) {
    const int tx0 = (p0.tile % tl.ntu) * T, ty0 = (p0.tile / tl.ntu) * T;
    const int nb = (int)min((uint32_t)MB, p0.end - p0.base);  // 0 for an empty tile
    if (lane < nb) {
      cpx<R> s = cur.s;
      if (v.weights) s = cmulc(s, cur.w);  // conj(G) = phase * K_u * K_v * conj(weight)
      s = scale(s, scl);
      R wu[JJU], wv[JJV];
      if constexpr (TAB) {
        tab_interp<JU>(ta, cur.c.ku, wu);
        tab_interp<JV>(tb, cur.c.kv, wv);
      } else if constexpr (POLY) {
        poly_taps<JU>(g.poly, cur.c.fu, wu);
        poly_taps<JV>(g.poly, cur.c.fv, wv);
        s = flip_phase(s, cur.c.ku, cur.c.kv);
      } else {
        taps_both<R, JU, JV>(g, kparams_v<JU, JV>(g), nullptr, cur.c.ku, cur.c.fu, cur.c.kv, cur.c.fv, Ju, Jv, wu,
                             wv);
      }
      int dx = wrap(cur.c.ku + 1 - tx0, g.Nu), dy = wrap(cur.c.kv + 1 - ty0, g.Nv);
      if (dx >= T) dx -= g.Nu;  // footprint starts left of / above the tile
      if (dy >= T) dy -= g.Nv;
      rd[(lane & 3) * MQ + (lane >> 2)] = (dx + 512) | ((dy + 512) << 16);
#pragma unroll
      for (int e = 0; e < JJV; ++e) {
        if (JV == 0 && e >= Jv) break;
        ra[lane * SV + e] = mk<R>(s.x * wv[e], s.y * wv[e]);
      }
#pragma unroll
      for (int e = 0; e < JJU; ++e) {
        if (JU == 0 && e >= Ju) break;
        rb[lane * SU + e] = wu[e];
      }
    } else {
      rd[(lane & 3) * MQ + (lane >> 2)] = (MFMA_NOTAP + 512) | ((MFMA_NOTAP + 512) << 16);
    }
    __syncthreads();
    // the next batch's table rows, the values two batches ahead and the entries three ahead go out
    // before this batch's MFMAs
    if (has(p1)) {
      if constexpr (TAB) {
        ta = tab_row<JU>(g.tab_u, nxt.c.fu);
        tb = tab_row<JV>(table_v<JU, JV>(g), nxt.c.fv);
      }
      cur = nxt;
    }
    if constexpr (ENT) {
      if (has(p2)) load_entry(nxt, v, yin, enext);
      if (has(p3)) enext = tl.ent[p3.base + lane];
    } else {
      if (has(p2)) load_entry(nxt, v, yin, inext, Ju, Jv);
      if (has(p3)) inext = tl.list[p3.base + lane];
    }
    // groups of four steps (16 entries; slots past nb hold no footprint): every LDS read of a group is
    // issued before its MFMAs
    const int ng = (nb + 15) >> 4;
    for (int q = 0; q < ng; ++q) {
      const int4 dq = *reinterpret_cast<const int4*>(&rd[lt * MQ + 4 * q]);
      const int dd[4] = {dq.x, dq.y, dq.z, dq.w};
      cpx<R> a[4];
      R b[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int t = 16 * q + 4 * r + lt;  // entry of step 4q + r; li is its row of A and column of B
        const unsigned ev = min((unsigned)(li + 512 - (dd[r] >> 16)), (unsigned)Jv);
        const unsigned eu = min((unsigned)(li + 512 - (dd[r] & 0xffff)), (unsigned)Ju);
        a[r] = ra[t * SV + (int)ev];
        b[r] = rb[t * SU + (int)eu];
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        cre = Mfma<R>::step(a[r].x, b[r], cre);
        cim = Mfma<R>::step(a[r].y, b[r], cim);
      }
    }
    __syncthreads();
    const bool tile_done = p1.tile != p0.tile;
    if constexpr (COMP) {
      if (++nacc == MFMA_COMP && !tile_done) {  // fold: hi += partial sum; carry the rounding error
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          cre[r] = two_sum(hre[r], (float)cre[r]);
          cim[r] = two_sum(him[r], (float)cim[r]);
        }
        nacc = 0;
      }
    }
    if (tile_done) {  // the tile's last batch: write its cells, start the next tile at zero
      const int col = lane & 15;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        cpx<R> val;
        if constexpr (COMP) {
          val = mk<R>((R)(hre[r] + (float)cre[r]), (R)(him[r] + (float)cim[r]));
          hre[r] = him[r] = 0.f;
        } else {
          val = mk<R>(cre[r], cim[r]);
        }
        grid[(int64_t)(ty0 + Mfma<R>::row(lane, r)) * g.Nu + tx0 + col] = val;
      }
      cre = acc_t{R(0), R(0), R(0), R(0)};
      cim = cre;
      nacc = 0;
    }
    p0 = p1;
    p1 = p2;
    p2 = p3;
    p3 = advance(p3);
  }
}

bool grid_mfma_ok(int Nu, int Nv, int Ju, int Jv) {
  return Nu % MFMA_TILE == 0 && Nv % MFMA_TILE == 0 && Nu >= MFMA_TILE + Ju - 1 && Nv >= MFMA_TILE + Jv - 1;
}

template <typename R, int A, int B, int TM>
static void degrid_launch(const DegridArgs<R>& a, dim3 blocks, size_t shm, hipStream_t s) {
  if (record_launch<DegridArgs<R>, k_degrid_tiles<R, A, B, TM, true>>(blocks, dim3(DG_THREADS), shm, true, a)) return;
  hipLaunchKernelGGL((k_degrid_tiles<R, A, B, TM>), blocks, dim3(DG_THREADS), shm, s, a, nullptr);
}

template <typename R>
void degrid(const GridGeom<R>& g, const VisData<R>& v, const TileLists& t, const cpx<R>* grid, cpx<R>* y, R scale,
            hipStream_t s) {
  if (v.M == 0 || t.n_own == 0) return;
  const dim3 blocks((unsigned)t.n_own);
  const size_t shm = sizeof(cpx<R>) * (size_t)(t.TU + g.Ju - 1) * (t.TV + g.Jv - 1);
  ProfScope prof("degrid", s);
  const int tm = sizeof(R) == 4 && g.Ju == g.Jv && g.Ju <= 8 ? (g.poly ? TAPS_POLY : g.tab_u ? TAPS_TABLE : TAPS_EVAL)
                                                            : TAPS_EVAL;
  const DegridArgs<R> a{g, v, t, grid, y, scale};
#define PF_DG(A, B)                                                                          \
  do {                                                                                       \
    if (tm == TAPS_POLY && A > 0) degrid_launch<R, A, B, TAPS_POLY>(a, blocks, shm, s);      \
    else if (tm == TAPS_TABLE && A > 0) degrid_launch<R, A, B, TAPS_TABLE>(a, blocks, shm, s); \
    else degrid_launch<R, A, B, TAPS_EVAL>(a, blocks, shm, s);                               \
  } while (0)
  if (g.Ju == g.Jv && fixed_len_ok(g)) {
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

bool grid_rc_ok(int Nu, int Nv, int TU, int TV, int Ju, int Jv) {
  return Ju <= 8 && Jv <= 8 && Nu >= TU + 8 && Nv >= TV + 8;
}

template <typename R, int A, int B, int TM>
static void grid_mfma_launch(const GridMfmaArgs<R>& a, dim3 blocks, hipStream_t s) {
  if (record_launch<GridMfmaArgs<R>, k_grid_mfma<R, A, B, TM, true>>(blocks, dim3(64), 0, true, a)) return;
  hipLaunchKernelGGL((k_grid_mfma<R, A, B, TM>), blocks, dim3(64), 0, s, a, nullptr);
}

template <typename R>
void grid_tiles(const GridGeom<R>& g, const VisData<R>& v, const TileLists& t, const TileLists& tm, const cpx<R>* y,
                cpx<R>* grid, R scale, hipStream_t s, int variant, int mfma_waves) {
  if (variant == GRID_MFMA && tm.slot && grid_mfma_ok(g.Nu, g.Nv, g.Ju, g.Jv)) {
    if (sizeof(R) == 4 && !tm.ent) throw Error(-5, "complex64 matrix-core gridding needs the entry records");
    const int nslots = tm.ntu * tm.ntv;
    const dim3 blocks((unsigned)(mfma_waves > 0 ? std::min(nslots, mfma_waves) : nslots));
    ProfScope prof("grid_tiles", s);
    // tap tables (complex64, equal supports <= 8): the pipelined table variant
    const int tmode = sizeof(R) == 4 && g.Ju == g.Jv && g.Ju <= 8 ? (g.poly ? TAPS_POLY : g.tab_u ? TAPS_TABLE : TAPS_EVAL)
                                                              : TAPS_EVAL;
    const GridMfmaArgs<R> a{g, v, tm, nslots, y, grid, scale};
#define PF_GM(A, B)                                                                             \
  do {                                                                                          \
    if (tmode == TAPS_POLY && A > 0) grid_mfma_launch<R, A, B, TAPS_POLY>(a, blocks, s);        \
    else if (tmode == TAPS_TABLE && A > 0) grid_mfma_launch<R, A, B, TAPS_TABLE>(a, blocks, s); \
    else grid_mfma_launch<R, A, B, TAPS_EVAL>(a, blocks, s);                                    \
  } while (0)
    if (g.Ju == g.Jv && fixed_len_ok(g)) {
      switch (g.Ju) {
        case 2: PF_GM(2, 2); break;
        case 3: PF_GM(3, 3); break;
        case 4: PF_GM(4, 4); break;
        case 5: PF_GM(5, 5); break;
        case 6: PF_GM(6, 6); break;
        case 7: PF_GM(7, 7); break;
        case 8: PF_GM(8, 8); break;
        default: PF_GM(0, 0);
      }
    } else {
      PF_GM(0, 0);
    }
#undef PF_GM
    PF_HIP(hipGetLastError());
    return;
  }
  const dim3 blocks((unsigned)(t.ntu * t.ntv));
  if (variant == GRID_RESIDUE && grid_rc_ok(g.Nu, g.Nv, t.TU, t.TV, g.Ju, g.Jv) && fixed_len_ok(g)) {
    const size_t shm = sizeof(cpx<R>) * ((size_t)(t.TU + 1) * t.TV + 64) + 2 * sizeof(R) * 2 * 64 * 8;
    ProfScope prof("grid_tiles", s);
#define PF_RC(A, B) hipLaunchKernelGGL((k_grid_rc<R, A, B>), blocks, dim3(64), shm, s, g, v, t, y, grid, scale)
#define PF_RC_V(A)             \
  switch (g.Jv) {              \
    case 1: PF_RC(A, 1); break; \
    case 2: PF_RC(A, 2); break; \
    case 3: PF_RC(A, 3); break; \
    case 4: PF_RC(A, 4); break; \
    case 5: PF_RC(A, 5); break; \
    case 6: PF_RC(A, 6); break; \
    case 7: PF_RC(A, 7); break; \
    default: PF_RC(A, 8);       \
  }
    switch (g.Ju) {
      case 1: PF_RC_V(1) break;
      case 2: PF_RC_V(2) break;
      case 3: PF_RC_V(3) break;
      case 4: PF_RC_V(4) break;
      case 5: PF_RC_V(5) break;
      case 6: PF_RC_V(6) break;
      case 7: PF_RC_V(7) break;
      default: PF_RC_V(8)
    }
#undef PF_RC_V
#undef PF_RC
    PF_HIP(hipGetLastError());
    return;
  }
  const size_t shm = 2 * sizeof(R) * (size_t)t.TU * t.TV;
  ProfScope prof("grid_tiles_atomic", s);
#define PF_GT(A, B) hipLaunchKernelGGL((k_grid_tiles<R, A, B>), blocks, dim3(64), shm, s, g, v, t, y, grid, scale)
  if (g.Ju == g.Jv && fixed_len_ok(g)) {
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

// ---------------------------------------------------------------------------------------------
// Permutation between the caller's order and the sorted order (VisPerm).  A random access costs one
// memory transaction whatever its size up to a line.  The two-pass plan moves each block's chunk or
// bucket through LDS: the chunk pass writes (reads) one contiguous run per bucket, the bucket pass
// writes (reads) its sorted range contiguously.  Every thread issues its loads before any LDS store.
// At M = 4e6: 8-byte values, scatter 52 us, two passes 28 us; 16-byte values, scatter 55 us, two
// passes 82 us (twice the bytes, and 489 blocks of 128 KiB LDS take two rounds of the 256 CUs).
// ---------------------------------------------------------------------------------------------
// one-pass: ys[inv[m]] = y[m] (coalesced reads, scattered writes)
template <typename R>
__global__ void __launch_bounds__(256) k_permute_scatter(const cpx<R>* __restrict__ y, const int32_t* __restrict__ inv,
                                                         cpx<R>* __restrict__ ys, int64_t M) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m < M) ys[inv[m]] = y[m];
}

// one-pass inverse: y[m] = ys[inv[m]] (scattered reads, coalesced writes)
template <typename R>
__global__ void __launch_bounds__(256) k_permute_gather(const cpx<R>* __restrict__ ys, const int32_t* __restrict__ inv,
                                                        cpx<R>* __restrict__ y, int64_t M) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m < M) y[m] = ys[inv[m]];
}

constexpr int PERM_THREADS = 1024;
template <typename R>
constexpr int perm_per_thread() { return (sizeof(R) == 4 ? 16384 : 8192) / PERM_THREADS; }

// TO_STAGING: chunk of the caller's vector -> staging runs; else staging runs -> caller's vector
template <typename R, bool TO_STAGING>
__global__ void __launch_bounds__(PERM_THREADS) k_perm_chunks(VisPerm p, const cpx<R>* __restrict__ src,
                                                              cpx<R>* __restrict__ dst) {
  extern __shared__ unsigned char smem_raw[];
  cpx<R>* sh = reinterpret_cast<cpx<R>*>(smem_raw);
  constexpr int E = perm_per_thread<R>();
  const int64_t c0 = (int64_t)blockIdx.x * p.S;
  const int n = (int)min((int64_t)p.S, p.M - c0);
  cpx<R> val[E];
  int idx[E];
  if (TO_STAGING) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = threadIdx.x + e * PERM_THREADS;
      if (k < n) {
        val[e] = src[c0 + k];
        idx[e] = p.l1[c0 + k];
      }
    }
#pragma unroll
    for (int e = 0; e < E; ++e)
      if (threadIdx.x + e * PERM_THREADS < n) sh[idx[e]] = val[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int j = threadIdx.x + e * PERM_THREADS;
      if (j < n) idx[e] = p.g1[c0 + j];
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int j = threadIdx.x + e * PERM_THREADS;
      if (j < n) dst[idx[e]] = sh[j];
    }
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int j = threadIdx.x + e * PERM_THREADS;
      if (j < n) idx[e] = p.g1[c0 + j];
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int j = threadIdx.x + e * PERM_THREADS;
      if (j < n) val[e] = src[idx[e]];
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int j = threadIdx.x + e * PERM_THREADS;
      if (j < n) sh[j] = val[e];
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = threadIdx.x + e * PERM_THREADS;
      if (k < n) idx[e] = p.l1[c0 + k];
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = threadIdx.x + e * PERM_THREADS;
      if (k < n) dst[c0 + k] = sh[idx[e]];
    }
  }
}

// TO_SORTED: bucket of the staging vector -> sorted range; else sorted range -> staging bucket
template <typename R, bool TO_SORTED>
__global__ void __launch_bounds__(PERM_THREADS) k_perm_buckets(VisPerm p, const cpx<R>* __restrict__ src,
                                                               cpx<R>* __restrict__ dst) {
  extern __shared__ unsigned char smem_raw[];
  cpx<R>* sh = reinterpret_cast<cpx<R>*>(smem_raw);
  constexpr int E = perm_per_thread<R>();
  const int64_t b0 = (int64_t)blockIdx.x * p.S;  // the bucket's staging range and sorted range start here
  const int n = (int)min((int64_t)p.S, p.M - b0);
  cpx<R> val[E];
  int idx[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = threadIdx.x + e * PERM_THREADS;
    if (k < n) {
      val[e] = src[b0 + k];
      idx[e] = p.p2[b0 + k];
    }
  }
  if (TO_SORTED) {
#pragma unroll
    for (int e = 0; e < E; ++e)
      if (threadIdx.x + e * PERM_THREADS < n) sh[idx[e]] = val[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = threadIdx.x + e * PERM_THREADS;
      if (k < n) dst[b0 + k] = sh[k];
    }
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e)
      if (threadIdx.x + e * PERM_THREADS < n) sh[threadIdx.x + e * PERM_THREADS] = val[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = threadIdx.x + e * PERM_THREADS;
      if (k < n) dst[b0 + k] = sh[idx[e]];
    }
  }
}

template <typename R>
static void check_perm(const VisPerm& p) {
  if (p.S <= 0 || p.S > perm_per_thread<R>() * PERM_THREADS || (p.S & (p.S - 1)) ||
      (int64_t)p.nb * p.S < p.M || (int64_t)(p.nb - 1) * p.S >= p.M)
    throw Error(-5, "permutation plan does not match the visibilities");  // PF_ERR_STATE
}

template <typename R>
void permute_to_sorted(const VisPerm& p, const cpx<R>* y, cpx<R>* tmp, cpx<R>* ys, hipStream_t s) {
  if (p.M == 0) return;
  ProfScope prof("permute_to_sorted", s);
  if (!p.two_pass) {
    if (!p.inv) throw Error(-5, "one-pass permutation without its index");  // PF_ERR_STATE
    hipLaunchKernelGGL((k_permute_scatter<R>), dim3((unsigned)((p.M + 255) / 256)), dim3(256), 0, s, y, p.inv, ys, p.M);
    PF_HIP(hipGetLastError());
    return;
  }
  check_perm<R>(p);
  const size_t shm = sizeof(cpx<R>) * (size_t)p.S;
  allow_lds((const void*)k_perm_chunks<R, true>, shm);
  allow_lds((const void*)k_perm_buckets<R, true>, shm);
  hipLaunchKernelGGL((k_perm_chunks<R, true>), dim3((unsigned)p.nb), dim3(PERM_THREADS), shm, s, p, y, tmp);
  hipLaunchKernelGGL((k_perm_buckets<R, true>), dim3((unsigned)p.nb), dim3(PERM_THREADS), shm, s, p, tmp, ys);
  PF_HIP(hipGetLastError());
}

template <typename R>
void permute_to_caller(const VisPerm& p, const cpx<R>* ys, cpx<R>* tmp, cpx<R>* y, hipStream_t s) {
  if (p.M == 0) return;
  ProfScope prof("permute_to_caller", s);
  if (!p.two_pass) {
    if (!p.inv) throw Error(-5, "one-pass permutation without its index");  // PF_ERR_STATE
    hipLaunchKernelGGL((k_permute_gather<R>), dim3((unsigned)((p.M + 255) / 256)), dim3(256), 0, s, ys, p.inv, y, p.M);
    PF_HIP(hipGetLastError());
    return;
  }
  check_perm<R>(p);
  const size_t shm = sizeof(cpx<R>) * (size_t)p.S;
  allow_lds((const void*)k_perm_chunks<R, false>, shm);
  allow_lds((const void*)k_perm_buckets<R, false>, shm);
  hipLaunchKernelGGL((k_perm_buckets<R, false>), dim3((unsigned)p.nb), dim3(PERM_THREADS), shm, s, p, ys, tmp);
  hipLaunchKernelGGL((k_perm_chunks<R, false>), dim3((unsigned)p.nb), dim3(PERM_THREADS), shm, s, p, tmp, y);
  PF_HIP(hipGetLastError());
}
template void permute_to_sorted<float>(const VisPerm&, const cpx<float>*, cpx<float>*, cpx<float>*, hipStream_t);
template void permute_to_sorted<double>(const VisPerm&, const cpx<double>*, cpx<double>*, cpx<double>*,
                                        hipStream_t);
template void permute_to_caller<float>(const VisPerm&, const cpx<float>*, cpx<float>*, cpx<float>*, hipStream_t);
template void permute_to_caller<double>(const VisPerm&, const cpx<double>*, cpx<double>*, cpx<double>*,
                                        hipStream_t);

template void degrid<float>(const GridGeom<float>&, const VisData<float>&, const TileLists&, const cpx<float>*,
                            cpx<float>*, float, hipStream_t);
template void degrid<double>(const GridGeom<double>&, const VisData<double>&, const TileLists&, const cpx<double>*,
                             cpx<double>*, double, hipStream_t);
template void grid_tiles<float>(const GridGeom<float>&, const VisData<float>&, const TileLists&, const TileLists&,
                                const cpx<float>*, cpx<float>*, float, hipStream_t, int, int);
template void grid_tiles<double>(const GridGeom<double>&, const VisData<double>&, const TileLists&, const TileLists&,
                                 const cpx<double>*, cpx<double>*, double, hipStream_t, int, int);

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

// x[k] = x[k] * h[k] * sc (the pointwise product of a convolution in Fourier space)
template <typename R>
__global__ void k_cmul_vec(cpx<R>* x, const cpx<R>* __restrict__ h, int64_t n, double sc) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    const cpx<R> a = cmul(x[k], h[k]);
    x[k] = mk<R>((R)((double)a.x * sc), (R)((double)a.y * sc));
  }
}
template <typename R>
void cmul_vec(cpx<R>* x, const cpx<R>* h, int64_t n, double sc, hipStream_t s) {
  hipLaunchKernelGGL((k_cmul_vec<R>), dim3(1024), dim3(256), 0, s, x, h, n, sc);
  PF_HIP(hipGetLastError());
}
template void cmul_vec<float>(cpx<float>*, const cpx<float>*, int64_t, double, hipStream_t);
template void cmul_vec<double>(cpx<double>*, const cpx<double>*, int64_t, double, hipStream_t);

template void norm2_fp64<float>(const cpx<float>*, int64_t, double*, double*, hipStream_t);
template void norm2_fp64<double>(const cpx<double>*, int64_t, double*, double*, hipStream_t);
template void scale_vec<float>(cpx<float>*, int64_t, double, hipStream_t);
template void scale_vec<double>(cpx<double>*, int64_t, double, hipStream_t);

}  // namespace pf
