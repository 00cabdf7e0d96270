// pf_build.hip -- operator construction on the device: the sort of the visibilities by owner tile, the
// gridding tile lists, the permutation plans and the per-visibility records the kernels read.
//
// Every order is fixed by a stable radix sort (rocPRIM) of keys whose ties come in ascending input
// order, so the result is the same order a serial host pass gives, run to run:
//   * visibilities by (owner tile, first-tap cell in the tile), ties by the caller's index;
//   * each tile's list by tile, ties by sorted index (so a list is in ascending sorted order);
//   * the two-pass permutation plan by (chunk, bucket) and (bucket, chunk), ties by the caller's index.
// Counts that only fix sizes (per-tile totals) come from atomic increments, whose totals are exact.
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/purify_amd.h"
#include "pf_internal.h"

namespace pf {
namespace {

__device__ inline int wrapd(int64_t a, int n) {
  const int64_t r = a % n;
  return (int)(r < 0 ? r + n : r);
}

unsigned nblk(int64_t n, int t = 256) { return (unsigned)std::max<int64_t>(1, (n + t - 1) / t); }

int bits_for(uint64_t nkeys) {
  int b = 1;
  while (b < 32 && ((uint64_t)1 << b) < nkeys) ++b;
  return b;
}

// key = owner tile * TU * TV + first-tap cell inside the tile; flag non-finite coordinates
__global__ void k_keys(const cpx<double>* __restrict__ uv, int64_t M, int Nu, int Nv, int Ku, int Kv, int TU, int TV,
                       int ntu, uint32_t* __restrict__ key, int32_t* __restrict__ idx, int* __restrict__ bad) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const cpx<double> c = uv[m];
  if (!isfinite(c.x) || !isfinite(c.y)) {
    atomicOr(bad, 1);
    key[m] = 0;
    idx[m] = (int32_t)m;
    return;
  }
  const int64_t ku = (int64_t)floor(c.x - Ku * 0.5), kv = (int64_t)floor(c.y - Kv * 0.5);  // operators.cc:34-35
  const int q0 = wrapd(ku + 1, Nu), p0 = wrapd(kv + 1, Nv);
  const uint32_t tile = (uint32_t)((p0 / TV) * ntu + q0 / TU);
  key[m] = tile * (uint32_t)(TU * TV) + (uint32_t)((p0 % TV) * TU + q0 % TU);
  idx[m] = (int32_t)m;
}

// Run bounds of a sorted key array: for the tile t = key / TT of each run of equal tiles, beg[t] = the
// run's first index and fin[t] = one past its last (both stay zero for a tile with no entries), so the
// tile's count is fin[t] - beg[t].  One writer per value: no atomics, which a dense tile would serialise.
__global__ void k_tile_runs(const uint32_t* __restrict__ key, int64_t M, uint32_t TT, uint32_t* __restrict__ beg,
                            uint32_t* __restrict__ fin) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const uint32_t t = key[i] / TT;
  if (i == 0 || key[i - 1] / TT != t) beg[t] = (uint32_t)i;
  if (i + 1 == M || key[i + 1] / TT != t) fin[t] = (uint32_t)(i + 1);
}

// the sorted visibilities' tap origins k = floor(u - K/2) and (complex64) fractions, (complex128) pixels,
// and weights, in the sorted order
template <typename R>
__global__ void k_gather(const cpx<double>* __restrict__ uv, const cpx<double>* __restrict__ wt,
                         const int32_t* __restrict__ perm, int64_t M, int Ku, int Kv, int2* __restrict__ kk,
                         cpx<float>* __restrict__ ff, cpx<double>* __restrict__ uvs, cpx<R>* __restrict__ ws) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const int64_t m = perm[i];
  const cpx<double> c = uv[m];
  const double ku = floor(c.x - Ku * 0.5), kv = floor(c.y - Kv * 0.5);
  if (kk) kk[i] = make_int2((int)ku, (int)kv);
  if (ff) ff[i] = mk<float>((float)(c.x - ku), (float)(c.y - kv));
  if (uvs) uvs[i] = c;
  if (ws) ws[i] = mk<R>((R)wt[m].x, (R)wt[m].y);
}

// distinct tile columns / rows met by a footprint of K cells from cell x0 (mod N)
__device__ inline int tile_set(int x0, int K, int N, int T, int* out) {
  int n = 0;
  for (int a = 0; a < K; ++a) {
    const int t = ((x0 + a) % N) / T;
    bool seen = false;
    for (int k = 0; k < n; ++k) seen |= out[k] == t;
    if (!seen) out[n++] = t;
  }
  return n;
}
constexpr int MAX_TILE_SPAN = 8;  // distinct tiles per axis a footprint can meet (K <= 7 T)

__global__ void k_touch_count(const int2* __restrict__ kk, int64_t M, int Nu, int Nv, int Ku, int Kv, int TU, int TV,
                              uint64_t* __restrict__ nt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const int2 k = kk[i];
  int tc[MAX_TILE_SPAN], tr[MAX_TILE_SPAN];
  const int nc = tile_set(wrapd((int64_t)k.x + 1, Nu), Ku, Nu, TU, tc);
  const int nr = tile_set(wrapd((int64_t)k.y + 1, Nv), Kv, Nv, TV, tr);
  nt[i] = (uint64_t)(nc * nr);
}
__global__ void k_touch_fill(const int2* __restrict__ kk, const uint64_t* __restrict__ off, int64_t M, int Nu, int Nv,
                             int Ku, int Kv, int TU, int TV, int ntu, uint32_t* __restrict__ tkey,
                             uint32_t* __restrict__ tval) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const int2 k = kk[i];
  int tc[MAX_TILE_SPAN], tr[MAX_TILE_SPAN];
  const int nc = tile_set(wrapd((int64_t)k.x + 1, Nu), Ku, Nu, TU, tc);
  const int nr = tile_set(wrapd((int64_t)k.y + 1, Nv), Kv, Nv, TV, tr);
  uint64_t o = off[i];
  for (int r = 0; r < nr; ++r)
    for (int c = 0; c < nc; ++c) {
      tkey[o] = (uint32_t)(tr[r] * ntu + tc[c]);
      tval[o] = (uint32_t)i;
      ++o;
    }
}

// complex64 matrix-core list records: (8 * sorted index, wrapped tap origin cu | cv << 16, fractions)
__global__ void k_entries(const uint32_t* __restrict__ list, int64_t n, const int2* __restrict__ kk,
                          const cpx<float>* __restrict__ ff, int Nu, int Nv, int4* __restrict__ ent) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t i = list[p];
  const int2 k = kk[i];
  const cpx<float> f = ff[i];
  const int cu = wrapd(k.x, Nu), cv = wrapd(k.y, Nv);
  ent[p] = make_int4((int32_t)(i * 8u), cu | (cv << 16), __float_as_int(f.x), __float_as_int(f.y));
}

__global__ void k_inverse(const int32_t* __restrict__ perm, int64_t M, int32_t* __restrict__ inv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < M) inv[perm[i]] = (int32_t)i;
}
// two-pass plan keys: (chunk, bucket) and (bucket, chunk) of caller index m
__global__ void k_plan_keys(const int32_t* __restrict__ inv, int64_t M, int S, uint32_t nb, uint32_t* __restrict__ k1,
                            uint32_t* __restrict__ k2, int32_t* __restrict__ idx) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const uint32_t c = (uint32_t)(m / S), b = (uint32_t)(inv[m] / S);
  k1[m] = c * nb + b;
  k2[m] = b * nb + c;
  idx[m] = (int32_t)m;
}
// after the (chunk, bucket) sort: the position j of m inside its chunk's LDS order
__global__ void k_plan_l1(const int32_t* __restrict__ order1, int64_t M, int S, uint16_t* __restrict__ l1,
                          int32_t* __restrict__ jpos) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= M) return;
  const int32_t m = order1[q];
  const int32_t j = (int32_t)(q - (int64_t)(m / S) * S);
  l1[m] = (uint16_t)j;
  jpos[m] = j;
}
// after the (bucket, chunk) sort: q is m's staging index
__global__ void k_plan_g1p2(const int32_t* __restrict__ order2, const int32_t* __restrict__ inv,
                            const int32_t* __restrict__ jpos, int64_t M, int S, int32_t* __restrict__ g1,
                            uint16_t* __restrict__ p2) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= M) return;
  const int32_t m = order2[q];
  const int64_t c = m / S, b = inv[m] / S;
  g1[c * S + jpos[m]] = (int32_t)q;
  p2[q] = (uint16_t)(inv[m] - b * S);
}

// device scratch of one construction, freed at the end
struct Scratch {
  std::vector<void*> bufs;
  ~Scratch() {
    cstream_sync();
    for (void* p : bufs) dev_free(p);
  }
  template <typename T>
  T* alloc(size_t n) {
    T* p = dev_alloc_n<T>(n);
    bufs.push_back(p);
    return p;
  }
};

// per-tile entry counts (host, nt tiles) of a sorted key array of n entries (tile = key / TT); synchronises s
void tile_counts(Scratch& sc, const uint32_t* key_s, int64_t n, uint32_t TT, int nt, std::vector<uint32_t>& count,
                 hipStream_t s) {
  count.assign((size_t)nt, 0);
  if (n == 0) return;
  uint32_t* beg = sc.alloc<uint32_t>((size_t)nt);
  uint32_t* fin = sc.alloc<uint32_t>((size_t)nt);
  PF_HIP(hipMemsetAsync(beg, 0, sizeof(uint32_t) * nt, s));
  PF_HIP(hipMemsetAsync(fin, 0, sizeof(uint32_t) * nt, s));
  hipLaunchKernelGGL(k_tile_runs, dim3(nblk(n)), dim3(256), 0, s, key_s, n, TT, beg, fin);
  PF_HIP(hipGetLastError());
  std::vector<uint32_t> hb((size_t)nt);
  PF_HIP(hipMemcpyAsync(hb.data(), beg, sizeof(uint32_t) * nt, hipMemcpyDeviceToHost, s));
  PF_HIP(hipMemcpyAsync(count.data(), fin, sizeof(uint32_t) * nt, hipMemcpyDeviceToHost, s));
  PF_HIP(hipStreamSynchronize(s));
  for (int t = 0; t < nt; ++t) count[t] -= hb[t];
}

// stable sort of (key, value) pairs on the device; keys below 2^bits
template <typename V>
void sort_pairs(Scratch& sc, const uint32_t* kin, uint32_t* kout, const V* vin, V* vout, int64_t n, int bits,
                hipStream_t s) {
  if (n == 0) return;
  size_t tmp = 0;
  PF_HIP(rocprim::radix_sort_pairs(nullptr, tmp, kin, kout, vin, vout, (size_t)n, 0, bits, s));
  void* t = sc.alloc<unsigned char>(tmp);
  PF_HIP(rocprim::radix_sort_pairs(t, tmp, kin, kout, vin, vout, (size_t)n, 0, bits, s));
}

}  // namespace

// The owner-tile sort: perm (sorted position -> caller's index, device, M entries, allocated by the
// caller) and the per-tile owner counts on the host.
void dev_sort_visibilities(const cpx<double>* d_uv, int64_t M, int Nu, int Nv, int Ku, int Kv, int TU, int TV,
                           int32_t* d_perm, std::vector<uint32_t>& owner_count) {
  hipStream_t s = cstream();
  Scratch sc;
  const int ntu = Nu / TU, ntv = Nv / TV, nt = ntu * ntv;
  owner_count.assign((size_t)nt, 0);
  if (M == 0) return;
  uint32_t* key = sc.alloc<uint32_t>(M);
  uint32_t* key_s = sc.alloc<uint32_t>(M);
  int32_t* idx = sc.alloc<int32_t>(M);
  int* bad = sc.alloc<int>(1);
  PF_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_keys, dim3(nblk(M)), dim3(256), 0, s, d_uv, M, Nu, Nv, Ku, Kv, TU, TV, ntu, key, idx, bad);
  PF_HIP(hipGetLastError());
  sort_pairs<int32_t>(sc, key, key_s, idx, d_perm, M, bits_for((uint64_t)nt * TU * TV), s);
  int hbad = 0;
  PF_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  tile_counts(sc, key_s, M, (uint32_t)(TU * TV), nt, owner_count, s);
  if (hbad) throw Error(PF_ERR_ARG, "non-finite u/v coordinate");
}

// Sorted records: kk (tap origins), ff (complex64 fractions), uvs (complex128 pixels), ws (weights in
// precision R); any may be null.
template <typename R>
void dev_gather_sorted(const cpx<double>* d_uv, const cpx<double>* d_wt, const int32_t* d_perm, int64_t M, int Ku,
                       int Kv, int2* kk, cpx<float>* ff, cpx<double>* uvs, cpx<R>* ws) {
  if (M == 0) return;
  hipLaunchKernelGGL((k_gather<R>), dim3(nblk(M)), dim3(256), 0, cstream(), d_uv, d_wt, d_perm, M, Ku, Kv, kk, ff, uvs,
                     ws);
  PF_HIP(hipGetLastError());
}
template void dev_gather_sorted<float>(const cpx<double>*, const cpx<double>*, const int32_t*, int64_t, int, int, int2*,
                                       cpx<float>*, cpx<double>*, cpx<float>*);
template void dev_gather_sorted<double>(const cpx<double>*, const cpx<double>*, const int32_t*, int64_t, int, int,
                                        int2*, cpx<float>*, cpx<double>*, cpx<double>*);

// Tile lists of TU x TV tiles for footprints Ku x Kv at the sorted tap origins kk: the device list
// (from dev_alloc; the caller owns it), its length, and the per-tile entry counts on the host.
uint32_t* dev_tile_lists(const int2* kk, int64_t M, int Nu, int Nv, int Ku, int Kv, int TU, int TV,
                         std::vector<uint32_t>& count, int64_t& len) {
  hipStream_t s = cstream();
  Scratch sc;
  const int ntu = Nu / TU, ntv = Nv / TV, nt = ntu * ntv;
  count.assign((size_t)nt, 0);
  len = 0;
  if ((Ku + TU - 1) / TU + 1 > MAX_TILE_SPAN || (Kv + TV - 1) / TV + 1 > MAX_TILE_SPAN)
    throw Error(PF_ERR_UNSUPPORTED, "footprint spans too many tiles");
  uint32_t* list = nullptr;
  if (M == 0) {
    return dev_alloc_n<uint32_t>(1);
  }
  uint64_t* nt_i = sc.alloc<uint64_t>(M);
  uint64_t* off = sc.alloc<uint64_t>(M);
  hipLaunchKernelGGL(k_touch_count, dim3(nblk(M)), dim3(256), 0, s, kk, M, Nu, Nv, Ku, Kv, TU, TV, nt_i);
  PF_HIP(hipGetLastError());
  {
    size_t tmp = 0;
    PF_HIP(rocprim::exclusive_scan(nullptr, tmp, nt_i, off, (uint64_t)0, (size_t)M, rocprim::plus<uint64_t>(), s));
    void* t = sc.alloc<unsigned char>(tmp);
    PF_HIP(rocprim::exclusive_scan(t, tmp, nt_i, off, (uint64_t)0, (size_t)M, rocprim::plus<uint64_t>(), s));
  }
  uint64_t last[2] = {0, 0};
  PF_HIP(hipMemcpyAsync(&last[0], off + (M - 1), sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  PF_HIP(hipMemcpyAsync(&last[1], nt_i + (M - 1), sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  PF_HIP(hipStreamSynchronize(s));
  const uint64_t total = last[0] + last[1];
  if (total > 0xffffffffULL) throw Error(PF_ERR_UNSUPPORTED, "gridding tile lists exceed 2^32 entries");
  len = (int64_t)total;
  uint32_t* tkey = sc.alloc<uint32_t>(total);
  uint32_t* tkey_s = sc.alloc<uint32_t>(total);
  uint32_t* tval = sc.alloc<uint32_t>(total);
  list = dev_alloc_n<uint32_t>((size_t)total);
  try {
    hipLaunchKernelGGL(k_touch_fill, dim3(nblk(M)), dim3(256), 0, s, kk, off, M, Nu, Nv, Ku, Kv, TU, TV, ntu, tkey,
                       tval);
    PF_HIP(hipGetLastError());
    sort_pairs<uint32_t>(sc, tkey, tkey_s, tval, list, (int64_t)total, bits_for((uint64_t)nt), s);
    tile_counts(sc, tkey_s, (int64_t)total, 1u, nt, count, s);
  } catch (...) {
    cstream_sync();
    dev_free(list);
    throw;
  }
  return list;
}

void dev_entries(const uint32_t* list, int64_t n, const int2* kk, const cpx<float>* ff, int Nu, int Nv, int4* ent) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_entries, dim3(nblk(n)), dim3(256), 0, cstream(), list, n, kk, ff, Nu, Nv, ent);
  PF_HIP(hipGetLastError());
}

void dev_inverse(const int32_t* perm, int64_t M, int32_t* inv) {
  if (M == 0) return;
  hipLaunchKernelGGL(k_inverse, dim3(nblk(M)), dim3(256), 0, cstream(), perm, M, inv);
  PF_HIP(hipGetLastError());
}

// The two-pass permutation plan (VisPerm) of perm with chunks and buckets of S: l1, g1 and p2 (device).
void dev_perm_plan(const int32_t* perm, int64_t M, int S, int nb, uint16_t* l1, int32_t* g1, uint16_t* p2) {
  if (M == 0) return;
  hipStream_t s = cstream();
  Scratch sc;
  int32_t* inv = sc.alloc<int32_t>(M);
  dev_inverse(perm, M, inv);
  uint32_t* k1 = sc.alloc<uint32_t>(M);
  uint32_t* k2 = sc.alloc<uint32_t>(M);
  uint32_t* ks = sc.alloc<uint32_t>(M);
  int32_t* idx = sc.alloc<int32_t>(M);
  int32_t* order = sc.alloc<int32_t>(M);
  int32_t* jpos = sc.alloc<int32_t>(M);
  hipLaunchKernelGGL(k_plan_keys, dim3(nblk(M)), dim3(256), 0, s, inv, M, S, (uint32_t)nb, k1, k2, idx);
  PF_HIP(hipGetLastError());
  const int bits = bits_for((uint64_t)nb * (uint64_t)nb);
  sort_pairs<int32_t>(sc, k1, ks, idx, order, M, bits, s);
  hipLaunchKernelGGL(k_plan_l1, dim3(nblk(M)), dim3(256), 0, s, order, M, S, l1, jpos);
  PF_HIP(hipGetLastError());
  sort_pairs<int32_t>(sc, k2, ks, idx, order, M, bits, s);
  hipLaunchKernelGGL(k_plan_g1p2, dim3(nblk(M)), dim3(256), 0, s, order, inv, jpos, M, S, g1, p2);
  PF_HIP(hipGetLastError());
  PF_HIP(hipStreamSynchronize(s));
}

}  // namespace pf
