// pf_noalloc.hip -- the unitary 2-D FFT without any device allocation, for callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops, harnesses that forbid it).
//
// pf_fft2d builds its twiddle table on the host and uploads it to a fresh allocation, and allocates a
// grid-sized workspace when none is given.  pf_fft2d_noalloc runs the same row and column kernels, with
// the twiddle table and the workspace in one file-scope device array, `scratch`: the table is computed
// on the device at every call (exp(-2 pi i m / L) from the first octant by symmetry, as the host table
// is, in FP64 sincospi), the workspace holds the row pass's output.  Each GPU holds its own copy of the
// array from the first call on it until the process ends.  Calls on one GPU share that copy, so calls
// that overlap on one GPU (different streams or threads) race; calls on one stream are ordered.
#include <mutex>

#include "pf_internal.h"
#include "purify_amd.h"

namespace pf {
namespace {

constexpr size_t NOALLOC_TW_BYTES = (size_t)MAX_AXIS * 16;  // twiddles of the longest axis, complex128
constexpr size_t NOALLOC_WS_BYTES = (size_t)PF_FFT2D_NOALLOC_MAX_BYTES;
constexpr size_t NOALLOC_BYTES = NOALLOC_TW_BYTES + NOALLOC_WS_BYTES;

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
void* scratch_dev[MAX_DEVICES] = {nullptr};

// The device address of this GPU's copy of `scratch` (the array's name is a host-side stand-in).
void* scratch_ptr() {
  int dev = 0;
  PF_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) throw Error(PF_ERR_UNSUPPORTED, "device index beyond the library's table");
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (!scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw Error(PF_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    scratch_dev[dev] = p;
  }
  return scratch_dev[dev];
}

// tw[m] = exp(-2 pi i m / L): the angle 2 m / L (units of pi) is reduced to [0, 1/4] by octant symmetry,
// so every mirror image of one value is the same rounding of it.
template <typename R>
__global__ void k_twiddles(cpx<R>* tw, int L) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= L) return;
  const int64_t num = 8LL * m;  // angle in units of pi/4 is num / L
  const int o = (int)(num / L);
  const int64_t rmd = num - (int64_t)o * L;
  double sn, c;
  const bool swap = o % 2 != 0;
  sincospi((double)(swap ? L - rmd : rmd) / (4.0 * (double)L), &sn, &c);
  if (swap) {
    const double t = c;
    c = sn;
    sn = t;
  }
  double cr, sr;
  switch ((o / 2) % 4) {
    case 0: cr = c; sr = sn; break;
    case 1: cr = -sn; sr = c; break;
    case 2: cr = -c; sr = -sn; break;
    default: cr = sn; sr = -c; break;
  }
  tw[m] = mk<R>((R)cr, (R)(-sr));
}

template <typename R>
void fft2d_noalloc(int Nv, int Nu, bool inv, const cpx<R>* in, cpx<R>* out, hipStream_t s) {
  Fft2dPlan p;
  fft_plan_setup(p, Nv, Nu, sizeof(R) == 8 ? 1 : 0);
  unsigned char* base = static_cast<unsigned char*>(scratch_ptr());
  cpx<R>* tw = reinterpret_cast<cpx<R>*>(base);
  cpx<R>* ws = reinterpret_cast<cpx<R>*>(base + NOALLOC_TW_BYTES);
  p.tw = tw;
  hipLaunchKernelGGL((k_twiddles<R>), dim3((p.Lmax + 255) / 256), dim3(256), 0, s, tw, p.Lmax);
  PF_HIP(hipGetLastError());
  SZ<R> sz;
  fft_rows<R>(p, ROW_PLAIN, inv, in, ws, Nv, sz, s);
  fft_cols<R>(p, inv, ws, out, RowWindow{0, 0}, false, false, s);
  scale_vec<R>(out, (int64_t)Nv * Nu, 1.0 / std::sqrt((double)Nv * Nu), s);
  p.tw = nullptr;  // the table belongs to the scratch array, not to the plan
}

bool pow2(int n) { return n >= 2 && (n & (n - 1)) == 0; }

}  // namespace
}  // namespace pf

using namespace pf;

extern "C" {

int pf_fft2d_noalloc(int32_t Nv, int32_t Nu, int32_t precision, int32_t sign, const void* in_dev, void* out_dev,
                     pf_stream stream) {
  return guarded([&] {
    if (sign != -1 && sign != 1) throw Error(PF_ERR_ARG, "sign must be -1 or +1");
    if (precision != PF_C64 && precision != PF_C128) throw Error(PF_ERR_ARG, "precision must be PF_C64 or PF_C128");
    if (!in_dev || !out_dev) throw Error(PF_ERR_ARG, "null buffer");
    if (!pow2(Nv) || !pow2(Nu)) throw Error(PF_ERR_UNSUPPORTED, "FFT axes must be powers of two >= 2");
    if (Nv > MAX_AXIS || Nu > MAX_AXIS)
      throw Error(PF_ERR_UNSUPPORTED, "FFT axis longer than " + std::to_string(MAX_AXIS) + " is not supported");
    const size_t bytes = (size_t)Nv * Nu * (precision == PF_C128 ? 16 : 8);
    if (bytes > NOALLOC_WS_BYTES)
      throw Error(PF_ERR_ARG, "grid of " + std::to_string(bytes) + " bytes exceeds the allocation-free limit of " +
                                  std::to_string(NOALLOC_WS_BYTES) + " bytes (PF_FFT2D_NOALLOC_MAX_BYTES)");
    hipStream_t s = (hipStream_t)stream;
    if (precision == PF_C128)
      fft2d_noalloc<double>(Nv, Nu, sign > 0, (const cpx<double>*)in_dev, (cpx<double>*)out_dev, s);
    else
      fft2d_noalloc<float>(Nv, Nu, sign > 0, (const cpx<float>*)in_dev, (cpx<float>*)out_dev, s);
  });
}

}  // extern "C"
