// pf_device.h -- device-side building blocks shared by the HIP kernels (gfx950, wave64).
//
//  * cpx<R>: interleaved complex of float (complex64 path) or double (complex128 path).
//  * KParams<R>: one 1-D gridding kernel of cpp/purify/kernels.cc:233-323 in evaluable form.
//    The Kaiser-Bessel kernel I0(alpha sqrt(t)) / I0(alpha), t = 1 - (2x/J)^2, is an entire function
//    of t (its power series is sum_k (alpha^2/4)^k t^k / (k!)^2 / I0(alpha)).  The host expands it in
//    Chebyshev polynomials of s = 2t - 1 on [0, 1] (coefficients from the long-double power series at
//    Chebyshev nodes) and the device sums the expansion with Clenshaw's recurrence: 13 terms reach
//    1e-9 and 20 terms 1e-17 at J = 6, against 22 and 33 power-series terms, and the fixed length
//    lets the loop unroll with the coefficients held in scalar registers.
//  * dft<R, N, INV>: in-register N-point DFT (N = 2..16) by radix-2 decimation in time.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pf {

template <typename R>
struct alignas(2 * sizeof(R)) cpx {
  R x, y;
};

template <typename R>
__host__ __device__ inline cpx<R> mk(R a, R b) {
  cpx<R> c;
  c.x = a;
  c.y = b;
  return c;
}
template <typename R>
__device__ inline cpx<R> operator+(cpx<R> a, cpx<R> b) {
  return mk(a.x + b.x, a.y + b.y);
}
template <typename R>
__device__ inline cpx<R> operator-(cpx<R> a, cpx<R> b) {
  return mk(a.x - b.x, a.y - b.y);
}
// The complex products state their fused multiply-adds explicitly: left to the compiler, a * b + c * d
// may be contracted as fma(a, b, c * d) at one inlined call site and fma(c, d, a * b) at another, and
// a kernel whose loop body is inlined twice would then round the same entry differently depending on
// which copy handled it.
__device__ inline float fmar(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fmar(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <typename R>
__device__ inline cpx<R> cmul(cpx<R> a, cpx<R> b) {
  return mk(fmar(a.x, b.x, -(a.y * b.y)), fmar(a.x, b.y, a.y * b.x));
}
// a * conj(b)
template <typename R>
__device__ inline cpx<R> cmulc(cpx<R> a, cpx<R> b) {
  return mk(fmar(a.x, b.x, a.y * b.y), fmar(a.y, b.x, -(a.x * b.y)));
}
template <typename R>
__device__ inline cpx<R> scale(cpx<R> a, R s) {
  return mk(a.x * s, a.y * s);
}

// Sum over the 64 lanes of a wave in a fixed butterfly order; every lane ends with the total.
template <typename R>
__device__ inline cpx<R> wave_sum(cpx<R> a) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    a.x += __shfl_xor(a.x, m, 64);
    a.y += __shfl_xor(a.y, m, 64);
  }
  return a;
}

// ------------------------------------------------------------------------------------------
// Gridding kernels (kernels.cc).  Kinds follow kernels::kernel (kernels.h:15).
// ------------------------------------------------------------------------------------------
enum KernelKind { K_KB = 0, K_GAUSS = 1, K_BOX = 2, K_PSWF = 3, K_KBMIN = 4, K_GAUSS_ALT = 5 };
constexpr int KB_MAX_COEF = 40;
// Fixed Clenshaw lengths: enough for supports up to 8 (checked on the host, which falls back to a
// runtime-length sum for longer expansions).
template <typename R>
struct KBLen;
template <>
struct KBLen<float> {
  static constexpr int N = 16;
};
template <>
struct KBLen<double> {
  static constexpr int N = 24;
};

template <typename R>
struct KParams {
  int kind;
  int J;         // support the kernel is evaluated with
  int ncoef;     // Kaiser-Bessel Chebyshev length (<= KBLen<R>::N uses the unrolled path)
  R half;        // J / 2
  R four_over_J2;  // 4 / J^2
  R two_over_J;  // 2 / J
  R inv_two_sigma2;  // Gaussian 1 / (2 sigma^2)
  R box;         // 1 / J
  R coef[KB_MAX_COEF];  // KB: Chebyshev coefficients of I0(alpha sqrt(t))/I0(alpha) in s = 2t - 1
};

// Clenshaw sum of c[0] T0(s) + ... + c[n-1] T_{n-1}(s).
template <typename R, int N>
__device__ inline R clenshaw_fixed(const R* c, R s) {
  const R s2 = s + s;
  R b1 = R(0), b2 = R(0);
#pragma unroll
  for (int k = N - 1; k >= 1; --k) {
    const R b0 = c[k] + s2 * b1 - b2;
    b2 = b1;
    b1 = b0;
  }
  return c[0] + s * b1 - b2;
}
template <typename R>
__device__ inline R clenshaw_rt(const R* c, int n, R s) {
  const R s2 = s + s;
  R b1 = R(0), b2 = R(0);
  for (int k = n - 1; k >= 1; --k) {
    const R b0 = c[k] + s2 * b1 - b2;
    b2 = b1;
    b1 = b0;
  }
  return c[0] + s * b1 - b2;
}

// Schwab's rational PSWF approximation (kernels.cc:71-105), alpha = 1, valid for J = 6 only.
template <typename R>
__device__ inline R pswf_eval(R eta0, int J) {
  if (J != 6) return R(0);
  const R a = eta0 < R(0) ? -eta0 : eta0;
  R num, den, e;
  if (a <= R(0.75)) {
    e = eta0 * eta0 - R(0.75 * 0.75);
    num = R(2.312756e-1);
    num = e * num + R(-5.335581e-1);
    num = e * num + R(6.278660e-1);
    num = e * num + R(-3.644705e-1);
    num = e * num + R(8.203343e-2);
    den = R(2.078043e-1);
    den = e * den + R(8.212018e-1);
    den = e * den + R(1.0);
  } else if (a <= R(1)) {
    e = eta0 * eta0 - R(1);
    num = R(6.412774e-2);
    num = e * num + R(-1.201436e-1);
    num = e * num + R(1.021332e-1);
    num = e * num + R(-3.697768e-2);
    num = e * num + R(4.028559e-3);
    den = R(2.918724e-1);
    den = e * den + R(9.599102e-1);
    den = e * den + R(1.0);
  } else {
    return R(0);
  }
  // The reference's numerator has a zero leading coefficient (p[5] = 0), omitted above.
  return num / den;
}

template <typename R>
__device__ inline R dexp(R x);
template <>
__device__ inline float dexp<float>(float x) {
  return expf(x);
}
template <>
__device__ inline double dexp<double>(double x) {
  return exp(x);
}

// Value of the gridding kernel at offset x (|x| <= J/2 by construction of the taps).  LEN: 1 = the
// fixed-length Kaiser-Bessel sum only (the caller guarantees ncoef <= KBLen<R>::N, which keeps the
// unused coefficients out of registers), 0 = runtime length, -1 = choose by ncoef.
template <typename R, int LEN = -1>
__device__ inline R kern1d(const KParams<R>& p, R x) {
  switch (p.kind) {
    case K_KB:
    case K_KBMIN: {
      // t = 1 - (2x/J)^2 written as (J/2 - x)(J/2 + x) * 4/J^2: no cancellation at the edges.
      R t = (p.half - x) * (p.half + x) * p.four_over_J2;
      // Outside the support the reference takes sqrt of a negative number and returns NaN
      // (kernels.cc:20); that happens when Ju != Jv because of the kernel-slot swap of operators.h:324.
      if (t < R(0)) return R(NAN);
      const R s = t + t - R(1);
      if constexpr (LEN == 1) {
        return clenshaw_fixed<R, KBLen<R>::N>(p.coef, s);
      } else if constexpr (LEN == 0) {
        return clenshaw_rt<R>(p.coef, p.ncoef, s);
      } else {
        if (p.ncoef <= KBLen<R>::N) return clenshaw_fixed<R, KBLen<R>::N>(p.coef, s);
        return clenshaw_rt<R>(p.coef, p.ncoef, s);
      }
    }
    case K_GAUSS:
    case K_GAUSS_ALT:
      return dexp<R>(-x * x * p.inv_two_sigma2);
    case K_PSWF: {
      const R eta = p.two_over_J * x;
      return pswf_eval<R>(eta, p.J) * (R(1) - eta * eta);
    }
    case K_BOX:
      return p.box;
    default:
      return R(0);
  }
}

// ------------------------------------------------------------------------------------------
// In-register DFT.  W16[k] = exp(-2 pi i k / 16); the inverse uses the conjugate.
// ------------------------------------------------------------------------------------------
constexpr double C16_1 = 0.92387953251128675613;  // cos(pi/8)
constexpr double C16_2 = 0.70710678118654752440;  // cos(pi/4)
constexpr double C16_3 = 0.38268343236508977173;  // cos(3 pi/8)

// v * exp(-+ 2 pi i K / 16), K in [0, 16): sign - for the forward transform, + for INV.
template <typename R, int K, bool INV>
__device__ inline cpx<R> tw16(cpx<R> v) {
  constexpr double cs[16] = {1.0,    C16_1,  C16_2,  C16_3,  0.0,    -C16_3, -C16_2, -C16_1,
                             -1.0,   -C16_1, -C16_2, -C16_3, 0.0,    C16_3,  C16_2,  C16_1};
  constexpr double sn[16] = {0.0,    C16_3,  C16_2,  C16_1,  1.0,    C16_1,  C16_2,  C16_3,
                             0.0,    -C16_3, -C16_2, -C16_1, -1.0,   -C16_1, -C16_2, -C16_3};
  if constexpr (K == 0) {
    return v;
  } else if constexpr (K == 8) {
    return mk<R>(-v.x, -v.y);
  } else if constexpr (K == 4) {
    // * (-i) forward, * (+i) inverse
    if constexpr (INV) return mk<R>(-v.y, v.x);
    else return mk<R>(v.y, -v.x);
  } else if constexpr (K == 12) {
    if constexpr (INV) return mk<R>(v.y, -v.x);
    else return mk<R>(-v.y, v.x);
  } else {
    const R c = R(cs[K]);
    const R s = INV ? R(sn[K]) : R(-sn[K]);
    return mk<R>(v.x * c - v.y * s, v.x * s + v.y * c);
  }
}

template <typename R, int N, bool INV>
struct DFT {
  template <int K>
  __device__ static inline void combine(cpx<R>* x, const cpx<R>* e, const cpx<R>* o) {
    if constexpr (K < N / 2) {
      const cpx<R> t = tw16<R, (16 / N) * K, INV>(o[K]);
      x[K] = e[K] + t;
      x[K + N / 2] = e[K] - t;
      combine<K + 1>(x, e, o);
    }
  }
  __device__ static inline void run(cpx<R>* x) {
    cpx<R> e[N / 2], o[N / 2];
#pragma unroll
    for (int i = 0; i < N / 2; ++i) {
      e[i] = x[2 * i];
      o[i] = x[2 * i + 1];
    }
    DFT<R, N / 2, INV>::run(e);
    DFT<R, N / 2, INV>::run(o);
    combine<0>(x, e, o);
  }
};
template <typename R, bool INV>
struct DFT<R, 1, INV> {
  __device__ static inline void run(cpx<R>*) {}
};

}  // namespace pf
