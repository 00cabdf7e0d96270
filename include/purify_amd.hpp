// purify_amd.hpp -- C++ host binding of the C-ABI in purify_amd.h: the reference's entry point
// measurementoperator::init_degrid_operator_2d (cpp/purify/operators.h:373-414) returning a
// LinearTransform whose direct and adjoint closures call pf_op_forward / pf_op_adjoint.
//
// LinearTransform and OperatorFunction follow the shape of SOPT's (used at operators.h:393-395):
// `*op * x` applies Phi, `op->adjoint() * y` applies Phi^dagger, and the callee sizes the output.  The
// vector type T is any contiguous container of std::complex<double> with data(), size() and
// resize() (std::vector, or Eigen's Vector<t_complex> where Eigen is available).  Construction
// errors surface as std::runtime_error carrying the reference's messages.
#ifndef PURIFY_AMD_HPP
#define PURIFY_AMD_HPP

#include <array>
#include <complex>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "purify_amd.h"

namespace purify_amd {

using t_complex = std::complex<double>;
using t_real = double;
using t_uint = unsigned int;

inline void check(int status) {
  if (status != PF_OK) throw std::runtime_error(pf_last_error());
}

// sopt::OperatorFunction<T>: output first, input second.
template <class T>
using OperatorFunction = std::function<void(T &, const T &)>;

// The direct / adjoint pair of sopt::LinearTransform<T>.  sizes {a, b, c}: the output length is
// a * N / b + c for an input of length N; the operators here use {0, 1, length}.
template <class T>
class LinearTransform {
 public:
  LinearTransform(OperatorFunction<T> direct, std::array<std::int64_t, 3> dsizes, OperatorFunction<T> indirect,
                  std::array<std::int64_t, 3> isizes)
      : direct_(std::move(direct)), indirect_(std::move(indirect)), dsizes_(dsizes), isizes_(isizes) {}
  T operator*(const T &x) const {
    T out;
    direct_(out, x);
    return out;
  }
  LinearTransform adjoint() const { return LinearTransform(indirect_, isizes_, direct_, dsizes_); }
  const std::array<std::int64_t, 3> &sizes() const { return dsizes_; }
  const std::array<std::int64_t, 3> &adjoint_sizes() const { return isizes_; }

 private:
  OperatorFunction<T> direct_, indirect_;
  std::array<std::int64_t, 3> dsizes_, isizes_;
};

namespace kernels {
// kernels::kernel (cpp/purify/kernels.h:15), same order as pf_kernel
enum class kernel { kb = PF_KB, gauss = PF_GAUSS, box = PF_BOX, pswf = PF_PSWF, kbmin = PF_KBMIN, gauss_alt = PF_GAUSS_ALT };
}  // namespace kernels

namespace measurementoperator {

// init_degrid_operator_2d(u, v, w, weights, imsizey, imsizex, oversample_ratio, power_iters, power_tol,
// kernel, Ju, Jv, w_term, Jw, cellx, celly): u, v in pixels (operators.h:373-396).  The extra
// `precision` selects the complex128 operator (the default: the reference's Vector<t_complex> is double
// precision, about 1e-15 relative error) or, with PF_C64, the complex64 one (about 1e-7 relative error,
// twice the speed); applies take and return host complex128 vectors either way.
template <class T = std::vector<t_complex>>
std::shared_ptr<LinearTransform<T> const> init_degrid_operator_2d(
    const std::vector<t_real> &u, const std::vector<t_real> &v, const std::vector<t_real> &w,
    const std::vector<t_complex> &weights, const t_uint &imsizey, const t_uint &imsizex,
    const t_real &oversample_ratio = 2, const t_uint &power_iters = 100, const t_real &power_tol = 1e-4,
    const kernels::kernel kernel = kernels::kernel::kb, const t_uint Ju = 4, const t_uint Jv = 4,
    const bool w_term = false, const t_uint Jw = 6, const t_real &cellx = 1, const t_real &celly = 1,
    const int precision = PF_C128) {
  if (u.size() != v.size())
    throw std::runtime_error("Size of u and v vectors are not the same for creating gridding matrix.");
  if (w_term && u.size() != w.size())
    throw std::runtime_error("Size of u and w vectors are not the same for creating gridding matrix.");
  if (!weights.empty() && u.size() != weights.size())
    throw std::runtime_error("Size of u and w vectors are not the same for creating gridding matrix.");
  pf_op_desc d;
  pf_op_desc_init(&d);
  d.imsizey = (int32_t)imsizey;
  d.imsizex = (int32_t)imsizex;
  d.oversample_ratio = oversample_ratio;
  d.kernel = (int32_t)kernel;
  d.Ju = (int32_t)Ju;
  d.Jv = (int32_t)Jv;
  d.w_term = w_term ? 1 : 0;
  d.Jw = (int32_t)Jw;
  d.cellx = cellx;
  d.celly = celly;
  d.units = PF_PIXELS;
  d.precision = precision;
  d.M = (int64_t)u.size();
  d.u = u.data();
  d.v = v.data();
  d.w = w_term ? w.data() : nullptr;
  d.weights = weights.empty() ? nullptr : reinterpret_cast<const double *>(weights.data());
  d.power_iters = (int32_t)power_iters;
  d.power_tol = power_tol;
  pf_op *raw = nullptr;
  check(pf_op_create(&d, &raw));
  std::shared_ptr<pf_op> op(raw, [](pf_op *p) { pf_op_destroy(p); });
  const std::int64_t M = d.M, N = (std::int64_t)imsizey * imsizex;
  auto direct = [op, M, N](T &out, const T &in) {
    if ((std::int64_t)in.size() != N) throw std::runtime_error("input image has the wrong size");
    out.resize(M);
    check(pf_op_forward(op.get(), in.data(), out.data(), PF_HOST, PF_HOST, nullptr));
  };
  auto indirect = [op, M, N](T &out, const T &in) {
    if ((std::int64_t)in.size() != M) throw std::runtime_error("input visibilities have the wrong size");
    out.resize(N);
    check(pf_op_adjoint(op.get(), in.data(), out.data(), PF_HOST, PF_HOST, nullptr));
  };
  return std::make_shared<LinearTransform<T> const>(direct, std::array<std::int64_t, 3>{0, 1, M}, indirect,
                                                     std::array<std::int64_t, 3>{0, 1, N});
}

}  // namespace measurementoperator
}  // namespace purify_amd

#endif
