// Exercises the C++ binding (include/purify_amd.hpp) the way the reference's tests use the operator
// (cpp/tests/operators.cc:142-154): *op * x, op->adjoint() * y, construction errors as runtime_error.
// Usage: test_wrapper cpu   -- error paths only (no GPU needed)
//        test_wrapper gpu   -- also applies and checks the adjoint dot product in FP64
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>

#include "purify_amd.hpp"

using namespace purify_amd;

static int failures = 0;
#define EXPECT(cond)                                                   \
  do {                                                                 \
    if (!(cond)) {                                                     \
      std::fprintf(stderr, "%s:%d: failed: %s\n", __FILE__, __LINE__, #cond); \
      ++failures;                                                      \
    }                                                                  \
  } while (0)

template <class F>
static std::string error_of(F &&f) {
  try {
    f();
  } catch (const std::runtime_error &e) {
    return e.what();
  }
  return "";
}

int main(int argc, char **argv) {
  const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
  std::vector<double> u(4, 0.0), v(4, 0.0), w;
  // kernels.cc:241-244 and operators.cc:18-20 messages
  EXPECT(error_of([&] {
           measurementoperator::init_degrid_operator_2d(u, v, w, {}, 16, 16, 2, 0, 1e-4, kernels::kernel::pswf, 4, 4);
         }).find("PSWF requires a support of 6") != std::string::npos);
  EXPECT(error_of([&] {
           measurementoperator::init_degrid_operator_2d(u, std::vector<double>(3), w, {}, 16, 16);
         }).find("Size of u and v") != std::string::npos);
  if (gpu) {
    const int n = 64, M = 3000;
    std::mt19937_64 rng(5);
    std::normal_distribution<double> nd(0.0, 30.0);
    std::vector<double> uu(M), vv(M);
    for (int m = 0; m < M; ++m) {
      uu[m] = nd(rng);
      vv[m] = nd(rng);
    }
    auto op = measurementoperator::init_degrid_operator_2d(uu, vv, w, {}, n, n, 2, 100, 1e-4, kernels::kernel::kb, 4, 4,
                                                           false, 6, 1, 1, PF_C128);
    std::vector<t_complex> x(n * n), y(M);
    for (auto &a : x) a = t_complex(nd(rng), nd(rng));
    for (auto &a : y) a = t_complex(nd(rng), nd(rng));
    const std::vector<t_complex> Px = *op * x;
    const std::vector<t_complex> Ay = op->adjoint() * y;
    EXPECT((int)Px.size() == M && (int)Ay.size() == n * n);
    t_complex lhs = 0, rhs = 0;
    for (int m = 0; m < M; ++m) lhs += std::conj(y[m]) * Px[m];
    for (int k = 0; k < n * n; ++k) rhs += std::conj(Ay[k]) * x[k];
    EXPECT(std::abs(lhs - rhs) / std::abs(lhs) < 1e-12);
    // the normalised operator has norm 1 (operators.cc:147-154): ||Phi x|| <= ||x|| for any x
    double nx = 0, ny = 0;
    for (auto &a : x) nx += std::norm(a);
    for (auto &a : Px) ny += std::norm(a);
    EXPECT(ny <= nx * (1 + 1e-3));
    EXPECT(op->sizes()[2] == M && op->adjoint().sizes()[2] == n * n);
  }
  if (failures) return 1;
  std::printf("ok\n");
  return 0;
}
