// Tensor contractions and the tensor::expr layer through the C++ header API, on the device
// (test/tensor_expressions_test.cpp, test/tensor_grid_mismatch_test.cpp and
// test/vector_tensor_test.cpp:346-420 of the reference): every device-materialised node and
// composite expression equals the host evalAt of the same tree bit for bit, the eager
// contractions equal their one-node expressions, integrate(expr) equals
// integrate(materialize(expr)), and mismatched grids raise the reference's messages.
#include <cmath>
#include <cstring>

#include <gridcalc/core/EigenAliases.hpp>
#include <gridcalc/core/Field.hpp>
#include <gridcalc/core/Grid.hpp>
#include <gridcalc/diff/Gradient.hpp>
#include <gridcalc/func/Densities.hpp>
#include <gridcalc/func/Functional.hpp>
#include <gridcalc/func/Integrate.hpp>
#include <gridcalc/tensor/Contraction.hpp>
#include <gridcalc/tensor/Expressions.hpp>

#include "check.hpp"

using gridcalc::core::Field;
using gridcalc::core::Grid;
using gridcalc::core::Mat3d;
using gridcalc::core::Vec3d;
namespace expr = gridcalc::tensor::expr;
namespace tensor = gridcalc::tensor;
namespace func = gridcalc::func;
namespace diff = gridcalc::diff;

namespace {

constexpr double kPi = 3.14159265358979323846;

Grid box(int n) {
  const double h = 2.0 * kPi / n;
  return Grid(n, n, n, Vec3d(h, h, h));
}

Mat3d matA(double x, double y, double z) {
  Mat3d m;
  m << std::sin(x), std::cos(y), std::sin(z), std::cos(x) * std::sin(y), std::sin(2.0 * x),
      std::cos(z), std::sin(x) * std::cos(z), std::cos(x) * std::sin(z), std::sin(x + y + z);
  return m;
}
Mat3d matB(double x, double y, double z) {
  Mat3d m;
  m << std::cos(x), std::sin(2.0 * y), std::cos(z + x), std::sin(x) * std::cos(y),
      std::cos(2.0 * z), std::sin(y), std::cos(y) * std::sin(z), std::sin(x) * std::cos(y),
      std::cos(x + y);
  return m;
}

template <class T>
bool bitEqual(const T& a, const T& b) {
  return std::memcmp(&a, &b, sizeof(T)) == 0;
}

// every site of the device materialisation equals the host evalAt of the same tree
template <class E>
bool deviceMatchesHost(const E& e) {
  const auto dev = expr::materialize(e);
  const Grid& g = e.grid();
  for (int k = 0; k < g.getNz(); ++k)
    for (int j = 0; j < g.getNy(); ++j)
      for (int i = 0; i < g.getNx(); ++i)
        if (!bitEqual(dev(i, j, k), e.evalAt(i, j, k))) return false;
  return true;
}

}  // namespace

TEST(TensorExpressions, EveryNodeMatchesHostEvaluation) {
  const Grid g = box(12);
  Field<Mat3d> A(g, matA), B(g, matB);
  Field<double> s(g, [](double x, double y, double z) { return std::sin(x) * std::cos(y - z); });
  Field<Vec3d> v(g, [](double x, double y, double z) { return Vec3d(std::sin(x), std::cos(y), x * z); });
  EXPECT_TRUE(deviceMatchesHost(expr::field(A)));
  EXPECT_TRUE(deviceMatchesHost(expr::identityField(g)));
  EXPECT_TRUE(deviceMatchesHost(2.5 * expr::field(A)));
  EXPECT_TRUE(deviceMatchesHost(expr::field(A) * 0.3));
  EXPECT_TRUE(deviceMatchesHost(expr::field(A) / 3.0));
  EXPECT_TRUE(deviceMatchesHost(-expr::field(A)));
  EXPECT_TRUE(deviceMatchesHost(expr::field(A) + expr::field(B)));
  EXPECT_TRUE(deviceMatchesHost(expr::field(A) - expr::field(B)));
  EXPECT_TRUE(deviceMatchesHost(expr::field(v) + 2.0 * expr::field(v)));
  EXPECT_TRUE(deviceMatchesHost(expr::field(s) * expr::field(s)));
  EXPECT_TRUE(deviceMatchesHost(expr::trace(expr::field(A))));
  EXPECT_TRUE(deviceMatchesHost(expr::sym(expr::field(A))));
  EXPECT_TRUE(deviceMatchesHost(expr::singleContract(expr::field(A), expr::field(B))));
  EXPECT_TRUE(deviceMatchesHost(expr::doubleContract(expr::field(A), expr::field(B))));
  // composites (tensor_expressions_test.cpp:231-335)
  EXPECT_TRUE(deviceMatchesHost(expr::trace(expr::sym(expr::field(A)))));
  EXPECT_TRUE(deviceMatchesHost(0.5 * expr::doubleContract(expr::field(A), expr::field(A)) -
                                2.0 * expr::doubleContract(expr::field(A), expr::field(B))));
  EXPECT_TRUE(deviceMatchesHost(expr::doubleContract(expr::identityField(g), expr::field(A))));
  EXPECT_TRUE(deviceMatchesHost(expr::singleContract(expr::identityField(g), expr::field(A))));
  EXPECT_TRUE(deviceMatchesHost(expr::trace(expr::field(A)) * expr::trace(expr::field(B)) +
                                expr::field(s)));
}

TEST(TensorExpressions, SymAndIdentityKnownValues) {
  const Grid g = box(6);
  Field<Mat3d> A(g, matA);
  const auto S = expr::materialize(expr::sym(expr::field(A)));
  const auto tI = expr::materialize(expr::trace(expr::identityField(g)));
  const auto AI = expr::materialize(expr::singleContract(expr::field(A), expr::identityField(g)));
  for (int k = 0; k < 6; ++k)
    for (int j = 0; j < 6; ++j)
      for (int i = 0; i < 6; ++i) {
        const Mat3d a = A(i, j, k);
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) EXPECT_EQ(S(i, j, k)(r, c), 0.5 * (a(r, c) + a(c, r)));
        EXPECT_EQ(tI(i, j, k), 3.0);
        // A . I = A in value (products by 1 and 0; a -0 entry plus +0 becomes +0)
        for (int q = 0; q < 9; ++q) EXPECT_EQ(AI(i, j, k)(q), a(q));
      }
}

TEST(TensorContraction, EagerEqualsExpression) {
  const Grid g = box(10);
  Field<Mat3d> A(g, matA), B(g, matB);
  const auto t = tensor::trace(A);
  const auto sc = tensor::singleContract(A, B);
  const auto dc = tensor::doubleContract(A, B);
  const auto te = expr::materialize(expr::trace(expr::field(A)));
  const auto sce = expr::materialize(expr::singleContract(expr::field(A), expr::field(B)));
  const auto dce = expr::materialize(expr::doubleContract(expr::field(A), expr::field(B)));
  EXPECT_TRUE(std::memcmp(t.data(), te.data(), t.getSize() * sizeof(double)) == 0);
  EXPECT_TRUE(std::memcmp(sc.data(), sce.data(), sc.getSize() * sizeof(Mat3d)) == 0);
  EXPECT_TRUE(std::memcmp(dc.data(), dce.data(), dc.getSize() * sizeof(double)) == 0);
  // A:A = |A|_F^2 (vector_tensor_test.cpp:402-420); A:B = tr(A^T B)
  const auto ff = tensor::doubleContract(A, A);
  for (std::size_t c = 0; c < ff.getSize(); ++c) {
    const Mat3d a = A.data()[c], b = B.data()[c];
    EXPECT_NEAR(ff.data()[c], a.squaredNorm(), 1e-14);
    EXPECT_NEAR(dc.data()[c], (a.transpose() * b).trace(), 1e-14);
  }
}

TEST(TensorExpressions, IntegrateExpression) {
  const Grid g = box(16);
  Field<Mat3d> A(g, matA), B(g, matB);
  const auto e = expr::doubleContract(expr::field(A), expr::field(B));
  EXPECT_EQ(func::integrate(e), func::integrate(expr::materialize(e)));
  EXPECT_EQ(func::integrate(e, func::Kahan{}), func::integrate(expr::materialize(e), func::Kahan{}));
}

TEST(TensorExpressions, ElasticEnergyThroughExpressions) {
  // elastic_energy_test.cpp:159-185: the Lame closed form through the expression layer agrees
  // with the fused density (the two associate the products differently)
  const Grid g = box(24);
  Field<Vec3d> u(g, [](double x, double y, double z) {
    return Vec3d(0.01 * std::sin(x + y), 0.02 * std::cos(z), 0.005 * std::sin(x - 2.0 * z));
  });
  const double lambda = 1.7, mu = 0.6;
  const auto G = diff::gradient(u);
  const auto eps = expr::sym(expr::field(G));
  const double F_et = func::integrate(0.5 * (lambda * (expr::trace(eps) * expr::trace(eps)) +
                                             2.0 * mu * expr::doubleContract(eps, eps)));
  const double F_d = func::evaluate(u, func::LinearElasticDensity{lambda, mu});
  EXPECT_LT(std::abs(F_et - F_d), 1e-13 * std::abs(F_d));
}

TEST(TensorGridMismatch, MessagesNameTheNode) {
  const Grid g = box(8);
  const Grid small(4, 4, 4, Vec3d(0.1, 0.1, 0.1));
  const Grid reshaped(4, 16, 8, g.getCellSize());
  const Grid spaced(8, 8, 8, Vec3d(0.2, 0.2, 0.2));
  Field<Mat3d> A(g, matA);
  Field<Mat3d> Bs(small, matB), Br(reshaped, matB), Bh(spaced, matB);
  Field<double> s(g, 1.0), ss(small, 1.0);
  EXPECT_THROW_MSG(tensor::singleContract(A, Bs), std::invalid_argument, "tensor::singleContract");
  EXPECT_THROW_MSG(tensor::singleContract(A, Br), std::invalid_argument, "tensor::singleContract");
  EXPECT_THROW_MSG(tensor::doubleContract(A, Bh), std::invalid_argument, "tensor::doubleContract");
  EXPECT_THROW_MSG(expr::field(A) + expr::field(Bs), std::invalid_argument, "tensor::expr::Plus");
  EXPECT_THROW_MSG(expr::field(A) - expr::field(Br), std::invalid_argument, "tensor::expr::Minus");
  EXPECT_THROW_MSG(expr::field(s) * expr::field(ss), std::invalid_argument, "tensor::expr::Mul");
  EXPECT_THROW_MSG(expr::singleContract(expr::field(A), expr::field(Bs)), std::invalid_argument,
                   "tensor::expr::SingleContract");
  EXPECT_THROW_MSG(expr::doubleContract(expr::field(A), expr::field(Bh)), std::invalid_argument,
                   "tensor::expr::DoubleContract");
  EXPECT_THROW_MSG(expr::field(A) + expr::identityField(small), std::invalid_argument,
                   "tensor::expr::Plus");
}

CHECK_MAIN()
