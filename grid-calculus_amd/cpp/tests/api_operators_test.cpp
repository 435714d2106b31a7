// Operators, Field residency, integrate() and evaluate() through the C++ header API, on
// the device.  Closed forms use the exact discrete symbols of the stencils on grid
// Fourier modes (so the tolerances are rounding-level), plus the continuum limits and
// known answers the reference's own tests check.
#include <algorithm>
#include <cmath>
#include <vector>

#include <gridcalc/core/EigenAliases.hpp>
#include <gridcalc/core/Field.hpp>
#include <gridcalc/core/Grid.hpp>
#include <gridcalc/diff/Biharmonic.hpp>
#include <gridcalc/diff/Divergence.hpp>
#include <gridcalc/diff/Gradient.hpp>
#include <gridcalc/diff/Laplacian.hpp>
#include <gridcalc/func/Functional.hpp>
#include <gridcalc/func/Integrate.hpp>
#include <gridcalc/fvm/CellLaplacian.hpp>

#include "check.hpp"

using gridcalc::core::Field;
using gridcalc::core::Grid;
using gridcalc::core::Mat3d;
using gridcalc::core::Vec3d;
namespace diff = gridcalc::diff;
namespace func = gridcalc::func;
namespace fvm = gridcalc::fvm;

namespace {

constexpr double kPi = 3.14159265358979323846;

Grid cube(int n) {
  const double h = 2.0 * kPi / n;
  return Grid(n, n, n, Vec3d(h, h, h));
}

template <class A, class B>
double maxDiff(const Field<A>& a, const Field<B>& b) {
  double m = 0.0;
  for (std::size_t c = 0; c < a.getSize(); ++c) m = std::max(m, std::abs(a.data()[c] - b.data()[c]));
  return m;
}

double sinsinsin(double x, double y, double z) { return std::sin(x) * std::sin(y) * std::sin(z); }

}  // namespace

TEST(Laplacian, DiscreteSymbolSecondOrder) {
  const int n = 32;
  const Grid g = cube(n);
  const double h = g.getCellSize()(0);
  Field<double> f(g, sinsinsin);
  const auto L = diff::laplacian(f);
  const double lam = 3.0 * (2.0 * std::cos(h) - 2.0) / (h * h);
  Field<double> want(g, [&](double x, double y, double z) { return lam * sinsinsin(x, y, z); });
  EXPECT_LT(maxDiff(L, want), 1e-12);
  Field<double> cont(g, [](double x, double y, double z) { return -3.0 * sinsinsin(x, y, z); });
  EXPECT_LT(maxDiff(L, cont), 1e-2);
}

TEST(Laplacian, FourthOrderConvergence) {
  std::vector<double> err;
  for (int n : {16, 32, 64}) {
    const Grid g = cube(n);
    Field<double> f(g, sinsinsin);
    const auto L = diff::laplacian<4>(f);
    Field<double> cont(g, [](double x, double y, double z) { return -3.0 * sinsinsin(x, y, z); });
    err.push_back(maxDiff(L, cont));
  }
  const double slope1 = std::log2(err[0] / err[1]), slope2 = std::log2(err[1] / err[2]);
  EXPECT_GT(slope1, 3.7);
  EXPECT_GT(slope2, 3.7);
  EXPECT_LT(slope2, 4.3);
}

TEST(Laplacian, FloatOverload) {
  const Grid g(48, 40, 1, Vec3d(0.5, 0.25, 1.0));
  Field<double> f(g, [](double x, double y, double) { return std::sin(0.7 * x) * std::cos(1.3 * y); });
  Field<float> f32(g);
  for (std::size_t c = 0; c < g.getSize(); ++c) f32.data()[c] = static_cast<float>(f.data()[c]);
  const auto L64 = diff::laplacian(f);
  const auto L32 = diff::laplacian(f32);
  double m = 0.0, s = 0.0;
  for (std::size_t c = 0; c < g.getSize(); ++c) {
    m = std::max(m, std::abs(double(L32.data()[c]) - L64.data()[c]));
    s = std::max(s, std::abs(L64.data()[c]));
  }
  EXPECT_LT(m / s, 1e-4);
}

TEST(Gradient, DiscreteSymbolBothOrders) {
  const int n = 24;
  const Grid g = cube(n);
  const double h = g.getCellSize()(0);
  Field<double> f(g, [](double x, double y, double z) { return std::sin(x + 2.0 * y - z); });
  const auto g2 = diff::gradient(f);
  const auto g4 = diff::gradient<4>(f);
  const int kk[3] = {1, 2, -1};
  double e2 = 0.0, e4 = 0.0;
  const auto& hs = g.getCellSize();
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const double c = std::cos(i * hs(0) + 2.0 * j * hs(1) - k * hs(2));
        for (int a = 0; a < 3; ++a) {
          const double th = kk[a] * h;
          const double s2 = std::sin(th) / h;
          const double s4 = (4.0 / 3.0 * std::sin(th) - 1.0 / 6.0 * std::sin(2.0 * th)) / h;
          e2 = std::max(e2, std::abs(g2(i, j, k)(a) - s2 * c));
          e4 = std::max(e4, std::abs(g4(i, j, k)(a) - s4 * c));
        }
      }
  EXPECT_LT(e2, 1e-12);
  EXPECT_LT(e4, 1e-12);
}

TEST(Divergence, OfGradientApproximatesLaplacian) {
  const Grid g = cube(32);
  Field<double> f(g, sinsinsin);
  const auto dg = diff::divergence(diff::gradient(f));
  Field<double> cont(g, [](double x, double y, double z) { return -3.0 * sinsinsin(x, y, z); });
  // the wide (2h) div-of-grad stencil differs from -3f by ~h^2 = 0.0385 at N = 32
  EXPECT_LT(maxDiff(dg, cont), 5e-2);
}

TEST(VectorTensor, GradientConventionAndRowDivergence) {
  const Grid g = cube(32);
  // v = (sin y, 0, 0): M(0, 1) = d_y v_0 = cos y is the only non-zero entry
  Field<Vec3d> v(g, [](double, double y, double) { return Vec3d(std::sin(y), 0.0, 0.0); });
  const auto M = diff::gradient(v);
  const auto& Mij = M(3, 5, 7);
  const double y = 5 * g.getCellSize()(1);
  EXPECT_NEAR(Mij(0, 1), std::cos(y), 1e-2);
  EXPECT_NEAR(Mij(0, 0), 0.0, 1e-14);
  EXPECT_NEAR(Mij(1, 1), 0.0, 1e-14);
  // div of M(x,y,z) = diag(a x, b y, c z) is (a, b, c) away from the periodic seam
  Field<Mat3d> D(g, [](double x, double y, double z) {
    Mat3d m = Mat3d::Zero();
    m(0, 0) = 2.0 * x;
    m(1, 1) = 3.0 * y;
    m(2, 2) = -1.0 * z;
    return m;
  });
  const auto dv = diff::divergence(D);
  const Vec3d c = dv(10, 11, 12);
  EXPECT_NEAR(c(0), 2.0, 1e-12);
  EXPECT_NEAR(c(1), 3.0, 1e-12);
  EXPECT_NEAR(c(2), -1.0, 1e-12);
}

TEST(Biharmonic, DiscreteSymbol) {
  const int n = 24;
  const Grid g = cube(n);
  const double h = g.getCellSize()(0);
  Field<double> f(g, sinsinsin);
  const auto B = diff::biharmonic(f);
  const double d2 = (2.0 * std::cos(h) - 2.0) / (h * h);
  const double d4 = (2.0 * std::cos(2.0 * h) - 8.0 * std::cos(h) + 6.0) / (h * h * h * h);
  const double lam = 3.0 * d4 + 2.0 * 3.0 * d2 * d2;
  Field<double> want(g, [&](double x, double y, double z) { return lam * sinsinsin(x, y, z); });
  EXPECT_LT(maxDiff(B, want), 1e-10);
  const auto B4 = diff::d4<4>(f);
  Field<double> cont(g, [](double x, double y, double z) { return 9.0 * sinsinsin(x, y, z); });
  EXPECT_LT(maxDiff(B4, cont), 2e-3);
}

TEST(Extent, AliasingAxisRejected) {
  const Grid g(4, 16, 16, Vec3d(1.0, 1.0, 1.0));
  Field<double> f(g, 1.0);
  EXPECT_THROW_MSG(diff::laplacian<4>(f), std::invalid_argument, "axis x");
  const Grid g2(16, 16, 5, Vec3d(1.0, 1.0, 1.0));
  Field<double> f2(g2, 1.0);
  EXPECT_THROW_MSG(diff::biharmonic<4>(f2), std::invalid_argument, "axis z");
  // degenerate axis accepted
  const Grid g3(16, 16, 1, Vec3d(1.0, 1.0, 1.0));
  Field<double> f3(g3, 1.0);
  EXPECT_LT(maxDiff(diff::laplacian<4>(f3), Field<double>(g3, 0.0)), 1e-15);
}

TEST(CellLaplacian, HarmonicMeanKnownAnswer) {
  const Grid g(4, 1, 1, Vec3d(1.0, 1.0, 1.0));
  Field<double> psi(g, 0.0), D(g, 0.0);
  const double p[4] = {0.0, 1.0, 1.0, 0.0}, d[4] = {1.0, 4.0, 4.0, 1.0};
  for (int i = 0; i < 4; ++i) {
    psi(i, 0, 0) = p[i];
    D(i, 0, 0) = d[i];
  }
  const auto r = fvm::cellLaplacian(psi, D);
  EXPECT_EQ(r(0, 0, 0), 1.6);
  EXPECT_EQ(r(1, 0, 0), -1.6);
  EXPECT_EQ(r(2, 0, 0), -1.6);
  EXPECT_EQ(r(3, 0, 0), 1.6);
}

TEST(CellLaplacian, ConstantDIsScaledLaplacian) {
  const Grid g(12, 10, 8, Vec3d(0.3, 0.2, 0.4));
  Field<double> psi(g, [](double x, double y, double z) { return std::sin(x) + std::cos(2.0 * y) * z; });
  const auto r = fvm::cellLaplacian(psi, Field<double>(g, 2.5));
  const auto L = diff::laplacian(psi);
  double m = 0.0, s = 0.0;
  for (std::size_t c = 0; c < g.getSize(); ++c) {
    m = std::max(m, std::abs(r.data()[c] - 2.5 * L.data()[c]));
    s = std::max(s, std::abs(r.data()[c]));
  }
  EXPECT_LT(m, 1e-12 * s);
}

TEST(CellLaplacian, PreconditionMessages) {
  const Grid g(8, 8, 8, Vec3d(1.0, 1.0, 1.0));
  Field<double> psi(g, 0.0), D(g, 1.0);
  D(5, 0, 0) = -0.25;
  EXPECT_THROW_MSG(fvm::cellLaplacian(psi, D), std::invalid_argument, "flat index 5 with D=-0.25");
  const Grid other(8, 8, 8, Vec3d(0.5, 1.0, 1.0));
  EXPECT_THROW_MSG(fvm::cellLaplacian(psi, Field<double>(other, 1.0)), std::invalid_argument,
                   "fvm::cellLaplacian: D");
}

TEST(Field, HostDeviceCoherence) {
  const Grid g(16, 16, 16, Vec3d(1.0, 1.0, 1.0));
  Field<double> f(g, 0.0);
  f(3, 4, 5) = 1.0;                      // host write
  auto L = diff::laplacian(f);           // device reads the uploaded copy
  EXPECT_EQ(L(3, 4, 5), -6.0);           // host read downloads
  L(3, 4, 5) = 100.0;                    // host write invalidates the device copy
  auto L2 = diff::laplacian(L);
  // the six neighbours of (3,4,5) hold 1: per axis 1 - 2*100 + 1 = -198
  EXPECT_EQ(L2(3, 4, 5), -594.0);
  Field<double> copy = L;                // copy keeps both views consistent
  EXPECT_EQ(copy(3, 4, 5), 100.0);
  Field<double> moved = std::move(copy);
  EXPECT_EQ(moved(3, 4, 5), 100.0);
  EXPECT_EQ(f(-1 + 4, 4, 5 + 16), 1.0);  // periodic wrap on host access
}

TEST(Integrate, KnownAnswersAndDeterminism) {
  const Grid g(8, 16, 4, Vec3d(0.25, 0.125, 0.5));
  Field<double> ones(g, 1.0);
  const double vol = (8 * 0.25) * (16 * 0.125) * (4 * 0.5);
  EXPECT_EQ(func::integrate(ones), vol);
  EXPECT_EQ(func::integrate(ones, func::Kahan{}), vol);
  const Grid c = cube(64);
  Field<double> f(c, [](double x, double y, double) { return 1.0 + std::sin(x) * std::cos(y); });
  const double p = func::integrate(f, func::Pairwise{}), k = func::integrate(f, func::Kahan{});
  EXPECT_LT(std::abs(p - k) / std::abs(p), 1e-13);
  EXPECT_EQ(p, func::integrate(f, func::Pairwise{}));
  EXPECT_NEAR(p, 8.0 * kPi * kPi * kPi, 1e-9);
}

TEST(Functional, AritiesAndFusedFreeEnergy) {
  const int n = 32;
  const Grid g = cube(n);
  Field<double> psi(g, [](double x, double, double) { return 0.5 * std::sin(x); });
  const double V = std::pow(2.0 * kPi, 3);
  // arity 1: integral of p^2 = V * 0.25/2
  const double I1 = func::evaluate(psi, [](double p) { return p * p; });
  EXPECT_NEAR(I1, V * 0.125, 1e-10);
  // arity 2, generic host density vs the fused Cahn-Hilliard kernel: identical
  const double kappa = 0.75;
  const double Fh = func::evaluate(psi, [kappa](double p, const Vec3d& gr) {
    return -0.5 * p * p + 0.25 * p * p * p * p + 0.5 * kappa * gr.squaredNorm();
  });
  const double Ff = func::evaluate(psi, func::CahnHilliardDensity{kappa});
  EXPECT_EQ(Fh, Ff);
  // arity 3 and 4 run: p * lap p integrates to -|grad p|^2 (summation by parts, exact
  // for the discrete symbols up to rounding at O2: sum p lap p = -sum |D+ p|^2)
  const double I3 = func::evaluate(psi, [](double p, const Vec3d&, double lap) { return p * lap; });
  EXPECT_LT(I3, 0.0);
  const double I4 = func::evaluate(psi, [](double p, const Vec3d&, double, double b) { return p * b; });
  EXPECT_GT(I4, 0.0);
}

CHECK_MAIN()
