// func::evaluate through the C++ header API: every density of func/Densities.hpp (one fused
// device pass) gives the bits of the same formula written as a caller's lambda (host path),
// pairwise and Kahan; and the reference's functional tests (test/functional_test.cpp,
// test/pfc_functional_test.cpp, test/func_evaluate_vector_test.cpp) pass on both paths.
#include <cmath>
#include <vector>

#include <gridcalc/core/EigenAliases.hpp>
#include <gridcalc/core/Field.hpp>
#include <gridcalc/core/Grid.hpp>
#include <gridcalc/func/Densities.hpp>
#include <gridcalc/func/Functional.hpp>
#include <gridcalc/func/Integrate.hpp>

#include "check.hpp"

using gridcalc::core::Field;
using gridcalc::core::Grid;
using gridcalc::core::Mat3d;
using gridcalc::core::Vec3d;
namespace func = gridcalc::func;

namespace {

constexpr double kPi = 3.14159265358979323846;
double pi3() { return kPi * kPi * kPi; }

Grid box(int n) {
  const double h = 2.0 * kPi / n;
  return Grid(n, n, n, Vec3d(h, h, h));
}

double sss(double x, double y, double z) { return std::sin(x) * std::sin(y) * std::sin(z); }

// a field with every Fourier mode present, so no term vanishes by symmetry
Field<double> rough(const Grid& g) {
  return Field<double>(g, [](double x, double y, double z) {
    return 0.6 * std::sin(x + 0.3) * std::cos(2.0 * y) + 0.3 * std::sin(3.0 * z + x) + 0.1 * std::cos(y - z);
  });
}

}  // namespace

TEST(Densities, DeviceEqualsHostLambda) {
  for (const Grid& g : {box(24), Grid(30, 20, 1, Vec3d(0.4, 0.3, 1.0)), Grid(17, 13, 11, Vec3d(0.5, 0.7, 0.9))}) {
    const auto psi = rough(g);
    const double kappa = 0.7, W = 1.3, q0 = 1.1, eps = 0.1;
    EXPECT_EQ(func::evaluate(psi, func::ValueDensity{}), func::evaluate(psi, [](double p) { return p; }));
    EXPECT_EQ(func::evaluate(psi, func::GradientEnergyDensity{}),
              func::evaluate(psi, [](double, const Vec3d& gr) { return gr.squaredNorm(); }));
    EXPECT_EQ(func::evaluate(psi, func::CahnHilliardDensity{kappa}),
              func::evaluate(psi, [&](double p, const Vec3d& gr) {
                return -0.5 * p * p + 0.25 * p * p * p * p + 0.5 * kappa * gr.squaredNorm();
              }));
    auto gl = [&](double p, const Vec3d& gr, double) {
      const double q = p * p - 1.0;
      return 0.5 * gr.squaredNorm() + W * q * q;
    };
    EXPECT_EQ(func::evaluate(psi, func::GinzburgLandauDensity{W}), func::evaluate(psi, gl));
    EXPECT_EQ(func::evaluate(psi, func::GinzburgLandauDensity{W}, func::Kahan{}),
              func::evaluate(psi, gl, func::Kahan{}));
    if (g.getNz() > 1) {
      auto pfc = [&](double p, const Vec3d&, double lap, double bih) {
        const double q2 = q0 * q0, q4 = q2 * q2, p2 = p * p;
        return 0.5 * (q4 - eps) * p2 + q2 * p * lap + 0.5 * p * bih + 0.25 * p2 * p2;
      };
      EXPECT_EQ(func::evaluate(psi, func::PfcDensity{q0, eps}), func::evaluate(psi, pfc));
      EXPECT_EQ(func::evaluate(psi, func::PfcDensity{q0, eps}, func::Kahan{}),
                func::evaluate(psi, pfc, func::Kahan{}));
    }
  }
}

TEST(Densities, LinearElasticDeviceEqualsHostLambda) {
  const Grid g = box(20);
  Field<Vec3d> u(g, [](double x, double y, double z) {
    return Vec3d(0.01 * std::sin(x + y), 0.02 * std::cos(z), 0.005 * std::sin(x - 2.0 * z));
  });
  const double lambda = 1.7, mu = 0.6;
  auto f = [&](const Vec3d&, const Mat3d& G) {
    const Mat3d e = 0.5 * (G + G.transpose());
    const double tr = e.trace();
    const double ee = (e.array() * e.array()).sum();
    return 0.5 * (lambda * tr * tr + 2.0 * mu * ee);
  };
  EXPECT_EQ(func::evaluate(u, func::LinearElasticDensity{lambda, mu}), func::evaluate(u, f));
  // uniaxial dilation wave u_x = alpha sin x: 2 pi^3 (lambda + 2 mu) alpha^2 in the continuum;
  // the central difference scales the strain by sin(h)/h, so the discrete value is exactly
  // that times (sin h / h)^2
  const double alpha = 0.01;
  const Grid g64 = box(64);
  Field<Vec3d> w(g64, [&](double x, double, double) { return Vec3d(alpha * std::sin(x), 0.0, 0.0); });
  const double F = func::evaluate(w, func::LinearElasticDensity{lambda, mu});
  const double h = g64.getCellSize()(0);
  const double ref = 2.0 * pi3() * (lambda + 2.0 * mu) * alpha * alpha;
  const double discrete = ref * (std::sin(h) / h) * (std::sin(h) / h);
  EXPECT_LT(std::abs(F - discrete) / ref, 1e-12);
  EXPECT_LT(std::abs(F - ref) / ref, 4e-3);
}

TEST(Densities, FloatFieldsUseTheDevicePath) {
  const Grid g = box(32);
  const auto psi = rough(g);
  Field<float> p32(g, [](double x, double y, double z) {
    return static_cast<float>(0.6 * std::sin(x + 0.3) * std::cos(2.0 * y) + 0.3 * std::sin(3.0 * z + x) +
                              0.1 * std::cos(y - z));
  });
  const double a = func::evaluate(psi, func::CahnHilliardDensity{0.5});
  const double b = func::evaluate(p32, func::CahnHilliardDensity{0.5});
  EXPECT_LT(std::abs(a - b), 1e-6 * std::abs(a));
}

// ---- the reference's functional tests ----------------------------------------------------

TEST(ReferenceFunctional, GinzburgLandauHandComputed) {
  const Grid g = box(64);
  Field<double> psi(g, sss);
  const double F = func::evaluate(psi, func::GinzburgLandauDensity{1.0});
  const double ref = (3.0 / 2.0 + 411.0 / 64.0) * pi3();
  EXPECT_LT(std::abs(F - ref) / ref, 1e-2);
}

TEST(ReferenceFunctional, PfcHandComputedAndSecondOrder) {
  const double ref = (0.5 * 0.9 - 3.0 + 4.5 + 27.0 / 256.0) * pi3();
  std::vector<double> err;
  for (int n : {16, 32, 64}) {
    Field<double> psi(box(n), sss);
    const double F = func::evaluate(psi, func::PfcDensity{1.0, 0.1});
    err.push_back(std::abs(F - ref) / std::abs(ref));
  }
  EXPECT_LT(err[2], 2e-2);
  for (int i = 1; i < 3; ++i) {
    EXPECT_GT(err[i - 1] / err[i], 3.5);
    EXPECT_LT(err[i - 1] / err[i], 4.5);
  }
}

TEST(ReferenceFunctional, ArityDispatch) {
  Field<double> psi(box(64), sss);
  EXPECT_NEAR(func::evaluate(psi, [](double p) { return p * p; }), pi3(), 1e-12);
  const double F3 = func::evaluate(psi, [](double p, const Vec3d&, double l) { return p * l; });
  EXPECT_LT(std::abs(F3 + 3.0 * pi3()) / (3.0 * pi3()), 1e-2);
  const double F2 = func::evaluate(psi, [](double, const Vec3d& g) { return 0.5 * g.squaredNorm(); });
  EXPECT_LT(std::abs(F2 - 1.5 * pi3()) / (1.5 * pi3()), 1e-2);
}

TEST(ReferenceFunctional, PairwiseVsKahan) {
  Field<double> psi(box(64), sss);
  const double a = func::evaluate(psi, func::GinzburgLandauDensity{1.0}, func::Pairwise{});
  const double b = func::evaluate(psi, func::GinzburgLandauDensity{1.0}, func::Kahan{});
  EXPECT_LT(std::abs(a - b) / std::max(std::abs(a), std::abs(b)), 1e-13);
}

CHECK_MAIN()
