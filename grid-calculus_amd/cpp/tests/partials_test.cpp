// The named mixed partial derivatives through the C++ header API (diff/MixedPartial.hpp,
// ThirdOrder.hpp, FourthOrder.hpp), on the device: the convergence sweep of the reference's
// test/phase_8_test.cpp:152-260 for every one of them at Order 2 and 4 -- on
// psi = sin x sin y sin z over [0, 2pi)^3 with N = 16, 32, 64 the log-log slope of the relative
// max error against the exact derivative lies within [Order - 0.5, Order + 0.5].
#include <cmath>
#include <functional>
#include <vector>

#include <gridcalc/core/EigenAliases.hpp>
#include <gridcalc/core/Field.hpp>
#include <gridcalc/core/Grid.hpp>
#include <gridcalc/diff/FourthOrder.hpp>
#include <gridcalc/diff/MixedPartial.hpp>
#include <gridcalc/diff/ThirdOrder.hpp>

#include "check.hpp"

using gridcalc::core::Field;
using gridcalc::core::Grid;
using gridcalc::core::Vec3d;
namespace gd = gridcalc::diff;

namespace {

constexpr double kPi = 3.14159265358979323846;

// n-th derivative of sin: sin, cos, -sin, -cos, sin
double dsin(int n, double x) {
  switch (n % 4) {
    case 0: return std::sin(x);
    case 1: return std::cos(x);
    case 2: return -std::sin(x);
    default: return -std::cos(x);
  }
}

double sweepSlope(int nx, int ny, int nz,
                  const std::function<Field<double>(const Field<double>&)>& op) {
  const std::vector<int> Ns = {16, 32, 64};
  std::vector<double> err;
  for (int N : Ns) {
    const double h = 2.0 * kPi / N;
    const Grid g(N, N, N, Vec3d(h, h, h));
    Field<double> psi(g, [](double x, double y, double z) {
      return std::sin(x) * std::sin(y) * std::sin(z);
    });
    Field<double> want(g, [&](double x, double y, double z) {
      return dsin(nx, x) * dsin(ny, y) * dsin(nz, z);
    });
    const auto got = op(psi);
    double e = 0.0, m = 0.0;
    for (std::size_t c = 0; c < got.getSize(); ++c) {
      e = std::max(e, std::abs(got.data()[c] - want.data()[c]));
      m = std::max(m, std::abs(want.data()[c]));
    }
    err.push_back(e / m);
  }
  // least-squares slope of log(err) against log(h)
  double sx = 0, sy = 0, sxy = 0, sxx = 0;
  for (std::size_t i = 0; i < Ns.size(); ++i) {
    const double x = std::log(2.0 * kPi / Ns[i]), y = std::log(err[i]);
    sx += x; sy += y; sxy += x * y; sxx += x * x;
  }
  const double n = static_cast<double>(Ns.size());
  return (n * sxy - sx * sy) / (n * sxx - sx * sx);
}

}  // namespace

#define PARTIAL_SWEEP(NAME, NX, NY, NZ)                                                    \
  TEST(PartialConvergence, NAME) {                                                         \
    const double s2 = sweepSlope(NX, NY, NZ, [](const Field<double>& f) { return gd::NAME<2>(f); }); \
    EXPECT_GT(s2, 1.5);                                                                    \
    EXPECT_LT(s2, 2.5);                                                                    \
    const double s4 = sweepSlope(NX, NY, NZ, [](const Field<double>& f) { return gd::NAME<4>(f); }); \
    EXPECT_GT(s4, 3.5);                                                                    \
    EXPECT_LT(s4, 4.5);                                                                    \
  }

PARTIAL_SWEEP(dxx, 2, 0, 0)
PARTIAL_SWEEP(dyy, 0, 2, 0)
PARTIAL_SWEEP(dzz, 0, 0, 2)
PARTIAL_SWEEP(dxy, 1, 1, 0)
PARTIAL_SWEEP(dxz, 1, 0, 1)
PARTIAL_SWEEP(dyz, 0, 1, 1)
PARTIAL_SWEEP(d3dx3, 3, 0, 0)
PARTIAL_SWEEP(d3dy3, 0, 3, 0)
PARTIAL_SWEEP(d3dz3, 0, 0, 3)
PARTIAL_SWEEP(d3dx2dy, 2, 1, 0)
PARTIAL_SWEEP(d3dxdy2, 1, 2, 0)
PARTIAL_SWEEP(d3dx2dz, 2, 0, 1)
PARTIAL_SWEEP(d3dxdz2, 1, 0, 2)
PARTIAL_SWEEP(d3dy2dz, 0, 2, 1)
PARTIAL_SWEEP(d3dydz2, 0, 1, 2)
PARTIAL_SWEEP(d3dxdydz, 1, 1, 1)
PARTIAL_SWEEP(d4dx4, 4, 0, 0)
PARTIAL_SWEEP(d4dy4, 0, 4, 0)
PARTIAL_SWEEP(d4dz4, 0, 0, 4)
PARTIAL_SWEEP(d4dx3dy, 3, 1, 0)
PARTIAL_SWEEP(d4dxdy3, 1, 3, 0)
PARTIAL_SWEEP(d4dx3dz, 3, 0, 1)
PARTIAL_SWEEP(d4dxdz3, 1, 0, 3)
PARTIAL_SWEEP(d4dy3dz, 0, 3, 1)
PARTIAL_SWEEP(d4dydz3, 0, 1, 3)
PARTIAL_SWEEP(d4dx2dy2, 2, 2, 0)
PARTIAL_SWEEP(d4dx2dz2, 2, 0, 2)
PARTIAL_SWEEP(d4dy2dz2, 0, 2, 2)
PARTIAL_SWEEP(d4dx2dydz, 2, 1, 1)
PARTIAL_SWEEP(d4dxdy2dz, 1, 2, 1)
PARTIAL_SWEEP(d4dxdydz2, 1, 1, 2)

TEST(PartialExtent, MessageNamesAxis) {
  Field<double> psi(Grid(8, 2, 16, Vec3d(1.0, 1.0, 1.0)), 0.0);
  EXPECT_THROW_MSG(gd::dxy(psi), std::invalid_argument, "axis y");
  (void)gd::dxz(psi);  // y is not differentiated: extent 2 is accepted
  EXPECT_THROW_MSG(gd::d3dz3<4>(Field<double>(Grid(8, 8, 6, Vec3d(1.0, 1.0, 1.0)), 0.0)),
                   std::invalid_argument, "axis z");
}

CHECK_MAIN()
