// The stencil axis-extent precondition of every operator, as the reference's
// test/stencil_axis_extent_test.cpp:116-409 exercises it: an axis with 2 <= N <= 2r (r the
// stencil radius on that axis) aliases the stencil's offsets under the periodic wrap and is
// rejected with std::invalid_argument naming the axis; N = 2r + 1 (the smallest valid
// extent) and N = 1 (a degenerate axis) are accepted and give finite values; an axis a mixed
// partial does not differentiate accepts any extent.
//
// Every operator is checked on every axis and every aliasing extent (the reference picks
// one axis and one extent per operator): each row below is (operator, radius per axis).
#include <gridcalc/core/Field.hpp>
#include <gridcalc/core/Grid.hpp>
#include <gridcalc/diff/Biharmonic.hpp>
#include <gridcalc/diff/Divergence.hpp>
#include <gridcalc/diff/Gradient.hpp>
#include <gridcalc/diff/Laplacian.hpp>
#include <gridcalc/diff/detail/MultiIndexDerivative.hpp>
#include <gridcalc/fvm/CellLaplacian.hpp>

#include <array>
#include <cmath>
#include <functional>
#include <stdexcept>
#include <string>

#include "check.hpp"

using gridcalc::core::Field;
using gridcalc::core::Grid;
using gridcalc::core::Mat3d;
using gridcalc::core::Vec3d;
namespace gd = gridcalc::diff;

namespace {

constexpr double kPi = 3.14159265358979323846;

// A periodic box of side 2 pi with n = (nx, ny, nz) cells.
Grid box(const std::array<int, 3>& n) {
  return Grid(n[0], n[1], n[2],
              Vec3d(2.0 * kPi / n[0], 2.0 * kPi / n[1], 2.0 * kPi / n[2]));
}

double smooth(double x, double y, double z) { return std::sin(x + 2.0 * y + 3.0 * z) + 0.5; }

bool finite(const Field<double>& f) {
  for (std::size_t n = 0; n < f.getSize(); ++n)
    if (!std::isfinite(f.data()[n])) return false;
  return true;
}
template <class V>
bool finite(const Field<V>& f) {
  for (std::size_t n = 0; n < f.getSize(); ++n)
    if (!f.data()[n].allFinite()) return false;
  return true;
}

Field<double> scalar(const Grid& g) { return Field<double>(g, smooth); }
Field<Vec3d> vector(const Grid& g) {
  return Field<Vec3d>(g, [](double x, double y, double z) {
    return Vec3d(smooth(x, y, z), smooth(y, z, x), smooth(z, x, y));
  });
}
Field<Mat3d> tensor(const Grid& g) {
  return Field<Mat3d>(g, [](double x, double y, double z) {
    Mat3d m;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) m(r, c) = smooth(x + r, y + c, z);
    return m;
  });
}

struct Op {
  const char* name;
  std::array<int, 3> radius;                  // stencil radius per axis (0: not differentiated)
  std::function<bool(const Grid&)> finite_on;  // applies the operator, true if all finite
};

template <int Nx, int Ny, int Nz, int O>
Op mixed(const char* name) {
  auto r = [](int n) { return n == 0 ? 0 : (n <= 2 ? O / 2 : O / 2 + 1); };
  return {name, {r(Nx), r(Ny), r(Nz)},
          [](const Grid& g) { return finite(gd::detail::mixedDerivative<Nx, Ny, Nz, O>(scalar(g))); }};
}

std::vector<Op> ops() {
  auto all = [](int r) { return std::array<int, 3>{r, r, r}; };
  return {
      {"laplacian<2>", all(1), [](const Grid& g) { return finite(gd::laplacian(scalar(g))); }},
      {"laplacian<4>", all(2), [](const Grid& g) { return finite(gd::laplacian<4>(scalar(g))); }},
      {"gradient<2>", all(1), [](const Grid& g) { return finite(gd::gradient(scalar(g))); }},
      {"gradient<4>", all(2), [](const Grid& g) { return finite(gd::gradient<4>(scalar(g))); }},
      {"divergence<2>(Vec3)", all(1), [](const Grid& g) { return finite(gd::divergence(vector(g))); }},
      {"divergence<4>(Vec3)", all(2), [](const Grid& g) { return finite(gd::divergence<4>(vector(g))); }},
      {"gradient<2>(Vec3)", all(1), [](const Grid& g) { return finite(gd::gradient(vector(g))); }},
      {"gradient<4>(Vec3)", all(2), [](const Grid& g) { return finite(gd::gradient<4>(vector(g))); }},
      {"divergence<2>(Mat3)", all(1), [](const Grid& g) { return finite(gd::divergence(tensor(g))); }},
      {"divergence<4>(Mat3)", all(2), [](const Grid& g) { return finite(gd::divergence<4>(tensor(g))); }},
      // the biharmonic's radius is that of its fourth-derivative stencil (one more than the
      // second derivative's)
      {"biharmonic<2>", all(2), [](const Grid& g) { return finite(gd::biharmonic(scalar(g))); }},
      {"biharmonic<4>", all(3), [](const Grid& g) { return finite(gd::biharmonic<4>(scalar(g))); }},
      {"cellLaplacian", all(1),
       [](const Grid& g) { return finite(gridcalc::fvm::cellLaplacian(scalar(g), Field<double>(g, 1.0))); }},
      mixed<1, 0, 0, 2>("mixed<1,0,0,2>"), mixed<0, 1, 0, 4>("mixed<0,1,0,4>"),
      mixed<3, 0, 0, 2>("mixed<3,0,0,2>"), mixed<4, 0, 0, 2>("mixed<4,0,0,2>"),
      mixed<0, 3, 0, 4>("mixed<0,3,0,4>"), mixed<0, 0, 4, 4>("mixed<0,0,4,4>"),
      mixed<3, 0, 0, 4>("mixed<3,0,0,4>"), mixed<4, 0, 0, 4>("mixed<4,0,0,4>"),
      mixed<1, 1, 0, 2>("mixed<1,1,0,2>"), mixed<2, 0, 2, 4>("mixed<2,0,2,4>"),
      mixed<1, 1, 1, 4>("mixed<1,1,1,4>"), mixed<1, 2, 1, 2>("mixed<1,2,1,2>"),
  };
}

// Smallest valid extent on every axis.
std::array<int, 3> smallest(const Op& op) {
  std::array<int, 3> n{};
  for (int a = 0; a < 3; ++a) n[a] = 2 * op.radius[a] + 1;
  return n;
}

const char* kAxis[3] = {"axis x", "axis y", "axis z"};

}  // namespace

TEST(StencilAxisExtent, AliasingExtentsRejectedNamingTheAxis) {
  for (const Op& op : ops()) {
    for (int a = 0; a < 3; ++a) {
      for (int N = 2; N <= 2 * op.radius[a]; ++N) {
        auto n = smallest(op);
        n[a] = N;
        bool thrown = false;
        try {
          (void)op.finite_on(box(n));
        } catch (const std::invalid_argument& e) {
          thrown = true;
          const std::string what = e.what();
          if (what.find(kAxis[a]) == std::string::npos)
            check::fail(__FILE__, __LINE__, std::string(op.name) + ": " + what, false);
          if (what.find("N=" + std::to_string(N)) == std::string::npos)
            check::fail(__FILE__, __LINE__, std::string(op.name) + " message lacks N: " + what, false);
        }
        if (!thrown)
          check::fail(__FILE__, __LINE__,
                      std::string(op.name) + " accepted N=" + std::to_string(N) + " on " + kAxis[a],
                      false);
      }
    }
  }
}

TEST(StencilAxisExtent, SmallestValidExtentAcceptedAndFinite) {
  for (const Op& op : ops()) {
    bool ok = false;
    try {
      ok = op.finite_on(box(smallest(op)));
    } catch (const std::exception& e) {
      check::fail(__FILE__, __LINE__, std::string(op.name) + " threw: " + e.what(), false);
      continue;
    }
    if (!ok) check::fail(__FILE__, __LINE__, std::string(op.name) + ": non-finite values", false);
  }
}

TEST(StencilAxisExtent, DegenerateAxisAccepted) {
  for (const Op& op : ops()) {
    for (int a = 0; a < 3; ++a) {
      auto n = smallest(op);
      n[a] = 1;
      bool ok = false;
      try {
        ok = op.finite_on(box(n));
      } catch (const std::exception& e) {
        check::fail(__FILE__, __LINE__, std::string(op.name) + " N=1 on " + kAxis[a] + ": " + e.what(),
                    false);
        continue;
      }
      if (!ok) check::fail(__FILE__, __LINE__, std::string(op.name) + ": non-finite values", false);
    }
  }
}

TEST(StencilAxisExtent, AxisNotDifferentiatedAcceptsAnyExtent) {
  for (const Op& op : ops()) {
    for (int a = 0; a < 3; ++a) {
      if (op.radius[a] != 0) continue;
      for (int N = 2; N <= 4; ++N) {
        auto n = smallest(op);
        n[a] = N;
        bool ok = false;
        try {
          ok = op.finite_on(box(n));
        } catch (const std::exception& e) {
          check::fail(__FILE__, __LINE__,
                      std::string(op.name) + " N=" + std::to_string(N) + " on " + kAxis[a] + ": " + e.what(),
                      false);
          continue;
        }
        if (!ok) check::fail(__FILE__, __LINE__, std::string(op.name) + ": non-finite values", false);
      }
    }
  }
}

CHECK_MAIN()
