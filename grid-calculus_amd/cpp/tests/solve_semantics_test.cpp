// Time-stepping semantics through the C++ header API, on the device:
//  - the right-hand-side grid check and its stage labels (test/integrator_grid_mismatch_test.cpp
//    :80-221 of the reference);
//  - explicitEuler == integrate(..., ExplicitEuler{}) bit for bit (integrator_test.cpp:255-285);
//  - the recognised CahnHilliardRhs (fused stage pipeline) == the same right-hand side written
//    as a host lambda the way the reference's example writes it, bit for bit, RK4 and Euler;
//  - heterogeneous-D diffusion == constant-D diffusion for a uniform D, and its CFL and D > 0
//    rejections (diffusion_test.cpp:216-275);
//  - the Cahn-Hilliard coarsening acceptance run of cahn_hilliard_test.cpp:101-184 (Lyapunov
//    decay, coarsening, mean conservation) through the unchanged lambda path;
//  - switching the library context in the middle of a computation.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include <gridcalc/core/EigenAliases.hpp>
#include <gridcalc/core/Field.hpp>
#include <gridcalc/core/Grid.hpp>
#include <gridcalc/diff/Laplacian.hpp>
#include <gridcalc/func/Functional.hpp>
#include <gridcalc/func/Integrate.hpp>
#include <gridcalc/solve/CahnHilliard.hpp>
#include <gridcalc/solve/Diffusion.hpp>
#include <gridcalc/solve/Integrator.hpp>
#include <gridcalc/tensor/Expressions.hpp>

#include "check.hpp"

using gridcalc::core::Field;
using gridcalc::core::Grid;
using gridcalc::core::Vec3d;
using gridcalc::solve::CahnHilliardRhs;
using gridcalc::solve::ExplicitEuler;
using gridcalc::solve::RK4;
using gridcalc::solve::integrate;
namespace diff = gridcalc::diff;
namespace func = gridcalc::func;
namespace solve = gridcalc::solve;

namespace {

constexpr double kPi = 3.14159265358979323846;

template <class T>
bool bitEqual(const Field<T>& a, const Field<T>& b) {
  if (a.getSize() != b.getSize()) return false;
  return std::memcmp(a.data(), b.data(), a.getSize() * sizeof(T)) == 0;
}

double maxAbsDiff(const Field<double>& a, const Field<double>& b) {
  double m = 0.0;
  for (std::size_t c = 0; c < a.getSize(); ++c) m = std::max(m, std::abs(a.data()[c] - b.data()[c]));
  return m;
}

// splitmix64 counter stream -> U(-amp, amp), mean removed
Field<double> noiseField(const Grid& g, std::uint64_t seed, double amp) {
  Field<double> f(g);
  double* p = f.data();
  double sum = 0.0;
  for (std::size_t c = 0; c < f.getSize(); ++c) {
    std::uint64_t z = seed + (c + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const double u = static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0);
    p[c] = amp * (2.0 * u - 1.0);
    sum += p[c];
  }
  const double mean = sum / static_cast<double>(f.getSize());
  for (std::size_t c = 0; c < f.getSize(); ++c) p[c] -= mean;
  return f;
}

// The Cahn-Hilliard right-hand side written as a caller's host lambda would write it:
// two device Laplacians with host loops for mu and the mobility scaling.
Field<double> lambdaChRhs(const Field<double>& psi, double M, double kappa) {
  auto lap_psi = diff::laplacian(psi);
  Field<double> mu(psi.getGrid());
  const std::size_t n = psi.getSize();
  const double* pp = psi.data();
  const double* pl = lap_psi.data();
  double* pm = mu.data();
  for (std::size_t c = 0; c < n; ++c) {
    const double p = pp[c];
    pm[c] = -p + p * p * p - kappa * pl[c];
  }
  auto rhs = diff::laplacian(mu);
  double* pr = rhs.data();
  for (std::size_t c = 0; c < n; ++c) pr[c] *= M;
  return rhs;
}

Field<double> makeStateField() { return Field<double>(Grid(8, 8, 8, Vec3d(0.1, 0.1, 0.1)), 0.0); }

struct DifferentTotalCountRhs {
  Grid wrong{4, 4, 4, Vec3d(0.1, 0.1, 0.1)};
  Field<double> operator()(const Field<double>&) const { return Field<double>(wrong, 0.0); }
};
struct SameSizeDifferentShapeRhs {
  Grid wrong{4, 16, 8, Vec3d(0.1, 0.1, 0.1)};
  Field<double> operator()(const Field<double>&) const { return Field<double>(wrong, 0.0); }
};
struct SameShapeDifferentCellSizeRhs {
  Grid wrong{8, 8, 8, Vec3d(0.2, 0.2, 0.2)};
  Field<double> operator()(const Field<double>&) const { return Field<double>(wrong, 0.0); }
};
struct FailOnNthCall {
  Grid wrong{4, 4, 4, Vec3d(0.1, 0.1, 0.1)};
  int n_good;
  mutable int calls = 0;
  Field<double> operator()(const Field<double>& psi) const {
    ++calls;
    if (calls <= n_good) return Field<double>(psi.getGrid(), 0.0);
    return Field<double>(wrong, 0.0);
  }
};

}  // namespace

// ---- right-hand-side grid check ----------------------------------------------------------

TEST(IntegratorGridMismatch, RK4RejectsEveryFlavour) {
  auto psi = makeStateField();
  EXPECT_THROW_MSG(integrate(psi, DifferentTotalCountRhs{}, 0.01, 1, RK4{}), std::invalid_argument, "k1");
  EXPECT_THROW_MSG(integrate(psi, SameSizeDifferentShapeRhs{}, 0.01, 1, RK4{}), std::invalid_argument, "k1");
  EXPECT_THROW_MSG(integrate(psi, SameShapeDifferentCellSizeRhs{}, 0.01, 1, RK4{}), std::invalid_argument,
                   "different Grid");
}

TEST(IntegratorGridMismatch, EulerRejectsEveryFlavour) {
  auto psi = makeStateField();
  EXPECT_THROW_MSG(integrate(psi, DifferentTotalCountRhs{}, 0.01, 1, ExplicitEuler{}), std::invalid_argument,
                   "ExplicitEuler): rhs(psi)");
  EXPECT_THROW_MSG(integrate(psi, SameSizeDifferentShapeRhs{}, 0.01, 1, ExplicitEuler{}),
                   std::invalid_argument, "ExplicitEuler): rhs(psi)");
  EXPECT_THROW_MSG(integrate(psi, SameShapeDifferentCellSizeRhs{}, 0.01, 1, ExplicitEuler{}),
                   std::invalid_argument, "ExplicitEuler): rhs(psi)");
}

TEST(IntegratorGridMismatch, AxpyFresh) {
  Field<double> a(Grid(8, 8, 8, Vec3d(0.1, 0.1, 0.1)), 1.0);
  Field<double> b(Grid(4, 4, 4, Vec3d(0.1, 0.1, 0.1)), 1.0);
  Field<double> c(Grid(8, 8, 8, Vec3d(0.2, 0.1, 0.1)), 1.0);
  EXPECT_THROW_MSG(solve::detail::axpyFresh(a, 0.5, b), std::invalid_argument, "axpyFresh");
  EXPECT_THROW_MSG(solve::detail::axpyFresh(a, 0.5, c), std::invalid_argument, "axpyFresh");
  Field<double> d(Grid(8, 8, 8, Vec3d(0.1, 0.1, 0.1)), 2.0);
  const auto out = solve::detail::axpyFresh(a, 0.5, d);
  ASSERT_EQ(out.getSize(), a.getSize());
  for (std::size_t i = 0; i < out.getSize(); ++i) EXPECT_EQ(out.data()[i], 2.0);
}

TEST(IntegratorGridMismatch, RK4StageLabels) {
  const char* labels[] = {"k1", "k2", "k3", "k4"};
  for (int n_good = 0; n_good < 4; ++n_good) {
    auto psi = makeStateField();
    bool thrown = false;
    try {
      integrate(psi, FailOnNthCall{Grid(4, 4, 4, Vec3d(0.1, 0.1, 0.1)), n_good}, 0.01, 1, RK4{});
    } catch (const std::invalid_argument& e) {
      thrown = true;
      const std::string msg = e.what();
      EXPECT_TRUE(msg.find(std::string("RK4): ") + labels[n_good]) != std::string::npos);
    }
    EXPECT_TRUE(thrown);
  }
}

TEST(IntegratorGridMismatch, HappyPath) {
  Field<double> psi(Grid(8, 8, 8, Vec3d(0.1, 0.1, 0.1)), 1.0);
  integrate(psi, [](const Field<double>& f) { return diff::laplacian(f); }, 0.0001, 3, RK4{});
  integrate(psi, [](const Field<double>& f) { return diff::laplacian(f); }, 0.0001, 3, ExplicitEuler{});
  for (std::size_t i = 0; i < psi.getSize(); ++i) EXPECT_EQ(psi.data()[i], 1.0);
}

// ---- integrator equivalences ---------------------------------------------------------------

TEST(Integrator, ExplicitEulerEqualsIntegrateEuler) {
  const int N = 16;
  const double h = 2.0 * kPi / N;
  const Grid grid(N, N, N, Vec3d(h, h, h));
  auto make = [&] {
    return Field<double>(grid, [](double x, double y, double z) { return std::sin(x) * std::cos(y) * std::sin(z); });
  };
  auto rhs = [](const Field<double>& f) {
    Field<double> out(f.getGrid());
    const double* p = f.data();
    double* o = out.data();
    for (std::size_t n = 0; n < out.getSize(); ++n) o[n] = -0.1 * p[n];
    return out;
  };
  auto a = make();
  auto b = make();
  solve::explicitEuler(a, rhs, 0.001, 50);
  integrate(b, rhs, 0.001, 50, ExplicitEuler{});
  EXPECT_TRUE(bitEqual(a, b));
  // and both equal the closed form (1 - 0.1 dt)^n psi0 to rounding
  const auto p0 = make();
  const double amp = std::pow(1.0 - 0.1 * 0.001, 50);
  double m = 0.0;
  for (std::size_t c = 0; c < a.getSize(); ++c) m = std::max(m, std::abs(a.data()[c] - amp * p0.data()[c]));
  EXPECT_LT(m, 1e-14);
}

TEST(CahnHilliard, FusedRhsEqualsLambdaRhs) {
  for (const Grid& g : {Grid(24, 20, 18, Vec3d(1.0, 0.9, 1.1)), Grid(64, 48, 1, Vec3d(1.0, 1.0, 1.0))}) {
    const auto psi = noiseField(g, 7, 0.6);
    const auto fused = CahnHilliardRhs{1.3, 0.6}(psi);
    const auto host = lambdaChRhs(psi, 1.3, 0.6);
    EXPECT_TRUE(bitEqual(fused, host));
  }
}

TEST(CahnHilliard, FusedIntegrateEqualsLambdaIntegrate) {
  for (const Grid& g : {Grid(24, 20, 18, Vec3d(1.0, 1.0, 1.0)), Grid(64, 64, 1, Vec3d(1.0, 1.0, 1.0))}) {
    const auto psi0 = noiseField(g, 11, 0.1);
    for (int euler = 0; euler < 2; ++euler) {
      auto a = psi0;
      auto b = psi0;
      auto rhs = [](const Field<double>& y) { return lambdaChRhs(y, 1.0, 0.5); };
      if (euler) {
        integrate(a, CahnHilliardRhs{1.0, 0.5}, 0.01, 21, ExplicitEuler{});
        integrate(b, rhs, 0.01, 21, ExplicitEuler{});
      } else {
        integrate(a, CahnHilliardRhs{1.0, 0.5}, 0.01, 21, RK4{});
        integrate(b, rhs, 0.01, 21, RK4{});
      }
      EXPECT_TRUE(bitEqual(a, b));
    }
  }
}

// ---- diffusion -----------------------------------------------------------------------------

TEST(Diffusion, HeterogeneousMatchesConstantOnUniformD) {
  const int N = 16;
  const double h = 2.0 * kPi / N;
  const Grid grid(N, N, N, Vec3d(h, h, h));
  const double D = 0.1, dt = 0.4 * h * h / (6.0 * D);
  Field<double> a(grid, [](double x, double y, double z) { return std::sin(x) * std::cos(y) * std::sin(z); });
  Field<double> b = a;
  Field<double> Df(grid, D);
  solve::diffuse(a, D, dt, 50);
  solve::diffuse(b, Df, dt, 50);
  EXPECT_LT(maxAbsDiff(a, b), 1e-12);
  Field<double> c = a;
  Field<double> e = a;
  solve::diffuse(c, D, dt, 20, RK4{});
  solve::diffuse(e, Df, dt, 20, RK4{});
  EXPECT_LT(maxAbsDiff(c, e), 1e-12);
}

TEST(Diffusion, HeterogeneousRejectsCflViolation) {
  const int N = 32;
  const double h = 2.0 * kPi / N;
  const Grid grid(N, N, N, Vec3d(h, h, h));
  Field<double> psi(grid, 1.0);
  Field<double> D(grid, [](double x, double, double) { return 0.5 + 0.5 * std::cos(x); });
  for (std::size_t i = 0; i < D.getSize(); ++i)
    if (D.data()[i] <= 0.0) D.data()[i] = 1e-6;
  EXPECT_THROW_MSG(solve::diffuse(psi, D, h * h / 3.0, 1), std::invalid_argument, "solve::diffuse");
  solve::diffuse(psi, D, 0.1 * h * h, 1);  // well inside the bound
}

TEST(Diffusion, HeterogeneousRejectsNonpositiveD) {
  const Grid grid(8, 8, 8, Vec3d(0.1, 0.1, 0.1));
  Field<double> psi(grid, 1.0);
  Field<double> D(grid, 1.0);
  D(3, 4, 5) = -0.01;
  EXPECT_THROW_MSG(solve::diffuse(psi, D, 1e-4, 1), std::invalid_argument, "flat index");
  D(3, 4, 5) = 0.0;
  EXPECT_THROW_MSG(solve::diffuse(psi, D, 1e-4, 1), std::invalid_argument, "flat index");
}

TEST(Diffusion, RejectsNegativeSteps) {
  Field<double> psi(Grid(4, 4, 4, Vec3d(0.1, 0.1, 0.1)), 0.0);
  EXPECT_THROW_MSG(solve::diffuse(psi, 0.1, 0.001, -1), std::invalid_argument, "n_steps");
  auto rhs = [](const Field<double>& f) { return Field<double>(f.getGrid(), 0.0); };
  EXPECT_THROW_MSG(solve::explicitEuler(psi, rhs, 0.001, -1), std::invalid_argument, "n_steps");
}

// ---- Cahn-Hilliard acceptance run, unchanged lambda path -------------------------------------

TEST(CahnHilliard, CoarseningAcceptanceLambdaPath) {
  // 24^3, dt = 0.02, kappa = 0.5, M = 1, amplitude 0.1, snapshots at t = 0, 10, 20, 30, 50
  const Grid grid(24, 24, 24, Vec3d(1.0, 1.0, 1.0));
  auto psi = noiseField(grid, 42, 0.1);
  auto rhs = [](const Field<double>& y) { return lambdaChRhs(y, 1.0, 0.5); };
  auto mean = [&](const Field<double>& f) {
    return func::evaluate(f, [](double p) { return p; }) /
           (static_cast<double>(grid.getSize()) * grid.getCellVolume());
  };
  auto F = [](const Field<double>& f) {
    return func::evaluate(f, [](double p, const Vec3d& g) {
      return -0.5 * p * p + 0.25 * p * p * p * p + 0.5 * 0.5 * g.squaredNorm();
    });
  };
  auto Eg = [](const Field<double>& f) {
    return func::evaluate(f, [](double, const Vec3d& g) { return g.squaredNorm(); });
  };
  std::vector<double> Fs{F(psi)}, Es{Eg(psi)}, means{mean(psi)};
  for (int seg : {500, 500, 500, 1000}) {
    integrate(psi, rhs, 0.02, seg, RK4{});
    Fs.push_back(F(psi));
    Es.push_back(Eg(psi));
    means.push_back(mean(psi));
  }
  const double tolF = 1e-6 * std::abs(Fs[0]) + 1e-9;
  for (std::size_t i = 0; i + 1 < Fs.size(); ++i) EXPECT_LE(Fs[i + 1], Fs[i] + tolF);
  EXPECT_GT((Es[3] - Es[4]) / Es[3], 0.03);
  EXPECT_LT(std::abs(means.back() - means.front()), 1e-10);
  // the fused density gives the lambda density's value
  EXPECT_EQ(func::evaluate(psi, func::CahnHilliardDensity{0.5}), Fs.back());
}

// ---- right-hand sides that stay on the device ------------------------------------------------

TEST(DeviceResidentRhs, ExpressionLambdaMatchesHostLoopsWithoutTransfers) {
  // the example's right-hand side with its loops written as tensor::expr expressions: the
  // same bits as the host loops, and not a single field transfer per step
  namespace expr = gridcalc::tensor::expr;
  for (const Grid& g : {Grid(32, 24, 20, Vec3d(1.0, 0.9, 1.1)), Grid(64, 48, 1, Vec3d(1.0, 1.0, 1.0))}) {
    const auto psi0 = noiseField(g, 3, 0.1);
    auto exprRhs = [](const Field<double>& y) {
      const auto lap = diff::laplacian(y);
      const auto p = expr::field(y);
      const auto mu = expr::materialize(-p + p * p * p - 0.5 * expr::field(lap));
      const auto r = diff::laplacian(mu);
      return expr::materialize(1.3 * expr::field(r));
    };
    auto a = psi0;
    auto b = psi0;
    auto c = psi0;
    gridcalc::hip::resetTransferStats();
    integrate(a, exprRhs, 0.01, 6, RK4{});
    const auto st = gridcalc::hip::transferStats();
    EXPECT_EQ(st.uploads, std::size_t(1));  // psi0's host copy, once
    EXPECT_EQ(st.downloads, std::size_t(0));
    integrate(b, [](const Field<double>& y) { return lambdaChRhs(y, 1.3, 0.5); }, 0.01, 6, RK4{});
    integrate(c, CahnHilliardRhs{1.3, 0.5}, 0.01, 6, RK4{});
    EXPECT_TRUE(bitEqual(a, b));
    EXPECT_TRUE(bitEqual(a, c));
  }
}

// ---- context switch in the middle of a computation ------------------------------------------

TEST(Runtime, SwitchContextMidComputation) {
  const Grid g(40, 36, 32, Vec3d(1.0, 1.0, 1.0));
  const auto psi0 = noiseField(g, 5, 0.3);
  // reference: everything on the default context
  auto ref = psi0;
  integrate(ref, CahnHilliardRhs{1.0, 0.7}, 0.01, 6, RK4{});
  const auto ref_lap = diff::laplacian(ref);
  // the same work with the context switched between operations, temporaries released and
  // reused across the switch
  gc_context* other = nullptr;
  ASSERT_EQ(gc_context_create(&other, nullptr), GC_OK);
  auto a = psi0;
  integrate(a, CahnHilliardRhs{1.0, 0.7}, 0.01, 3, RK4{});
  { const auto tmp = diff::laplacian(a); (void)tmp; }  // a block returns to the pool
  gridcalc::hip::setContext(other);
  integrate(a, CahnHilliardRhs{1.0, 0.7}, 0.01, 3, RK4{});
  const auto lap_other = diff::laplacian(a);
  gridcalc::hip::setContext(nullptr);
  EXPECT_TRUE(bitEqual(a, ref));
  EXPECT_TRUE(bitEqual(lap_other, ref_lap));
  gc_context_destroy(other);
}

CHECK_MAIN()
