// RK4 steps of Cahn-Hilliard through solve::integrate with three right-hand sides of the same
// mathematics, timed on the device and checked to give the same bits:
//   host   -- the caller's lambda of examples/common/cahn_hilliard.hpp:38-59 (two device
//             Laplacians, host loops for mu and for the mobility scaling): every stage moves
//             fields between host and device;
//   expr   -- the same lambda with the loops written as tensor::expr expressions
//             (-psi + psi*psi*psi - kappa*lap, then M*rhs): every operation on the device;
//   fused  -- solve::CahnHilliardRhs{M, kappa}: the fused stage pipeline.
// Prints one JSON object per grid: ms per RK4 step and field transfers per step of each path.
// Usage: rhs_paths [n2d n3d steps_host steps_expr steps_fused]
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include <gridcalc/core/EigenAliases.hpp>
#include <gridcalc/core/Field.hpp>
#include <gridcalc/core/Grid.hpp>
#include <gridcalc/diff/Laplacian.hpp>
#include <gridcalc/hip/Runtime.hpp>
#include <gridcalc/solve/CahnHilliard.hpp>
#include <gridcalc/solve/Integrator.hpp>
#include <gridcalc/tensor/Expressions.hpp>

using gridcalc::core::Field;
using gridcalc::core::Grid;
using gridcalc::core::Vec3d;
namespace diff = gridcalc::diff;
namespace expr = gridcalc::tensor::expr;
namespace hip = gridcalc::hip;
namespace solve = gridcalc::solve;

namespace {

constexpr double kM = 1.0, kKappa = 1.0, kDt = 0.01;

Field<double> initial(const Grid& g, std::uint64_t seed) {
  Field<double> f(g);
  double* p = f.data();
  for (std::size_t c = 0; c < f.getSize(); ++c) {
    std::uint64_t z = seed + (c + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    p[c] = 0.05 * (2.0 * (static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0);
  }
  return f;
}

Field<double> hostRhs(const Field<double>& psi) {
  auto lap_psi = diff::laplacian(psi);
  Field<double> mu(psi.getGrid());
  const std::size_t n = psi.getSize();
  const double* pp = psi.data();
  const double* pl = lap_psi.data();
  double* pm = mu.data();
  for (std::size_t c = 0; c < n; ++c) {
    const double p = pp[c];
    pm[c] = -p + p * p * p - kKappa * pl[c];
  }
  auto rhs = diff::laplacian(mu);
  double* pr = rhs.data();
  for (std::size_t c = 0; c < n; ++c) pr[c] *= kM;
  return rhs;
}

Field<double> exprRhs(const Field<double>& psi) {
  const auto lap_psi = diff::laplacian(psi);
  const auto p = expr::field(psi);
  const auto mu = expr::materialize(-p + p * p * p - kKappa * expr::field(lap_psi));
  const auto rhs = diff::laplacian(mu);
  return expr::materialize(kM * expr::field(rhs));
}

template <class Rhs>
double msPerStep(Field<double> psi, Rhs rhs, int steps, double* transfers) {
  solve::integrate(psi, rhs, kDt, 1, solve::RK4{});  // first use: pool and tuner warm-up
  hip::synchronize();
  hip::resetTransferStats();
  const auto t0 = std::chrono::steady_clock::now();
  solve::integrate(psi, rhs, kDt, steps, solve::RK4{});
  hip::synchronize();
  const auto t1 = std::chrono::steady_clock::now();
  const auto st = hip::transferStats();
  *transfers = static_cast<double>(st.uploads + st.downloads) / steps;
  return std::chrono::duration<double, std::milli>(t1 - t0).count() / steps;
}

bool sameBits(const Field<double>& a, const Field<double>& b) {
  return std::memcmp(a.data(), b.data(), a.getSize() * sizeof(double)) == 0;
}

void run(const Grid& g, const char* name, int sh, int se, int sf) {
  const auto psi0 = initial(g, 42);
  // the three paths from the same state over 8 steps: identical bits
  auto a = psi0, b = psi0, c = psi0;
  solve::integrate(a, hostRhs, kDt, 8, solve::RK4{});
  solve::integrate(b, exprRhs, kDt, 8, solve::RK4{});
  solve::integrate(c, solve::CahnHilliardRhs{kM, kKappa}, kDt, 8, solve::RK4{});
  const bool same = sameBits(a, b) && sameBits(a, c);
  double th = 0, te = 0, tf = 0;
  const double mh = msPerStep(psi0, hostRhs, sh, &th);
  const double me = msPerStep(psi0, exprRhs, se, &te);
  const double mf = msPerStep(psi0, solve::CahnHilliardRhs{kM, kKappa}, sf, &tf);
  std::printf("{\"grid\": \"%s\", \"sites\": %zu, \"identical_bits\": %s, "
              "\"ms_per_step\": {\"host_lambda\": %.6f, \"expr_lambda\": %.6f, \"fused\": %.6f}, "
              "\"transfers_per_step\": {\"host_lambda\": %.2f, \"expr_lambda\": %.2f, \"fused\": %.2f}, "
              "\"speedup_vs_host_lambda\": {\"expr_lambda\": %.2f, \"fused\": %.2f}}\n",
              name, g.getSize(), same ? "true" : "false", mh, me, mf, th, te, tf, mh / me, mh / mf);
  std::fflush(stdout);
}

}  // namespace

int main(int argc, char** argv) {
  const int n2 = argc > 1 ? std::atoi(argv[1]) : 128;
  const int n3 = argc > 2 ? std::atoi(argv[2]) : 64;
  const int sh = argc > 3 ? std::atoi(argv[3]) : 20;
  const int se = argc > 4 ? std::atoi(argv[4]) : 100;
  const int sf = argc > 5 ? std::atoi(argv[5]) : 500;
  try {
    run(Grid(n2, n2, 1, Vec3d(1.0, 1.0, 1.0)), (std::to_string(n2) + "^2").c_str(), sh, se, sf);
    run(Grid(n3, n3, n3, Vec3d(1.0, 1.0, 1.0)), (std::to_string(n3) + "^3").c_str(), sh, se, sf);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "rhs_paths: %s\n", e.what());
    return 1;
  }
  return 0;
}
