// The Cahn-Hilliard demo run (examples/cahn_hilliard.cpp of the reference: 3-D cubic grid,
// random initial state of amplitude 0.05 with its mean removed, RK4, diagnostics and a VTK
// snapshot every --snapshot-every steps) on the fused right-hand side and the fused
// diagnostics, with the snapshots written by io::AsyncVtkWriter while the stepping goes on.
// Same options (--n-x, --dt, --n-steps, --snapshot-every, --kappa, --mobility, --seed,
// --out-dir), same diagnostic lines, same psi_NNNNNN.vtk files byte for byte.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <iomanip>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>

#include <gridcalc/core/EigenAliases.hpp>
#include <gridcalc/core/Field.hpp>
#include <gridcalc/core/Grid.hpp>
#include <gridcalc/func/Densities.hpp>
#include <gridcalc/func/Functional.hpp>
#include <gridcalc/io/AsyncVtkWriter.hpp>
#include <gridcalc/solve/CahnHilliard.hpp>
#include <gridcalc/solve/Integrator.hpp>

using gridcalc::core::Field;
using gridcalc::core::Grid;
using gridcalc::core::Vec3d;
namespace func = gridcalc::func;

namespace {

struct Options {
  int nx = 64, n_steps = 10000, snapshot_every = 500;
  double dt = 0.01, kappa = 1.0, mobility = 1.0;
  std::uint64_t seed = 42;
  std::string out_dir = "./out";
};

Options parse(int argc, char** argv) {
  Options o;
  for (int i = 1; i + 1 < argc; i += 2) {
    const std::string f = argv[i], v = argv[i + 1];
    if (f == "--n-x") o.nx = std::stoi(v);
    else if (f == "--dt") o.dt = std::stod(v);
    else if (f == "--n-steps") o.n_steps = std::stoi(v);
    else if (f == "--snapshot-every") o.snapshot_every = std::stoi(v);
    else if (f == "--kappa") o.kappa = std::stod(v);
    else if (f == "--mobility") o.mobility = std::stod(v);
    else if (f == "--seed") o.seed = std::stoull(v);
    else if (f == "--out-dir") o.out_dir = v;
    else throw std::invalid_argument("ch_snapshots: unknown option " + f);
  }
  return o;
}

// The demo's initial state: psi ~ U(-amp, amp) from std::mt19937_64(seed), drawn in flat
// index order, then its arithmetic mean subtracted.
Field<double> randomIc(const Grid& g, std::uint64_t seed, double amp) {
  Field<double> psi(g);
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> dist(-amp, amp);
  double sum = 0.0;
  double* p = psi.data();
  for (std::size_t i = 0; i < g.getSize(); ++i) {
    p[i] = dist(rng);
    sum += p[i];
  }
  const double mean = sum / static_cast<double>(g.getSize());
  for (std::size_t i = 0; i < g.getSize(); ++i) p[i] -= mean;
  return psi;
}

void diagnostics(int step, double t, const Field<double>& psi, double kappa) {
  const auto& g = psi.getGrid();
  const double F = func::evaluate(psi, func::CahnHilliardDensity{kappa});
  const double E = func::evaluate(psi, func::GradientEnergyDensity{});
  const double mean = func::evaluate(psi, func::ValueDensity{}) /
                      (static_cast<double>(g.getSize()) * g.getCellVolume());
  double m = 0.0;
  const double* p = psi.data();
  for (std::size_t i = 0; i < g.getSize(); ++i) m = std::abs(p[i]) > m ? std::abs(p[i]) : m;
  std::printf("step=%6d  t=%9.4f  F=%+.6e  E_grad=%.6e  <psi>=%+.3e  max|psi|=%.4f\n", step, t, F,
              E, mean, m);
  std::fflush(stdout);
}

std::string snapshotPath(const std::string& dir, int step) {
  std::ostringstream name;
  name << "psi_" << std::setw(6) << std::setfill('0') << step << ".vtk";
  return (std::filesystem::path(dir) / name.str()).string();
}

}  // namespace

int main(int argc, char** argv) {
  try {
    const Options o = parse(argc, argv);
    std::filesystem::create_directories(o.out_dir);
    const Grid grid(o.nx, o.nx, o.nx, Vec3d(1.0, 1.0, 1.0));
    Field<double> psi = randomIc(grid, o.seed, 0.05);
    gridcalc::io::AsyncVtkWriter writer;
    diagnostics(0, 0.0, psi, o.kappa);
    writer.write(psi, snapshotPath(o.out_dir, 0), "psi");
    const auto t0 = std::chrono::steady_clock::now();
    int step = 0;
    while (step < o.n_steps) {
      const int chunk = std::min(o.snapshot_every, o.n_steps - step);
      gridcalc::solve::integrate(psi, gridcalc::solve::CahnHilliardRhs{o.mobility, o.kappa}, o.dt,
                                 chunk, gridcalc::solve::RK4{});
      step += chunk;
      diagnostics(step, static_cast<double>(step) * o.dt, psi, o.kappa);
      writer.write(psi, snapshotPath(o.out_dir, step), "psi");
    }
    writer.flush();
    const auto t1 = std::chrono::steady_clock::now();
    std::printf("wall_clock=%gs\n", std::chrono::duration<double>(t1 - t0).count());
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "ch_snapshots: %s\n", e.what());
    return 1;
  }
}
