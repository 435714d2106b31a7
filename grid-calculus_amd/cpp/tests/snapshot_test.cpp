// io::AsyncVtkWriter through the C++ header API, on the device:
//  - its files equal, byte for byte, the legacy ASCII VTK the reference's
//    examples/common/vtk_writer.hpp:34-67 writes (restated below with std::ofstream and
//    precision(17)), for device-current and host-current fields and awkward values;
//  - a snapshot holds the field as it was when write() was called, while the stepping that
//    follows at once changes the field;
//  - an unwritable path surfaces as std::runtime_error from flush().
#include <cmath>
#include <filesystem>
#include <fstream>
#include <limits>
#include <sstream>
#include <string>

#include <unistd.h>

#include <gridcalc/core/EigenAliases.hpp>
#include <gridcalc/core/Field.hpp>
#include <gridcalc/core/Grid.hpp>
#include <gridcalc/diff/Laplacian.hpp>
#include <gridcalc/io/AsyncVtkWriter.hpp>
#include <gridcalc/solve/CahnHilliard.hpp>
#include <gridcalc/solve/Integrator.hpp>

#include "check.hpp"

using gridcalc::core::Field;
using gridcalc::core::Grid;
using gridcalc::core::Vec3d;
namespace fs = std::filesystem;

namespace {

// the reference writer's format, written the way it writes it
void referenceWrite(const Field<double>& f, const std::string& path, const std::string& name) {
  const Grid& g = f.getGrid();
  const auto& h = g.getCellSize();
  std::ofstream out(path);
  out.precision(17);
  out << "# vtk DataFile Version 3.0\n";
  out << "gridcalc field: " << name << "\n";
  out << "ASCII\n";
  out << "DATASET STRUCTURED_POINTS\n";
  out << "DIMENSIONS " << g.getNx() << ' ' << g.getNy() << ' ' << g.getNz() << '\n';
  out << "ORIGIN 0 0 0\n";
  out << "SPACING " << h(0) << ' ' << h(1) << ' ' << h(2) << '\n';
  out << "POINT_DATA " << g.getSize() << '\n';
  out << "SCALARS " << name << " double 1\n";
  out << "LOOKUP_TABLE default\n";
  const double* d = f.data();
  for (std::size_t i = 0; i < g.getSize(); ++i) out << d[i] << '\n';
}

std::string slurp(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  std::ostringstream s;
  s << in.rdbuf();
  return s.str();
}

fs::path scratchDir() {
  const fs::path d = fs::temp_directory_path() / ("gc_snapshot_test_" + std::to_string(::getpid()));
  fs::create_directories(d);
  return d;
}

}  // namespace

TEST(AsyncVtkWriter, SameBytesAsTheReferenceWriter) {
  const fs::path dir = scratchDir();
  const Grid g(7, 5, 3, Vec3d(0.1, 1.0 / 3.0, 2.5));
  Field<double> f(g, [](double x, double y, double z) { return std::sin(37.0 * x + y) * std::exp(3.0 * z); });
  double* p = f.data();
  p[0] = -0.0;
  p[1] = 1e-5;
  p[2] = 3.0;
  p[3] = -1e300;
  p[4] = 0.1;
  p[5] = std::numeric_limits<double>::denorm_min();
  p[6] = 123456789012345678.0;
  {
    gridcalc::io::AsyncVtkWriter w;
    w.write(f, (dir / "host.vtk").string(), "psi");         // host-current field
    const auto lap = gridcalc::diff::laplacian(f);           // device-only result
    w.write(lap, (dir / "device.vtk").string(), "lap");
    w.flush();
    referenceWrite(f, (dir / "host_ref.vtk").string(), "psi");
    referenceWrite(lap, (dir / "device_ref.vtk").string(), "lap");
  }
  EXPECT_TRUE(slurp((dir / "host.vtk").string()) == slurp((dir / "host_ref.vtk").string()));
  EXPECT_TRUE(slurp((dir / "device.vtk").string()) == slurp((dir / "device_ref.vtk").string()));
  fs::remove_all(dir);
}

TEST(AsyncVtkWriter, SnapshotIsTakenAtWrite) {
  const fs::path dir = scratchDir();
  const Grid g(48, 40, 24, Vec3d(1.0, 1.0, 1.0));
  Field<double> psi(g, [](double x, double y, double z) { return 0.05 * std::sin(0.4 * x + 0.2 * y) * std::cos(0.3 * z); });
  const gridcalc::solve::CahnHilliardRhs rhs{1.0, 1.0};
  gridcalc::solve::integrate(psi, rhs, 0.01, 3, gridcalc::solve::RK4{});  // device-current
  std::vector<Field<double>> states;
  {
    gridcalc::io::AsyncVtkWriter w(2);
    for (int s = 0; s < 5; ++s) {
      states.push_back(psi);  // a device copy of the state being written
      w.write(psi, (dir / ("s" + std::to_string(s) + ".vtk")).string(), "psi");
      gridcalc::solve::integrate(psi, rhs, 0.01, 4, gridcalc::solve::RK4{});  // at once
    }
  }  // the destructor flushes
  for (int s = 0; s < 5; ++s) {
    referenceWrite(states[s], (dir / ("r" + std::to_string(s) + ".vtk")).string(), "psi");
    EXPECT_TRUE(slurp((dir / ("s" + std::to_string(s) + ".vtk")).string()) ==
                slurp((dir / ("r" + std::to_string(s) + ".vtk")).string()));
  }
  fs::remove_all(dir);
}

TEST(AsyncVtkWriter, UnwritablePathRaisesOnFlush) {
  Field<double> f(Grid(4, 4, 4, Vec3d(1.0, 1.0, 1.0)), 1.0);
  gridcalc::io::AsyncVtkWriter w;
  w.write(f, "/nonexistent_dir_for_gridcalc/psi.vtk", "psi");
  EXPECT_THROW_MSG(w.flush(), std::runtime_error, "cannot open");
}

CHECK_MAIN()
