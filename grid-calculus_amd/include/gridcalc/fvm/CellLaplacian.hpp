// fvm::cellLaplacian(psi, D) = div(D grad psi) with harmonic-mean face diffusivities
// 2ab/(a+b) (fvm/CellLaplacian.hpp:48-143 of the reference), on the device.  The
// preconditions run as in the reference: radius-1 extents, same Grid for D, and D > 0 at
// every cell (first violation reported by flat index and value).
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include <gridcalc/core/Field.hpp>
#include <gridcalc/core/Grid.hpp>
#include <gridcalc/diff/detail/PreconditionAxisExtent.hpp>
#include <gridcalc/hip/Runtime.hpp>
#include <gridcalc/solve/detail/RhsGridCheck.hpp>

namespace gridcalc::fvm {

namespace detail {

inline double harmonicMean(double a, double b) noexcept { return 2.0 * a * b / (a + b); }

// The reference's message for the first cell with D <= 0 (flat index `bad`).
template <class T>
[[noreturn]] inline void throwNonPositiveD(const core::Field<T>& D, int64_t bad) {
  T value{};
  hip::download(&value, D.deviceData() + bad, sizeof(T));
  throw std::invalid_argument(
      "fvm::cellLaplacian: D must be strictly positive at every "
      "cell (the harmonic-mean face-D averaging is undefined "
      "otherwise); first violation at flat index " +
      std::to_string(static_cast<std::size_t>(bad)) + " with D=" +
      std::to_string(static_cast<double>(value)));
}

template <class T>
inline void requirePositiveD(const core::Field<T>& D) {
  const gc_grid g = hip::toGrid(D.getGrid());
  int64_t bad = -1;
  hip::check(gc_first_nonpositive(hip::context(), &g, hip::dtype_of<T>::value, D.deviceData(),
                                  &bad));
  if (bad >= 0) throwNonPositiveD(D, bad);
}

template <class T>
inline core::Field<T> cellLaplacianImpl(const core::Field<T>& psi, const core::Field<T>& D) {
  const auto& grid = psi.getGrid();
  diff::detail::requireAxisExtent("x", grid.getNx(), 1);
  diff::detail::requireAxisExtent("y", grid.getNy(), 1);
  diff::detail::requireAxisExtent("z", grid.getNz(), 1);
  solve::detail::requireSameGrid(psi, D, "fvm::cellLaplacian: D");
  core::Field<T> result(grid, core::detail::device_only);
  const gc_grid g = hip::toGrid(grid);
  int64_t bad = -1;  // the D > 0 check runs inside the operator's kernel
  const gc_status st = gc_cell_laplacian(hip::context(), &g, hip::dtype_of<T>::value,
                                         psi.deviceData(), D.deviceData(),
                                         result.deviceDataDiscard(), &bad);
  if (st == GC_ERR_NONPOS_D && bad >= 0) throwNonPositiveD(D, bad);
  hip::check(st);
  return result;
}

}  // namespace detail

inline core::Field<double> cellLaplacian(const core::Field<double>& psi,
                                         const core::Field<double>& D) {
  return detail::cellLaplacianImpl(psi, D);
}

inline core::Field<float> cellLaplacian(const core::Field<float>& psi,
                                        const core::Field<float>& D) {
  return detail::cellLaplacianImpl(psi, D);
}

}  // namespace gridcalc::fvm
