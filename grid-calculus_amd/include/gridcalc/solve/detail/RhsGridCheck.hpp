// requireSameGrid (solve/detail/RhsGridCheck.hpp:42-64 of the reference): bit-exact Grid
// comparison, with the reference's message naming the producing stage.
#pragma once

#include <stdexcept>
#include <string>

#include <gridcalc/core/Field.hpp>
#include <gridcalc/core/Grid.hpp>

namespace gridcalc::solve::detail {

inline void requireSameGrid(const core::Grid& a, const core::Grid& b, const char* context) {
  if (a == b) return;
  const auto& ha = a.getCellSize();
  const auto& hb = b.getCellSize();
  throw std::invalid_argument(
      std::string("gridcalc: ") + context +
      " produced a Field on a different Grid than expected. "
      "Expected (Nx, Ny, Nz, hx, hy, hz) = (" +
      std::to_string(a.getNx()) + ", " + std::to_string(a.getNy()) + ", " +
      std::to_string(a.getNz()) + ", " + std::to_string(ha(0)) + ", " + std::to_string(ha(1)) +
      ", " + std::to_string(ha(2)) + "); got (" + std::to_string(b.getNx()) + ", " +
      std::to_string(b.getNy()) + ", " + std::to_string(b.getNz()) + ", " +
      std::to_string(hb(0)) + ", " + std::to_string(hb(1)) + ", " + std::to_string(hb(2)) +
      "). The integrator loops over psi.getSize() against the returned "
      "field's data() pointer; a Grid mismatch silently misaligns "
      "(or reads past the end of) the buffer.");
}

template <class A, class B>
inline void requireSameGrid(const core::Field<A>& expected, const core::Field<B>& actual,
                            const char* context) {
  requireSameGrid(expected.getGrid(), actual.getGrid(), context);
}

}  // namespace gridcalc::solve::detail
