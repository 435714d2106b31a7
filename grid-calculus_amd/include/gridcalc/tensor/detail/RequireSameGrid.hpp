// tensor::detail::requireSameGrid (tensor/detail/RequireSameGrid.hpp of the reference):
// bit-exact Grid comparison for the contraction operands, with the reference's message.
#pragma once

#include <stdexcept>
#include <string>

#include <gridcalc/core/Field.hpp>
#include <gridcalc/core/Grid.hpp>

namespace gridcalc::tensor::detail {

inline std::string gridText(const core::Grid& g) {
  const auto& h = g.getCellSize();
  return "(" + std::to_string(g.getNx()) + ", " + std::to_string(g.getNy()) + ", " +
         std::to_string(g.getNz()) + ", " + std::to_string(h(0)) + ", " + std::to_string(h(1)) +
         ", " + std::to_string(h(2)) + ")";
}

inline void requireSameGrid(const core::Grid& lhs, const core::Grid& rhs, const char* context) {
  if (lhs == rhs) return;
  throw std::invalid_argument(std::string("gridcalc: ") + context +
                              " requires both operands to share the same Grid, but they do not. "
                              "lhs (Nx, Ny, Nz, hx, hy, hz) = " +
                              gridText(lhs) + "; rhs = " + gridText(rhs) + ".");
}

template <class T>
inline void requireSameGrid(const core::Field<T>& lhs, const core::Field<T>& rhs,
                            const char* context) {
  if (lhs.getGrid() == rhs.getGrid()) return;
  throw std::invalid_argument(std::string("gridcalc: ") + context +
                              " requires both operands to share the same Grid, but they do not. "
                              "lhs (Nx, Ny, Nz, hx, hy, hz) = " +
                              gridText(lhs.getGrid()) + "; rhs = " + gridText(rhs.getGrid()) +
                              ". The flat-index loop would silently misalign or read past the "
                              "end of the smaller buffer.");
}

}  // namespace gridcalc::tensor::detail
