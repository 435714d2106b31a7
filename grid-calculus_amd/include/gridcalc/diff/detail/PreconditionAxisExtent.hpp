// requireAxisExtent (diff/detail/PreconditionAxisExtent.hpp:64-79 of the reference): a
// periodic axis of extent 2 <= N <= 2r aliases the [-r, r] offsets; N = 1 (degenerate
// axis, e.g. 2-D as Nz = 1) and N >= 2r+1 are accepted.  Same message as the reference.
#pragma once

#include <stdexcept>
#include <string>

namespace gridcalc::diff::detail {

inline void requireAxisExtent(const char* axis_label, int N, int radius) {
  if (radius == 0) return;
  if (N == 1) return;
  const int min_N = 2 * radius + 1;
  if (N < min_N) {
    throw std::invalid_argument(
        "gridcalc: axis " + std::string(axis_label) + " has N=" + std::to_string(N) +
        " but the requested stencil has radius " + std::to_string(radius) +
        ", which requires N>=" + std::to_string(min_N) +
        " on a periodic grid (or N=1 for a degenerate axis); "
        "intermediate extents alias the [-r, r] offsets under the "
        "periodic wrap and the stencil silently degrades. Use a larger "
        "grid or a lower-Order stencil.");
  }
}

}  // namespace gridcalc::diff::detail
