// diff::detail::mixedDerivative<Nx, Ny, Nz, Order> (diff/detail/MultiIndexDerivative.hpp:
// 96-136 of the reference): the tensor product of one axis stencil per axis (identity, first,
// second, third or fourth central difference at accuracy Order), summed z-outer, y, x-inner as
// wx * (wy * wz) * psi and scaled by 1/h_x^Nx 1/h_y^Ny 1/h_z^Nz, computed on the device by
// gc_mixed_partial.  Double results equal the reference's bit for bit; the float overload is
// this library's single-precision form.
#pragma once

#include <gridcalc/core/Field.hpp>
#include <gridcalc/hip/Runtime.hpp>

namespace gridcalc::diff::detail {

template <int Nx, int Ny, int Nz, int Order, class T>
inline core::Field<T> mixedDerivative(const core::Field<T>& psi) {
  static_assert(Nx >= 0 && Ny >= 0 && Nz >= 0 && Nx + Ny + Nz >= 1 && Nx + Ny + Nz <= 4,
                "mixedDerivative: derivative orders must total 1 to 4");
  static_assert(Order == 2 || Order == 4, "mixedDerivative: Order must be 2 or 4");
  const auto& grid = psi.getGrid();
  core::Field<T> result(grid, core::detail::device_only);
  const gc_grid g = hip::toGrid(grid);
  hip::check(gc_mixed_partial(hip::context(), &g, hip::dtype_of<T>::value, Nx, Ny, Nz, Order,
                              psi.deviceData(), result.deviceDataDiscard()));
  return result;
}

}  // namespace gridcalc::diff::detail
