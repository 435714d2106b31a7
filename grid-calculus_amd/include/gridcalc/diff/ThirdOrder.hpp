// The named third-order partials (diff/ThirdOrder.hpp of the reference): each is
// detail::mixedDerivative<Nx, Ny, Nz, Order> on the device (gc_mixed_partial), Order = 2 by
// default as in the reference, for Field<double> (bit-identical) and Field<float>.
#pragma once

#include <gridcalc/core/Field.hpp>
#include <gridcalc/diff/detail/MultiIndexDerivative.hpp>

namespace gridcalc::diff {

template <int Order = 2, class T>
inline core::Field<T> d3dx3(const core::Field<T>& psi) {
  return detail::mixedDerivative<3, 0, 0, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d3dy3(const core::Field<T>& psi) {
  return detail::mixedDerivative<0, 3, 0, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d3dz3(const core::Field<T>& psi) {
  return detail::mixedDerivative<0, 0, 3, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d3dx2dy(const core::Field<T>& psi) {
  return detail::mixedDerivative<2, 1, 0, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d3dxdy2(const core::Field<T>& psi) {
  return detail::mixedDerivative<1, 2, 0, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d3dx2dz(const core::Field<T>& psi) {
  return detail::mixedDerivative<2, 0, 1, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d3dxdz2(const core::Field<T>& psi) {
  return detail::mixedDerivative<1, 0, 2, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d3dy2dz(const core::Field<T>& psi) {
  return detail::mixedDerivative<0, 2, 1, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d3dydz2(const core::Field<T>& psi) {
  return detail::mixedDerivative<0, 1, 2, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d3dxdydz(const core::Field<T>& psi) {
  return detail::mixedDerivative<1, 1, 1, Order>(psi);
}

}  // namespace gridcalc::diff
