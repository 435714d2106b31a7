// The named fourth-order partials (diff/FourthOrder.hpp of the reference): each is
// detail::mixedDerivative<Nx, Ny, Nz, Order> on the device (gc_mixed_partial), Order = 2 by
// default as in the reference, for Field<double> (bit-identical) and Field<float>.
#pragma once

#include <gridcalc/core/Field.hpp>
#include <gridcalc/diff/detail/MultiIndexDerivative.hpp>

namespace gridcalc::diff {

template <int Order = 2, class T>
inline core::Field<T> d4dx4(const core::Field<T>& psi) {
  return detail::mixedDerivative<4, 0, 0, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d4dy4(const core::Field<T>& psi) {
  return detail::mixedDerivative<0, 4, 0, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d4dz4(const core::Field<T>& psi) {
  return detail::mixedDerivative<0, 0, 4, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d4dx3dy(const core::Field<T>& psi) {
  return detail::mixedDerivative<3, 1, 0, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d4dxdy3(const core::Field<T>& psi) {
  return detail::mixedDerivative<1, 3, 0, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d4dx3dz(const core::Field<T>& psi) {
  return detail::mixedDerivative<3, 0, 1, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d4dxdz3(const core::Field<T>& psi) {
  return detail::mixedDerivative<1, 0, 3, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d4dy3dz(const core::Field<T>& psi) {
  return detail::mixedDerivative<0, 3, 1, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d4dydz3(const core::Field<T>& psi) {
  return detail::mixedDerivative<0, 1, 3, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d4dx2dy2(const core::Field<T>& psi) {
  return detail::mixedDerivative<2, 2, 0, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d4dx2dz2(const core::Field<T>& psi) {
  return detail::mixedDerivative<2, 0, 2, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d4dy2dz2(const core::Field<T>& psi) {
  return detail::mixedDerivative<0, 2, 2, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d4dx2dydz(const core::Field<T>& psi) {
  return detail::mixedDerivative<2, 1, 1, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d4dxdy2dz(const core::Field<T>& psi) {
  return detail::mixedDerivative<1, 2, 1, Order>(psi);
}

template <int Order = 2, class T>
inline core::Field<T> d4dxdydz2(const core::Field<T>& psi) {
  return detail::mixedDerivative<1, 1, 2, Order>(psi);
}

}  // namespace gridcalc::diff
