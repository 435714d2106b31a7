// func::integrate (func/Integrate.hpp:43-135 of the reference): cellVolume * sum over the
// field, Pairwise (the reference's recursion tree, reproduced exactly on the device) or
// Kahan (compensated; a parallel Neumaier/TwoSum reduction in double).  As in the
// reference the Field overloads are noexcept: a device failure terminates the program.
#pragma once

#include <gridcalc/core/Field.hpp>
#include <gridcalc/hip/Runtime.hpp>

namespace gridcalc::func {

struct Pairwise {};
struct Kahan {};

namespace detail {

template <class T>
inline double deviceIntegrate(const core::Field<T>& field, int method) noexcept {
  const gc_grid g = hip::toGrid(field.getGrid());
  double r = 0.0;
  hip::check(gc_integrate(hip::context(), &g, hip::dtype_of<T>::value, method,
                          field.deviceData(), &r));
  return r;
}

}  // namespace detail

inline double integrate(const core::Field<double>& field, Pairwise tag = {}) noexcept {
  (void)tag;
  return detail::deviceIntegrate(field, GC_PAIRWISE);
}

inline double integrate(const core::Field<double>& field, Kahan tag) noexcept {
  (void)tag;
  return detail::deviceIntegrate(field, GC_KAHAN);
}

// Single-precision fields are reduced in double precision.
inline double integrate(const core::Field<float>& field, Pairwise tag = {}) noexcept {
  (void)tag;
  return detail::deviceIntegrate(field, GC_PAIRWISE);
}

inline double integrate(const core::Field<float>& field, Kahan tag) noexcept {
  (void)tag;
  return detail::deviceIntegrate(field, GC_KAHAN);
}

}  // namespace gridcalc::func
