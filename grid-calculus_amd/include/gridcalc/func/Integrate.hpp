// func::integrate (func/Integrate.hpp:43-180 of the reference): cellVolume * sum over the
// field, Pairwise (the reference's recursion tree, reproduced exactly on the device) or
// Kahan (compensated; a parallel Neumaier/TwoSum reduction in double); and over a
// double-valued tensor::expr expression, integrate(materialize(expr)) as the reference defines
// it (the expression is evaluated on the device in one pass).  As in the reference the Field
// overloads are noexcept: a device failure prints the library's message and aborts.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <exception>
#include <type_traits>

#include <gridcalc/core/Field.hpp>
#include <gridcalc/hip/Runtime.hpp>
#include <gridcalc/tensor/Expressions.hpp>

namespace gridcalc::func {

struct Pairwise {};
struct Kahan {};

namespace detail {

template <class T>
inline double deviceIntegrate(const core::Field<T>& field, int method) noexcept {
  try {
    const gc_grid g = hip::toGrid(field.getGrid());
    double r = 0.0;
    hip::check(gc_integrate(hip::context(), &g, hip::dtype_of<T>::value, method,
                            field.deviceData(), &r));
    return r;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "func::integrate: %s\n", e.what());
    std::abort();
  }
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

// func/Integrate.hpp:172-180: the integral of a double-valued expression
template <class E, class Tag = Pairwise,
          std::enable_if_t<tensor::expr::is_expr_v<E> &&
                               std::is_same_v<typename std::decay_t<E>::value_type, double>,
                           int> = 0>
inline double integrate(const E& expr, Tag tag = {}) {
  static_assert(std::is_same_v<Tag, Pairwise> || std::is_same_v<Tag, Kahan>,
                "func::integrate(expr): Tag must be func::Pairwise or func::Kahan");
  return integrate(tensor::expr::materialize(expr), tag);
}

}  // namespace gridcalc::func
