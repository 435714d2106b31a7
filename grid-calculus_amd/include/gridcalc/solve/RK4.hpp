// Classic RK4 (solve/RK4.hpp:30-154 of the reference).  For a general right-hand side the
// stage inputs (axpyFresh) and the final combination psi += dt/6 (k1+k4) + dt/3 (k2+k3)
// run on the device with the reference's roundings; a recognised solve::CahnHilliardRhs
// runs the fused stage pipeline (one kernel per stage) instead.
#pragma once

#include <stdexcept>
#include <string>
#include <type_traits>

#include <gridcalc/core/Field.hpp>
#include <gridcalc/hip/Runtime.hpp>
#include <gridcalc/solve/CahnHilliard.hpp>
#include <gridcalc/solve/ExplicitEuler.hpp>
#include <gridcalc/solve/detail/RhsGridCheck.hpp>

namespace gridcalc::solve {

struct RK4 {
  static constexpr double diffusionCFLLimit = 0.6963;
};

namespace detail {

// out = a + scale * b (RK4.hpp:51-64)
template <class T>
inline core::Field<T> axpyFresh(const core::Field<T>& a, double scale, const core::Field<T>& b) {
  requireSameGrid(a, b, "solve::detail::axpyFresh");
  core::Field<T> out(a.getGrid(), core::detail::device_only);
  hip::check(gc_axpy(hip::context(), hip::dtype_of<T>::value, out.getSize(), a.deviceData(),
                     scale, b.deviceData(), out.deviceDataDiscard()));
  return out;
}

}  // namespace detail

template <class T, class Rhs, std::enable_if_t<detail::is_supported_scalar_v<T>, int> = 0>
inline void integrate(core::Field<T>& psi, Rhs&& rhs, double dt, int n_steps, RK4 tag) {
  (void)tag;
  static_assert(std::is_invocable_r_v<core::Field<T>, Rhs&, const core::Field<T>&>,
                "solve::integrate(... RK4): rhs must be invocable as "
                "Field<double> rhs(const Field<double>&)");
  if (n_steps < 0)
    throw std::invalid_argument("solve::integrate(... RK4): n_steps must be >= 0; got " +
                                std::to_string(n_steps));
  if (dt < 0.0)
    throw std::invalid_argument("solve::integrate(... RK4): dt must be >= 0; got " +
                                std::to_string(dt));
  if constexpr (std::is_same_v<std::decay_t<Rhs>, CahnHilliardRhs>) {
    detail::integrateCahnHilliard(psi, rhs, dt, n_steps, GC_RK4);
  } else {
    const double half_dt = 0.5 * dt;
    for (int step = 0; step < n_steps; ++step) {
      core::Field<T> k1 = rhs(static_cast<const core::Field<T>&>(psi));
      detail::requireSameGrid(psi, k1, "solve::integrate(... RK4): k1");
      core::Field<T> y2 = detail::axpyFresh(psi, half_dt, k1);
      core::Field<T> k2 = rhs(static_cast<const core::Field<T>&>(y2));
      detail::requireSameGrid(psi, k2, "solve::integrate(... RK4): k2");
      core::Field<T> y3 = detail::axpyFresh(psi, half_dt, k2);
      core::Field<T> k3 = rhs(static_cast<const core::Field<T>&>(y3));
      detail::requireSameGrid(psi, k3, "solve::integrate(... RK4): k3");
      core::Field<T> y4 = detail::axpyFresh(psi, dt, k3);
      core::Field<T> k4 = rhs(static_cast<const core::Field<T>&>(y4));
      detail::requireSameGrid(psi, k4, "solve::integrate(... RK4): k4");
      hip::check(gc_rk4_combine(hip::context(), hip::dtype_of<T>::value, psi.getSize(),
                                psi.deviceDataForWrite(), dt, k1.deviceData(), k2.deviceData(),
                                k3.deviceData(), k4.deviceData()));
    }
  }
}

}  // namespace gridcalc::solve
