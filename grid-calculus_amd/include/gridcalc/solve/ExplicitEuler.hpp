// Explicit Euler (solve/ExplicitEuler.hpp:37-135 of the reference): psi += dt * rhs(psi)
// per step.  The update runs on the device (gc_euler_update); a recognised
// solve::CahnHilliardRhs runs the fused Cahn-Hilliard pipeline instead.
#pragma once

#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

#include <gridcalc/core/Field.hpp>
#include <gridcalc/hip/Runtime.hpp>
#include <gridcalc/solve/CahnHilliard.hpp>
#include <gridcalc/solve/detail/RhsGridCheck.hpp>

namespace gridcalc::solve {

struct ExplicitEuler {
  static constexpr double diffusionCFLLimit = 0.5;
};

namespace detail {
template <class>
struct deferred_false : std::false_type {};

template <class T>
inline constexpr bool is_supported_scalar_v = std::is_same_v<T, double> || std::is_same_v<T, float>;
}  // namespace detail

template <class T, class Rhs, std::enable_if_t<detail::is_supported_scalar_v<T>, int> = 0>
inline void integrate(core::Field<T>& psi, Rhs&& rhs, double dt, int n_steps, ExplicitEuler tag) {
  (void)tag;
  static_assert(std::is_invocable_r_v<core::Field<T>, Rhs&, const core::Field<T>&>,
                "solve::integrate(... ExplicitEuler): rhs must be invocable as "
                "Field<double> rhs(const Field<double>&)");
  if (n_steps < 0)
    throw std::invalid_argument("solve::integrate(... ExplicitEuler): n_steps must be >= 0; got " +
                                std::to_string(n_steps));
  if (dt < 0.0)
    throw std::invalid_argument("solve::integrate(... ExplicitEuler): dt must be >= 0; got " +
                                std::to_string(dt));
  if constexpr (std::is_same_v<std::decay_t<Rhs>, CahnHilliardRhs>) {
    detail::integrateCahnHilliard(psi, rhs, dt, n_steps, GC_EULER);
  } else {
    const std::size_t N = psi.getSize();
    for (int step = 0; step < n_steps; ++step) {
      core::Field<T> tmp = rhs(static_cast<const core::Field<T>&>(psi));
      detail::requireSameGrid(psi, tmp, "solve::integrate(... ExplicitEuler): rhs(psi)");
      hip::check(gc_euler_update(hip::context(), hip::dtype_of<T>::value, N,
                                 psi.deviceDataForWrite(), dt, tmp.deviceData()));
    }
  }
}

template <class T, class Rhs>
inline void explicitEuler(core::Field<T>& psi, Rhs&& rhs, double dt, int n_steps) {
  integrate(psi, std::forward<Rhs>(rhs), dt, n_steps, ExplicitEuler{});
}

}  // namespace gridcalc::solve
