// func::evaluate (func/Functional.hpp:88-220 of the reference): integrate a density
// f(psi, grad psi, lap psi, biharm psi) (arity 4/3/2/1, detected at compile time) or
// f(u, grad u) / f(u) for a Vec3 field.  A density of func/Densities.hpp (value, gradient
// energy, Cahn-Hilliard, Ginzburg-Landau, phase-field crystal, linear elastic) is evaluated
// on the device in one fused pass (derivatives + density + reduction), with the bits the
// host evaluation of the same formula gives.  Any other callable runs on the host: the
// Order-2 derivatives are computed on the device, the integrand is assembled on the host and
// reduced on the device.
#pragma once

#include <cstddef>
#include <type_traits>

#include <gridcalc/core/EigenAliases.hpp>
#include <gridcalc/core/Field.hpp>
#include <gridcalc/diff/Biharmonic.hpp>
#include <gridcalc/diff/Gradient.hpp>
#include <gridcalc/diff/Laplacian.hpp>
#include <gridcalc/func/Densities.hpp>
#include <gridcalc/func/Integrate.hpp>

namespace gridcalc::func {

namespace detail {
template <class>
struct deferred_false : std::false_type {};

template <class Tag>
constexpr int methodOf() {
  return std::is_same_v<Tag, Kahan> ? GC_KAHAN : GC_PAIRWISE;
}

template <class Tag>
inline double evaluateOnDevice(const core::Grid& grid, gc_dtype dt, const gc_density& d,
                               const void* field) {
  static_assert(std::is_same_v<Tag, Pairwise> || std::is_same_v<Tag, Kahan>,
                "func::evaluate: Tag must be func::Pairwise or func::Kahan");
  const gc_grid g = hip::toGrid(grid);
  double r = 0.0;
  hip::check(gc_evaluate(hip::context(), &g, dt, &d, methodOf<Tag>(), field, &r));
  return r;
}
}  // namespace detail

// Recognised scalar densities, FP64 or FP32 fields: one fused device pass.
template <class T, class D, class Tag = Pairwise,
          std::enable_if_t<is_device_density_v<D>, int> = 0>
inline double evaluate(const core::Field<T>& psi, const D& f, Tag tag = {}) {
  (void)tag;
  return detail::evaluateOnDevice<Tag>(psi.getGrid(), hip::dtype_of<T>::value, f.desc(),
                                       psi.deviceData());
}

// The linear elastic energy density of a displacement field: one fused device pass.
template <class Tag = Pairwise>
inline double evaluate(const core::Field<core::Vec3d>& u, const LinearElasticDensity& f,
                       Tag tag = {}) {
  (void)tag;
  return detail::evaluateOnDevice<Tag>(u.getGrid(), GC_F64, f.desc(), u.deviceData());
}

template <class F, class Tag = Pairwise, std::enable_if_t<!is_device_density_v<F>, int> = 0>
inline double evaluate(const core::Field<double>& psi, F&& f, Tag tag = {}) {
  static_assert(std::is_same_v<Tag, Pairwise> || std::is_same_v<Tag, Kahan>,
                "func::evaluate: Tag must be func::Pairwise or func::Kahan");
  const auto& grid = psi.getGrid();
  const std::size_t n = grid.getSize();
  core::Field<double> integrand(grid);
  double* out = integrand.data();
  const double* p = psi.data();
  if constexpr (std::is_invocable_r_v<double, F&, double, const core::Vec3d&, double, double>) {
    const auto grad = diff::gradient(psi);
    const auto lap = diff::laplacian(psi);
    const auto biharm = diff::biharmonic(psi);
    const core::Vec3d* gp = grad.data();
    const double* lp = lap.data();
    const double* bp = biharm.data();
    for (std::size_t c = 0; c < n; ++c) out[c] = static_cast<double>(f(p[c], gp[c], lp[c], bp[c]));
  } else if constexpr (std::is_invocable_r_v<double, F&, double, const core::Vec3d&, double>) {
    const auto grad = diff::gradient(psi);
    const auto lap = diff::laplacian(psi);
    const core::Vec3d* gp = grad.data();
    const double* lp = lap.data();
    for (std::size_t c = 0; c < n; ++c) out[c] = static_cast<double>(f(p[c], gp[c], lp[c]));
  } else if constexpr (std::is_invocable_r_v<double, F&, double, const core::Vec3d&>) {
    const auto grad = diff::gradient(psi);
    const core::Vec3d* gp = grad.data();
    for (std::size_t c = 0; c < n; ++c) out[c] = static_cast<double>(f(p[c], gp[c]));
  } else if constexpr (std::is_invocable_r_v<double, F&, double>) {
    for (std::size_t c = 0; c < n; ++c) out[c] = static_cast<double>(f(p[c]));
  } else {
    static_assert(detail::deferred_false<F>::value,
                  "func::evaluate: f must be invocable as f(double), "
                  "f(double, const Vec3d&), f(double, const Vec3d&, double), "
                  "or f(double, const Vec3d&, double, double), "
                  "with a return type convertible to double");
  }
  return integrate(integrand, tag);
}

template <class F, class Tag = Pairwise,
          std::enable_if_t<!std::is_same_v<std::decay_t<F>, LinearElasticDensity>, int> = 0>
inline double evaluate(const core::Field<core::Vec3d>& u, F&& f, Tag tag = {}) {
  static_assert(std::is_same_v<Tag, Pairwise> || std::is_same_v<Tag, Kahan>,
                "func::evaluate: Tag must be func::Pairwise or func::Kahan");
  const auto& grid = u.getGrid();
  const std::size_t n = grid.getSize();
  core::Field<double> integrand(grid);
  double* out = integrand.data();
  const core::Vec3d* up = u.data();
  if constexpr (std::is_invocable_r_v<double, F&, const core::Vec3d&, const core::Mat3d&>) {
    const auto grad_u = diff::gradient(u);
    const core::Mat3d* gp = grad_u.data();
    for (std::size_t c = 0; c < n; ++c) out[c] = static_cast<double>(f(up[c], gp[c]));
  } else if constexpr (std::is_invocable_r_v<double, F&, const core::Vec3d&>) {
    for (std::size_t c = 0; c < n; ++c) out[c] = static_cast<double>(f(up[c]));
  } else {
    static_assert(detail::deferred_false<F>::value,
                  "func::evaluate: f must be invocable as f(const Vec3d&) "
                  "or f(const Vec3d&, const Mat3d&), with a return type "
                  "convertible to double");
  }
  return integrate(integrand, tag);
}

}  // namespace gridcalc::func
