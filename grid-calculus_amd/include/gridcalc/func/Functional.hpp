// func::evaluate (func/Functional.hpp:88-220 of the reference): integrate a density
// f(psi, grad psi, lap psi, biharm psi) (arity 4/3/2/1, detected at compile time) or
// f(u, grad u) / f(u) for a Vec3 field.  The derivatives (Order 2) are computed on the
// device; a general density is a host callable, so the integrand is assembled on the host
// and reduced on the device.  func::CahnHilliardDensity{kappa} -- the free-energy density
// -p^2/2 + p^4/4 + kappa/2 |grad p|^2 of examples/common/cahn_hilliard.hpp:64-68 -- is
// recognised and evaluated by one fused kernel (gradient + density + reduction).
#pragma once

#include <cstddef>
#include <type_traits>

#include <gridcalc/core/EigenAliases.hpp>
#include <gridcalc/core/Field.hpp>
#include <gridcalc/diff/Biharmonic.hpp>
#include <gridcalc/diff/Gradient.hpp>
#include <gridcalc/diff/Laplacian.hpp>
#include <gridcalc/func/Integrate.hpp>

namespace gridcalc::func {

namespace detail {
template <class>
struct deferred_false : std::false_type {};

template <class Tag>
constexpr int methodOf() {
  return std::is_same_v<Tag, Kahan> ? GC_KAHAN : GC_PAIRWISE;
}
}  // namespace detail

// Cahn-Hilliard free-energy density with gradient coefficient kappa.
struct CahnHilliardDensity {
  double kappa;
  double operator()(double p, const core::Vec3d& g) const {
    return -0.5 * p * p + 0.25 * p * p * p * p + 0.5 * kappa * g.squaredNorm();
  }
};

template <class T, class Tag = Pairwise>
inline double evaluate(const core::Field<T>& psi, const CahnHilliardDensity& f, Tag tag = {}) {
  static_assert(std::is_same_v<Tag, Pairwise> || std::is_same_v<Tag, Kahan>,
                "func::evaluate: Tag must be func::Pairwise or func::Kahan");
  (void)tag;
  const auto& grid = psi.getGrid();
  diff::detail::requireAxisExtent("x", grid.getNx(), 1);
  diff::detail::requireAxisExtent("y", grid.getNy(), 1);
  diff::detail::requireAxisExtent("z", grid.getNz(), 1);
  const gc_grid g = hip::toGrid(grid);
  double F = 0.0;
  hip::check(gc_ch_free_energy(hip::context(), &g, hip::dtype_of<T>::value, f.kappa,
                               detail::methodOf<Tag>(), psi.deviceData(), &F));
  return F;
}

template <class F, class Tag = Pairwise,
          std::enable_if_t<!std::is_same_v<std::decay_t<F>, CahnHilliardDensity>, int> = 0>
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

template <class F, class Tag = Pairwise>
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
