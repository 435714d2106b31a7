// solve::CahnHilliardRhs{M, kappa}: the Cahn-Hilliard right-hand side
// M * lap(-psi + psi^3 - kappa * lap psi) with Order-2 Laplacians, as
// examples/common/cahn_hilliard.hpp:38-59 of the reference computes it.  Called directly
// it returns the right-hand side (one fused kernel); passed to solve::integrate it selects
// the fused stage pipeline, in which each RK4 stage (or Euler step) is one kernel and one
// pass over the fields.  Results are bit-identical to the reference's arithmetic in double
// precision.
#pragma once

#include <stdexcept>
#include <string>

#include <gridcalc/core/Field.hpp>
#include <gridcalc/diff/detail/PreconditionAxisExtent.hpp>
#include <gridcalc/hip/Runtime.hpp>

namespace gridcalc::solve {

struct CahnHilliardRhs {
  double M = 1.0;
  double kappa = 1.0;

  template <class T>
  core::Field<T> operator()(const core::Field<T>& psi) const {
    const auto& grid = psi.getGrid();
    diff::detail::requireAxisExtent("x", grid.getNx(), 1);
    diff::detail::requireAxisExtent("y", grid.getNy(), 1);
    diff::detail::requireAxisExtent("z", grid.getNz(), 1);
    core::Field<T> out(grid, core::detail::device_only);
    const gc_grid g = hip::toGrid(grid);
    const gc_ch_params p{M, kappa};
    hip::check(gc_ch_rhs(hip::context(), &g, hip::dtype_of<T>::value, &p, psi.deviceData(),
                         out.deviceDataDiscard()));
    return out;
  }
};

namespace detail {
template <class T>
inline void integrateCahnHilliard(core::Field<T>& psi, const CahnHilliardRhs& rhs, double dt,
                                  int n_steps, gc_integrator integrator) {
  if (n_steps == 0) return;
  const auto& grid = psi.getGrid();
  diff::detail::requireAxisExtent("x", grid.getNx(), 1);
  diff::detail::requireAxisExtent("y", grid.getNy(), 1);
  diff::detail::requireAxisExtent("z", grid.getNz(), 1);
  const gc_grid g = hip::toGrid(grid);
  const gc_ch_params p{rhs.M, rhs.kappa};
  hip::check(gc_ch_integrate(hip::context(), &g, hip::dtype_of<T>::value, &p,
                             psi.deviceDataForWrite(), dt, n_steps, integrator));
}
}  // namespace detail

}  // namespace gridcalc::solve
