// solve::diffuse (solve/Diffusion.hpp:45-216 of the reference).
//  - diffuse(psi, double D, dt, n, tag): d psi/dt = D lap psi, with D folded into the step
//    (the right-hand side is lap psi and the step D*dt), run by the fused stage pipeline.
//  - diffuse(psi, Field D, dt, n, tag): d psi/dt = div(D grad psi) through the cell
//    Laplacian, fused likewise.
// The explicit-diffusion CFL check (D*dt*sum 1/h^2 <= Tag::diffusionCFLLimit, the largest
// harmonic-mean face D for a field) and every precondition run as in the reference, with
// the same messages.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <typeinfo>

#include <gridcalc/core/Field.hpp>
#include <gridcalc/diff/Laplacian.hpp>
#include <gridcalc/fvm/CellLaplacian.hpp>
#include <gridcalc/hip/Runtime.hpp>
#include <gridcalc/solve/Integrator.hpp>
#include <gridcalc/solve/detail/RhsGridCheck.hpp>

namespace gridcalc::solve {

namespace detail {

template <class Tag>
inline void checkDiffusionCFL(const core::Grid& grid, double D, double dt) {
  const auto& h = grid.getCellSize();
  const double inv_sum = 1.0 / (h(0) * h(0)) + 1.0 / (h(1) * h(1)) + 1.0 / (h(2) * h(2));
  const double cfl = D * dt * inv_sum;
  if (cfl > Tag::diffusionCFLLimit) {
    throw std::invalid_argument(
        std::string("solve::diffuse: explicit-diffusion CFL violated for ") + typeid(Tag).name() +
        ": D=" + std::to_string(D) + ", dt=" + std::to_string(dt) + ", h=(" +
        std::to_string(h(0)) + ", " + std::to_string(h(1)) + ", " + std::to_string(h(2)) +
        "); D*dt*(1/hx^2 + 1/hy^2 + 1/hz^2) = " + std::to_string(cfl) + " > " +
        std::to_string(Tag::diffusionCFLLimit));
  }
}

template <class Tag>
constexpr gc_integrator integratorOf() {
  return std::is_same_v<Tag, RK4> ? GC_RK4 : GC_EULER;
}

inline void requireRadius1(const core::Grid& grid) {
  diff::detail::requireAxisExtent("x", grid.getNx(), 1);
  diff::detail::requireAxisExtent("y", grid.getNy(), 1);
  diff::detail::requireAxisExtent("z", grid.getNz(), 1);
}

}  // namespace detail

template <class Integrator = ExplicitEuler, class T = double,
          std::enable_if_t<detail::is_supported_scalar_v<T>, int> = 0>
inline void diffuse(core::Field<T>& psi, double D, double dt, int n_steps,
                    Integrator tag = {}) {
  (void)tag;
  if (D < 0.0)
    throw std::invalid_argument("solve::diffuse: D must be >= 0; got " + std::to_string(D));
  if (dt < 0.0)
    throw std::invalid_argument("solve::diffuse: dt must be >= 0; got " + std::to_string(dt));
  if (n_steps < 0)
    throw std::invalid_argument("solve::diffuse: n_steps must be >= 0; got " +
                                std::to_string(n_steps));
  detail::checkDiffusionCFL<Integrator>(psi.getGrid(), D, dt);
  if (n_steps == 0) return;
  detail::requireRadius1(psi.getGrid());
  const gc_grid g = hip::toGrid(psi.getGrid());
  hip::check(gc_diffuse_const(hip::context(), &g, hip::dtype_of<T>::value,
                              psi.deviceDataForWrite(), D, dt, n_steps,
                              detail::integratorOf<Integrator>()));
}

template <class Integrator = ExplicitEuler, class T = double,
          std::enable_if_t<detail::is_supported_scalar_v<T>, int> = 0>
inline void diffuse(core::Field<T>& psi, const core::Field<T>& D, double dt, int n_steps,
                    Integrator tag = {}) {
  (void)tag;
  if (dt < 0.0)
    throw std::invalid_argument("solve::diffuse: dt must be >= 0; got " + std::to_string(dt));
  if (n_steps < 0)
    throw std::invalid_argument("solve::diffuse: n_steps must be >= 0; got " +
                                std::to_string(n_steps));
  detail::requireSameGrid(psi, D, "solve::diffuse: D");
  const gc_grid g = hip::toGrid(psi.getGrid());
  double D_max_face = 0.0;
  hip::check(gc_max_face_d(hip::context(), &g, hip::dtype_of<T>::value, D.deviceData(),
                           &D_max_face));
  detail::checkDiffusionCFL<Integrator>(psi.getGrid(), D_max_face, dt);
  if (n_steps == 0) return;
  // the first right-hand-side evaluation (cellLaplacian) runs these checks in the reference
  detail::requireRadius1(psi.getGrid());
  fvm::detail::requirePositiveD(D);
  int64_t bad = -1;
  hip::check(gc_diffuse_field(hip::context(), &g, hip::dtype_of<T>::value,
                              psi.deviceDataForWrite(), D.deviceData(), dt, n_steps,
                              detail::integratorOf<Integrator>(), &bad));
}

}  // namespace gridcalc::solve
