// Densities that func::evaluate recognises and integrates on the device in one fused pass
// (gc_evaluate: the Order-2 derivatives each site needs, the density and the reduction).
// Each struct is also an ordinary callable of the arity func::evaluate dispatches on
// (func/Functional.hpp:88-148 and :183-220 of the reference), and its operator() performs
// the device kernel's operations in the same order, so passing the struct or a lambda of the
// same expression gives the same result bits.
//
//   ValueDensity{}                 f(p) = p                                       (arity 1)
//   GradientEnergyDensity{}        f(p, g) = |g|^2                                (arity 2)
//   CahnHilliardDensity{kappa}     -p^2/2 + p^4/4 + kappa/2 |g|^2                 (arity 2)
//   GinzburgLandauDensity{W}       |g|^2/2 + W (p^2 - 1)^2                        (arity 3)
//   PfcDensity{q0, eps}            (q0^4 - eps)/2 p^2 + q0^2 p lap + p bih/2 + p^4/4 (arity 4)
//   LinearElasticDensity{lambda, mu}  e = sym(grad u): (lambda tr(e)^2 + 2 mu e:e)/2 (Vec3)
//
// The formulas are those of examples/common/cahn_hilliard.hpp:64-76, test/functional_test.cpp
// :45-48, test/pfc_functional_test.cpp:44-50 and examples/common/elastic_energy.hpp:81-88.
#pragma once

#include <type_traits>

#include <gridcalc/core/EigenAliases.hpp>
#include <gridcalc_hip.h>

namespace gridcalc::func {

struct ValueDensity {
  static constexpr gc_density_kind kind = GC_DENSITY_VALUE;
  gc_density desc() const { return gc_density{kind, 0.0, 0.0}; }
  double operator()(double p) const { return p; }
};

struct GradientEnergyDensity {
  static constexpr gc_density_kind kind = GC_DENSITY_GRADIENT_ENERGY;
  gc_density desc() const { return gc_density{kind, 0.0, 0.0}; }
  double operator()(double, const core::Vec3d& g) const { return g.squaredNorm(); }
};

struct CahnHilliardDensity {
  static constexpr gc_density_kind kind = GC_DENSITY_CAHN_HILLIARD;
  double kappa;
  gc_density desc() const { return gc_density{kind, kappa, 0.0}; }
  double operator()(double p, const core::Vec3d& g) const {
    return -0.5 * p * p + 0.25 * p * p * p * p + 0.5 * kappa * g.squaredNorm();
  }
};

struct GinzburgLandauDensity {
  static constexpr gc_density_kind kind = GC_DENSITY_GINZBURG_LANDAU;
  double W;
  gc_density desc() const { return gc_density{kind, W, 0.0}; }
  double operator()(double p, const core::Vec3d& g, double) const {
    const double q = p * p - 1.0;
    return 0.5 * g.squaredNorm() + W * q * q;
  }
};

struct PfcDensity {
  static constexpr gc_density_kind kind = GC_DENSITY_PFC;
  double q0, eps;
  gc_density desc() const { return gc_density{kind, q0, eps}; }
  double operator()(double p, const core::Vec3d&, double lap, double bih) const {
    const double q2 = q0 * q0;
    const double q4 = q2 * q2;
    const double p2 = p * p;
    return 0.5 * (q4 - eps) * p2 + q2 * p * lap + 0.5 * p * bih + 0.25 * p2 * p2;
  }
};

struct LinearElasticDensity {
  static constexpr gc_density_kind kind = GC_DENSITY_LINEAR_ELASTIC;
  double lambda, mu;
  gc_density desc() const { return gc_density{kind, lambda, mu}; }
  double operator()(const core::Vec3d&, const core::Mat3d& grad_u) const {
    const core::Mat3d e = 0.5 * (grad_u + grad_u.transpose());
    const double tr = e.trace();
    const double ee = (e.array() * e.array()).sum();
    return 0.5 * (lambda * tr * tr + 2.0 * mu * ee);
  }
};

template <class F>
struct is_device_density
    : std::bool_constant<std::is_same_v<F, ValueDensity> ||
                         std::is_same_v<F, GradientEnergyDensity> ||
                         std::is_same_v<F, CahnHilliardDensity> ||
                         std::is_same_v<F, GinzburgLandauDensity> ||
                         std::is_same_v<F, PfcDensity>> {};
template <class F>
inline constexpr bool is_device_density_v = is_device_density<std::decay_t<F>>::value;

}  // namespace gridcalc::func
