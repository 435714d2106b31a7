// Both integrators (solve/Integrator.hpp of the reference).
#pragma once

#include <gridcalc/solve/ExplicitEuler.hpp>
#include <gridcalc/solve/RK4.hpp>
