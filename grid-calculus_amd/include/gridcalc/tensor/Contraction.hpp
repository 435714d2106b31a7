// The eager contractions of tensor/Contraction.hpp (gridcalc 0.15): pointwise trace of a
// Field<Mat3d>, the single contraction (matrix product) and the double contraction A:B of two,
// each a one-node expression run on the device (tensor::expr::materialize), with the
// reference's grid checks and messages.
#pragma once

#include <gridcalc/core/EigenAliases.hpp>
#include <gridcalc/core/Field.hpp>
#include <gridcalc/tensor/Expressions.hpp>
#include <gridcalc/tensor/detail/RequireSameGrid.hpp>

namespace gridcalc::tensor {

inline core::Field<double> trace(const core::Field<core::Mat3d>& m) {
  return expr::materialize(expr::trace(expr::field(m)));
}

inline core::Field<core::Mat3d> singleContract(const core::Field<core::Mat3d>& a,
                                               const core::Field<core::Mat3d>& b) {
  detail::requireSameGrid(a, b, "tensor::singleContract");
  return expr::materialize(expr::singleContract(expr::field(a), expr::field(b)));
}

inline core::Field<double> doubleContract(const core::Field<core::Mat3d>& a,
                                          const core::Field<core::Mat3d>& b) {
  detail::requireSameGrid(a, b, "tensor::doubleContract");
  return expr::materialize(expr::doubleContract(expr::field(a), expr::field(b)));
}

}  // namespace gridcalc::tensor
