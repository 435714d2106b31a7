// Periodic index policy (core/IndexPolicy.hpp:18-29 of the reference).
#pragma once

namespace gridcalc::core::IndexPolicy {

struct Periodic {
  // ((i mod N) + N) mod N: any integer index folds into [0, N)
  static constexpr int wrap(int i, int N) noexcept { return ((i % N) + N) % N; }
};

}  // namespace gridcalc::core::IndexPolicy
