// tensor::expr (tensor/Expressions.hpp of the reference): lazy expressions over Field<double>,
// Field<Vec3d> and Field<Mat3d> -- leaves, the identity field, scaling, sums, differences,
// scalar products, trace, symmetric part, single and double contraction -- with the
// reference's node types, operators, grid checks and messages.
//
// materialize(E) lowers the tree to a register program (gc_expr_insn) and runs it on the
// device in one pass (gc_expr_eval): every site evaluates the whole tree with its
// intermediates in on-chip memory, so no rank-2 temporary field is ever written.  The host
// evalAt(i, j, k) of every node stays available and performs the same operations in the same
// order, so both give the same bits.
#pragma once

#include <cstddef>
#include <stdexcept>
#include <type_traits>
#include <utility>
#include <vector>

#include <gridcalc/core/EigenAliases.hpp>
#include <gridcalc/core/Field.hpp>
#include <gridcalc/core/Grid.hpp>
#include <gridcalc/hip/Runtime.hpp>
#include <gridcalc/tensor/detail/RequireSameGrid.hpp>

namespace gridcalc::tensor::expr {

template <class T>
struct is_expr : std::false_type {};

template <class T>
inline constexpr bool is_expr_v = is_expr<std::remove_cv_t<std::remove_reference_t<T>>>::value;

namespace detail {

template <class T>
constexpr int components() {
  if constexpr (std::is_same_v<T, double>) return 1;
  else if constexpr (std::is_same_v<T, core::Vec3d>) return 3;
  else {
    static_assert(std::is_same_v<T, core::Mat3d>, "tensor::expr: value types are double, Vec3d, Mat3d");
    return 9;
  }
}

// A register program under construction: instructions and the free value slots.
class Program {
 public:
  static constexpr int kSlots = 8, kMaxInsn = 32;
  int alloc() {
    for (int s = 0; s < kSlots; ++s)
      if (!(used_ & (1u << s))) {
        used_ |= 1u << s;
        return s;
      }
    throw std::invalid_argument("tensor::expr: expression needs more than 8 live values");
  }
  void release(int s) { used_ &= ~(1u << s); }
  void emit(gc_expr_op op, int dst, int a, int b, int nc, double c = 0.0, const void* ptr = nullptr) {
    if ((int)insn_.size() == kMaxInsn)
      throw std::invalid_argument("tensor::expr: expression longer than 32 operations");
    insn_.push_back(gc_expr_insn{op, dst, a, b, nc, c, ptr});
  }
  const std::vector<gc_expr_insn>& insns() const { return insn_; }

 private:
  std::vector<gc_expr_insn> insn_;
  unsigned used_ = 0;
};

}  // namespace detail

// ---- leaves -----------------------------------------------------------------------------

template <class T>
class Leaf {
 public:
  using value_type = T;
  explicit Leaf(const core::Field<T>& f) noexcept : _field(&f) {}
  explicit Leaf(core::Field<T>&&) = delete;
  const core::Grid& grid() const noexcept { return _field->getGrid(); }
  const T& evalAt(int i, int j, int k) const { return (*_field)(i, j, k); }
  int lower(detail::Program& p) const {
    const int s = p.alloc();
    p.emit(GC_EXPR_LOAD, s, 0, 0, detail::components<T>(), 0.0, _field->deviceData());
    return s;
  }

 private:
  const core::Field<T>* _field;
};
template <class T>
struct is_expr<Leaf<T>> : std::true_type {};

template <class T>
inline Leaf<T> field(const core::Field<T>& f) noexcept {
  return Leaf<T>{f};
}
template <class T>
Leaf<T> field(core::Field<T>&&) = delete;

class IdentityField {
 public:
  using value_type = core::Mat3d;
  explicit IdentityField(const core::Grid& g) noexcept : _grid(&g) {}
  explicit IdentityField(core::Grid&&) = delete;
  const core::Grid& grid() const noexcept { return *_grid; }
  core::Mat3d evalAt(int, int, int) const noexcept { return core::Mat3d::Identity(); }
  int lower(detail::Program& p) const {
    const int s = p.alloc();
    p.emit(GC_EXPR_IDENTITY, s, 0, 0, 9);
    return s;
  }

 private:
  const core::Grid* _grid;
};
template <>
struct is_expr<IdentityField> : std::true_type {};

inline IdentityField identityField(const core::Grid& g) noexcept { return IdentityField{g}; }
IdentityField identityField(core::Grid&&) = delete;

// ---- combinators ------------------------------------------------------------------------

template <class E>
class Scaled {
 public:
  using value_type = typename E::value_type;
  Scaled(double c, E inner) noexcept(std::is_nothrow_move_constructible_v<E>)
      : _c(c), _inner(std::move(inner)) {}
  const core::Grid& grid() const noexcept { return _inner.grid(); }
  value_type evalAt(int i, int j, int k) const { return _c * _inner.evalAt(i, j, k); }
  int lower(detail::Program& p) const {
    const int s = _inner.lower(p);
    p.emit(GC_EXPR_SCALE, s, s, 0, detail::components<value_type>(), _c);
    return s;
  }

 private:
  double _c;
  E _inner;
};
template <class E>
struct is_expr<Scaled<E>> : std::true_type {};

// Binary pointwise node: Plus, Minus (same value types) and Mul (scalars).
template <class L, class R, gc_expr_op OP>
class Binary {
 public:
  static_assert(std::is_same_v<typename L::value_type, typename R::value_type>,
                "tensor::expr: operands must have the same value_type");
  static_assert(OP != GC_EXPR_MUL || std::is_same_v<typename L::value_type, double>,
                "tensor::expr::Mul: operands must have value_type == double");
  using value_type = typename L::value_type;
  Binary(L lhs, R rhs) : _lhs(std::move(lhs)), _rhs(std::move(rhs)) {
    tensor::detail::requireSameGrid(_lhs.grid(), _rhs.grid(), name());
  }
  static const char* name() {
    return OP == GC_EXPR_ADD ? "tensor::expr::Plus"
           : OP == GC_EXPR_SUB ? "tensor::expr::Minus"
                               : "tensor::expr::Mul";
  }
  const core::Grid& grid() const noexcept { return _lhs.grid(); }
  value_type evalAt(int i, int j, int k) const {
    if constexpr (OP == GC_EXPR_ADD) return _lhs.evalAt(i, j, k) + _rhs.evalAt(i, j, k);
    else if constexpr (OP == GC_EXPR_SUB) return _lhs.evalAt(i, j, k) - _rhs.evalAt(i, j, k);
    else return _lhs.evalAt(i, j, k) * _rhs.evalAt(i, j, k);
  }
  int lower(detail::Program& p) const {
    const int a = _lhs.lower(p);
    const int b = _rhs.lower(p);
    p.emit(OP, a, a, b, detail::components<value_type>());
    p.release(b);
    return a;
  }

 private:
  L _lhs;
  R _rhs;
};
template <class L, class R>
using Plus = Binary<L, R, GC_EXPR_ADD>;
template <class L, class R>
using Minus = Binary<L, R, GC_EXPR_SUB>;
template <class L, class R>
using Mul = Binary<L, R, GC_EXPR_MUL>;
template <class L, class R, gc_expr_op OP>
struct is_expr<Binary<L, R, OP>> : std::true_type {};

// ---- rank-2 primitives ------------------------------------------------------------------

template <class E>
class Trace {
 public:
  static_assert(std::is_same_v<typename E::value_type, core::Mat3d>,
                "tensor::expr::Trace requires a Mat3d-valued expression");
  using value_type = double;
  explicit Trace(E inner) noexcept(std::is_nothrow_move_constructible_v<E>)
      : _inner(std::move(inner)) {}
  const core::Grid& grid() const noexcept { return _inner.grid(); }
  double evalAt(int i, int j, int k) const { return _inner.evalAt(i, j, k).trace(); }
  int lower(detail::Program& p) const {
    const int s = _inner.lower(p);
    p.emit(GC_EXPR_TRACE, s, s, 0, 1);
    return s;
  }

 private:
  E _inner;
};
template <class E>
struct is_expr<Trace<E>> : std::true_type {};

template <class E, std::enable_if_t<is_expr_v<E>, int> = 0>
inline auto trace(E&& e) {
  return Trace<std::decay_t<E>>{std::forward<E>(e)};
}

template <class E>
class Sym {
 public:
  static_assert(std::is_same_v<typename E::value_type, core::Mat3d>,
                "tensor::expr::Sym requires a Mat3d-valued expression");
  using value_type = core::Mat3d;
  explicit Sym(E inner) noexcept(std::is_nothrow_move_constructible_v<E>)
      : _inner(std::move(inner)) {}
  const core::Grid& grid() const noexcept { return _inner.grid(); }
  core::Mat3d evalAt(int i, int j, int k) const {
    const core::Mat3d m = _inner.evalAt(i, j, k);
    return 0.5 * (m + m.transpose());
  }
  int lower(detail::Program& p) const {
    const int s = _inner.lower(p);
    p.emit(GC_EXPR_SYM, s, s, 0, 9);
    return s;
  }

 private:
  E _inner;
};
template <class E>
struct is_expr<Sym<E>> : std::true_type {};

template <class E, std::enable_if_t<is_expr_v<E>, int> = 0>
inline auto sym(E&& e) {
  return Sym<std::decay_t<E>>{std::forward<E>(e)};
}

template <class L, class R, gc_expr_op OP>
class Contract {
 public:
  static_assert(std::is_same_v<typename L::value_type, core::Mat3d> &&
                    std::is_same_v<typename R::value_type, core::Mat3d>,
                "tensor::expr: contractions take Mat3d-valued operands");
  using value_type = std::conditional_t<OP == GC_EXPR_MATMUL, core::Mat3d, double>;
  Contract(L lhs, R rhs) : _lhs(std::move(lhs)), _rhs(std::move(rhs)) {
    tensor::detail::requireSameGrid(_lhs.grid(), _rhs.grid(),
                                    OP == GC_EXPR_MATMUL ? "tensor::expr::SingleContract"
                                                         : "tensor::expr::DoubleContract");
  }
  const core::Grid& grid() const noexcept { return _lhs.grid(); }
  value_type evalAt(int i, int j, int k) const {
    const core::Mat3d a = _lhs.evalAt(i, j, k);
    const core::Mat3d b = _rhs.evalAt(i, j, k);
    if constexpr (OP == GC_EXPR_MATMUL) return a * b;
    else return (a.array() * b.array()).sum();
  }
  int lower(detail::Program& p) const {
    const int a = _lhs.lower(p);
    const int b = _rhs.lower(p);
    p.emit(OP, a, a, b, detail::components<value_type>());
    p.release(b);
    return a;
  }

 private:
  L _lhs;
  R _rhs;
};
template <class L, class R>
using SingleContract = Contract<L, R, GC_EXPR_MATMUL>;
template <class L, class R>
using DoubleContract = Contract<L, R, GC_EXPR_DDOT>;
template <class L, class R, gc_expr_op OP>
struct is_expr<Contract<L, R, OP>> : std::true_type {};

template <class L, class R, std::enable_if_t<is_expr_v<L> && is_expr_v<R>, int> = 0>
inline auto singleContract(L&& a, R&& b) {
  return SingleContract<std::decay_t<L>, std::decay_t<R>>{std::forward<L>(a), std::forward<R>(b)};
}

template <class L, class R, std::enable_if_t<is_expr_v<L> && is_expr_v<R>, int> = 0>
inline auto doubleContract(L&& a, R&& b) {
  return DoubleContract<std::decay_t<L>, std::decay_t<R>>{std::forward<L>(a), std::forward<R>(b)};
}

// ---- operators --------------------------------------------------------------------------

template <class E, std::enable_if_t<is_expr_v<E>, int> = 0>
inline auto operator-(E&& e) {
  return Scaled<std::decay_t<E>>{-1.0, std::forward<E>(e)};
}
template <class E, std::enable_if_t<is_expr_v<E>, int> = 0>
inline auto operator*(double c, E&& e) {
  return Scaled<std::decay_t<E>>{c, std::forward<E>(e)};
}
template <class E, std::enable_if_t<is_expr_v<E>, int> = 0>
inline auto operator*(E&& e, double c) {
  return Scaled<std::decay_t<E>>{c, std::forward<E>(e)};
}
template <class E, std::enable_if_t<is_expr_v<E>, int> = 0>
inline auto operator/(E&& e, double c) {
  return Scaled<std::decay_t<E>>{1.0 / c, std::forward<E>(e)};
}
template <class L, class R, std::enable_if_t<is_expr_v<L> && is_expr_v<R>, int> = 0>
inline auto operator+(L&& l, R&& r) {
  return Plus<std::decay_t<L>, std::decay_t<R>>{std::forward<L>(l), std::forward<R>(r)};
}
template <class L, class R, std::enable_if_t<is_expr_v<L> && is_expr_v<R>, int> = 0>
inline auto operator-(L&& l, R&& r) {
  return Minus<std::decay_t<L>, std::decay_t<R>>{std::forward<L>(l), std::forward<R>(r)};
}
template <class L, class R,
          std::enable_if_t<is_expr_v<L> && is_expr_v<R> &&
                               std::is_same_v<typename std::decay_t<L>::value_type, double> &&
                               std::is_same_v<typename std::decay_t<R>::value_type, double>,
                           int> = 0>
inline auto operator*(L&& l, R&& r) {
  return Mul<std::decay_t<L>, std::decay_t<R>>{std::forward<L>(l), std::forward<R>(r)};
}

// ---- materialisation --------------------------------------------------------------------

// Evaluates the expression at every site on the device; returns a device-resident field.
template <class E, std::enable_if_t<is_expr_v<E>, int> = 0>
inline core::Field<typename std::decay_t<E>::value_type> materialize(const E& expr) {
  using T = typename std::decay_t<E>::value_type;
  const core::Grid& g = expr.grid();
  detail::Program p;
  const int out_slot = expr.lower(p);
  core::Field<T> out(g, core::detail::device_only);
  const gc_grid gg = hip::toGrid(g);
  hip::check(gc_expr_eval(hip::context(), &gg, p.insns().data(), (int)p.insns().size(), out_slot,
                          detail::components<T>(), out.deviceDataDiscard()));
  return out;
}

}  // namespace gridcalc::tensor::expr
