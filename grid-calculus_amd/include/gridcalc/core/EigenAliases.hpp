// Vec3d / Mat3d: fixed-size 3-vectors and 3x3 matrices of double
// (core/EigenAliases.hpp:18,32 of the reference alias Eigen::Vector3d / Matrix3d).
//
// This build carries its own small dense type with the subset of that interface the
// library and its callers use: element access, arithmetic, products, norms, trace,
// transpose, array-wise products, reductions, comparisons and the comma initializer.
// Storage is column-major and unpadded (3 or 9 contiguous doubles), so a Field of these
// types is an array-of-structures the device kernels read directly.
#pragma once

#include <cmath>
#include <cstddef>
#include <initializer_list>
#include <type_traits>

namespace gridcalc::core {

template <int R, int C>
class SmallMat;

template <int R, int C>
class SmallArray;

namespace detail {
template <int R, int C>
struct CommaInit;
}

template <int R, int C>
class SmallMat {
 public:
  static constexpr int Rows = R, Cols = C, Size = R * C;

  constexpr SmallMat() : v_{} {}
  template <int RR = R, int CC = C, std::enable_if_t<RR * CC == 3, int> = 0>
  constexpr SmallMat(double x, double y, double z) : v_{x, y, z} {}

  static constexpr SmallMat Zero() { return SmallMat(); }
  static constexpr SmallMat Constant(double c) {
    SmallMat m;
    for (int i = 0; i < Size; ++i) m.v_[i] = c;
    return m;
  }
  static constexpr SmallMat Identity() {
    SmallMat m;
    for (int i = 0; i < (R < C ? R : C); ++i) m(i, i) = 1.0;
    return m;
  }
  static constexpr SmallMat Ones() { return Constant(1.0); }
  static constexpr SmallMat UnitX() { SmallMat m; m.v_[0] = 1.0; return m; }
  static constexpr SmallMat UnitY() { SmallMat m; m.v_[1] = 1.0; return m; }
  static constexpr SmallMat UnitZ() { SmallMat m; m.v_[2] = 1.0; return m; }

  constexpr double& operator()(int r, int c) { return v_[r + R * c]; }
  constexpr const double& operator()(int r, int c) const { return v_[r + R * c]; }
  constexpr double& operator()(int i) { return v_[i]; }
  constexpr const double& operator()(int i) const { return v_[i]; }
  constexpr double& operator[](int i) { return v_[i]; }
  constexpr const double& operator[](int i) const { return v_[i]; }
  constexpr double& x() { return v_[0]; }
  constexpr double& y() { return v_[1]; }
  constexpr double& z() { return v_[2]; }
  constexpr double x() const { return v_[0]; }
  constexpr double y() const { return v_[1]; }
  constexpr double z() const { return v_[2]; }

  double* data() { return v_; }
  const double* data() const { return v_; }
  static constexpr int rows() { return R; }
  static constexpr int cols() { return C; }
  static constexpr int size() { return Size; }

  SmallMat& setZero() { *this = Zero(); return *this; }
  SmallMat& setConstant(double c) { *this = Constant(c); return *this; }
  SmallMat& setIdentity() { *this = Identity(); return *this; }

  // sums run over the stored coefficients in storage order, left to right
  double sum() const {
    double s = v_[0];
    for (int i = 1; i < Size; ++i) s += v_[i];
    return s;
  }
  double squaredNorm() const {
    double s = v_[0] * v_[0];
    for (int i = 1; i < Size; ++i) s += v_[i] * v_[i];
    return s;
  }
  double norm() const { return std::sqrt(squaredNorm()); }
  SmallMat normalized() const { return *this / norm(); }
  double maxCoeff() const {
    double m = v_[0];
    for (int i = 1; i < Size; ++i) m = v_[i] > m ? v_[i] : m;
    return m;
  }
  double minCoeff() const {
    double m = v_[0];
    for (int i = 1; i < Size; ++i) m = v_[i] < m ? v_[i] : m;
    return m;
  }
  double prod() const {
    double p = v_[0];
    for (int i = 1; i < Size; ++i) p *= v_[i];
    return p;
  }
  double mean() const { return sum() / Size; }
  double trace() const {
    double t = (*this)(0, 0);
    for (int i = 1; i < (R < C ? R : C); ++i) t += (*this)(i, i);
    return t;
  }
  SmallMat cwiseAbs() const {
    SmallMat m;
    for (int i = 0; i < Size; ++i) m.v_[i] = std::abs(v_[i]);
    return m;
  }
  SmallMat cwiseProduct(const SmallMat& o) const {
    SmallMat m;
    for (int i = 0; i < Size; ++i) m.v_[i] = v_[i] * o.v_[i];
    return m;
  }
  SmallMat<C, R> transpose() const {
    SmallMat<C, R> t;
    for (int r = 0; r < R; ++r)
      for (int c = 0; c < C; ++c) t(c, r) = (*this)(r, c);
    return t;
  }
  double dot(const SmallMat& o) const {
    double s = v_[0] * o.v_[0];
    for (int i = 1; i < Size; ++i) s += v_[i] * o.v_[i];
    return s;
  }
  template <int RR = R, int CC = C, std::enable_if_t<RR == 3 && CC == 1, int> = 0>
  SmallMat cross(const SmallMat& o) const {
    return SmallMat(v_[1] * o.v_[2] - v_[2] * o.v_[1], v_[2] * o.v_[0] - v_[0] * o.v_[2],
                    v_[0] * o.v_[1] - v_[1] * o.v_[0]);
  }
  SmallMat<R, 1> col(int c) const {
    SmallMat<R, 1> v;
    for (int r = 0; r < R; ++r) v(r) = (*this)(r, c);
    return v;
  }
  SmallMat<1, C> row(int r) const {
    SmallMat<1, C> v;
    for (int c = 0; c < C; ++c) v(c) = (*this)(r, c);
    return v;
  }
  SmallArray<R, C> array() const;

  bool allFinite() const {
    for (int i = 0; i < Size; ++i)
      if (!std::isfinite(v_[i])) return false;
    return true;
  }
  bool hasNaN() const {
    for (int i = 0; i < Size; ++i)
      if (std::isnan(v_[i])) return true;
    return false;
  }
  // |a - b| <= prec * min(|a|, |b|) in the Frobenius norm
  bool isApprox(const SmallMat& o, double prec = 1e-12) const {
    const double d = (*this - o).squaredNorm();
    const double a = squaredNorm(), b = o.squaredNorm();
    return d <= prec * prec * (a < b ? a : b);
  }
  bool isZero(double prec = 1e-12) const {
    for (int i = 0; i < Size; ++i)
      if (std::abs(v_[i]) > prec) return false;
    return true;
  }

  SmallMat& operator+=(const SmallMat& o) {
    for (int i = 0; i < Size; ++i) v_[i] += o.v_[i];
    return *this;
  }
  SmallMat& operator-=(const SmallMat& o) {
    for (int i = 0; i < Size; ++i) v_[i] -= o.v_[i];
    return *this;
  }
  SmallMat& operator*=(double s) {
    for (int i = 0; i < Size; ++i) v_[i] *= s;
    return *this;
  }
  SmallMat& operator/=(double s) {
    for (int i = 0; i < Size; ++i) v_[i] /= s;
    return *this;
  }
  friend SmallMat operator+(SmallMat a, const SmallMat& b) { return a += b; }
  friend SmallMat operator-(SmallMat a, const SmallMat& b) { return a -= b; }
  friend SmallMat operator-(SmallMat a) {
    for (int i = 0; i < Size; ++i) a.v_[i] = -a.v_[i];
    return a;
  }
  friend SmallMat operator*(SmallMat a, double s) { return a *= s; }
  friend SmallMat operator*(double s, SmallMat a) {
    for (int i = 0; i < Size; ++i) a.v_[i] = s * a.v_[i];
    return a;
  }
  friend SmallMat operator/(SmallMat a, double s) { return a /= s; }
  friend bool operator==(const SmallMat& a, const SmallMat& b) {
    for (int i = 0; i < Size; ++i)
      if (!(a.v_[i] == b.v_[i])) return false;
    return true;
  }
  friend bool operator!=(const SmallMat& a, const SmallMat& b) { return !(a == b); }

  // comma initializer: m << a, b, c, ... fills row by row
  detail::CommaInit<R, C> operator<<(double first);

 private:
  double v_[Size];
};

// Matrix product (K-sum in index order).
template <int R, int K, int C>
SmallMat<R, C> operator*(const SmallMat<R, K>& a, const SmallMat<K, C>& b) {
  SmallMat<R, C> m;
  for (int r = 0; r < R; ++r)
    for (int c = 0; c < C; ++c) {
      double s = a(r, 0) * b(0, c);
      for (int k = 1; k < K; ++k) s += a(r, k) * b(k, c);
      m(r, c) = s;
    }
  return m;
}

// Coefficient-wise view: a.array() * b.array() multiplies element by element.
template <int R, int C>
class SmallArray {
 public:
  explicit SmallArray(const SmallMat<R, C>& m) : m_(m) {}
  friend SmallArray operator*(const SmallArray& a, const SmallArray& b) {
    return SmallArray(a.m_.cwiseProduct(b.m_));
  }
  friend SmallArray operator+(const SmallArray& a, const SmallArray& b) {
    return SmallArray(a.m_ + b.m_);
  }
  friend SmallArray operator-(const SmallArray& a, const SmallArray& b) {
    return SmallArray(a.m_ - b.m_);
  }
  SmallArray abs() const { return SmallArray(m_.cwiseAbs()); }
  SmallArray square() const { return SmallArray(m_.cwiseProduct(m_)); }
  double sum() const { return m_.sum(); }
  double maxCoeff() const { return m_.maxCoeff(); }
  double minCoeff() const { return m_.minCoeff(); }
  const SmallMat<R, C>& matrix() const { return m_; }

 private:
  SmallMat<R, C> m_;
};

template <int R, int C>
SmallArray<R, C> SmallMat<R, C>::array() const {
  return SmallArray<R, C>(*this);
}

namespace detail {
template <int R, int C>
struct CommaInit {
  SmallMat<R, C>& m;
  int n;
  CommaInit& operator,(double v) {
    if (n < R * C) m(n / C, n % C) = v;
    ++n;
    return *this;
  }
};
}  // namespace detail

template <int R, int C>
detail::CommaInit<R, C> SmallMat<R, C>::operator<<(double first) {
  (*this)(0, 0) = first;
  return detail::CommaInit<R, C>{*this, 1};
}

using Vec3d = SmallMat<3, 1>;
using Mat3d = SmallMat<3, 3>;

static_assert(sizeof(Vec3d) == 3 * sizeof(double) && std::is_standard_layout_v<Vec3d>,
              "Vec3d must be three packed doubles");
static_assert(sizeof(Mat3d) == 9 * sizeof(double) && std::is_standard_layout_v<Mat3d>,
              "Mat3d must be nine packed doubles");

}  // namespace gridcalc::core
