// gc_mixed.hip -- mixed partial derivatives d^(nx+ny+nz) psi / dx^nx dy^ny dz^nz
// (diff/detail/MultiIndexDerivative.hpp:96-136 of the reference; the named second-, third- and
// fourth-order partials of diff/MixedPartial.hpp, ThirdOrder.hpp and FourthOrder.hpp).
//
// The reference forms a tensor product of one axis stencil per axis (the identity for order 0,
// then the first, second, third and fourth central differences) and sums, per site,
//   for z offsets: for y offsets: wyz = wy * wz; for x offsets: sum += wx * wyz * psi
// including zero weights, then multiplies by 1/h_x^nx 1/h_y^ny 1/h_z^nz (formed by repeated
// division).  Here the operator runs on the z-march LDS engine with a ring radius equal to
// the largest axis radius and keeps that loop order, so double-precision results equal the
// reference's bit for bit.
#include "gc_common.hpp"
#include "gc_zmarch.hpp"

namespace gc {
namespace {

// The axis stencil for an n-th derivative (MultiIndexDerivative.hpp:46-84).
template <int N, int O>
struct AxisSt;
template <int O>
struct AxisSt<0, O> {
  static constexpr int R = 0, N = 1;
  __host__ __device__ static constexpr double w(int) { return 1.0; }
};
template <int O>
struct AxisSt<1, O> : D1<O> {};
template <int O>
struct AxisSt<2, O> : D2<O> {};
template <int O>
struct AxisSt<3, O> : D3<O> {};
template <int O>
struct AxisSt<4, O> : D4<O> {};

constexpr int cmax(int a, int b) { return a > b ? a : b; }

template <class T, int NX, int NY, int NZ, int O>
struct OpMixed {
  using SX = AxisSt<NX, O>;
  using SY = AxisSt<NY, O>;
  using SZ = AxisSt<NZ, O>;
  static constexpr int R = cmax(cmax(SX::R, SY::R), cmax(SZ::R, 1));
  T scale;
  template <class A>
  __device__ __forceinline__ void operator()(const A& F, T* __restrict__ out, size_t c) const {
    T sum = 0;
#pragma unroll
    for (int mz = 0; mz < SZ::N; ++mz) {
      const int oz = mz - SZ::R;
      const double wz = SZ::w(mz);
#pragma unroll
      for (int my = 0; my < SY::N; ++my) {
        const int oy = my - SY::R;
        const T wyz = T(SY::w(my) * wz);
#pragma unroll
        for (int mx = 0; mx < SX::N; ++mx) {
          const int ox = mx - SX::R;
          sum += T(SX::w(mx)) * wyz * F(0, ox, oy, oz);
        }
      }
    }
    out[c] = sum * scale;
  }
};

template <class T, int NX, int NY, int NZ, int O>
gc_status run_mixed(gc_context* ctx, const gc_grid* g, const void* in, void* out) {
  using Op = OpMixed<T, NX, NY, NZ, O>;
  GC_TRY(check_extent3(g, Op::SX::R, Op::SY::R, Op::SZ::R));
  // MultiIndexDerivative.hpp:104-108
  double scale = 1.0;
  for (int n = 0; n < NX; ++n) scale /= g->hx;
  for (int n = 0; n < NY; ++n) scale /= g->hy;
  for (int n = 0; n < NZ; ++n) scale /= g->hz;
  zmarch_launch<T>(ctx, dims_of(g), Op{T(scale)}, (const T*)in, (T*)out);
  GC_LAUNCHED();
  return GC_OK;
}

template <int NX, int NY, int NZ>
gc_status run_mixed_any(gc_context* ctx, const gc_grid* g, gc_dtype dt, int order, const void* in,
                        void* out) {
  if (dt == GC_F64)
    return order == 2 ? run_mixed<double, NX, NY, NZ, 2>(ctx, g, in, out)
                      : run_mixed<double, NX, NY, NZ, 4>(ctx, g, in, out);
  return order == 2 ? run_mixed<float, NX, NY, NZ, 2>(ctx, g, in, out)
                    : run_mixed<float, NX, NY, NZ, 4>(ctx, g, in, out);
}

}  // namespace
}  // namespace gc

using namespace gc;

extern "C" {

gc_status gc_mixed_partial(gc_context* ctx, const gc_grid* g, gc_dtype dt, int nx, int ny, int nz,
                           int order, const void* in, void* out) {
  GC_TRY(check_ctx(ctx));
  GC_TRY(check_grid(g));
  GC_TRY(check_dtype(dt));
  if (order != 2 && order != 4)
    return set_error(GC_ERR_INVALID, "gridcalc: Order must be 2 or 4; got %d", order);
  if (!in || !out) return set_error(GC_ERR_INVALID, "gridcalc: null field pointer");
  if (in == out) return set_error(GC_ERR_INVALID, "gridcalc: output must not alias the input");
  const int total = nx + ny + nz;
  if (nx < 0 || ny < 0 || nz < 0 || total < 1 || total > 4)
    return set_error(GC_ERR_INVALID,
                     "gridcalc: mixed partial orders (%d, %d, %d) must be >= 0 with a total of "
                     "1 to 4",
                     nx, ny, nz);
#define GC_MIXED_CASE(X, Y, Z) \
  if (nx == X && ny == Y && nz == Z) return run_mixed_any<X, Y, Z>(ctx, g, dt, order, in, out);
  // total 1
  GC_MIXED_CASE(1, 0, 0) GC_MIXED_CASE(0, 1, 0) GC_MIXED_CASE(0, 0, 1)
  // total 2 (diff/MixedPartial.hpp)
  GC_MIXED_CASE(2, 0, 0) GC_MIXED_CASE(0, 2, 0) GC_MIXED_CASE(0, 0, 2)
  GC_MIXED_CASE(1, 1, 0) GC_MIXED_CASE(1, 0, 1) GC_MIXED_CASE(0, 1, 1)
  // total 3 (diff/ThirdOrder.hpp)
  GC_MIXED_CASE(3, 0, 0) GC_MIXED_CASE(0, 3, 0) GC_MIXED_CASE(0, 0, 3)
  GC_MIXED_CASE(2, 1, 0) GC_MIXED_CASE(1, 2, 0) GC_MIXED_CASE(2, 0, 1)
  GC_MIXED_CASE(1, 0, 2) GC_MIXED_CASE(0, 2, 1) GC_MIXED_CASE(0, 1, 2)
  GC_MIXED_CASE(1, 1, 1)
  // total 4 (diff/FourthOrder.hpp)
  GC_MIXED_CASE(4, 0, 0) GC_MIXED_CASE(0, 4, 0) GC_MIXED_CASE(0, 0, 4)
  GC_MIXED_CASE(3, 1, 0) GC_MIXED_CASE(1, 3, 0) GC_MIXED_CASE(3, 0, 1)
  GC_MIXED_CASE(1, 0, 3) GC_MIXED_CASE(0, 3, 1) GC_MIXED_CASE(0, 1, 3)
  GC_MIXED_CASE(2, 2, 0) GC_MIXED_CASE(2, 0, 2) GC_MIXED_CASE(0, 2, 2)
  GC_MIXED_CASE(2, 1, 1) GC_MIXED_CASE(1, 2, 1) GC_MIXED_CASE(1, 1, 2)
#undef GC_MIXED_CASE
  return set_error(GC_ERR_INVALID, "gridcalc: mixed partial orders (%d, %d, %d) not handled", nx,
                   ny, nz);
}

}  // extern "C"
