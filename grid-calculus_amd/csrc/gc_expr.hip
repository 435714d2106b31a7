// gc_expr.hip -- the tensor contraction expressions of tensor/Expressions.hpp and the eager
// contractions of tensor/Contraction.hpp (gridcalc 0.15), evaluated on the device.
//
// The C++ expression tree (gridcalc/tensor/Expressions.hpp) is lowered on the host to a short
// register program; one kernel runs it per site.  Each thread keeps its value slots (up to 9
// doubles each: scalar, Vec3 or column-major Mat3) in its own LDS column, so the program's
// rank-2 intermediates never touch HBM and no thread reads another's slots (no barriers).
// Per-site arithmetic follows the host evalAt of the same nodes operation for operation:
//   SCALE c*v, ADD a+b, SUB a-b, MUL a*b (per component), TRACE (m00 + m11) + m22,
//   SYM 0.5*(m(r,c) + m(c,r)), MATMUL (a*b)(r,c) = a(r,0)b(0,c) + a(r,1)b(1,c) + a(r,2)b(2,c)
//   left to right, DDOT sum of a_i*b_i over the 9 entries in storage order.
#include "gc_common.hpp"

#include <cstring>

namespace gc {
namespace {

constexpr int kExprThreads = 64;
constexpr int kExprSlots = 8;
constexpr int kExprMaxInsn = 32;

struct ExprProg {
  gc_expr_insn insn[kExprMaxInsn];
  int n;
};

__global__ __launch_bounds__(kExprThreads) void k_expr(const ExprProg prog, size_t n, int out_slot,
                                                       int out_nc, double* __restrict__ out) {
  __shared__ double lds[kExprSlots * 9 * kExprThreads];
  const int t = threadIdx.x;
  auto S = [&](int slot, int q) -> double& { return lds[(slot * 9 + q) * kExprThreads + t]; };
  const size_t stride = (size_t)gridDim.x * kExprThreads;
  for (size_t c = (size_t)blockIdx.x * kExprThreads + t; c < n; c += stride) {
    for (int p = 0; p < prog.n; ++p) {
      const gc_expr_insn I = prog.insn[p];
      switch (I.op) {
        case GC_EXPR_LOAD: {
          const double* src = (const double*)I.ptr + c * (size_t)I.nc;
          for (int q = 0; q < I.nc; ++q) S(I.dst, q) = src[q];
          break;
        }
        case GC_EXPR_IDENTITY:
          for (int q = 0; q < 9; ++q) S(I.dst, q) = (q % 4 == 0) ? 1.0 : 0.0;
          break;
        case GC_EXPR_SCALE:
          for (int q = 0; q < I.nc; ++q) S(I.dst, q) = I.c * S(I.a, q);
          break;
        case GC_EXPR_ADD:
          for (int q = 0; q < I.nc; ++q) S(I.dst, q) = S(I.a, q) + S(I.b, q);
          break;
        case GC_EXPR_SUB:
          for (int q = 0; q < I.nc; ++q) S(I.dst, q) = S(I.a, q) - S(I.b, q);
          break;
        case GC_EXPR_MUL:
          for (int q = 0; q < I.nc; ++q) S(I.dst, q) = S(I.a, q) * S(I.b, q);
          break;
        case GC_EXPR_TRACE: {
          const double tr = S(I.a, 0) + S(I.a, 4) + S(I.a, 8);
          S(I.dst, 0) = tr;
          break;
        }
        case GC_EXPR_SYM: {
          double m[9];
          for (int q = 0; q < 9; ++q) m[q] = S(I.a, q);
#pragma unroll
          for (int col = 0; col < 3; ++col)
#pragma unroll
            for (int r = 0; r < 3; ++r) S(I.dst, r + 3 * col) = 0.5 * (m[r + 3 * col] + m[col + 3 * r]);
          break;
        }
        case GC_EXPR_MATMUL: {
          double a[9], b[9];
          for (int q = 0; q < 9; ++q) {
            a[q] = S(I.a, q);
            b[q] = S(I.b, q);
          }
#pragma unroll
          for (int col = 0; col < 3; ++col)
#pragma unroll
            for (int r = 0; r < 3; ++r) {
              double s = a[r] * b[3 * col];
              s += a[r + 3] * b[1 + 3 * col];
              s += a[r + 6] * b[2 + 3 * col];
              S(I.dst, r + 3 * col) = s;
            }
          break;
        }
        case GC_EXPR_DDOT: {
          double s = S(I.a, 0) * S(I.b, 0);
          for (int q = 1; q < 9; ++q) s += S(I.a, q) * S(I.b, q);
          S(I.dst, 0) = s;
          break;
        }
        default: break;
      }
    }
    double* o = out + c * (size_t)out_nc;
    for (int q = 0; q < out_nc; ++q) o[q] = S(out_slot, q);
  }
}

bool valid_nc(int nc) { return nc == 1 || nc == 3 || nc == 9; }

}  // namespace
}  // namespace gc

using namespace gc;

extern "C" {

gc_status gc_expr_eval(gc_context* ctx, const gc_grid* g, const gc_expr_insn* prog, int n_insn,
                       int out_slot, int out_nc, void* out) {
  GC_TRY(check_ctx(ctx));
  GC_TRY(check_grid(g));
  if (!prog || !out) return set_error(GC_ERR_INVALID, "gridcalc: null expression program or output");
  if (n_insn < 1 || n_insn > kExprMaxInsn)
    return set_error(GC_ERR_INVALID, "gridcalc: expression program of %d instructions (1 to %d)",
                     n_insn, kExprMaxInsn);
  if (out_slot < 0 || out_slot >= kExprSlots || !valid_nc(out_nc))
    return set_error(GC_ERR_INVALID, "gridcalc: bad expression output slot %d / width %d", out_slot,
                     out_nc);
  ExprProg P;
  std::memset(&P, 0, sizeof P);
  P.n = n_insn;
  for (int i = 0; i < n_insn; ++i) {
    const gc_expr_insn& I = prog[i];
    const bool slots_ok = I.dst >= 0 && I.dst < kExprSlots && I.a >= 0 && I.a < kExprSlots &&
                          I.b >= 0 && I.b < kExprSlots;
    if (!slots_ok || !valid_nc(I.nc) || I.op < GC_EXPR_LOAD || I.op > GC_EXPR_DDOT)
      return set_error(GC_ERR_INVALID, "gridcalc: bad expression instruction %d", i);
    if (I.op == GC_EXPR_LOAD && !I.ptr)
      return set_error(GC_ERR_INVALID, "gridcalc: expression instruction %d loads a null field", i);
    if (I.op == GC_EXPR_LOAD && I.ptr == out)
      return set_error(GC_ERR_INVALID, "gridcalc: expression output must not alias an operand");
    P.insn[i] = I;
  }
  const size_t n = sites(g);
  const size_t want = (n + kExprThreads - 1) / kExprThreads;
  const unsigned nb = (unsigned)(want < 16384 ? want : 16384);
  hipLaunchKernelGGL(k_expr, dim3(nb), dim3(kExprThreads), 0, ctx->stream, P, n, out_slot, out_nc,
                     (double*)out);
  GC_LAUNCHED();
  return GC_OK;
}

}  // extern "C"
