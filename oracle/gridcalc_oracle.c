/*
 * gridcalc_oracle.c -- CPU restatement of gridcalc's periodic-grid operator and
 * time-stepping path (gridcalc 0.15.1), used ONLY as the test oracle and as the
 * single-core CPU baseline in the benchmark.  It is test infrastructure: the
 * shipped HIP library never calls, links or loads anything in this directory.
 *
 * Every function follows the reference's loops in the reference's arithmetic
 * order (operand order of every + and *, loop nesting k,j,i, per-axis partial
 * sums accumulated in stencil-offset order).  It is compiled with
 * -ffp-contract=off so no multiply-add is fused, which makes it bit-exact with
 * the reference built by a compiler that does not contract (g++ on x86-64
 * without -mfma, the reference's Release preset).
 *
 * Each function names the reference file:line it restates.  How each part is
 * pinned (closed forms, known answers from the reference's tests, invariants)
 * is recorded in the test files that exercise it (tests/test_oracle.py).
 *
 * Parity notes:
 *  - or_free_energy: Eigen's Vector3d::squaredNorm() summation order is taken
 *    as (gx*gx + gy*gy) + gz*gz.  On 2-D grids (Nz = 1) gz is exactly 0 and
 *    both possible orders agree bit-for-bit; on 3-D grids the order is parity
 *    unpinned at the level of one rounding of the gradient-energy term.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef OR_CFLAGS
#define OR_CFLAGS "unknown flags"
#endif

/* The compiler and flags this library was built with (recorded beside CPU timings). */
const char* or_build_info(void) { return "gcc " __VERSION__ "; " OR_CFLAGS; }

typedef struct {
  int nx, ny, nz;
  double hx, hy, hz;
} or_grid;

/* status codes shared with the Python wrapper */
enum { OR_OK = 0, OR_EXTENT = 1, OR_ARG = 2, OR_CFL = 3, OR_NONPOS_D = 4, OR_NOMEM = 5 };

/* core/IndexPolicy.hpp:28 -- Periodic::wrap */
static inline int or_wrap(int i, int n) { return ((i % n) + n) % n; }

/* core/Grid.hpp:132-137 + core/Field.hpp:100-104 (wrapped access) */
static inline size_t at(const or_grid* g, int i, int j, int k) {
  return (size_t)or_wrap(i, g->nx) +
         (size_t)g->nx * ((size_t)or_wrap(j, g->ny) + (size_t)g->ny * (size_t)or_wrap(k, g->nz));
}

static inline size_t gsize(const or_grid* g) {
  return (size_t)g->nx * (size_t)g->ny * (size_t)g->nz;
}

/* ---- stencil tables: stencil/CentralDifference.hpp:41-74,
 *      stencil/FirstDerivative.hpp:51-86, stencil/FourthDerivative.hpp:49-83 ---- */
static const int OFF3[3] = {-1, 0, 1};
static const int OFF5[5] = {-2, -1, 0, 1, 2};
static const int OFF7[7] = {-3, -2, -1, 0, 1, 2, 3};
static const double D2_O2[3] = {1.0, -2.0, 1.0};
static const double D2_O4[5] = {-1.0 / 12.0, 4.0 / 3.0, -5.0 / 2.0, 4.0 / 3.0, -1.0 / 12.0};
static const double D1_O2[3] = {-0.5, 0.0, 0.5};
static const double D1_O4[5] = {1.0 / 12.0, -2.0 / 3.0, 0.0, 2.0 / 3.0, -1.0 / 12.0};
static const double D4_O2[5] = {1.0, -4.0, 6.0, -4.0, 1.0};
static const double D4_O4[7] = {-1.0 / 6.0, 2.0, -13.0 / 2.0, 28.0 / 3.0, -13.0 / 2.0, 2.0, -1.0 / 6.0};

typedef struct {
  int n, radius;
  const int* off;
  const double* w;
} or_stencil;

static int stencil_d2(int order, or_stencil* s) {
  if (order == 2) { s->n = 3; s->radius = 1; s->off = OFF3; s->w = D2_O2; return 0; }
  if (order == 4) { s->n = 5; s->radius = 2; s->off = OFF5; s->w = D2_O4; return 0; }
  return -1;
}
static int stencil_d1(int order, or_stencil* s) {
  if (order == 2) { s->n = 3; s->radius = 1; s->off = OFF3; s->w = D1_O2; return 0; }
  if (order == 4) { s->n = 5; s->radius = 2; s->off = OFF5; s->w = D1_O4; return 0; }
  return -1;
}
/* stencil/ThirdDerivative.hpp */
static const double D3_O2[5] = {-1.0 / 2.0, 1.0, 0.0, -1.0, 1.0 / 2.0};
static const double D3_O4[7] = {1.0 / 8.0, -1.0, 13.0 / 8.0, 0.0, -13.0 / 8.0, 1.0, -1.0 / 8.0};
static int stencil_d3(int order, or_stencil* s) {
  if (order == 2) { s->n = 5; s->radius = 2; s->off = OFF5; s->w = D3_O2; return 0; }
  if (order == 4) { s->n = 7; s->radius = 3; s->off = OFF7; s->w = D3_O4; return 0; }
  return -1;
}
static int stencil_d4(int order, or_stencil* s) {
  if (order == 2) { s->n = 5; s->radius = 2; s->off = OFF5; s->w = D4_O2; return 0; }
  if (order == 4) { s->n = 7; s->radius = 3; s->off = OFF7; s->w = D4_O4; return 0; }
  return -1;
}

/* Weight-table accessor for tests: kind 1/2/3/4 = first/second/third/fourth derivative. */
int or_stencil_table(int kind, int order, int* n, int* radius, int* offsets, double* weights) {
  or_stencil s;
  int rc = kind == 1 ? stencil_d1(order, &s) : kind == 2 ? stencil_d2(order, &s)
         : kind == 3 ? stencil_d3(order, &s) : kind == 4 ? stencil_d4(order, &s) : -1;
  if (rc) return OR_ARG;
  *n = s.n; *radius = s.radius;
  for (int m = 0; m < s.n; ++m) { offsets[m] = s.off[m]; weights[m] = s.w[m]; }
  return OR_OK;
}

/* diff/detail/PreconditionAxisExtent.hpp:64-79 */
static int axis_ok(int n, int radius) {
  if (radius == 0 || n == 1) return 1;
  return n >= 2 * radius + 1;
}
static int extents_ok(const or_grid* g, int radius) {
  return axis_ok(g->nx, radius) && axis_ok(g->ny, radius) && axis_ok(g->nz, radius);
}

/* ---------------- diff operators ---------------- */

/* diff/Laplacian.hpp:44-75 */
int or_laplacian(const or_grid* g, int order, const double* f, double* out) {
  or_stencil s;
  if (stencil_d2(order, &s)) return OR_ARG;
  if (!extents_ok(g, s.radius)) return OR_EXTENT;
  const double ihx2 = 1.0 / (g->hx * g->hx);
  const double ihy2 = 1.0 / (g->hy * g->hy);
  const double ihz2 = 1.0 / (g->hz * g->hz);
  for (int k = 0; k < g->nz; ++k)
    for (int j = 0; j < g->ny; ++j)
      for (int i = 0; i < g->nx; ++i) {
        double dxx = 0.0, dyy = 0.0, dzz = 0.0;
        for (int m = 0; m < s.n; ++m) {
          const int o = s.off[m];
          const double w = s.w[m];
          dxx += w * f[at(g, i + o, j, k)];
          dyy += w * f[at(g, i, j + o, k)];
          dzz += w * f[at(g, i, j, k + o)];
        }
        out[at(g, i, j, k)] = dxx * ihx2 + dyy * ihy2 + dzz * ihz2;
      }
  return OR_OK;
}

/* diff/Gradient.hpp:43-74 -- scalar -> Vec3 (AoS, 3 doubles per site) */
int or_gradient(const or_grid* g, int order, const double* f, double* out3) {
  or_stencil s;
  if (stencil_d1(order, &s)) return OR_ARG;
  if (!extents_ok(g, s.radius)) return OR_EXTENT;
  const double ihx = 1.0 / g->hx, ihy = 1.0 / g->hy, ihz = 1.0 / g->hz;
  for (int k = 0; k < g->nz; ++k)
    for (int j = 0; j < g->ny; ++j)
      for (int i = 0; i < g->nx; ++i) {
        double dx = 0.0, dy = 0.0, dz = 0.0;
        for (int m = 0; m < s.n; ++m) {
          const int o = s.off[m];
          const double w = s.w[m];
          dx += w * f[at(g, i + o, j, k)];
          dy += w * f[at(g, i, j + o, k)];
          dz += w * f[at(g, i, j, k + o)];
        }
        const size_t c = at(g, i, j, k);
        out3[3 * c + 0] = dx * ihx;
        out3[3 * c + 1] = dy * ihy;
        out3[3 * c + 2] = dz * ihz;
      }
  return OR_OK;
}

/* diff/Gradient.hpp:100-137 -- Vec3 -> Mat3 (column-major 3x3: M(r,c) at r + 3c),
 * M(comp, axis) = d_axis v_comp */
int or_gradient_vec(const or_grid* g, int order, const double* v3, double* out9) {
  or_stencil s;
  if (stencil_d1(order, &s)) return OR_ARG;
  if (!extents_ok(g, s.radius)) return OR_EXTENT;
  const double ihx = 1.0 / g->hx, ihy = 1.0 / g->hy, ihz = 1.0 / g->hz;
  for (int k = 0; k < g->nz; ++k)
    for (int j = 0; j < g->ny; ++j)
      for (int i = 0; i < g->nx; ++i) {
        const size_t c = at(g, i, j, k);
        for (int comp = 0; comp < 3; ++comp) {
          double dx = 0.0, dy = 0.0, dz = 0.0;
          for (int m = 0; m < s.n; ++m) {
            const int o = s.off[m];
            const double w = s.w[m];
            dx += w * v3[3 * at(g, i + o, j, k) + comp];
            dy += w * v3[3 * at(g, i, j + o, k) + comp];
            dz += w * v3[3 * at(g, i, j, k + o) + comp];
          }
          out9[9 * c + comp + 0] = dx * ihx;
          out9[9 * c + comp + 3] = dy * ihy;
          out9[9 * c + comp + 6] = dz * ihz;
        }
      }
  return OR_OK;
}

/* diff/Divergence.hpp:43-74 -- Vec3 -> scalar */
int or_divergence(const or_grid* g, int order, const double* v3, double* out) {
  or_stencil s;
  if (stencil_d1(order, &s)) return OR_ARG;
  if (!extents_ok(g, s.radius)) return OR_EXTENT;
  const double ihx = 1.0 / g->hx, ihy = 1.0 / g->hy, ihz = 1.0 / g->hz;
  for (int k = 0; k < g->nz; ++k)
    for (int j = 0; j < g->ny; ++j)
      for (int i = 0; i < g->nx; ++i) {
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int m = 0; m < s.n; ++m) {
          const int o = s.off[m];
          const double w = s.w[m];
          ax += w * v3[3 * at(g, i + o, j, k) + 0];
          ay += w * v3[3 * at(g, i, j + o, k) + 1];
          az += w * v3[3 * at(g, i, j, k + o) + 2];
        }
        out[at(g, i, j, k)] = ax * ihx + ay * ihy + az * ihz;
      }
  return OR_OK;
}

/* diff/Divergence.hpp:100-135 -- Mat3 (column-major) -> Vec3 */
int or_divergence_mat(const or_grid* g, int order, const double* m9, double* out3) {
  or_stencil s;
  if (stencil_d1(order, &s)) return OR_ARG;
  if (!extents_ok(g, s.radius)) return OR_EXTENT;
  const double ihx = 1.0 / g->hx, ihy = 1.0 / g->hy, ihz = 1.0 / g->hz;
  for (int k = 0; k < g->nz; ++k)
    for (int j = 0; j < g->ny; ++j)
      for (int i = 0; i < g->nx; ++i) {
        const size_t c = at(g, i, j, k);
        for (int comp = 0; comp < 3; ++comp) {
          double dx = 0.0, dy = 0.0, dz = 0.0;
          for (int m = 0; m < s.n; ++m) {
            const int o = s.off[m];
            const double w = s.w[m];
            dx += w * m9[9 * at(g, i + o, j, k) + comp + 0];
            dy += w * m9[9 * at(g, i, j + o, k) + comp + 3];
            dz += w * m9[9 * at(g, i, j, k + o) + comp + 6];
          }
          out3[3 * c + comp] = dx * ihx + dy * ihy + dz * ihz;
        }
      }
  return OR_OK;
}

/* diff/Biharmonic.hpp:48-104 -- direct stencil: S4 loop, then the mixed loop
 * (mb outer, ma inner), then the final expression as written. */
int or_biharmonic(const or_grid* g, int order, const double* f, double* out) {
  or_stencil s2, s4;
  if (stencil_d2(order, &s2) || stencil_d4(order, &s4)) return OR_ARG;
  const int rmax = s4.radius > s2.radius ? s4.radius : s2.radius;
  if (!extents_ok(g, rmax)) return OR_EXTENT;
  const double ihx2 = 1.0 / (g->hx * g->hx);
  const double ihy2 = 1.0 / (g->hy * g->hy);
  const double ihz2 = 1.0 / (g->hz * g->hz);
  const double ihx4 = ihx2 * ihx2, ihy4 = ihy2 * ihy2, ihz4 = ihz2 * ihz2;
  for (int k = 0; k < g->nz; ++k)
    for (int j = 0; j < g->ny; ++j)
      for (int i = 0; i < g->nx; ++i) {
        double dx4 = 0.0, dy4 = 0.0, dz4 = 0.0;
        for (int m = 0; m < s4.n; ++m) {
          const int o = s4.off[m];
          const double w = s4.w[m];
          dx4 += w * f[at(g, i + o, j, k)];
          dy4 += w * f[at(g, i, j + o, k)];
          dz4 += w * f[at(g, i, j, k + o)];
        }
        double dx2y2 = 0.0, dx2z2 = 0.0, dy2z2 = 0.0;
        for (int mb = 0; mb < s2.n; ++mb) {
          const int ob = s2.off[mb];
          const double wb = s2.w[mb];
          for (int ma = 0; ma < s2.n; ++ma) {
            const int oa = s2.off[ma];
            const double wab = s2.w[ma] * wb;
            dx2y2 += wab * f[at(g, i + oa, j + ob, k)];
            dx2z2 += wab * f[at(g, i + oa, j, k + ob)];
            dy2z2 += wab * f[at(g, i, j + oa, k + ob)];
          }
        }
        out[at(g, i, j, k)] = dx4 * ihx4 + dy4 * ihy4 + dz4 * ihz4 +
                              2.0 * (dx2y2 * ihx2 * ihy2 + dx2z2 * ihx2 * ihz2 +
                                     dy2z2 * ihy2 * ihz2);
      }
  return OR_OK;
}

/* ---------------- finite volume ---------------- */

/* fvm/CellLaplacian.hpp:48-50 */
static inline double hmean(double a, double b) { return 2.0 * a * b / (a + b); }

/* diff/detail/MultiIndexDerivative.hpp:46-84: the axis stencil of an n-th derivative
 * (order 0 is the identity) */
static const int OFF1[1] = {0};
static const double D0_W[1] = {1.0};
static int axis_stencil(int n, int order, or_stencil* s) {
  if (n == 0) { s->n = 1; s->radius = 0; s->off = OFF1; s->w = D0_W; return 0; }
  if (n == 1) return stencil_d1(order, s);
  if (n == 2) return stencil_d2(order, s);
  if (n == 3) return stencil_d3(order, s);
  if (n == 4) return stencil_d4(order, s);
  return -1;
}

/* diff/detail/MultiIndexDerivative.hpp:96-136 */
int or_mixed_partial(const or_grid* g, int nx, int ny, int nz, int order, const double* f,
                     double* out) {
  or_stencil sx, sy, sz;
  if (nx < 0 || ny < 0 || nz < 0 || nx + ny + nz < 1 || nx + ny + nz > 4) return OR_ARG;
  if (axis_stencil(nx, order, &sx) || axis_stencil(ny, order, &sy) ||
      axis_stencil(nz, order, &sz))
    return OR_ARG;
  if (!axis_ok(g->nx, sx.radius) || !axis_ok(g->ny, sy.radius) || !axis_ok(g->nz, sz.radius))
    return OR_EXTENT;
  double scale = 1.0;
  for (int n = 0; n < nx; ++n) scale /= g->hx;
  for (int n = 0; n < ny; ++n) scale /= g->hy;
  for (int n = 0; n < nz; ++n) scale /= g->hz;
  for (int k = 0; k < g->nz; ++k)
    for (int j = 0; j < g->ny; ++j)
      for (int i = 0; i < g->nx; ++i) {
        double sum = 0.0;
        for (int mz = 0; mz < sz.n; ++mz) {
          const int oz = sz.off[mz];
          const double wz = sz.w[mz];
          for (int my = 0; my < sy.n; ++my) {
            const int oy = sy.off[my];
            const double wy = sy.w[my];
            const double wyz = wy * wz;
            for (int mx = 0; mx < sx.n; ++mx)
              sum += sx.w[mx] * wyz * f[at(g, i + sx.off[mx], j + oy, k + oz)];
          }
        }
        out[at(g, i, j, k)] = sum * scale;
      }
  return OR_OK;
}

/* fvm/CellLaplacian.hpp:91-143.  On a non-positive D returns OR_NONPOS_D and
 * stores the first violating flat index in *bad_index (CellLaplacian.hpp:99-109). */
int or_cell_laplacian(const or_grid* g, const double* psi, const double* D, double* out,
                      long long* bad_index) {
  if (!extents_ok(g, 1)) return OR_EXTENT;
  const size_t n = gsize(g);
  for (size_t idx = 0; idx < n; ++idx) {
    if (!(D[idx] > 0.0)) {
      if (bad_index) *bad_index = (long long)idx;
      return OR_NONPOS_D;
    }
  }
  const double ihx2 = 1.0 / (g->hx * g->hx);
  const double ihy2 = 1.0 / (g->hy * g->hy);
  const double ihz2 = 1.0 / (g->hz * g->hz);
  for (int k = 0; k < g->nz; ++k)
    for (int j = 0; j < g->ny; ++j)
      for (int i = 0; i < g->nx; ++i) {
        const double pc = psi[at(g, i, j, k)];
        const double Dc = D[at(g, i, j, k)];
        const double dxp = hmean(Dc, D[at(g, i + 1, j, k)]);
        const double dxm = hmean(D[at(g, i - 1, j, k)], Dc);
        const double dyp = hmean(Dc, D[at(g, i, j + 1, k)]);
        const double dym = hmean(D[at(g, i, j - 1, k)], Dc);
        const double dzp = hmean(Dc, D[at(g, i, j, k + 1)]);
        const double dzm = hmean(D[at(g, i, j, k - 1)], Dc);
        const double rx = (dxp * (psi[at(g, i + 1, j, k)] - pc) -
                           dxm * (pc - psi[at(g, i - 1, j, k)])) * ihx2;
        const double ry = (dyp * (psi[at(g, i, j + 1, k)] - pc) -
                           dym * (pc - psi[at(g, i, j - 1, k)])) * ihy2;
        const double rz = (dzp * (psi[at(g, i, j, k + 1)] - pc) -
                           dzm * (pc - psi[at(g, i, j, k - 1)])) * ihz2;
        out[at(g, i, j, k)] = rx + ry + rz;
      }
  return OR_OK;
}

/* solve/Diffusion.hpp:187-207 -- largest harmonic-mean face D over +x,+y,+z faces */
double or_max_face_d(const or_grid* g, const double* D) {
  double dmax = 0.0;
  for (int k = 0; k < g->nz; ++k) {
    const int kp = (k + 1) % g->nz;
    for (int j = 0; j < g->ny; ++j) {
      const int jp = (j + 1) % g->ny;
      for (int i = 0; i < g->nx; ++i) {
        const int ip = (i + 1) % g->nx;
        const double Dc = D[at(g, i, j, k)];
        const double fx = hmean(Dc, D[at(g, ip, j, k)]);
        const double fy = hmean(Dc, D[at(g, i, jp, k)]);
        const double fz = hmean(Dc, D[at(g, i, j, kp)]);
        if (fx > dmax) dmax = fx;
        if (fy > dmax) dmax = fy;
        if (fz > dmax) dmax = fz;
      }
    }
  }
  return dmax;
}

/* ---------------- reductions ---------------- */

/* func/Integrate.hpp:65-74 */
double or_pairwise_sum(const double* x, size_t n) {
  if (n <= 64) {
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += x[i];
    return s;
  }
  const size_t mid = n / 2;
  return or_pairwise_sum(x, mid) + or_pairwise_sum(x + mid, n - mid);
}

/* func/Integrate.hpp:85-99 */
double or_neumaier_sum(const double* x, size_t n) {
  double sum = 0.0, c = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double v = x[i];
    const double t = sum + v;
    if (fabs(sum) >= fabs(v)) c += (sum - t) + v;
    else c += (v - t) + sum;
    sum = t;
  }
  return sum + c;
}

/* core/Grid.hpp:104-106 */
static inline double cell_volume(const or_grid* g) { return g->hx * g->hy * g->hz; }

/* func/Integrate.hpp:116-120 (method 0) and :131-135 (method 1) */
double or_integrate(const or_grid* g, int method, const double* x) {
  const size_t n = gsize(g);
  return cell_volume(g) * (method == 0 ? or_pairwise_sum(x, n) : or_neumaier_sum(x, n));
}

/* func/Functional.hpp:122-130 (2-argument density f(psi, grad psi), gradient at
 * Order 2) with the Cahn-Hilliard density of examples/common/cahn_hilliard.hpp:64-68:
 * -0.5*p*p + 0.25*p*p*p*p + 0.5*kappa*|g|^2, reduced by method 0/1. */
int or_free_energy(const or_grid* g, double kappa, int method, const double* psi, double* F) {
  const size_t n = gsize(g);
  double* grad = (double*)malloc(3 * n * sizeof(double));
  double* dens = (double*)malloc(n * sizeof(double));
  if (!grad || !dens) { free(grad); free(dens); return OR_NOMEM; }
  int rc = or_gradient(g, 2, psi, grad);
  if (rc) { free(grad); free(dens); return rc; }
  for (size_t c = 0; c < n; ++c) {
    const double p = psi[c];
    const double gx = grad[3 * c], gy = grad[3 * c + 1], gz = grad[3 * c + 2];
    const double sq = gx * gx + gy * gy + gz * gz;
    dens[c] = -0.5 * p * p + 0.25 * p * p * p * p + 0.5 * kappa * sq;
  }
  *F = or_integrate(g, method, dens);
  free(grad);
  free(dens);
  return OR_OK;
}

/* examples/common/cahn_hilliard.hpp:73-77 -- integral of |grad psi|^2 */
int or_gradient_energy(const or_grid* g, int method, const double* psi, double* E) {
  const size_t n = gsize(g);
  double* grad = (double*)malloc(3 * n * sizeof(double));
  double* dens = (double*)malloc(n * sizeof(double));
  if (!grad || !dens) { free(grad); free(dens); return OR_NOMEM; }
  int rc = or_gradient(g, 2, psi, grad);
  if (rc) { free(grad); free(dens); return rc; }
  for (size_t c = 0; c < n; ++c) {
    const double gx = grad[3 * c], gy = grad[3 * c + 1], gz = grad[3 * c + 2];
    dens[c] = gx * gx + gy * gy + gz * gz;
  }
  *E = or_integrate(g, method, dens);
  free(grad);
  free(dens);
  return OR_OK;
}

/* func/Functional.hpp:88-148 (scalar field; the Order-2 gradient, Laplacian and biharmonic
 * the density needs are materialised first, in that order, as the reference does) and
 * :183-220 (Vec3 field; its Order-2 rank-2 gradient), then the density site by site and
 * or_integrate.  kind: 0 p; 1 |g|^2; 2 Cahn-Hilliard (a = kappa); 3 Ginzburg-Landau
 * (a = W); 4 phase-field crystal (a = q0, b = eps); 5 linear elastic (Vec3 field,
 * a = lambda, b = mu).  The formulas are those of examples/common/cahn_hilliard.hpp:64-76,
 * test/functional_test.cpp:45-48, test/pfc_functional_test.cpp:44-50 and
 * examples/common/elastic_energy.hpp:81-88, with |g|^2 = (gx*gx + gy*gy) + gz*gz and the
 * strain sums over the 3x3 entries in column-major order. */
int or_evaluate(const or_grid* g, int kind, double a, double b, int method, const double* f,
                double* result) {
  const size_t n = gsize(g);
  if (kind < 0 || kind > 5) return OR_ARG;
  if (kind != 0 && !extents_ok(g, 1)) return OR_EXTENT;
  double* dens = (double*)malloc(n * sizeof(double));
  double* d1 = (double*)malloc(9 * n * sizeof(double));
  double* d2 = (double*)malloc(n * sizeof(double));
  double* d4 = (double*)malloc(n * sizeof(double));
  int rc = (!dens || !d1 || !d2 || !d4) ? OR_NOMEM : OR_OK;
  if (!rc && kind == 5) rc = or_gradient_vec(g, 2, f, d1);
  if (!rc && kind >= 1 && kind <= 4) rc = or_gradient(g, 2, f, d1);
  if (!rc && (kind == 3 || kind == 4)) rc = or_laplacian(g, 2, f, d2);
  if (!rc && kind == 4) rc = or_biharmonic(g, 2, f, d4);
  if (!rc) {
    for (size_t c = 0; c < n; ++c) {
      if (kind == 5) {
        const double* G = d1 + 9 * c; /* G(r, col) at r + 3 col */
        double e[9];
        for (int col = 0; col < 3; ++col)
          for (int r = 0; r < 3; ++r) e[r + 3 * col] = 0.5 * (G[r + 3 * col] + G[col + 3 * r]);
        const double tr = e[0] + e[4] + e[8];
        double ee = e[0] * e[0];
        for (int q = 1; q < 9; ++q) ee += e[q] * e[q];
        dens[c] = 0.5 * (a * tr * tr + 2.0 * b * ee);
        continue;
      }
      const double p = f[c];
      if (kind == 0) {
        dens[c] = p;
        continue;
      }
      const double gx = d1[3 * c], gy = d1[3 * c + 1], gz = d1[3 * c + 2];
      const double sq = gx * gx + gy * gy + gz * gz;
      if (kind == 1) {
        dens[c] = sq;
      } else if (kind == 2) {
        dens[c] = -0.5 * p * p + 0.25 * p * p * p * p + 0.5 * a * sq;
      } else if (kind == 3) {
        const double q = p * p - 1.0;
        dens[c] = 0.5 * sq + a * q * q;
      } else {
        const double q2 = a * a;
        const double q4 = q2 * q2;
        const double p2 = p * p;
        dens[c] = 0.5 * (q4 - b) * p2 + q2 * p * d2[c] + 0.5 * p * d4[c] + 0.25 * p2 * p2;
      }
    }
    *result = or_integrate(g, method, dens);
  }
  free(dens);
  free(d1);
  free(d2);
  free(d4);
  return rc;
}

/* ---------------- right-hand sides ---------------- */

typedef int (*or_rhs_fn)(const void* ctx, const or_grid* g, const double* in, double* out);

typedef struct { double M, kappa; } or_ch_params;

/* examples/common/cahn_hilliard.hpp:38-59 -- M * lap(-p + p^3 - kappa*lap p), Order 2 */
int or_ch_rhs(const void* ctx, const or_grid* g, const double* psi, double* out) {
  const or_ch_params* P = (const or_ch_params*)ctx;
  const size_t n = gsize(g);
  double* lap = (double*)malloc(n * sizeof(double));
  double* mu = (double*)malloc(n * sizeof(double));
  if (!lap || !mu) { free(lap); free(mu); return OR_NOMEM; }
  int rc = or_laplacian(g, 2, psi, lap);
  if (!rc) {
    for (size_t c = 0; c < n; ++c) {
      const double p = psi[c];
      mu[c] = -p + p * p * p - P->kappa * lap[c];
    }
    rc = or_laplacian(g, 2, mu, out);
  }
  if (!rc)
    for (size_t c = 0; c < n; ++c) out[c] *= P->M;
  free(lap);
  free(mu);
  return rc;
}

/* solve/Diffusion.hpp:108 -- rhs = laplacian<2>(f) */
static int lap2_rhs(const void* ctx, const or_grid* g, const double* in, double* out) {
  (void)ctx;
  return or_laplacian(g, 2, in, out);
}

/* solve/Diffusion.hpp:212 -- rhs = cellLaplacian(f, D) */
typedef struct { const double* D; long long bad_index; } or_fvm_ctx;
static int fvm_rhs(const void* ctx, const or_grid* g, const double* in, double* out) {
  or_fvm_ctx* c = (or_fvm_ctx*)ctx;
  return or_cell_laplacian(g, in, c->D, out, &c->bad_index);
}

/* ---------------- integrators ---------------- */

/* solve/ExplicitEuler.hpp:79-112 */
int or_euler(const or_grid* g, double* psi, or_rhs_fn rhs, const void* ctx, double dt, int n_steps) {
  if (n_steps < 0 || dt < 0.0) return OR_ARG;
  const size_t N = gsize(g);
  double* r = (double*)malloc(N * sizeof(double));
  if (!r) return OR_NOMEM;
  for (int step = 0; step < n_steps; ++step) {
    int rc = rhs(ctx, g, psi, r);
    if (rc) { free(r); return rc; }
    for (size_t c = 0; c < N; ++c) psi[c] += dt * r[c];
  }
  free(r);
  return OR_OK;
}

/* solve/RK4.hpp:51-64 */
static void axpy_fresh(size_t N, const double* a, double scale, const double* b, double* out) {
  for (size_t n = 0; n < N; ++n) out[n] = a[n] + scale * b[n];
}

/* solve/RK4.hpp:99-154 */
int or_rk4(const or_grid* g, double* psi, or_rhs_fn rhs, const void* ctx, double dt, int n_steps) {
  if (n_steps < 0 || dt < 0.0) return OR_ARG;
  const size_t N = gsize(g);
  const double half_dt = 0.5 * dt;
  const double sixth_dt = dt / 6.0;
  const double third_dt = dt / 3.0;
  double* buf = (double*)malloc(5 * N * sizeof(double));
  if (!buf) return OR_NOMEM;
  double *k1 = buf, *k2 = buf + N, *k3 = buf + 2 * N, *k4 = buf + 3 * N, *y = buf + 4 * N;
  int rc = OR_OK;
  for (int step = 0; step < n_steps && !rc; ++step) {
    if ((rc = rhs(ctx, g, psi, k1))) break;
    axpy_fresh(N, psi, half_dt, k1, y);
    if ((rc = rhs(ctx, g, y, k2))) break;
    axpy_fresh(N, psi, half_dt, k2, y);
    if ((rc = rhs(ctx, g, y, k3))) break;
    axpy_fresh(N, psi, dt, k3, y);
    if ((rc = rhs(ctx, g, y, k4))) break;
    for (size_t n = 0; n < N; ++n)
      psi[n] += sixth_dt * (k1[n] + k4[n]) + third_dt * (k2[n] + k3[n]);
  }
  free(buf);
  return rc;
}

/* Cahn-Hilliard run as examples/cahn_hilliard.cpp:171-179 drives it: RK4 (method 1)
 * or explicit Euler (method 0) over the computeRhs right-hand side. */
int or_ch_integrate(const or_grid* g, double M, double kappa, double* psi, double dt,
                    int n_steps, int method) {
  or_ch_params P = {M, kappa};
  return method == 1 ? or_rk4(g, psi, or_ch_rhs, &P, dt, n_steps)
                     : or_euler(g, psi, or_ch_rhs, &P, dt, n_steps);
}

/* solve/Diffusion.hpp:45-59 -- returns 1 when the CFL bound is violated */
int or_cfl_violated(const or_grid* g, double D, double dt, double limit) {
  const double inv_sum = 1.0 / (g->hx * g->hx) + 1.0 / (g->hy * g->hy) + 1.0 / (g->hz * g->hz);
  const double cfl = D * dt * inv_sum;
  return cfl > limit;
}

static double cfl_limit(int method) { return method == 1 ? 0.6963 : 0.5; }

/* solve/Diffusion.hpp:89-112 -- constant D folded into the step: integrate with D*dt */
int or_diffuse_const(const or_grid* g, double* psi, double D, double dt, int n_steps, int method) {
  if (D < 0.0 || dt < 0.0 || n_steps < 0) return OR_ARG;
  if (or_cfl_violated(g, D, dt, cfl_limit(method))) return OR_CFL;
  return method == 1 ? or_rk4(g, psi, lap2_rhs, NULL, D * dt, n_steps)
                     : or_euler(g, psi, lap2_rhs, NULL, D * dt, n_steps);
}

/* solve/Diffusion.hpp:161-216 -- variable D via the FVM cell Laplacian */
int or_diffuse_field(const or_grid* g, double* psi, const double* D, double dt, int n_steps,
                     int method, long long* bad_index) {
  if (dt < 0.0 || n_steps < 0) return OR_ARG;
  const double dmax = or_max_face_d(g, D);
  if (or_cfl_violated(g, dmax, dt, cfl_limit(method))) return OR_CFL;
  or_fvm_ctx c = {D, -1};
  int rc = method == 1 ? or_rk4(g, psi, fvm_rhs, &c, dt, n_steps)
                       : or_euler(g, psi, fvm_rhs, &c, dt, n_steps);
  if (bad_index) *bad_index = c.bad_index;
  return rc;
}
