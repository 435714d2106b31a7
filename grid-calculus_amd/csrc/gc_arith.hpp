// gc_arith.hpp -- the per-site arithmetic of the Cahn-Hilliard and diffusion right-hand sides
// and of the RK4 / Euler stage updates, shared by every kernel family (one-stage, two-stage,
// untiled, ensemble) so that all tile shapes and families give the same bits.
//
// FP64: the reference's operations in the reference's order, every + and * rounded as written
//   (the library is compiled with -ffp-contract=off):
//     Laplacian           diff/Laplacian.hpp:60-70 (per-axis sums from 0 in offset order)
//     mu, M lap(mu)       examples/common/cahn_hilliard.hpp:41-57
//     cell Laplacian      fvm/CellLaplacian.hpp:119-138
//     stage updates       solve/RK4.hpp:127-153, solve/ExplicitEuler.hpp:107-126
//   so FP64 results are bit-identical to the reference compiled without contraction.
// FP32: this library's own arithmetic.  kappa, M and the inverse squared cell sizes are folded
//   into the weights once on the host, each axis' two neighbours are summed first and every
//   weight is applied with a fused multiply-add: about half the operations of the FP64 form.
//   FP32 results are compared with the FP64 oracle within stated bounds, never bit for bit.
#pragma once

#include <hip/hip_runtime.h>

namespace gc {

// ---- Cahn-Hilliard coefficients -------------------------------------------------------------

template <class T>
struct CHCoef {
  T M, kappa, ihx2, ihy2, ihz2;
  // FP32 form: mu = c^3 - c + Σ_axis (n- + n+) ua + c u0, k = Σ_axis (n- + n+) ka + c k0
  T ux, uy, uz, u0, u02;  // ua = -kappa/h_a^2, u0 = 2 kappa Σ 1/h^2 (u02: x and y only)
  T kx, ky, kz, k0, k02;  // ka = M/h_a^2, k0 = -2 M Σ 1/h^2 (k02: x and y only)
};

// Folded weights computed in double and rounded once to T.
template <class T>
__host__ __device__ inline CHCoef<T> make_ch_coef(T M, T kappa, T ihx2, T ihy2, T ihz2) {
  const double m = M, k = kappa, x = ihx2, y = ihy2, z = ihz2;
  CHCoef<T> c{M, kappa, ihx2, ihy2, ihz2};
  c.ux = T(-k * x);
  c.uy = T(-k * y);
  c.uz = T(-k * z);
  c.u0 = T(2.0 * k * (x + y + z));
  c.u02 = T(2.0 * k * (x + y));
  c.kx = T(m * x);
  c.ky = T(m * y);
  c.kz = T(m * z);
  c.k0 = T(-2.0 * m * (x + y + z));
  c.k02 = T(-2.0 * m * (x + y));
  return c;
}

// Order-2 Laplacian of the reference from the centre and its six neighbours (x-, x+, y-, y+,
// z-, z+): per-axis sums from 0 over offsets -1, 0, +1, then dxx ihx2 + dyy ihy2 + dzz ihz2.
__device__ __forceinline__ double lap7_ref(double c, double xm, double xp, double ym, double yp,
                                           double zm, double zp, double ihx2, double ihy2,
                                           double ihz2) {
  double dxx = 0, dyy = 0, dzz = 0;
  dxx += xm;
  dyy += ym;
  dzz += zm;
  dxx += -2.0 * c;
  dyy += -2.0 * c;
  dzz += -2.0 * c;
  dxx += xp;
  dyy += yp;
  dzz += zp;
  return dxx * ihx2 + dyy * ihy2 + dzz * ihz2;
}

// mu = -c + c^3 - kappa lap(c)
__device__ __forceinline__ double ch_mu(const CHCoef<double>& k, double c, double xm, double xp,
                                        double ym, double yp, double zm, double zp) {
  const double l = lap7_ref(c, xm, xp, ym, yp, zm, zp, k.ihx2, k.ihy2, k.ihz2);
  return -c + c * c * c - k.kappa * l;
}
__device__ __forceinline__ float ch_mu(const CHCoef<float>& k, float c, float xm, float xp,
                                       float ym, float yp, float zm, float zp) {
  float t = fmaf(c * c, c, -c);
  t = fmaf(c, k.u0, t);
  t = fmaf(zm + zp, k.uz, t);
  t = fmaf(ym + yp, k.uy, t);
  return fmaf(xm + xp, k.ux, t);
}

// k = lap(mu) M
__device__ __forceinline__ double ch_k(const CHCoef<double>& k, double c, double xm, double xp,
                                       double ym, double yp, double zm, double zp) {
  return lap7_ref(c, xm, xp, ym, yp, zm, zp, k.ihx2, k.ihy2, k.ihz2) * k.M;
}
__device__ __forceinline__ float ch_k(const CHCoef<float>& k, float c, float xm, float xp,
                                      float ym, float yp, float zm, float zp) {
  float t = c * k.k0;
  t = fmaf(zm + zp, k.kz, t);
  t = fmaf(ym + yp, k.ky, t);
  return fmaf(xm + xp, k.kx, t);
}

// Nz = 1.  FP64: the reference forms the z terms from the centre value (its wrapped access),
// exactly as the 3-D form with z- = z+ = c.  FP32: the z terms are left out.
__device__ __forceinline__ double ch_mu2(const CHCoef<double>& k, double c, double xm, double xp,
                                         double ym, double yp) {
  return ch_mu(k, c, xm, xp, ym, yp, c, c);
}
__device__ __forceinline__ float ch_mu2(const CHCoef<float>& k, float c, float xm, float xp,
                                        float ym, float yp) {
  float t = fmaf(c * c, c, -c);
  t = fmaf(c, k.u02, t);
  t = fmaf(ym + yp, k.uy, t);
  return fmaf(xm + xp, k.ux, t);
}
__device__ __forceinline__ double ch_k2(const CHCoef<double>& k, double c, double xm, double xp,
                                        double ym, double yp) {
  return ch_k(k, c, xm, xp, ym, yp, c, c);
}
__device__ __forceinline__ float ch_k2(const CHCoef<float>& k, float c, float xm, float xp,
                                       float ym, float yp) {
  float t = c * k.k02;
  t = fmaf(ym + yp, k.ky, t);
  return fmaf(xm + xp, k.kx, t);
}

// ---- constant-D diffusion: k = lap(y) ---------------------------------------------------------

template <class T>
struct LapCoef {
  T ihx2, ihy2, ihz2;
  T c0;  // FP32 form: -2 (ihx2 + ihy2 + ihz2)
};
template <class T>
inline LapCoef<T> make_lap_coef(T ihx2, T ihy2, T ihz2) {
  return LapCoef<T>{ihx2, ihy2, ihz2, T(-2.0 * ((double)ihx2 + (double)ihy2 + (double)ihz2))};
}
__device__ __forceinline__ double lap7(const LapCoef<double>& k, double c, double xm, double xp,
                                       double ym, double yp, double zm, double zp) {
  return lap7_ref(c, xm, xp, ym, yp, zm, zp, k.ihx2, k.ihy2, k.ihz2);
}
__device__ __forceinline__ float lap7(const LapCoef<float>& k, float c, float xm, float xp,
                                      float ym, float yp, float zm, float zp) {
  float t = c * k.c0;
  t = fmaf(zm + zp, k.ihz2, t);
  t = fmaf(ym + yp, k.ihy2, t);
  return fmaf(xm + xp, k.ihx2, t);
}

// ---- heterogeneous-D diffusion: the cell Laplacian ------------------------------------------

// Face coefficient between cells of diffusivity a and b, for an axis with inverse squared
// cell size ih.  FP64: the harmonic mean 2 a b / (a + b) (fvm/CellLaplacian.hpp:60-64); the
// axis factor is applied per cell later, in the reference's order.  FP32: the harmonic mean
// with the hardware reciprocal (1 ulp) and the axis factor folded in.
__device__ __forceinline__ double face_coef(double a, double b, double /*ih*/) {
  return 2.0 * a * b / (a + b);
}
__device__ __forceinline__ float face_coef(float a, float b, float ih) {
  return (2.0f * ih) * (a * b) * __builtin_amdgcn_rcpf(a + b);
}

// k = cell Laplacian at a cell of value c from its neighbours and the six face coefficients
// (x-, x+, y-, y+, z-, z+ faces; each formed by face_coef).
__device__ __forceinline__ double fvm_k(double c, double xm, double xp, double ym, double yp,
                                        double zm, double zp, double fxm, double fxp, double fym,
                                        double fyp, double fzm, double fzp, double ihx2,
                                        double ihy2, double ihz2) {
  const double rx = (fxp * (xp - c) - fxm * (c - xm)) * ihx2;
  const double ry = (fyp * (yp - c) - fym * (c - ym)) * ihy2;
  const double rz = (fzp * (zp - c) - fzm * (c - zm)) * ihz2;
  return rx + ry + rz;
}
__device__ __forceinline__ float fvm_k(float c, float xm, float xp, float ym, float yp, float zm,
                                       float zp, float fxm, float fxp, float fym, float fyp,
                                       float fzm, float fzp, float, float, float) {
  float t = fzp * (zp - c);
  t = fmaf(fzm, zm - c, t);
  t = fmaf(fyp, yp - c, t);
  t = fmaf(fym, ym - c, t);
  t = fmaf(fxp, xp - c, t);
  return fmaf(fxm, xm - c, t);
}

// ---- stage updates ------------------------------------------------------------------------

// base + h k (the stage values y2 = p + dt/2 k1, y3 = p + dt/2 k2, y4 = p + dt k3 and the
// Euler step p + dt f(p))
__device__ __forceinline__ double stage_axpy(double base, double h, double k) {
  return base + h * k;
}
__device__ __forceinline__ float stage_axpy(float base, float h, float k) {
  return fmaf(h, k, base);
}

// The RK4 update p + (dt/6 (k1 + k4) + dt/3 (k2 + k3)), k23 = k2 + k3 formed by the caller.
__device__ __forceinline__ double rk4_final(double p, double k1, double k23, double k4,
                                            double sixth_dt, double third_dt) {
  return p + (sixth_dt * (k1 + k4) + third_dt * k23);
}
__device__ __forceinline__ float rk4_final(float p, float k1, float k23, float k4, float sixth_dt,
                                           float third_dt) {
  return fmaf(sixth_dt, k1 + k4, fmaf(third_dt, k23, p));
}

}  // namespace gc
