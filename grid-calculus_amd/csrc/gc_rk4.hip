// gc_rk4.hip -- explicit time stepping: the pointwise integrator steps of the generic
// right-hand-side path and the fused stage pipeline for the Cahn-Hilliard, constant-D
// diffusion and heterogeneous-D (FVM) right-hand sides.
//
// Stage pipeline (solve/RK4.hpp:122-153).  The reference computes per step
//   k1 = f(p); k2 = f(p + dt/2 k1); k3 = f(p + dt/2 k2); k4 = f(p + dt k3)
//   p += dt/6 (k1 + k4) + dt/3 (k2 + k3)
// Here each stage is ONE kernel that evaluates the right-hand side at a site and
// applies that stage's update in the same pass:
//   S1: y=p  -> a = k1,          yA = p + dt/2 k1
//   S2: y=yA -> b = k2,          yB = p + dt/2 k2
//   S3: y=yB -> b = b + k3,      yA = p + dt k3
//   S4: y=yA -> p = p + (dt/6 (a + k4) + dt/3 b)
// which performs exactly the reference's roundings (k2 + k3 is formed once, as the
// reference forms p2 + p3).  HBM traffic per step: 3 + 4 + 5 + 5 = 17 words per site.
#include "gc_common.hpp"
#include "gc_arith.hpp"

#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

namespace gc {

gc_status launch_laplacian(gc_context* ctx, const gc_grid* g, gc_dtype dt, int order,
                           const void* in, void* out);
gc_status first_nonpositive(gc_context* ctx, const gc_grid* g, gc_dtype dt, const void* D,
                            int64_t* first_bad);
gc_status max_face_d(gc_context* ctx, const gc_grid* g, gc_dtype dt, const void* D, double* dmax);

namespace {

constexpr int TX = 64, TY = 4;

// ---- pointwise steps --------------------------------------------------------------------

template <class T>
__global__ __launch_bounds__(256) void k_axpy(size_t n, const T* __restrict__ a, T scale,
                                              const T* __restrict__ b, T* __restrict__ out) {
  const size_t G = (size_t)gridDim.x * blockDim.x;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += G)
    out[c] = a[c] + scale * b[c];
}

template <class T>
__global__ __launch_bounds__(256) void k_euler_update(size_t n, T* __restrict__ p, T dt,
                                                      const T* __restrict__ r) {
  const size_t G = (size_t)gridDim.x * blockDim.x;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += G)
    p[c] += dt * r[c];
}

template <class T>
__global__ __launch_bounds__(256) void k_rk4_combine(size_t n, T* __restrict__ p, T sixth,
                                                     T third, const T* __restrict__ k1,
                                                     const T* __restrict__ k2,
                                                     const T* __restrict__ k3,
                                                     const T* __restrict__ k4) {
  const size_t G = (size_t)gridDim.x * blockDim.x;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += G)
    p[c] += sixth * (k1[c] + k4[c]) + third * (k2[c] + k3[c]);
}

inline unsigned pw_blocks(size_t n) {
  const size_t want = (n + 255) / 256;
  return (unsigned)(want < 8192 ? (want == 0 ? 1 : want) : 8192);
}

// ---- right-hand sides ---------------------------------------------------------------------

struct Site {
  int i, j, k;
};

__device__ __forceinline__ size_t lin(const Dims& d, int i, int j, int k) {
  return (size_t)i + (size_t)d.nx * (size_t)j + (size_t)d.plane * (size_t)k;
}

// The seven values of the order-2 Laplacian stencil at (i,j,k): centre, x-, x+, y-, y+, z-, z+.
template <class T>
struct Nb7 {
  T c, xm, xp, ym, yp, zm, zp;
};
template <class T>
__device__ __forceinline__ Nb7<T> nb7_at(const T* __restrict__ y, const Dims& d, int i, int j,
                                         int k) {
  const int im = wrap(i - 1, d.nx), ip = wrap(i + 1, d.nx);
  const int jm = wrap(j - 1, d.ny), jp = wrap(j + 1, d.ny);
  const int km = wrap(k - 1, d.nz), kp = wrap(k + 1, d.nz);
  return Nb7<T>{y[lin(d, i, j, k)],  y[lin(d, im, j, k)], y[lin(d, ip, j, k)],
                y[lin(d, i, jm, k)], y[lin(d, i, jp, k)], y[lin(d, i, j, km)],
                y[lin(d, i, j, kp)]};
}

// Cahn-Hilliard: M * lap(mu), mu = -y + y^3 - kappa*lap(y)
// (examples/common/cahn_hilliard.hpp:41-57; arithmetic in gc_arith.hpp).
template <class T>
struct CHRhs {
  CHCoef<T> cf;
  __device__ __forceinline__ T mu(const T* __restrict__ y, const Dims& d, int i, int j,
                                  int k) const {
    const Nb7<T> n = nb7_at(y, d, i, j, k);
    return ch_mu(cf, n.c, n.xm, n.xp, n.ym, n.yp, n.zm, n.zp);
  }
  __device__ __forceinline__ T operator()(const T* __restrict__ y, const Dims& d, int i, int j,
                                          int k) const {
    const T mc = mu(y, d, i, j, k);
    const T mxm = mu(y, d, wrap(i - 1, d.nx), j, k), mxp = mu(y, d, wrap(i + 1, d.nx), j, k);
    const T mym = mu(y, d, i, wrap(j - 1, d.ny), k), myp = mu(y, d, i, wrap(j + 1, d.ny), k);
    const T mzm = mu(y, d, i, j, wrap(k - 1, d.nz)), mzp = mu(y, d, i, j, wrap(k + 1, d.nz));
    return ch_k(cf, mc, mxm, mxp, mym, myp, mzm, mzp);
  }
};

// solve/Diffusion.hpp:108 -- laplacian<2>
template <class T>
struct LapRhs {
  LapCoef<T> lc;
  __device__ __forceinline__ T operator()(const T* __restrict__ y, const Dims& d, int i, int j,
                                          int k) const {
    const Nb7<T> n = nb7_at(y, d, i, j, k);
    return lap7(lc, n.c, n.xm, n.xp, n.ym, n.yp, n.zm, n.zp);
  }
};

// solve/Diffusion.hpp:212 -- fvm::cellLaplacian (fvm/CellLaplacian.hpp:119-138)
template <class T>
struct FvmRhs {
  const T* __restrict__ D;
  LapCoef<T> lc;
  __device__ __forceinline__ T operator()(const T* __restrict__ y, const Dims& d, int i, int j,
                                          int k) const {
    const size_t c = lin(d, i, j, k);
    const size_t xp = lin(d, wrap(i + 1, d.nx), j, k), xm = lin(d, wrap(i - 1, d.nx), j, k);
    const size_t yp = lin(d, i, wrap(j + 1, d.ny), k), ym = lin(d, i, wrap(j - 1, d.ny), k);
    const size_t zp = lin(d, i, j, wrap(k + 1, d.nz)), zm = lin(d, i, j, wrap(k - 1, d.nz));
    const T Dc = D[c];
    return fvm_k(y[c], y[xm], y[xp], y[ym], y[yp], y[zm], y[zp], face_coef(D[xm], Dc, lc.ihx2),
                 face_coef(Dc, D[xp], lc.ihx2), face_coef(D[ym], Dc, lc.ihy2),
                 face_coef(Dc, D[yp], lc.ihy2), face_coef(D[zm], Dc, lc.ihz2),
                 face_coef(Dc, D[zp], lc.ihz2), lc.ihx2, lc.ihy2, lc.ihz2);
  }
};

// ---- fused stage kernel ---------------------------------------------------------------------

enum Stage : int { kRhsOnly = 0, kS1 = 1, kS2 = 2, kS3 = 3, kS4 = 4, kEuler = 5 };

template <class T>
struct StageCoef {
  T half_dt, dt, sixth_dt, third_dt;
};

// The per-site stage update (see the file comment); kv is the right-hand side at site c.
template <class T>
__device__ __forceinline__ void stage_update_regs(int stage, size_t c, T kv, T pv, T av, T bv,
                                                  T* psi, T* ynext, T* a, T* b,
                                                  const StageCoef<T>& co) {
  switch (stage) {
    case kRhsOnly:
      ynext[c] = kv;
      break;
    case kS1:
      a[c] = kv;
      ynext[c] = stage_axpy(pv, co.half_dt, kv);
      break;
    case kS2:
      b[c] = kv;
      ynext[c] = stage_axpy(pv, co.half_dt, kv);
      break;
    case kS3:
      b[c] = bv + kv;
      ynext[c] = stage_axpy(pv, co.dt, kv);
      break;
    case kS4:
      psi[c] = rk4_final(pv, av, bv, kv, co.sixth_dt, co.third_dt);
      break;
    default:  // explicit Euler: ynext = y + dt*f(y), pv = y[c]
      ynext[c] = stage_axpy(pv, co.dt, kv);
      break;
  }
}

template <class T>
__device__ __forceinline__ void stage_update(int stage, size_t c, T kv, const T* y, T* psi,
                                             T* ynext, T* a, T* b, const StageCoef<T>& co) {
  const T pv = (stage >= kS1 && stage <= kS4) ? psi[c] : (stage == kEuler ? y[c] : T(0));
  const T av = stage == kS4 ? a[c] : T(0);
  const T bv = (stage == kS3 || stage == kS4) ? b[c] : T(0);
  stage_update_regs(stage, c, kv, pv, av, bv, psi, ynext, a, b, co);
}

template <class T, class Rhs>
__global__ __launch_bounds__(256) void k_stage(Rhs rhs, int stage, const T* __restrict__ y,
                                               T* psi, T* __restrict__ ynext, T* __restrict__ a,
                                               T* __restrict__ b, Dims d, int nbx, int nby,
                                               StageCoef<T> co) {
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int bx = (int)(bid % (unsigned)nbx);
  const unsigned r = bid / (unsigned)nbx;
  const int by = (int)(r % (unsigned)nby);
  const int k = (int)(r / (unsigned)nby);
  const int i = bx * TX + (int)(threadIdx.x % TX);
  const int j = by * TY + (int)(threadIdx.x / TX);
  if (i >= d.nx || j >= d.ny) return;
  const size_t c = lin(d, i, j, k);
  const T kv = rhs(y, d, i, j, k);
  stage_update(stage, c, kv, y, psi, ynext, a, b, co);
}

// ---- LDS-tiled Cahn-Hilliard stage kernels ----------------------------------------------
//
// 3-D (z-march).  A block owns an OX x OY tile of xy-columns and a chunk of ZC planes.
// It streams y planes (tile + 2-site halo, periodic wrap applied while filling) into a
// 3-plane LDS ring, computes mu once per site on the tile + 1-site halo into a second
// 3-plane ring, and evaluates M*lap(mu) plus the stage update on the tile of the plane
// two steps behind.  Per output plane: one y plane read (+halo, mostly L2 hits) and the
// stage's point reads/writes -- the HBM minimum for the stage.  The next y plane and the
// point operands are prefetched into registers one iteration ahead.
//
// 2-D (Nz = 1): the same with a single plane; the degenerate z terms of both Laplacians
// are formed from the centre value, exactly as the reference's wrapped access does.

// The order-2 stencil around LDS offset o: in-plane neighbours from plane `c` (row width W),
// z neighbours from planes `lo` (k-1) and `hi` (k+1) at the same offset.
#define GC_NB_LDS(pc, plo, phi, o, W) \
  (pc)[o], (pc)[(o) - 1], (pc)[(o) + 1], (pc)[(o) - (W)], (pc)[(o) + (W)], (plo)[o], (phi)[o]

template <class T, int OX, int OY, int ZC>
struct Ch3Cfg {
  static constexpr int NT = 256;
  static constexpr int YW = OX + 4, YH = OY + 4, YN = YW * YH;  // y plane, halo 2
  static constexpr int MW = OX + 2, MH = OY + 2, MN = MW * MH;  // mu plane, halo 1
  static constexpr int YL = (YN + NT - 1) / NT;                 // y loads per thread
  static constexpr int ML = (MN + NT - 1) / NT;                 // mu sites per thread
  static constexpr int OR = OX * OY / NT;                       // output sites per thread
  static_assert(OX % 64 == 0 || NT % OX == 0, "tile width must divide the block");
  static_assert((OX * OY) % NT == 0, "tile must be a multiple of the block");
};

// Stage operands and results as the tiled kernels move them: operands read once and
// results written once per stage bypass cache residency (non-temporal), so the caches keep
// the y planes whose halos neighbouring tiles re-read.
template <int STAGE, class T>
__device__ __forceinline__ void stage_in(const T* psi, const T* a, const T* b, unsigned c, T& pv,
                                         T& av, T& bv) {
  if constexpr (STAGE >= kS2 && STAGE <= kS4) pv = __builtin_nontemporal_load(psi + c);
  if constexpr (STAGE == kS4) av = __builtin_nontemporal_load(a + c);
  if constexpr (STAGE == kS3 || STAGE == kS4) bv = __builtin_nontemporal_load(b + c);
}

template <int STAGE, class T>
__device__ __forceinline__ void stage_out(unsigned c, T kv, T pv, T av, T bv, T* psi, T* ynext,
                                          T* a, T* b, const StageCoef<T>& co) {
  if constexpr (STAGE == kRhsOnly) {
    __builtin_nontemporal_store(kv, ynext + c);
  } else if constexpr (STAGE == kS1) {
    __builtin_nontemporal_store(kv, a + c);
    __builtin_nontemporal_store(stage_axpy(pv, co.half_dt, kv), ynext + c);
  } else if constexpr (STAGE == kS2) {
    __builtin_nontemporal_store(kv, b + c);
    __builtin_nontemporal_store(stage_axpy(pv, co.half_dt, kv), ynext + c);
  } else if constexpr (STAGE == kS3) {
    __builtin_nontemporal_store(bv + kv, b + c);
    __builtin_nontemporal_store(stage_axpy(pv, co.dt, kv), ynext + c);
  } else if constexpr (STAGE == kS4) {
    __builtin_nontemporal_store(rk4_final(pv, av, bv, kv, co.sixth_dt, co.third_dt), psi + c);
  } else {  // explicit Euler: ynext = y + dt*f(y)
    __builtin_nontemporal_store(stage_axpy(pv, co.dt, kv), ynext + c);
  }
}

// S1 and Euler take the stage's base value from the stencil input itself (y = psi), which
// the kernels already hold in LDS.
template <int STAGE>
constexpr bool base_from_y() {
  return STAGE == kS1 || STAGE == kEuler;
}

template <class T, int OX, int OY, int ZC, int STAGE>
__global__ __launch_bounds__(256) void k_ch_stage3d(CHCoef<T> cf, const T* __restrict__ y, T* psi,
                                                    T* __restrict__ ynext, T* __restrict__ a,
                                                    T* __restrict__ b, Dims d, int ntx, int nty,
                                                    StageCoef<T> co) {
  using C = Ch3Cfg<T, OX, OY, ZC>;
  __shared__ T ys[3][C::YN];
  __shared__ T ms[3][C::MN];
  const int t = threadIdx.x;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int tx = (int)(bid % (unsigned)ntx);
  const unsigned r = bid / (unsigned)ntx;
  const int ty = (int)(r % (unsigned)nty);
  const int zc = (int)(r / (unsigned)nty);
  const int x0 = tx * OX, y0 = ty * OY, z0 = zc * ZC;
  const int nout = min(ZC, d.nz - z0);

  // xy part of the (32-bit) offset of every y-plane slot this thread fills; only the
  // last slot can fall outside the halo tile
  unsigned yoff[C::YL];
#pragma unroll
  for (int s = 0; s < C::YL; ++s) {
    const int idx = t + s * C::NT;
    const int px = idx % C::YW, py = idx / C::YW;
    yoff[s] = (s + 1 < C::YL || idx < C::YN)
                  ? (unsigned)(wrap_any(x0 + px - 2, d.nx) + d.nx * wrap_any(y0 + py - 2, d.ny))
                  : 0u;
  }
  constexpr bool kLastYPartial = (C::YN % C::NT) != 0;
  // output sites: column ox, rows oy0 + q*ROWS
  constexpr int ROWS = C::NT / OX;
  const int ox = t % OX, oy0 = t / OX;
  unsigned ooff[C::OR];
  bool oin[C::OR];
#pragma unroll
  for (int q = 0; q < C::OR; ++q) {
    const int gy = y0 + oy0 + q * ROWS;
    oin[q] = (x0 + ox < d.nx) && (gy < d.ny);
    ooff[q] = oin[q] ? (unsigned)(x0 + ox + d.nx * gy) : 0u;
  }

  T yreg[C::YL];
  auto load_y = [&](int plane) {
    const T* yp = y + (size_t)d.plane * (size_t)wrap_any(plane, d.nz);
#pragma unroll
    for (int s = 0; s < C::YL; ++s)
      if (s + 1 < C::YL || !kLastYPartial || t + s * C::NT < C::YN) yreg[s] = yp[yoff[s]];
  };
  T pv[C::OR] = {}, av[C::OR] = {}, bv[C::OR] = {};

  // Iteration p stores y plane z0-2+p into ring slot p%3, computes mu plane z0+p-3 from y
  // planes p-2..p, then the output plane z0+p-4 from mu planes p-2..p.  The two barriers
  // separate the y store from the mu pass and the mu pass from the output pass.  A y slot
  // written in iteration p was last read before the second barrier of iteration p-1, and a
  // mu slot before the first barrier of iteration p; the output pass reads no y slot.
  load_y(z0 - 2);
  const int last = nout + 3;
  for (int p = 0; p <= last; ++p) {
    const int s0 = p % 3, s1 = (p + 2) % 3, s2 = (p + 1) % 3;  // slots of p, p-1, p-2
    T* yw = ys[s0];
#pragma unroll
    for (int s = 0; s < C::YL; ++s)
      if (s + 1 < C::YL || !kLastYPartial || t + s * C::NT < C::YN) yw[t + s * C::NT] = yreg[s];
    if (p < last) load_y(z0 - 1 + p);
    if (p >= 4) {
      const size_t kb = (size_t)d.plane * (size_t)(z0 + p - 4);
#pragma unroll
      for (int q = 0; q < C::OR; ++q)
        if (oin[q]) stage_in<STAGE>(psi + kb, a + kb, b + kb, ooff[q], pv[q], av[q], bv[q]);
    }
    __syncthreads();
    // S1/Euler: the base value of output plane z0+p-4 is y plane p-2 (slot s2), which the
    // store of iteration p+1 overwrites; read it before the second barrier.
    if constexpr (base_from_y<STAGE>()) {
      if (p >= 4) {
#pragma unroll
        for (int q = 0; q < C::OR; ++q)
          if (oin[q]) pv[q] = ys[s2][(oy0 + q * ROWS + 2) * C::YW + (ox + 2)];
      }
    }
    if (p >= 2) {  // mu plane z0+p-3 from y planes p-2, p-1, p
      const T* yc = ys[s1];
      const T* yb = ys[s2];
      const T* ya = ys[s0];
      T* mw = ms[s0];
#pragma unroll
      for (int s = 0; s < C::ML; ++s) {
        const int idx = t + s * C::NT;
        if (s + 1 < C::ML || idx < C::MN) {
          const int mx = idx % C::MW, my = idx / C::MW;
          const int o = (my + 1) * C::YW + (mx + 1);
          mw[idx] = ch_mu(cf, GC_NB_LDS(yc, yb, ya, o, C::YW));
        }
      }
    }
    __syncthreads();
    if (p >= 4) {  // output plane z0+p-4 from mu planes p-2, p-1, p
      const size_t kb = (size_t)d.plane * (size_t)(z0 + p - 4);
      const T* mc = ms[s1];
      const T* mb = ms[s2];
      const T* ma = ms[s0];
#pragma unroll
      for (int q = 0; q < C::OR; ++q) {
        if (oin[q]) {
          const int oy = oy0 + q * ROWS;
          const int o = (oy + 1) * C::MW + (ox + 1);
          const T kv = ch_k(cf, GC_NB_LDS(mc, mb, ma, o, C::MW));
          stage_out<STAGE>(ooff[q], kv, pv[q], av[q], bv[q], psi + kb, ynext + kb, a + kb,
                           b + kb, co);
        }
      }
    }
  }
}

template <class T, int OX, int OY, int STAGE>
__global__ __launch_bounds__(256) void k_ch_stage2d(CHCoef<T> cf, const T* __restrict__ y, T* psi,
                                                    T* __restrict__ ynext, T* __restrict__ a,
                                                    T* __restrict__ b, Dims d, int ntx,
                                                    StageCoef<T> co) {
  using C = Ch3Cfg<T, OX, OY, 1>;
  __shared__ T ys[C::YN];
  __shared__ T ms[C::MN];
  const int t = threadIdx.x;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int tx = (int)(bid % (unsigned)ntx);
  const int ty = (int)(bid / (unsigned)ntx);
  const int x0 = tx * OX, y0 = ty * OY;
  constexpr int ROWS = C::NT / OX;
  const int ox = t % OX, oy0 = t / OX;

#pragma unroll
  for (int s = 0; s < C::YL; ++s) {
    const int idx = t + s * C::NT;
    if (s + 1 < C::YL || idx < C::YN) {
      const int px = idx % C::YW, py = idx / C::YW;
      ys[idx] = y[wrap_any(x0 + px - 2, d.nx) + d.nx * wrap_any(y0 + py - 2, d.ny)];
    }
  }
  T pv[C::OR] = {}, av[C::OR] = {}, bv[C::OR] = {};
  unsigned ooff[C::OR];
  bool oin[C::OR];
#pragma unroll
  for (int q = 0; q < C::OR; ++q) {
    const int gy = y0 + oy0 + q * ROWS;
    oin[q] = (x0 + ox < d.nx) && (gy < d.ny);
    ooff[q] = oin[q] ? (unsigned)(x0 + ox + d.nx * gy) : 0u;
    if (oin[q]) stage_in<STAGE>(psi, a, b, ooff[q], pv[q], av[q], bv[q]);
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < C::ML; ++s) {
    const int idx = t + s * C::NT;
    if (s + 1 < C::ML || idx < C::MN) {
      const int mx = idx % C::MW, my = idx / C::MW;
      const int o = (my + 1) * C::YW + (mx + 1);
      // Nz = 1 (ch_mu2: the z neighbours are the centre itself)
      ms[idx] = ch_mu2(cf, ys[o], ys[o - 1], ys[o + 1], ys[o - C::YW], ys[o + C::YW]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < C::OR; ++q) {
    if (oin[q]) {
      const int oy = oy0 + q * ROWS;
      const int o = (oy + 1) * C::MW + (ox + 1);
      const T kv = ch_k2(cf, ms[o], ms[o - 1], ms[o + 1], ms[o - C::MW], ms[o + C::MW]);
      if constexpr (base_from_y<STAGE>()) pv[q] = ys[(oy + 2) * C::YW + (ox + 2)];
      stage_out<STAGE>(ooff[q], kv, pv[q], av[q], bv[q], psi, ynext, a, b, co);
    }
  }
}

template <class T, class Rhs>
gc_status launch_stage_untiled(gc_context* ctx, const gc_grid* g, const Rhs& rhs, int stage,
                               const T* y, T* psi, T* ynext, T* a, T* b,
                               const StageCoef<T>& co) {
  const Dims d = dims_of(g);
  const int nbx = (d.nx + TX - 1) / TX, nby = (d.ny + TY - 1) / TY;
  const dim3 grid((unsigned)(nbx * nby * d.nz));
  hipLaunchKernelGGL((k_stage<T, Rhs>), grid, dim3(TX * TY), 0, ctx->stream, rhs, stage, y, psi,
                     ynext, a, b, d, nbx, nby, co);
  GC_LAUNCHED();
  return GC_OK;
}

// ---- LDS-tiled diffusion stage kernel (3-D) ----------------------------------------------
//
// The constant-D (laplacian<2>) and heterogeneous-D (fvm::cellLaplacian) right-hand sides
// have radius 1: a block keeps the planes z-1, z, z+1 of the stage input (and of D) for its
// 64 x 8 tile + 1-site halo in a 3-plane LDS ring and marches a chunk of planes; the next
// plane and the stage operands are fetched while the current plane is formed.  Arithmetic as
// LapRhs / FvmRhs.

template <class T, int OX, int OY, int ZC, int STAGE, bool FVM>
__global__ __launch_bounds__(256) void k_diff_stage3d(const T* __restrict__ Dfield,
                                                      LapCoef<T> lc, const T* __restrict__ y,
                                                      T* psi, T* __restrict__ ynext,
                                                      T* __restrict__ a, T* __restrict__ b,
                                                      Dims d, int ntx, int nty,
                                                      StageCoef<T> co) {
  constexpr int NT = 256, PW = OX + 2, PN = PW * (OY + 2), PL = (PN + NT - 1) / NT;
  constexpr int ROWS = NT / OX, OR = OX * OY / NT;
  static_assert(NT % OX == 0 && (OX * OY) % NT == 0, "tile must match the block");
  __shared__ T ys[3][PN];
  __shared__ T ds[FVM ? 3 : 1][FVM ? PN : 1];
  const int t = threadIdx.x;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int tx = (int)(bid % (unsigned)ntx);
  const unsigned r = bid / (unsigned)ntx;
  const int ty = (int)(r % (unsigned)nty);
  const int zc = (int)(r / (unsigned)nty);
  const int x0 = tx * OX, y0 = ty * OY, z0 = zc * ZC;
  const int nout = min(ZC, d.nz - z0);

  unsigned poff[PL];
#pragma unroll
  for (int s = 0; s < PL; ++s) {
    const int idx = t + s * NT;
    const int px = idx % PW, py = idx / PW;
    poff[s] = idx < PN ? (unsigned)(wrap_any(x0 + px - 1, d.nx) + d.nx * wrap_any(y0 + py - 1, d.ny)) : 0u;
  }
  const int ox = t % OX, oy0 = t / OX;
  unsigned ooff[OR];
  bool oin[OR];
#pragma unroll
  for (int q = 0; q < OR; ++q) {
    const int gy = y0 + oy0 + q * ROWS;
    oin[q] = (x0 + ox < d.nx) && (gy < d.ny);
    ooff[q] = oin[q] ? (unsigned)(x0 + ox + d.nx * gy) : 0u;
  }
  T yreg[PL], dreg[FVM ? PL : 1];
  auto load = [&](int plane) {
    const size_t zo = (size_t)d.plane * (size_t)wrap_any(plane, d.nz);
#pragma unroll
    for (int s = 0; s < PL; ++s)
      if (t + s * NT < PN) {
        yreg[s] = y[zo + poff[s]];
        if constexpr (FVM) dreg[s] = Dfield[zo + poff[s]];
      }
  };
  auto store = [&](int slot) {
#pragma unroll
    for (int s = 0; s < PL; ++s)
      if (t + s * NT < PN) {
        ys[slot][t + s * NT] = yreg[s];
        if constexpr (FVM) ds[slot][t + s * NT] = dreg[s];
      }
  };
  T pv[OR] = {}, av[OR] = {}, bv[OR] = {};

  // iteration p stores plane z0-1+p in slot p%3 and forms output plane z0+p-2
  load(z0 - 1);
  const int last = nout + 1;
  for (int p = 0; p <= last; ++p) {
    const int s0 = p % 3, s1 = (p + 2) % 3, s2 = (p + 1) % 3;  // slots of p, p-1, p-2
    store(s0);
    if (p < last) load(z0 + p);
    if (p >= 2) {
      const size_t kb = (size_t)d.plane * (size_t)(z0 + p - 2);
#pragma unroll
      for (int q = 0; q < OR; ++q)
        if (oin[q]) stage_in<STAGE>(psi + kb, a + kb, b + kb, ooff[q], pv[q], av[q], bv[q]);
    }
    __syncthreads();
    if (p >= 2) {
      const size_t kb = (size_t)d.plane * (size_t)(z0 + p - 2);
#pragma unroll
      for (int q = 0; q < OR; ++q) {
        if (oin[q]) {
          const int o = (oy0 + q * ROWS + 1) * PW + (ox + 1);
          const T* yc = ys[s1];
          const T* ylo = ys[s2];
          const T* yhi = ys[s0];
          T kv;
          if constexpr (FVM) {
            const T* dc = ds[s1];
            const T Dc = dc[o];
            kv = fvm_k(GC_NB_LDS(yc, ylo, yhi, o, PW), face_coef(dc[o - 1], Dc, lc.ihx2),
                       face_coef(Dc, dc[o + 1], lc.ihx2), face_coef(dc[o - PW], Dc, lc.ihy2),
                       face_coef(Dc, dc[o + PW], lc.ihy2), face_coef(ds[s2][o], Dc, lc.ihz2),
                       face_coef(Dc, ds[s0][o], lc.ihz2), lc.ihx2, lc.ihy2, lc.ihz2);
          } else {
            kv = lap7(lc, GC_NB_LDS(yc, ylo, yhi, o, PW));
          }
          if constexpr (base_from_y<STAGE>()) pv[q] = yc[o];
          stage_out<STAGE>(ooff[q], kv, pv[q], av[q], bv[q], psi + kb, ynext + kb, a + kb,
                           b + kb, co);
        }
      }
    }
    __syncthreads();
  }
}

// Two diffusion stages per pass (PAIR 0 = S1+S2, PAIR 1 = S3+S4), the scheme of the
// Cahn-Hilliard two-stage kernels below with radius-1 right-hand sides: the first stage on
// the tile + 1-site halo (its stage value stays in LDS), the second on the tile; stage input
// and D read with a 2-site halo.  Per site and step 9 W (+ 2 W of D) instead of 17 W (+ 4 W).
// Iteration p stores input plane z0-2+p, forms k_A on plane z0-3+p and the outputs on plane
// z0-4+p; rings: input and y' 3 planes each.  Heterogeneous D: the harmonic mean of each face
// depends on D only, so it is formed once per plane and used by both cells and both stages
// (x and y faces in 3-plane LDS rings, z faces in per-thread queues along the columns; D
// itself in a 2-plane ring): about 4 FP64 divisions per site and step pair instead of 13.

template <int OX, int OY, int NT>
struct DPairCfg {
  static constexpr int W2 = OX + 4, H2 = OY + 4, N2 = W2 * H2;  // stage input, D
  static constexpr int W1 = OX + 2, N1 = W1 * (OY + 2);  // k_A, y'
  static constexpr int L2 = (N2 + NT - 1) / NT;
  static constexpr int NR = 2 * W1 + 2 * OY;              // tile+1 sites outside the tile
  static constexpr int LR = (NR + NT - 1) / NT;
  static constexpr int OR = OX * OY / NT, ROWS = NT / OX;
  static_assert(NT % OX == 0 && (OX * OY) % NT == 0, "tile must match the block");
};

// Ring index r -> (mx, my) in tile+1 coordinates: the row above, the row below, the columns
// either side.
template <class C, int OY>
__device__ __forceinline__ void ring1_site(int r, int& mx, int& my) {
  if (r < C::W1) {
    mx = r;
    my = 0;
  } else if (r < 2 * C::W1) {
    mx = r - C::W1;
    my = OY + 1;
  } else {
    r -= 2 * C::W1;
    my = 1 + (r >> 1);
    mx = (r & 1) ? C::W1 - 1 : 0;
  }
}

template <class T, int OX, int OY, int ZC, int NT, int PAIR, bool FVM>
__global__ __launch_bounds__(NT) void k_diff_pair3d(const T* __restrict__ Dfield, LapCoef<T> lc,
                                                    const T* __restrict__ y,
                                                    const T* __restrict__ psi,
                                                    T* __restrict__ out, T* __restrict__ a,
                                                    T* __restrict__ b, Dims d, int ntx, int nty,
                                                    StageCoef<T> co) {
  using C = DPairCfg<OX, OY, NT>;
  constexpr int ND = FVM ? 2 : 1, NF = FVM ? 3 : 1, NFN = FVM ? C::N2 : 1;
  __shared__ T ys[3][C::N2];
  __shared__ T ds[ND][FVM ? C::N2 : 1];
  __shared__ T fx[NF][NFN], fy[NF][NFN];  // FVM: face harmonic means of the last 3 planes
  __shared__ T y2s[3][C::N1];
  const int t = threadIdx.x;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int tx = (int)(bid % (unsigned)ntx);
  const unsigned r = bid / (unsigned)ntx;
  const int ty = (int)(r % (unsigned)nty);
  const int zc = (int)(r / (unsigned)nty);
  const int x0 = tx * OX, y0 = ty * OY, z0 = zc * ZC;
  const int nout = min(ZC, d.nz - z0);

  const int ox = t % OX, oy0 = t / OX;
  // Constant D (kQ): a thread's region sites are slots 0 .. OR-1 its tile sites (the output
  // sites below), then halo sites -- the tile+1 ring in ring1_site order (so ring site t + s NT
  // is this thread's slot OR + s, as below), then the outer ring; PAIR 1 then finds the psi
  // and k2 of its own sites from earlier planes in registers.  Heterogeneous D: region site
  // t + s NT in row-major order, and PAIR 1 reloads those values (from L2): the queues would
  // take the FP32 kernels past 64 VGPRs, and with them from two workgroups per CU to one
  // (256^3, one box: FP32 0.190 -> 0.241 ms per step, FP64 0.381 -> 0.396 ms).
  constexpr bool kQ = !FVM;
  constexpr int NH2 = C::N2 - OX * OY, NRO = NH2 - C::NR;  // halo sites, outer-ring sites
  constexpr int LS = kQ ? C::OR + (NH2 + NT - 1) / NT : C::L2;  // slots
  unsigned yoff[LS];
  int ro[kQ ? LS : 1];     // kQ: region index (LDS planes)
  bool have[kQ ? LS : 1];  // kQ: the slot holds a site
  // region index of slot s and whether the slot holds a site
  auto rix = [&](int s) -> int {
    if constexpr (kQ) return ro[s];
    else return t + s * NT;
  };
  auto has = [&](int s) -> bool {
    if constexpr (kQ) return have[s];
    else return t + s * NT < C::N2;
  };
#pragma unroll
  for (int s = 0; s < LS; ++s) {
    int px = 0, py = 0;
    bool hs = true;
    if (!kQ) {
      const int idx = t + s * NT;
      hs = idx < C::N2;
      px = idx % C::W2;
      py = idx / C::W2;
    } else if (s < C::OR) {
      px = ox + 2;
      py = oy0 + s * C::ROWS + 2;
    } else {
      const int h = t + (s - C::OR) * NT;
      hs = h < NH2;
      if (h < C::NR) {
        int mx = 0, my = 0;
        ring1_site<C, OY>(h, mx, my);
        px = mx + 1;
        py = my + 1;
      } else {
        const int g = h - C::NR;  // rows 0 and H2-1, then columns 0 and W2-1 of rows 1..H2-2
        if (g < 2 * C::W2) {
          px = g % C::W2;
          py = g < C::W2 ? 0 : C::H2 - 1;
        } else {
          const int r = g - 2 * C::W2;
          py = 1 + (r >> 1);
          px = (r & 1) ? C::W2 - 1 : 0;
        }
      }
    }
    if constexpr (kQ) {
      ro[s] = py * C::W2 + px;
      have[s] = hs;
    }
    yoff[s] = hs ? (unsigned)(wrap_any(x0 + px - 2, d.nx) + d.nx * wrap_any(y0 + py - 2, d.ny))
                      : 0u;
  }
  static_assert(NRO == 2 * C::W2 + 2 * (C::H2 - 2), "outer ring");
  static_assert(!kQ || C::LR == 1 || C::OR + C::LR <= LS, "ring slots");
  unsigned ooff[C::OR];  // tile sites, wrapped; oin: the site is an output
  bool oin[C::OR];
#pragma unroll
  for (int q = 0; q < C::OR; ++q) {
    const int gy = y0 + oy0 + q * C::ROWS;
    oin[q] = (x0 + ox < d.nx) && (gy < d.ny);
    ooff[q] = (unsigned)(wrap_any(x0 + ox, d.nx) + d.nx * wrap_any(gy, d.ny));
  }
  // the 1-site ring around the tile: E1 coordinates
  int rmx[C::LR], rmy[C::LR];
  constexpr bool kReload = PAIR == 1 && !kQ;
  unsigned roff[kReload ? C::LR : 1];  // reload path: psi at the ring sites of stage A
#pragma unroll
  for (int s = 0; s < C::LR; ++s) {
    const int rr = t + s * NT;
    int mx = 0, my = 0;
    if (rr < C::NR) ring1_site<C, OY>(rr, mx, my);
    rmx[s] = rr < C::NR ? mx : -1;
    rmy[s] = my;
    if constexpr (kReload)
      roff[s] = rr < C::NR ? (unsigned)(wrap_any(x0 + mx - 1, d.nx) + d.nx * wrap_any(y0 + my - 1, d.ny))
                           : 0u;
  }

  // S1+S2 does not store y3 = psi + dt/2 k2; S3+S4 forms it from psi and k2 as it loads them
  // (stage_axpy, the same bits) -- 7 words per site and step instead of 9 (+ D)
  T yreg[LS], dreg[FVM ? LS : 1], kreg[PAIR == 1 ? LS : 1];
  auto load = [&](int plane) {
    const size_t zo = (size_t)d.plane * (size_t)wrap_any(plane, d.nz);
#pragma unroll
    for (int s = 0; s < LS; ++s)
      if (has(s)) {
        if constexpr (PAIR == 1) {
          yreg[s] = psi[zo + yoff[s]];
          kreg[s] = b[zo + yoff[s]];
        } else {
          yreg[s] = y[zo + yoff[s]];
        }
        if constexpr (FVM) dreg[s] = Dfield[zo + yoff[s]];
      }
  };
  // PAIR 1: psi and k2 of the stored planes as loaded (newest first): tile slots keep three
  // planes (k_A's plane is the last but one, the output plane the one before), ring slots one
  // (the reload path: pB, rB hold psi of k_A's plane, pC and kC2 psi and k2 of the output plane)
  constexpr int NQ = PAIR == 1 ? C::OR : 1, NRQ = PAIR == 1 ? C::LR : 1;
  T pA[NQ] = {}, pB[NQ] = {}, pC[NQ] = {}, kA2[NQ] = {}, kB2[NQ] = {}, kC2[NQ] = {};
  T rA[NRQ] = {}, rB[NRQ] = {};
  T k0r[PAIR == 1 ? C::OR : 1] = {}, k1r[PAIR == 1 ? C::OR : 1] = {};  // k_A, newest first
  T av[PAIR == 1 ? C::OR : 1] = {};  // PAIR 1: k1 of the output site
  // FVM: z-face means along this thread's columns: tile sites keep the last three planes'
  // (newest first), ring sites the last one
  T fzt0[FVM ? C::OR : 1] = {}, fzt1[FVM ? C::OR : 1] = {}, fzt2[FVM ? C::OR : 1] = {};
  T fzr[FVM ? C::LR : 1] = {};

  load(z0 - 2);
  const int last = nout + 3;
  for (int p = 0; p <= last; ++p) {
    const int s0 = p % 3, s1 = (p + 2) % 3, s2 = (p + 1) % 3;  // slots of p, p-1, p-2
    const int d0 = p % ND, d1 = (p + ND - 1) % ND;               // D slots of p, p-1
    if constexpr (PAIR == 1 && kQ) {  // the plane stored now becomes the newest
#pragma unroll
      for (int q = 0; q < C::OR; ++q) {
        pC[q] = pB[q];
        pB[q] = pA[q];
        pA[q] = yreg[q];
        kC2[q] = kB2[q];
        kB2[q] = kA2[q];
        kA2[q] = kreg[q];
      }
#pragma unroll
      for (int s = 0; s < C::LR; ++s) {
        rB[s] = rA[s];
        rA[s] = yreg[C::OR + s];
      }
    }
#pragma unroll
    for (int s = 0; s < LS; ++s)
      if (has(s)) {
        if constexpr (PAIR == 1) ys[s0][rix(s)] = stage_axpy(yreg[s], co.half_dt, kreg[s]);
        else ys[s0][rix(s)] = yreg[s];
        if constexpr (FVM) ds[d0][rix(s)] = dreg[s];
      }
    if (p < last) load(z0 - 1 + p);
    if constexpr (PAIR == 1) {
      // psi of k_A's plane z0-3+p (tile: pB, ring: rB) and psi, k2 of the output plane
      // z0-4+p (pC, kC2) are in registers (kQ) or reloaded; k1 of the output plane comes
      // from HBM
      if constexpr (kReload) {
        if (p >= 2) {
          const T* pp = psi + (size_t)d.plane * (size_t)wrap_any(z0 - 3 + p, d.nz);
#pragma unroll
          for (int s = 0; s < C::LR; ++s)
            if (t + s * NT < C::NR) rB[s] = pp[roff[s]];
#pragma unroll
          for (int q = 0; q < C::OR; ++q) pB[q] = pp[ooff[q]];
        }
      }
      if (p >= 4) {
        const size_t kb = (size_t)d.plane * (size_t)(z0 - 4 + p);
#pragma unroll
        for (int q = 0; q < C::OR; ++q)
          if (oin[q]) {
            if constexpr (kReload) pC[q] = __builtin_nontemporal_load(psi + kb + ooff[q]);
            av[q] = __builtin_nontemporal_load(a + kb + ooff[q]);
            if constexpr (kReload) kC2[q] = __builtin_nontemporal_load(b + kb + ooff[q]);
          }
      }
    }
    __syncthreads();
    if constexpr (FVM) {
      // x- and y-face means of the plane just stored (fvm/CellLaplacian.hpp:48-50): the face
      // between a site and its +x (+y) neighbour, shared by both cells and both stages
      const T* dp = ds[d0];
#pragma unroll
      for (int s = 0; s < LS; ++s) {
        const int idx = rix(s);
        if (has(s)) {
          const int mx = idx % C::W2, my = idx / C::W2;
          if (mx < C::W2 - 1) fx[s0][idx] = face_coef(dp[idx], dp[idx + 1], lc.ihx2);
          if (my < C::H2 - 1) fy[s0][idx] = face_coef(dp[idx], dp[idx + C::W2], lc.ihy2);
        }
      }
      if (p == 1) {  // z-faces of plane z0-2 (under the first k_A plane)
#pragma unroll
        for (int q = 0; q < C::OR; ++q) {
          const int e2 = (oy0 + q * C::ROWS + 2) * C::W2 + (ox + 2);
          fzt0[q] = face_coef(ds[d1][e2], ds[d0][e2], lc.ihz2);
        }
#pragma unroll
        for (int s = 0; s < C::LR; ++s)
          if (rmx[s] >= 0) {
            const int e2 = (rmy[s] + 1) * C::W2 + (rmx[s] + 1);
            fzr[s] = face_coef(ds[d1][e2], ds[d0][e2], lc.ihz2);
          }
      }
      __syncthreads();
    }
    if (p >= 2) {  // k_A and y' on plane z0-3+p, tile+1
      const int zA = z0 - 3 + p;
      const bool own = zA >= z0 && zA < z0 + nout;
      // right-hand side at E1 site (mx, my); FVM: fzp / fzm are the z-face means above / below
      auto kA = [&](int mx, int my, T fzp, T fzm) {
        const int o = (my + 1) * C::W2 + (mx + 1);
        if constexpr (FVM) {
          return fvm_k(GC_NB_LDS(ys[s1], ys[s2], ys[s0], o, C::W2), fx[s1][o - 1], fx[s1][o],
                       fy[s1][o - C::W2], fy[s1][o], fzm, fzp, lc.ihx2, lc.ihy2, lc.ihz2);
        } else {
          return lap7(lc, GC_NB_LDS(ys[s1], ys[s2], ys[s0], o, C::W2));
        }
      };
#pragma unroll
      for (int q = 0; q < C::OR; ++q) {  // tile sites: the output owner
        const int mx = ox + 1, my = oy0 + q * C::ROWS + 1;
        T fzp = T(0), fzm = T(0);
        if constexpr (FVM) {
          const int o = (my + 1) * C::W2 + (mx + 1);
          fzp = face_coef(ds[d1][o], ds[d0][o], lc.ihz2);  // planes zA, zA+1
          fzm = fzt0[q];
          fzt2[q] = fzt1[q];
          fzt1[q] = fzt0[q];
          fzt0[q] = fzp;
        }
        const T kv = kA(mx, my, fzp, fzm);
        const int e = my * C::W1 + mx;
        if constexpr (PAIR == 0) {
          y2s[s0][e] = stage_axpy(ys[s1][(my + 1) * C::W2 + (mx + 1)], co.half_dt, kv);
          if (own && oin[q])
            __builtin_nontemporal_store(kv, a + (size_t)d.plane * (size_t)zA + ooff[q]);
        } else {
          y2s[s0][e] = stage_axpy(pB[q], co.dt, kv);
          k1r[q] = k0r[q];
          k0r[q] = kv;
        }
      }
#pragma unroll
      for (int s = 0; s < C::LR; ++s) {  // the 1-site ring around the tile
        if (rmx[s] >= 0) {
          const int mx = rmx[s], my = rmy[s];
          T fzp = T(0), fzm = T(0);
          if constexpr (FVM) {
            const int o = (my + 1) * C::W2 + (mx + 1);
            fzp = face_coef(ds[d1][o], ds[d0][o], lc.ihz2);
            fzm = fzr[s];
            fzr[s] = fzp;
          }
          const T kv = kA(mx, my, fzp, fzm);
          const int e = my * C::W1 + mx;
          if constexpr (PAIR == 0) y2s[s0][e] = stage_axpy(ys[s1][(my + 1) * C::W2 + (mx + 1)], co.half_dt, kv);
          else y2s[s0][e] = stage_axpy(rB[s], co.dt, kv);
        }
      }
    }
    __syncthreads();
    if (p >= 4) {  // outputs on plane z0-4+p from y' planes p-2..p
      const size_t kb = (size_t)d.plane * (size_t)(z0 - 4 + p);
#pragma unroll
      for (int q = 0; q < C::OR; ++q) {
        if (oin[q]) {
          const int oy = oy0 + q * C::ROWS;
          const int o1 = (oy + 1) * C::W1 + (ox + 1);
          T kv;
          if constexpr (FVM) {
            const int o2 = (oy + 2) * C::W2 + (ox + 2);  // the same site in the face planes
            kv = fvm_k(GC_NB_LDS(y2s[s1], y2s[s2], y2s[s0], o1, C::W1), fx[s2][o2 - 1], fx[s2][o2],
                       fy[s2][o2 - C::W2], fy[s2][o2], fzt2[q], fzt1[q], lc.ihx2, lc.ihy2,
                       lc.ihz2);
          } else {
            kv = lap7(lc, GC_NB_LDS(y2s[s1], y2s[s2], y2s[s0], o1, C::W1));
          }
          if constexpr (PAIR == 0) {
            __builtin_nontemporal_store(kv, b + kb + ooff[q]);
          } else {
            const T bsum = kC2[q] + k1r[q];  // k3 of this plane (k_A pass of iteration p-1)
            __builtin_nontemporal_store(rk4_final(pC[q], av[q], bsum, kv, co.sixth_dt, co.third_dt),
                                        out + kb + ooff[q]);
          }
        }
      }
    }
    __syncthreads();
  }
}

#define GC_STAGE_SWITCH(LAUNCH)      \
  switch (stage) {                   \
    case kRhsOnly: LAUNCH(kRhsOnly); break; \
    case kS1: LAUNCH(kS1); break;    \
    case kS2: LAUNCH(kS2); break;    \
    case kS3: LAUNCH(kS3); break;    \
    case kS4: LAUNCH(kS4); break;    \
    default: LAUNCH(kEuler); break;  \
  }

constexpr int kDiffOX = 64, kDiffOY = 8, kDiffZC = 32;

// Grids with Nz >= 4 run the tiled kernel; flatter grids the untiled one.
template <class T, bool FVM>
gc_status launch_diff_stage(gc_context* ctx, const gc_grid* g, const T* D, const LapCoef<T>& lc,
                            int stage, const T* y, T* psi, T* ynext, T* a, T* b,
                            const StageCoef<T>& co) {
  const Dims d = dims_of(g);
  const int ntx = (d.nx + kDiffOX - 1) / kDiffOX, nty = (d.ny + kDiffOY - 1) / kDiffOY;
  const int nzc = (d.nz + kDiffZC - 1) / kDiffZC;
  const dim3 grid((unsigned)(ntx * nty * nzc));
#define GC_LD(S)                                                                               \
  hipLaunchKernelGGL((k_diff_stage3d<T, kDiffOX, kDiffOY, kDiffZC, S, FVM>), grid, dim3(256), 0, \
                     ctx->stream, D, lc, y, psi, ynext, a, b, d, ntx, nty, co)
  GC_STAGE_SWITCH(GC_LD)
#undef GC_LD
  GC_LAUNCHED();
  return GC_OK;
}

// Two-stage diffusion kernels: 64 x 16 tiles, 1024 threads, 32-plane chunks; grids with
// Nz >= 8.
constexpr int kDpOX = 64, kDpOY = 16, kDpZC = 32, kDpNT = 1024;

template <class T, bool FVM>
gc_status diff_pair_step(gc_context* ctx, const gc_grid* g, const T* D, const LapCoef<T>& lc,
                         const T* psi_in, T* psi_out, T* y3, T* a, T* b, const StageCoef<T>& co) {
  const Dims d = dims_of(g);
  const int ntx = (d.nx + kDpOX - 1) / kDpOX, nty = (d.ny + kDpOY - 1) / kDpOY;
  const int nzc = (d.nz + kDpZC - 1) / kDpZC;
  const dim3 grid((unsigned)(ntx * nty * nzc));
  // y3 is not stored: S3+S4 forms it from psi and k2 (7 words per step instead of 9, plus D);
  // 256^3 on one box: constant D FP64 0.2305 -> 0.2220 ms, FP32 0.1429 -> 0.1369 ms, field D
  // FP32 0.194 -> 0.190 ms, FP64 unchanged (0.382 ms)
  (void)y3;
  hipLaunchKernelGGL((k_diff_pair3d<T, kDpOX, kDpOY, kDpZC, kDpNT, 0, FVM>), grid,
                     dim3(kDpNT), 0, ctx->stream, D, lc, psi_in, psi_in, nullptr, a, b, d, ntx, nty,
                     co);
  GC_LAUNCHED();
  hipLaunchKernelGGL((k_diff_pair3d<T, kDpOX, kDpOY, kDpZC, kDpNT, 1, FVM>), grid,
                     dim3(kDpNT), 0, ctx->stream, D, lc, psi_in, psi_in, psi_out, a, b, d, ntx, nty,
                     co);
  GC_LAUNCHED();
  return GC_OK;
}

template <class T>
gc_status pair_step(gc_context* ctx, const gc_grid* g, const LapRhs<T>& rhs, const T* psi_in,
                    T* psi_out, T* y3, T* a, T* b, const StageCoef<T>& co) {
  return diff_pair_step<T, false>(ctx, g, nullptr, rhs.lc, psi_in, psi_out, y3, a, b, co);
}

template <class T>
gc_status pair_step(gc_context* ctx, const gc_grid* g, const FvmRhs<T>& rhs, const T* psi_in,
                    T* psi_out, T* y3, T* a, T* b, const StageCoef<T>& co) {
  return diff_pair_step<T, true>(ctx, g, rhs.D, rhs.lc, psi_in, psi_out, y3, a, b, co);
}

template <class T>
gc_status launch_stage(gc_context* ctx, const gc_grid* g, const LapRhs<T>& rhs, int stage,
                       const T* y, T* psi, T* ynext, T* a, T* b, const StageCoef<T>& co) {
  if (g->nz < 4) return launch_stage_untiled<T>(ctx, g, rhs, stage, y, psi, ynext, a, b, co);
  return launch_diff_stage<T, false>(ctx, g, nullptr, rhs.lc, stage, y, psi, ynext, a, b, co);
}

template <class T>
gc_status launch_stage(gc_context* ctx, const gc_grid* g, const FvmRhs<T>& rhs, int stage,
                       const T* y, T* psi, T* ynext, T* a, T* b, const StageCoef<T>& co) {
  if (g->nz < 4) return launch_stage_untiled<T>(ctx, g, rhs, stage, y, psi, ynext, a, b, co);
  return launch_diff_stage<T, true>(ctx, g, rhs.D, rhs.lc, stage, y, psi, ynext, a, b, co);
}

// ---- two RK4 stages per pass (3-D) ---------------------------------------------------------
//
// PAIR 0 runs S1+S2, PAIR 1 runs S3+S4.  The first stage of a pair is evaluated on the tile
// plus a 2-site halo, so its stage value y' = base + c*k stays in LDS and the second stage
// runs on the tile from it; the stencil input is read with a 4-site halo.  Per site and
// step: PAIR 0 reads psi and writes k1, k2 and y3 (4 W); PAIR 1 reads y3, psi, k1 and k2 and
// writes the new state into the other state buffer (5 W) -- 9 W per step instead of 17:
//   PAIR 0: y = psi; k1 on tile+2; y2 = psi + dt/2 k1 (LDS); k2 on tile;
//           a = k1, b = k2, out = psi + dt/2 k2
//   PAIR 1: y = y3; k3 on tile+2; y4 = psi + dt k3 (LDS); k4 on tile;
//           out = psi + (dt/6 (k1 + k4) + dt/3 (k2 + k3))
// Every value is formed with the single-stage kernels' arithmetic, so the results are the
// same bits.  PAIR 1 must not write psi in place: neighbouring tiles and z-chunks still read
// it as their halo.  The diffusion kernels (k_diff_pair3d) and the FP32 register-queue
// Cahn-Hilliard kernels (k_ch_pairq3d) drop y3: PAIR 0 writes k1 and k2 (3 W), PAIR 1 reads
// psi and k2 with the halo and forms y3 = psi + dt/2 k2 as it loads them (4 W) -- 7 W per
// step.
//
// Iteration p of the z-march stores stencil-input plane z0-4+p, forms mu_A on plane z0-5+p
// (tile+3), k_A and y' on plane z0-6+p (tile+2), mu_B on plane z0-7+p (tile+1) and the
// outputs on plane z0-8+p.  Each of y, mu_A, y', mu_B lives in a 3-plane LDS ring; four
// barriers per iteration separate the passes, and every ring slot written in a pass was last
// read before the preceding barrier.  k_A on the tile is computed by the thread that owns
// the output site, which keeps it in registers for the two iterations until PAIR 1 needs
// it; the 2-site ring around the tile is spread over the block.

template <int OX, int OY, int NT>
struct PairCfg {
  static constexpr int W4 = OX + 8, H4 = OY + 8, N4 = W4 * H4;  // stencil input
  static constexpr int W3 = OX + 6, H3 = OY + 6, N3 = W3 * H3;  // mu of stage A
  static constexpr int W2 = OX + 4, H2 = OY + 4, N2 = W2 * H2;  // k_A and y'
  static constexpr int W1 = OX + 2, H1 = OY + 2, N1 = W1 * H1;  // mu of stage B
  static constexpr int L4 = (N4 + NT - 1) / NT, L3 = (N3 + NT - 1) / NT;
  static constexpr int L2 = (N2 + NT - 1) / NT, L1 = (N1 + NT - 1) / NT;
  static constexpr int OR = OX * OY / NT;  // output sites per thread
  static constexpr int ROWS = NT / OX;
  static constexpr int NR = 4 * W2 + 4 * OY;  // tile+2 sites outside the tile
  static constexpr int LR = (NR + NT - 1) / NT;
  static_assert(NT % OX == 0 && (OX * OY) % NT == 0, "tile must match the block");
};

// Ring index r -> (mx, my) in tile+2 coordinates: two rows above, two below, two columns
// either side.
template <class C, int OY>
__device__ __forceinline__ void ring_site(int r, int& mx, int& my) {
  if (r < 2 * C::W2) {
    mx = r % C::W2;
    my = r / C::W2;
  } else if (r < 4 * C::W2) {
    r -= 2 * C::W2;
    mx = r % C::W2;
    my = OY + 2 + r / C::W2;
  } else {
    r -= 4 * C::W2;
    const int col = r & 3;
    my = 2 + (r >> 2);
    mx = col < 2 ? col : C::W2 - 4 + col;
  }
}

template <class T, int OX, int OY, int ZC, int NT, int PAIR>
__global__ __launch_bounds__(NT) void k_ch_pair3d(CHCoef<T> cf, const T* __restrict__ y,
                                                  const T* __restrict__ psi,
                                                  T* __restrict__ out, T* __restrict__ a,
                                                  T* __restrict__ b, Dims d, int ntx, int nty,
                                                  StageCoef<T> co) {
  using C = PairCfg<OX, OY, NT>;
  __shared__ T ys[3][C::N4];
  __shared__ T mas[3][C::N3];
  __shared__ T y2s[3][C::N2];
  __shared__ T mbs[3][C::N1];
  const int t = threadIdx.x;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int tx = (int)(bid % (unsigned)ntx);
  const unsigned r = bid / (unsigned)ntx;
  const int ty = (int)(r % (unsigned)nty);
  const int zc = (int)(r / (unsigned)nty);
  const int x0 = tx * OX, y0 = ty * OY, z0 = zc * ZC;
  const int nout = min(ZC, d.nz - z0);

  unsigned yoff[C::L4];
#pragma unroll
  for (int s = 0; s < C::L4; ++s) {
    const int idx = t + s * NT;
    const int px = idx % C::W4, py = idx / C::W4;
    yoff[s] = idx < C::N4 ? (unsigned)(wrap_any(x0 + px - 4, d.nx) + d.nx * wrap_any(y0 + py - 4, d.ny))
                          : 0u;
  }
  const int ox = t % OX, oy0 = t / OX;
  // PAIR 1: psi at the ring sites of stage A (the tile sites use ooff)
  unsigned roff[PAIR == 1 ? C::LR : 1];
  if constexpr (PAIR == 1) {
#pragma unroll
    for (int s = 0; s < C::LR; ++s) {
      const int r = t + s * NT;
      int mx = 0, my = 0;
      if (r < C::NR) ring_site<C, OY>(r, mx, my);
      roff[s] = r < C::NR ? (unsigned)(wrap_any(x0 + mx - 2, d.nx) + d.nx * wrap_any(y0 + my - 2, d.ny))
                          : 0u;
    }
  }
  unsigned ooff[C::OR];  // tile sites, wrapped (a tile site past the grid edge is the halo
  bool oin[C::OR];       // of a site inside it); oin: the site is an output
#pragma unroll
  for (int q = 0; q < C::OR; ++q) {
    const int gy = y0 + oy0 + q * C::ROWS;
    oin[q] = (x0 + ox < d.nx) && (gy < d.ny);
    ooff[q] = (unsigned)(wrap_any(x0 + ox, d.nx) + d.nx * wrap_any(gy, d.ny));
  }

  // LDS offsets of this thread's sites in every pass (the same on every plane)
  int oA[C::L3];  // mu_A: centre in the y plane (tile+4 layout); -1: no site
#pragma unroll
  for (int s = 0; s < C::L3; ++s) {
    const int idx = t + s * NT;
    oA[s] = idx < C::N3 ? (idx / C::W3 + 1) * C::W4 + idx % C::W3 + 1 : -1;
  }
  // k_A ring: centre in the mu_A plane, index in the y' plane, base in the y plane
  int rk[C::LR], re[C::LR], rb[C::LR];
#pragma unroll
  for (int s = 0; s < C::LR; ++s) {
    const int rr = t + s * NT;
    int mx = 0, my = 0;
    if (rr < C::NR) ring_site<C, OY>(rr, mx, my);
    rk[s] = rr < C::NR ? (my + 1) * C::W3 + (mx + 1) : -1;
    re[s] = my * C::W2 + mx;
    rb[s] = (my + 2) * C::W4 + (mx + 2);
  }
  int oB[C::L1];  // mu_B: centre in the y' plane; -1: no site
#pragma unroll
  for (int s = 0; s < C::L1; ++s) {
    const int idx = t + s * NT;
    oB[s] = idx < C::N1 ? (idx / C::W1 + 1) * C::W2 + idx % C::W1 + 1 : -1;
  }

  T yreg[C::L4];
  auto load_y = [&](int plane) {
    const T* yp = y + (size_t)d.plane * (size_t)wrap_any(plane, d.nz);
#pragma unroll
    for (int s = 0; s < C::L4; ++s)
      if (t + s * NT < C::N4) yreg[s] = yp[yoff[s]];
  };
  T pe[PAIR == 1 ? C::LR : 1];    // PAIR 1: psi at this thread's ring sites
  T pt[PAIR == 1 ? C::OR : 1];    // PAIR 1: psi at this thread's tile sites, plane of k_A
  // PAIR 1: k_A (= k3) at the tile sites of the last three k_A planes, newest first
  T k0r[PAIR == 1 ? C::OR : 1] = {}, k1r[PAIR == 1 ? C::OR : 1] = {}, k2r[PAIR == 1 ? C::OR : 1] = {};
  T pv[C::OR] = {}, av[C::OR] = {}, bv[C::OR] = {};

  load_y(z0 - 4);
  const int last = nout + 7;
  for (int p = 0; p <= last; ++p) {
    const int s0 = p % 3, s1 = (p + 2) % 3, s2 = (p + 1) % 3;  // slots of p, p-1, p-2
#pragma unroll
    for (int s = 0; s < C::L4; ++s)
      if (t + s * NT < C::N4) ys[s0][t + s * NT] = yreg[s];
    if (p < last) load_y(z0 - 3 + p);
    if constexpr (PAIR == 1) {
      if (p >= 4) {
        const T* pp = psi + (size_t)d.plane * (size_t)wrap_any(z0 - 6 + p, d.nz);
#pragma unroll
        for (int s = 0; s < C::LR; ++s)
          if (t + s * NT < C::NR) pe[s] = pp[roff[s]];
#pragma unroll
        for (int q = 0; q < C::OR; ++q) pt[q] = pp[ooff[q]];
      }
    }
    if (p >= 8) {
      const size_t kb = (size_t)d.plane * (size_t)(z0 - 8 + p);
#pragma unroll
      for (int q = 0; q < C::OR; ++q) {
        if (oin[q]) {
          pv[q] = __builtin_nontemporal_load(psi + kb + ooff[q]);
          if constexpr (PAIR == 1) {
            av[q] = __builtin_nontemporal_load(a + kb + ooff[q]);
            bv[q] = __builtin_nontemporal_load(b + kb + ooff[q]);
          }
        }
      }
    }
    __syncthreads();
    if (p >= 2) {  // mu_A on plane z0-5+p, tile+3
#pragma unroll
      for (int s = 0; s < C::L3; ++s) {
        const int o = oA[s];
        if (o >= 0) {
          mas[s0][t + s * NT] = ch_mu(cf, GC_NB_LDS(ys[s1], ys[s2], ys[s0], o, C::W4));
        }
      }
    }
    __syncthreads();
    if (p >= 4) {  // k_A and y' on plane z0-6+p, tile+2
      const int zA = z0 - 6 + p;
      const bool own = zA >= z0 && zA < z0 + nout;
      auto kA = [&](int mx, int my) {
        const int o = (my + 1) * C::W3 + (mx + 1);
        return ch_k(cf, GC_NB_LDS(mas[s1], mas[s2], mas[s0], o, C::W3));
      };
      // tile sites: the output owner
#pragma unroll
      for (int q = 0; q < C::OR; ++q) {
        const int mx = ox + 2, my = oy0 + q * C::ROWS + 2;
        const T kv = kA(mx, my);
        const int e = my * C::W2 + mx;
        if constexpr (PAIR == 0) {
          y2s[s0][e] = stage_axpy(ys[s2][(my + 2) * C::W4 + (mx + 2)], co.half_dt, kv);
          if (own && oin[q])
            __builtin_nontemporal_store(kv, a + (size_t)d.plane * (size_t)zA + ooff[q]);
        } else {
          y2s[s0][e] = stage_axpy(pt[q], co.dt, kv);
          k2r[q] = k1r[q];
          k1r[q] = k0r[q];
          k0r[q] = kv;
        }
      }
      // the 2-site ring around the tile
#pragma unroll
      for (int s = 0; s < C::LR; ++s) {
        const int o = rk[s];
        if (o >= 0) {
          const T kv = ch_k(cf, GC_NB_LDS(mas[s1], mas[s2], mas[s0], o, C::W3));
          const int e = re[s];
          if constexpr (PAIR == 0) y2s[s0][e] = stage_axpy(ys[s2][rb[s]], co.half_dt, kv);
          else y2s[s0][e] = stage_axpy(pe[s], co.dt, kv);
        }
      }
    }
    __syncthreads();
    if (p >= 6) {  // mu_B on plane z0-7+p, tile+1
#pragma unroll
      for (int s = 0; s < C::L1; ++s) {
        const int o = oB[s];
        if (o >= 0) {
          mbs[s0][t + s * NT] = ch_mu(cf, GC_NB_LDS(y2s[s1], y2s[s2], y2s[s0], o, C::W2));
        }
      }
    }
    __syncthreads();
    if (p >= 8) {  // outputs on plane z0-8+p
      const size_t kb = (size_t)d.plane * (size_t)(z0 - 8 + p);
#pragma unroll
      for (int q = 0; q < C::OR; ++q) {
        if (oin[q]) {
          const int oy = oy0 + q * C::ROWS;
          const int o = (oy + 1) * C::W1 + (ox + 1);
          const T kv = ch_k(cf, GC_NB_LDS(mbs[s1], mbs[s2], mbs[s0], o, C::W1));
          if constexpr (PAIR == 0) {
            __builtin_nontemporal_store(kv, b + kb + ooff[q]);
            __builtin_nontemporal_store(stage_axpy(pv[q], co.half_dt, kv), out + kb + ooff[q]);
          } else {
            const T bsum = bv[q] + k2r[q];  // k3 of this plane (k_A pass of iteration p-2)
            __builtin_nontemporal_store(rk4_final(pv[q], av[q], bsum, kv, co.sixth_dt, co.third_dt),
                                        out + kb + ooff[q]);
          }
        }
      }
    }
  }
}

// The same two stages with the z direction in registers.  A thread owns columns (x, y) of
// the tile+4 region (column t + s NT, row-major) and keeps, per column, the last three planes
// of each quantity the stages form there -- y, mu_A, y', mu_B -- as register queues; LDS holds
// one plane of each quantity, for the in-plane neighbours only.  A Laplacian then reads 4 LDS
// values (x and y neighbours) instead of 7 and takes its centre and z neighbours from the
// queue; the LDS addresses are loop invariant and each plane is staged once.  Iteration p
// loads plane P_p = z0 - 4 + p and runs two phases separated by barriers, each pairing two
// independent pieces of work (a plane later for mu_B and the outputs, so that every input a
// phase needs was formed in an earlier phase):
//   X: mu_A on P_{p-1} (columns at least 1 from the region edge), mu_B on P_{p-4} (at
//      least 3)                               [stores mu_A(P_{p-2}), mu_B(P_{p-5})]
//   Y: k_A and y' on P_{p-2} (at least 2), the outputs on P_{p-5} (the tile)
//                                             [stores y(P_p), y'(P_{p-3})]
// Each LDS plane is written in the phase after its last reader and read after the next
// barrier, so one slot per quantity suffices.  The arithmetic (operands, order) is that of
// k_ch_pair3d.
// The order-2 stencil with the centre and z neighbours from registers (cc, zlo, zhi) and the
// in-plane neighbours from LDS plane `pl` (row width W) around offset o.
#define GC_NB_Q(pl, o, W, cc, zlo, zhi) \
  cc, (pl)[(o) - 1], (pl)[(o) + 1], (pl)[(o) - (W)], (pl)[(o) + (W)], zlo, zhi

template <class T, int OX, int OY, int ZC, int NT, int PAIR>
__global__ __launch_bounds__(NT) void k_ch_pairq3d(CHCoef<T> cf, const T* __restrict__ y,
                                                   const T* __restrict__ psi,
                                                   T* __restrict__ out, T* __restrict__ a,
                                                   T* __restrict__ b, Dims d, int ntx, int nty,
                                                   StageCoef<T> co) {
  using C = PairCfg<OX, OY, NT>;
  constexpr int W = C::W4, H = C::H4, N = C::N4;
  // A thread's columns: slots 0 .. OR-1 are its tile sites (the outputs), the rest halo sites
  // (rows 0-3, rows H-4 .. H-1, then the four columns each side of the tile rows), so the
  // output-only queues and operands need OR slots instead of L
  constexpr int OR = C::OR, NH = N - OX * OY, L = OR + (NH + NT - 1) / NT;
  __shared__ T ys[N], ms[N], ws[N], bs[N];  // y, mu_A, y', mu_B: one plane each
  const int t = threadIdx.x;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int tx = (int)(bid % (unsigned)ntx);
  const unsigned r = bid / (unsigned)ntx;
  const int ty = (int)(r % (unsigned)nty);
  const int zc = (int)(r / (unsigned)nty);
  const int x0 = tx * OX, y0 = ty * OY, z0 = zc * ZC;
  const int nout = min(ZC, d.nz - z0);

  unsigned goff[L];  // the column's site in a plane (wrapped)
  int dist[L];       // distance from the region edge; -1: no column
  int ro[L];         // the column's index in the region (LDS planes)
  bool oin[L];       // the column is a tile site inside the grid (an output)
#pragma unroll
  for (int s = 0; s < L; ++s) {
    int px = 0, py = 0;
    bool have = true;
    if (s < OR) {
      const int i = t + s * NT;
      px = 4 + i % OX;
      py = 4 + i / OX;
    } else {
      const int h = t + (s - OR) * NT;
      have = h < NH;
      if (h < 4 * W) {
        px = h % W;
        py = h / W;
      } else if (h < 8 * W) {
        px = (h - 4 * W) % W;
        py = H - 4 + (h - 4 * W) / W;
      } else {
        const int c = (h - 8 * W) % 8;
        px = c < 4 ? c : W - 8 + c;
        py = 4 + (h - 8 * W) / 8;
      }
    }
    if (have) {
      goff[s] = (unsigned)(wrap_any(x0 + px - 4, d.nx) + d.nx * wrap_any(y0 + py - 4, d.ny));
      dist[s] = min(min(px, py), min(W - 1 - px, H - 1 - py));
      ro[s] = py * W + px;
      oin[s] = s < OR && x0 + px - 4 < d.nx && y0 + py - 4 < d.ny;
    } else {
      goff[s] = 0u;
      dist[s] = -1;
      ro[s] = 0;
      oin[s] = false;
    }
  }
  // Queues of the last three planes of y, mu_A, y', mu_B (and, PAIR 1, k_A): physical slot
  // R = p mod 3 takes the value formed in iteration p, so slot (R + 2) mod 3 holds iteration
  // p-1's and (R + 1) mod 3 iteration p-2's.  The loop is unrolled by three so that R is a
  // compile-time constant and the queues rotate by renaming instead of register moves.
  T yq[3][L] = {}, mq[3][L] = {}, wq[3][L] = {}, bq[3][L] = {};
  T kq[3][PAIR == 1 ? OR : 1] = {};  // PAIR 1: k3 of the tile sites
  // FP32: S2 does not store y3 = psi + dt/2 k2; S3+S4 forms it from psi and k2 as it loads
  // them (stage_axpy, the same bits), 7 words per site and step instead of 9.  FP64 keeps the
  // stored y3: its S3+S4 kernel has no registers for the second load stream (1024-thread
  // blocks cap it at 128 VGPRs; with it, 0.2943 ms per step at 256^3 against 0.2830 ms
  // without, on one box).
  constexpr bool kY3 = sizeof(T) == 4;
  T yn[L] = {};  // the next plane of y (kY3, PAIR 1: of psi), in flight
  T kn[PAIR == 1 && kY3 ? L : 1] = {};  // kY3, PAIR 1: the next plane of k2, in flight
  T pe[PAIR == 1 ? L : 1] = {};  // PAIR 1: psi on P_{p-2} (base of y')
  // FP64: the output's psi operand (plane P_{p-5}) from registers instead of HBM.  PAIR 1:
  // pe of the tile slots over the last three iterations (iteration p's output plane is
  // iteration p-3's P_{p-2}).  PAIR 0: the tile slots' y (= psi) planes P_{p-3} .. P_{p-5},
  // taken from the y queue as they leave it.  (The FP32 tiles hold two output sites per
  // thread, and the queue would spill.)
  constexpr bool kPQ = sizeof(T) == 8 && (PAIR == 1 || !kY3);
  T pq[3][kPQ ? OR : 1] = {};
  // the output site's point operands: PAIR 1 psi, k1, k2; PAIR 0 psi (for y3) unless kY3
  constexpr bool kPV = PAIR == 1 || !kY3;
  T pv[kPV ? OR : 1] = {}, av[PAIR == 1 ? OR : 1] = {}, bv[PAIR == 1 ? OR : 1] = {};
  // They are requested one iteration ahead (after that iteration's y plane, so waiting for
  // them never waits for the newer y plane): they have a whole iteration to arrive instead of
  // phase X only.  The S3+S4 kernels of FP64 and of tiles with more than three columns per
  // thread have no registers to spare for them (they would spill) and request them in the
  // iteration that uses them.
  constexpr bool kAhead = kPV && (PAIR == 0 || (sizeof(T) == 4 && L <= 3));
  T pen[PAIR == 1 && kAhead ? L : 1] = {};
  T pvn[kAhead ? OR : 1] = {}, avn[kAhead && PAIR == 1 ? OR : 1] = {},
    bvn[kAhead && PAIR == 1 ? OR : 1] = {};

  auto load_y = [&](int plane) {
    const size_t zo = (size_t)d.plane * (size_t)wrap_any(plane, d.nz);
    if constexpr (PAIR == 1 && kY3) {
#pragma unroll
      for (int s = 0; s < L; ++s)
        if (dist[s] >= 0) {
          yn[s] = psi[zo + goff[s]];
          kn[s] = b[zo + goff[s]];
        }
    } else {
      const T* yp = y + zo;
#pragma unroll
      for (int s = 0; s < L; ++s)
        if (dist[s] >= 0) yn[s] = yp[goff[s]];
    }
  };
  const int last = nout + 8;
  auto iter = [&](auto rc, int p) {
    constexpr int R0 = decltype(rc)::value, R1 = (R0 + 2) % 3, R2 = (R0 + 1) % 3;
    if constexpr (kPQ && PAIR == 0) {  // P_{p-3}, leaving the y queue
#pragma unroll
      for (int s = 0; s < OR; ++s) pq[R0][s] = yq[R0][s];
    }
#pragma unroll
    for (int s = 0; s < L; ++s) {
      if constexpr (PAIR == 1 && kY3) yq[R0][s] = stage_axpy(yn[s], co.half_dt, kn[s]);
      else yq[R0][s] = yn[s];
    }
    if (p < last) load_y(z0 - 3 + p);
    if constexpr (PAIR == 1 && kAhead) {
      if (p >= 4) {
#pragma unroll
        for (int s = 0; s < L; ++s) pe[s] = pen[s];
      }
      if (p >= 3 && p < last) {  // psi on P_{p-1}, for iteration p+1
        const T* pp = psi + (size_t)d.plane * (size_t)wrap_any(z0 - 5 + p, d.nz);
#pragma unroll
        for (int s = 0; s < L; ++s)
          if (dist[s] >= 2) pen[s] = pp[goff[s]];
      }
    } else if constexpr (PAIR == 1) {
      if (p >= 4) {
        const T* pp = psi + (size_t)d.plane * (size_t)wrap_any(z0 - 6 + p, d.nz);
#pragma unroll
        for (int s = 0; s < L; ++s)
          if (dist[s] >= 2) pe[s] = pp[goff[s]];
      }
    }
    if constexpr (kAhead) {
      if (p >= 9) {
#pragma unroll
        for (int s = 0; s < OR; ++s) {
          if constexpr (!kPQ) pv[s] = pvn[s];
          if constexpr (PAIR == 1) {
            av[s] = avn[s];
            bv[s] = bvn[s];
          }
        }
      }
      if (p >= 8 && p < last) {  // the outputs of iteration p+1: plane z0-8+p
        const size_t kb = (size_t)d.plane * (size_t)(z0 - 8 + p);
#pragma unroll
        for (int s = 0; s < OR; ++s) {
          if (oin[s]) {
            if constexpr (!kPQ) pvn[s] = __builtin_nontemporal_load(psi + kb + goff[s]);
            if constexpr (PAIR == 1) {
              avn[s] = __builtin_nontemporal_load(a + kb + goff[s]);
              bvn[s] = __builtin_nontemporal_load(b + kb + goff[s]);
            }
          }
        }
      }
    } else if (PAIR == 1 && p >= 9) {
      const size_t kb = (size_t)d.plane * (size_t)(z0 - 9 + p);
#pragma unroll
      for (int s = 0; s < OR; ++s) {
        if (oin[s]) {
          if constexpr (!kPQ) pv[s] = __builtin_nontemporal_load(psi + kb + goff[s]);
          av[s] = __builtin_nontemporal_load(a + kb + goff[s]);
          bv[s] = __builtin_nontemporal_load(b + kb + goff[s]);
        }
      }
    }
    // X: mu_A on P_{p-1}, mu_B on P_{p-4}; store mu_A(P_{p-2}) and mu_B(P_{p-5}) (formed in
    // phase X of iteration p-1)
#pragma unroll
    for (int s = 0; s < L; ++s) {
      const int o = ro[s];
      if (p >= 3 && dist[s] >= 1) ms[o] = mq[R1][s];
      if (p >= 8 && dist[s] >= 3) bs[o] = bq[R1][s];
      if (p >= 2 && dist[s] >= 1)
        mq[R0][s] = ch_mu(cf, GC_NB_Q(ys, o, W, yq[R1][s], yq[R2][s], yq[R0][s]));
      if (p >= 7 && dist[s] >= 3)  // y' on P_{p-3}, P_{p-4}, P_{p-5}: slots R1, R2, R0
        bq[R0][s] = ch_mu(cf, GC_NB_Q(ws, o, W, wq[R2][s], wq[R0][s], wq[R1][s]));
    }
    __syncthreads();
    // Y: the outputs on P_{p-5}, k_A and y' on P_{p-2}; store y(P_p) and y'(P_{p-3}) (formed
    // in phase Y of iteration p-1)
    {
      const size_t kb = (size_t)d.plane * (size_t)(z0 - 9 + p);
      const int zA = z0 - 6 + p;
      const bool own = zA >= z0 && zA < z0 + nout;
#pragma unroll
      for (int s = 0; s < L; ++s) {
        const int o = ro[s];
        if (dist[s] >= 0) ys[o] = yq[R0][s];
        if (p >= 5 && dist[s] >= 2) ws[o] = wq[R1][s];
        if (s < OR && p >= 9 && oin[s]) {  // mu_B on P_{p-4}, P_{p-5}, P_{p-6}: slots R0, R1, R2
          const T kv = ch_k(cf, GC_NB_Q(bs, o, W, bq[R1][s], bq[R2][s], bq[R0][s]));
          if constexpr (PAIR == 0) {
            __builtin_nontemporal_store(kv, b + kb + goff[s]);
            if constexpr (!kY3)  // psi on P_{p-5}: saved in iteration p-2, slot R2
              __builtin_nontemporal_store(stage_axpy(kPQ ? pq[R2][s] : pv[s], co.half_dt, kv),
                                          out + kb + goff[s]);
          } else {
            // k3 of P_{p-5}: phase Y of iteration p-3, slot R0 (read before it is refilled
            // below)
            const T bsum = bv[s] + kq[R0][s];
            __builtin_nontemporal_store(
                rk4_final(kPQ ? pq[R0][s] : pv[s], av[s], bsum, kv, co.sixth_dt, co.third_dt),
                out + kb + goff[s]);
          }
        }
        if (p >= 4 && dist[s] >= 2) {
          const T kv = ch_k(cf, GC_NB_Q(ms, o, W, mq[R1][s], mq[R2][s], mq[R0][s]));
          if constexpr (PAIR == 0) {
            wq[R0][s] = stage_axpy(yq[R2][s], co.half_dt, kv);
            if (own && oin[s])
              __builtin_nontemporal_store(kv, a + (size_t)d.plane * (size_t)zA + goff[s]);
          } else {
            wq[R0][s] = stage_axpy(pe[s], co.dt, kv);
            if (s < OR) kq[R0][s] = kv;
          }
        }
      }
      if constexpr (kPQ && PAIR == 1) {  // this iteration's P_{p-2}, the output plane 3 on
#pragma unroll
        for (int s = 0; s < OR; ++s) pq[R0][s] = pe[s];
      }
    }
    __syncthreads();
  };
  load_y(z0 - 4);
  for (int p = 0; p <= last; p += 3) {
    iter(std::integral_constant<int, 0>{}, p);
    if (p + 1 <= last) iter(std::integral_constant<int, 1>{}, p + 1);
    if (p + 2 <= last) iter(std::integral_constant<int, 2>{}, p + 2);
  }
}

// ---- one kernel per RK4 step (2-D, temporal blocking) -------------------------------------
//
// A block owns an OX x OY tile of a 2-D grid, loads it with an 8-site periodic halo into LDS and
// runs all four stages there: stage s forms mu on the tile + (9 - 2s) halo and k_s on the tile
// + (8 - 2s) halo, so the halo shrinks by two per stage and k4 lands exactly on the tile; the
// new state goes to the other state buffer.  A step reads the state once (plus halos) and
// writes it once, in one launch instead of four.  Every site value is formed with the stage
// kernels' arithmetic (ch_mu2, ch_k2, stage_axpy, rk4_final), so the results are theirs bit
// for bit -- in FP64 the reference's (solve/RK4.hpp:127-153 with the right-hand side of
// examples/common/cahn_hilliard.hpp:41-57).
template <int OX, int OY>
struct Tb2Cfg {
  static constexpr int H = 8, W = OX + 2 * H, HH = OY + 2 * H, N = W * HH, NT = 256;
};

template <class T, int OX, int OY>
__global__ __launch_bounds__(256) void k_ch_tb2d(CHCoef<T> cf, const T* __restrict__ src,
                                                 T* __restrict__ dst, Dims d, int ntx,
                                                 StageCoef<T> co) {
  using C = Tb2Cfg<OX, OY>;
  constexpr int W = C::W, H = C::H;
  __shared__ T P[C::N];   // psi on the tile + 8 halo
  __shared__ T Y[C::N];   // stage input y_s (s >= 2)
  __shared__ T MU[C::N];  // mu of the current stage
  __shared__ T K1[OX * OY], K23[OX * OY];
  const int t = threadIdx.x;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int x0 = (int)(bid % (unsigned)ntx) * OX, y0 = (int)(bid / (unsigned)ntx) * OY;
  for (int idx = t; idx < C::N; idx += C::NT) {
    const int px = idx % W, py = idx / W;
    P[idx] = src[(size_t)wrap_any(x0 + px - H, d.nx) + (size_t)d.nx * (size_t)wrap_any(y0 + py - H, d.ny)];
  }
  __syncthreads();
#pragma unroll
  for (int s = 1; s <= 4; ++s) {
    const T* __restrict__ in = (s == 1) ? P : Y;
    {  // mu on the tile + rm halo
      const int rm = 9 - 2 * s, wm = OX + 2 * rm, nm = wm * (OY + 2 * rm);
      for (int idx = t; idx < nm; idx += C::NT) {
        const int o = (idx / wm + H - rm) * W + (idx % wm + H - rm);
        MU[o] = ch_mu2(cf, in[o], in[o - 1], in[o + 1], in[o - W], in[o + W]);
      }
    }
    __syncthreads();
    {  // k_s on the tile + rk halo, then the stage update
      const int rk = 8 - 2 * s, wk = OX + 2 * rk, nk = wk * (OY + 2 * rk);
      for (int idx = t; idx < nk; idx += C::NT) {
        const int kx = idx % wk, ky = idx / wk;
        const int o = (ky + H - rk) * W + (kx + H - rk);
        const T kv = ch_k2(cf, MU[o], MU[o - 1], MU[o + 1], MU[o - W], MU[o + W]);
        const int ix = kx - rk, iy = ky - rk;
        const bool interior = ix >= 0 && ix < OX && iy >= 0 && iy < OY;
        const int ii = iy * OX + ix;
        if (s == 1) {
          Y[o] = stage_axpy(P[o], co.half_dt, kv);
          if (interior) K1[ii] = kv;
        } else if (s == 2) {
          Y[o] = stage_axpy(P[o], co.half_dt, kv);
          if (interior) K23[ii] = kv;
        } else if (s == 3) {
          Y[o] = stage_axpy(P[o], co.dt, kv);
          if (interior) K23[ii] = K23[ii] + kv;
        } else {
          const int gx = x0 + ix, gy = y0 + iy;
          if (gx < d.nx && gy < d.ny)
            dst[(size_t)gx + (size_t)d.nx * (size_t)gy] =
                rk4_final(P[o], K1[ii], K23[ii], kv, co.sixth_dt, co.third_dt);
        }
      }
    }
    __syncthreads();
  }
}

// ---- tile shapes of the Cahn-Hilliard stage kernels and their selection ------------------
//
// Candidate shapes (tile width OX, tile height OY, z-chunk ZC).  The default is used until
// gc_ch_tune has measured the candidates for a grid's extents and dtype; the winner is
// then remembered for that (extents, dtype) for the life of the process.

struct TileCfg {
  int ox, oy, zc;
};
constexpr TileCfg kCh3Cfgs[] = {{64, 4, 16}, {64, 4, 32},  {64, 8, 8},   {64, 8, 16}, {64, 8, 32},
                                {64, 16, 8}, {64, 16, 16}, {64, 16, 32}, {64, 16, 64}};
constexpr TileCfg kCh2Cfgs[] = {{64, 4, 1},  {64, 8, 1},   {64, 16, 1},
                                {64, 32, 1}, {128, 8, 1}, {128, 16, 1}};
constexpr int kN3 = sizeof(kCh3Cfgs) / sizeof(kCh3Cfgs[0]);
constexpr int kN2 = sizeof(kCh2Cfgs) / sizeof(kCh2Cfgs[0]);
// one-stage defaults: the measured best one-stage shapes on 256^3 (FP32: 64x16x16, FP64:
// 64x8x32) and 2-D 1024^2; used for Euler and single right-hand-side evaluations
constexpr int kStage3F32 = 6, kStage3F64 = 4, kDefault2 = 2;

// Grids with 1 < Nz < 8 use the untiled kernel (the z-march would mostly reload halos).
inline bool ch_tiled(const Dims& d) { return d.nz == 1 || d.nz >= 8; }

// 3-D two-stage (fused pair) shapes: ids kN3 .. kN3 + kNP - 1; block size NT threads; q:
// register-queue kernel (k_ch_pairq3d) instead of the plane-ring kernel (k_ch_pair3d).  The
// ring kernel's 128 x 16 and 64 x 32 tiles need more LDS than a workgroup has in FP64, and
// the queue kernel's FP64 queues fit the registers of a 1024-thread block only for 64 x 16
// tiles; FP64 runs 64 x 16 x ZC there.
struct PairTile {
  int ox, oy, zc, nt, q;
};
constexpr PairTile kChPairCfgs[] = {
    {64, 8, 32, 512, 0},    {64, 16, 16, 1024, 0},  {64, 16, 32, 1024, 0},  {64, 16, 64, 1024, 0},
    {128, 16, 16, 1024, 0}, {128, 16, 32, 1024, 0}, {128, 16, 64, 1024, 0}, {64, 32, 16, 1024, 0},
    {64, 32, 32, 1024, 0},  {64, 8, 32, 512, 1},    {64, 16, 32, 1024, 1},  {64, 16, 64, 1024, 1},
    {64, 32, 32, 1024, 1},  {64, 32, 64, 1024, 1},  {128, 16, 32, 1024, 1}};
constexpr int kNP = sizeof(kChPairCfgs) / sizeof(kChPairCfgs[0]);
inline bool is_pair(const Dims& d, int id) { return d.nz > 1 && id >= kN3; }
// 2-D one-launch-per-step (temporal blocking) shapes: ids kN2 .. kN2 + kNT2 - 1
constexpr TileCfg kChTb2Cfgs[] = {{32, 16, 1}, {32, 32, 1}, {64, 16, 1}, {64, 32, 1}};
constexpr int kNT2 = sizeof(kChTb2Cfgs) / sizeof(kChTb2Cfgs[0]);
inline bool is_tb2(const Dims& d, int id) { return d.nz == 1 && id >= kN2; }
// 3-D RK4 defaults: the measured best on 256^3 (register-queue two-stage kernels; FP32
// 64x32x32, FP64 64x16x64)
constexpr int kDefault3F32 = kN3 + 12, kDefault3F64 = kN3 + 11;


template <class T, int OX, int OY, int ZC>
void launch_ch3(gc_context* ctx, const Dims& d, const CHCoef<T>& cf, int stage, const T* y,
                T* psi, T* ynext, T* a, T* b, const StageCoef<T>& co) {
  const int ntx = (d.nx + OX - 1) / OX, nty = (d.ny + OY - 1) / OY, nzc = (d.nz + ZC - 1) / ZC;
  const dim3 grid((unsigned)(ntx * nty * nzc));
#define GC_L3(S)                                                                         \
  hipLaunchKernelGGL((k_ch_stage3d<T, OX, OY, ZC, S>), grid, dim3(256), 0, ctx->stream, cf, y, \
                     psi, ynext, a, b, d, ntx, nty, co)
  GC_STAGE_SWITCH(GC_L3)
#undef GC_L3
}

template <class T, int OX, int OY>
void launch_ch2(gc_context* ctx, const Dims& d, const CHCoef<T>& cf, int stage, const T* y,
                T* psi, T* ynext, T* a, T* b, const StageCoef<T>& co) {
  const int ntx = (d.nx + OX - 1) / OX, nty = (d.ny + OY - 1) / OY;
  const dim3 grid((unsigned)(ntx * nty));
#define GC_L2(S)                                                                         \
  hipLaunchKernelGGL((k_ch_stage2d<T, OX, OY, S>), grid, dim3(256), 0, ctx->stream, cf, y, psi, \
                     ynext, a, b, d, ntx, co)
  GC_STAGE_SWITCH(GC_L2)
#undef GC_L2
}

template <class T>
void launch_ch_cfg(gc_context* ctx, int id, const Dims& d, const CHCoef<T>& cf, int stage,
                   const T* y, T* psi, T* ynext, T* a, T* b, const StageCoef<T>& co) {
#define GC_CH3(i, OX, OY, ZC) \
  case i: launch_ch3<T, OX, OY, ZC>(ctx, d, cf, stage, y, psi, ynext, a, b, co); break;
#define GC_CH2(i, OX, OY) \
  case i: launch_ch2<T, OX, OY>(ctx, d, cf, stage, y, psi, ynext, a, b, co); break;
  if (d.nz == 1) {
    switch (id) {
      GC_CH2(0, 64, 4)
      GC_CH2(1, 64, 8)
      GC_CH2(2, 64, 16)
      GC_CH2(3, 64, 32)
      GC_CH2(4, 128, 8)
      default: launch_ch2<T, 128, 16>(ctx, d, cf, stage, y, psi, ynext, a, b, co); break;
    }
  } else {
    switch (id) {
      GC_CH3(0, 64, 4, 16)
      GC_CH3(1, 64, 4, 32)
      GC_CH3(2, 64, 8, 8)
      GC_CH3(3, 64, 8, 16)
      GC_CH3(4, 64, 8, 32)
      GC_CH3(5, 64, 16, 8)
      GC_CH3(6, 64, 16, 16)
      GC_CH3(7, 64, 16, 32)
      default: launch_ch3<T, 64, 16, 64>(ctx, d, cf, stage, y, psi, ynext, a, b, co); break;
    }
  }
#undef GC_CH3
#undef GC_CH2
}

template <class T, int OX, int OY, int ZC, int NT, bool Q = false>
void launch_pair3(gc_context* ctx, const Dims& d, const CHCoef<T>& cf, int pair, const T* y,
                  const T* psi, T* out, T* a, T* b, const StageCoef<T>& co) {
  const int ntx = (d.nx + OX - 1) / OX, nty = (d.ny + OY - 1) / OY, nzc = (d.nz + ZC - 1) / ZC;
  const dim3 grid((unsigned)(ntx * nty * nzc));
  if constexpr (Q) {
    if (pair == 0)
      hipLaunchKernelGGL((k_ch_pairq3d<T, OX, OY, ZC, NT, 0>), grid, dim3(NT), 0, ctx->stream, cf,
                         y, psi, out, a, b, d, ntx, nty, co);
    else
      hipLaunchKernelGGL((k_ch_pairq3d<T, OX, OY, ZC, NT, 1>), grid, dim3(NT), 0, ctx->stream, cf,
                         y, psi, out, a, b, d, ntx, nty, co);
  } else {
    if (pair == 0)
      hipLaunchKernelGGL((k_ch_pair3d<T, OX, OY, ZC, NT, 0>), grid, dim3(NT), 0, ctx->stream, cf,
                         y, psi, out, a, b, d, ntx, nty, co);
    else
      hipLaunchKernelGGL((k_ch_pair3d<T, OX, OY, ZC, NT, 1>), grid, dim3(NT), 0, ctx->stream, cf,
                         y, psi, out, a, b, d, ntx, nty, co);
  }
}

template <class T>
void launch_pair_cfg(gc_context* ctx, int id, const Dims& d, const CHCoef<T>& cf, int pair,
                     const T* y, const T* psi, T* out, T* a, T* b, const StageCoef<T>& co) {
  constexpr bool f32 = sizeof(T) == 4;
  switch (id - kN3) {
    case 0: launch_pair3<T, 64, 8, 32, 512>(ctx, d, cf, pair, y, psi, out, a, b, co); break;
    case 1: launch_pair3<T, 64, 16, 16, 1024>(ctx, d, cf, pair, y, psi, out, a, b, co); break;
    case 2: launch_pair3<T, 64, 16, 32, 1024>(ctx, d, cf, pair, y, psi, out, a, b, co); break;
    case 3: launch_pair3<T, 64, 16, 64, 1024>(ctx, d, cf, pair, y, psi, out, a, b, co); break;
    case 4:
      if constexpr (f32) launch_pair3<T, 128, 16, 16, 1024>(ctx, d, cf, pair, y, psi, out, a, b, co);
      else launch_pair3<T, 64, 16, 16, 1024>(ctx, d, cf, pair, y, psi, out, a, b, co);
      break;
    case 5:
      if constexpr (f32) launch_pair3<T, 128, 16, 32, 1024>(ctx, d, cf, pair, y, psi, out, a, b, co);
      else launch_pair3<T, 64, 16, 32, 1024>(ctx, d, cf, pair, y, psi, out, a, b, co);
      break;
    case 6:
      if constexpr (f32) launch_pair3<T, 128, 16, 64, 1024>(ctx, d, cf, pair, y, psi, out, a, b, co);
      else launch_pair3<T, 64, 16, 64, 1024>(ctx, d, cf, pair, y, psi, out, a, b, co);
      break;
    case 7:
      if constexpr (f32) launch_pair3<T, 64, 32, 16, 1024>(ctx, d, cf, pair, y, psi, out, a, b, co);
      else launch_pair3<T, 64, 16, 16, 1024>(ctx, d, cf, pair, y, psi, out, a, b, co);
      break;
    case 8:
      if constexpr (f32) launch_pair3<T, 64, 32, 32, 1024>(ctx, d, cf, pair, y, psi, out, a, b, co);
      else launch_pair3<T, 64, 16, 32, 1024>(ctx, d, cf, pair, y, psi, out, a, b, co);
      break;
    case 9: launch_pair3<T, 64, 8, 32, 512, true>(ctx, d, cf, pair, y, psi, out, a, b, co); break;
    case 10: launch_pair3<T, 64, 16, 32, 1024, true>(ctx, d, cf, pair, y, psi, out, a, b, co); break;
    case 11: launch_pair3<T, 64, 16, 64, 1024, true>(ctx, d, cf, pair, y, psi, out, a, b, co); break;
    case 12:
      if constexpr (f32) launch_pair3<T, 64, 32, 32, 1024, true>(ctx, d, cf, pair, y, psi, out, a, b, co);
      else launch_pair3<T, 64, 16, 32, 1024, true>(ctx, d, cf, pair, y, psi, out, a, b, co);
      break;
    case 13:
      if constexpr (f32) launch_pair3<T, 64, 32, 64, 1024, true>(ctx, d, cf, pair, y, psi, out, a, b, co);
      else launch_pair3<T, 64, 16, 64, 1024, true>(ctx, d, cf, pair, y, psi, out, a, b, co);
      break;
    default:
      if constexpr (f32) launch_pair3<T, 128, 16, 32, 1024, true>(ctx, d, cf, pair, y, psi, out, a, b, co);
      else launch_pair3<T, 64, 16, 32, 1024, true>(ctx, d, cf, pair, y, psi, out, a, b, co);
      break;
  }
}

// Process-wide choice of tile shape per (extents, dtype).
struct CfgKey {
  int nx, ny, nz, dt;
  bool operator<(const CfgKey& o) const {
    return std::tie(nx, ny, nz, dt) < std::tie(o.nx, o.ny, o.nz, o.dt);
  }
};
std::mutex g_cfg_mu;
std::map<CfgKey, int> g_cfg;

inline long tile_blocks(const Dims& d, int ox, int oy, int zc) {
  return (long)((d.nx + ox - 1) / ox) * ((d.ny + oy - 1) / oy) * ((d.nz + zc - 1) / zc);
}

// The shape used before (or without) gc_ch_tune.  3-D: the two-stage default measured best on
// 256^3 while it gives every CU a workgroup; smaller grids are latency-bound (a workgroup
// marches its z-chunk plane by plane), so they take the one-stage shape with the most
// workgroups and, among those, the shortest z-chunk (measured: 64x8x8 is the fastest
// candidate on 64^3 and 128^3 in both precisions, 3-5x the two-stage default there).
int default_config(const Dims& d, int dt) {
  if (d.nz == 1) return kN2;  // 2-D RK4: one launch per step, 32 x 16 tiles
  const int pair = dt == GC_F64 ? kDefault3F64 : kDefault3F32;
  const PairTile& pt = kChPairCfgs[pair - kN3];
  const int poy = (dt == GC_F64 && pt.q) ? 16 : pt.oy;  // FP64 queue kernels run 64 x 16
  if (tile_blocks(d, pt.ox, poy, pt.zc) >= 256) return pair;
  int best = 0;
  for (int i = 1; i < kN3; ++i) {
    const TileCfg& c = kCh3Cfgs[i];
    const TileCfg& b = kCh3Cfgs[best];
    const long nc = tile_blocks(d, c.ox, c.oy, c.zc), nb = tile_blocks(d, b.ox, b.oy, b.zc);
    if (nc > nb || (nc == nb && c.zc < b.zc)) best = i;
  }
  return best;
}

int ch_config(const Dims& d, int dt) {
  std::lock_guard<std::mutex> lock(g_cfg_mu);
  auto it = g_cfg.find(CfgKey{d.nx, d.ny, d.nz, dt});
  if (it != g_cfg.end()) return it->second;
  return default_config(d, dt);
}

template <class T>
gc_status launch_stage(gc_context* ctx, const gc_grid* g, const CHRhs<T>& rhs, int stage,
                       const T* y, T* psi, T* ynext, T* a, T* b, const StageCoef<T>& co,
                       int cfg = -1) {
  const Dims d = dims_of(g);
  const CHCoef<T>& cf = rhs.cf;
  if (ch_tiled(d)) {
    int id = cfg >= 0 ? cfg : ch_config(d, sizeof(T) == 8 ? GC_F64 : GC_F32);
    if (is_pair(d, id)) id = sizeof(T) == 8 ? kStage3F64 : kStage3F32;  // one stage only
    if (is_tb2(d, id)) id = kDefault2;
    launch_ch_cfg<T>(ctx, id, d, cf, stage, y, psi, ynext, a, b, co);
  } else {
    const int nbx = (d.nx + TX - 1) / TX, nby = (d.ny + TY - 1) / TY;
    hipLaunchKernelGGL((k_stage<T, CHRhs<T>>), dim3((unsigned)(nbx * nby * d.nz)),
                       dim3(TX * TY), 0, ctx->stream, rhs, stage, y, psi, ynext, a, b, d, nbx,
                       nby, co);
  }
  GC_LAUNCHED();
  return GC_OK;
}

// One 2-D RK4 step in one launch (k_ch_tb2d): psi_in -> psi_out (distinct buffers).
template <class T>
gc_status tb2_step(gc_context* ctx, const gc_grid* g, const CHRhs<T>& rhs, int id,
                   const T* psi_in, T* psi_out, const StageCoef<T>& co) {
  const Dims d = dims_of(g);
  const TileCfg c = kChTb2Cfgs[id - kN2];
#define GC_TB(OX, OY)                                                                       \
  do {                                                                                      \
    const int ntx = (d.nx + OX - 1) / OX, nty = (d.ny + OY - 1) / OY;                       \
    hipLaunchKernelGGL((k_ch_tb2d<T, OX, OY>), dim3((unsigned)(ntx * nty)), dim3(256), 0,  \
                       ctx->stream, rhs.cf, psi_in, psi_out, d, ntx, co);                   \
  } while (0)
  if (c.ox == 32 && c.oy == 16) GC_TB(32, 16);
  else if (c.ox == 32) GC_TB(32, 32);
  else if (c.oy == 16) GC_TB(64, 16);
  else GC_TB(64, 32);
#undef GC_TB
  GC_LAUNCHED();
  return GC_OK;
}

// One RK4 step with the two-stage kernels: psi_in -> psi_out (distinct buffers).
template <class T>
gc_status pair_step(gc_context* ctx, const gc_grid* g, const CHRhs<T>& rhs, int id,
                    const T* psi_in, T* psi_out, T* y3, T* a, T* b, const StageCoef<T>& co) {
  const Dims d = dims_of(g);
  const CHCoef<T>& cf = rhs.cf;
  launch_pair_cfg<T>(ctx, id, d, cf, 0, psi_in, psi_in, y3, a, b, co);
  GC_LAUNCHED();
  launch_pair_cfg<T>(ctx, id, d, cf, 1, y3, psi_in, psi_out, a, b, co);
  GC_LAUNCHED();
  return GC_OK;
}

// Runs n_steps of RK4 or explicit Euler with the fused stage kernels.  psi in place.
template <class T, class Rhs>
gc_status run_steps(gc_context* ctx, const gc_grid* g, const Rhs& rhs, T* psi, double dt,
                    int n_steps, int integrator) {
  if (n_steps == 0) return GC_OK;
  const size_t n = sites(g);
  StageCoef<T> co;
  // solve/RK4.hpp:123-125 (in double, then rounded to T)
  co.half_dt = T(0.5 * dt);
  co.dt = T(dt);
  co.sixth_dt = T(dt / 6.0);
  co.third_dt = T(dt / 3.0);
  void* scr = nullptr;
  if (integrator == GC_RK4) {
    GC_TRY(scratch(ctx, 4 * n * sizeof(T), &scr));
    T* yA = (T*)scr;
    T* yB = yA + n;
    T* a = yB + n;
    T* b = a + n;
    if constexpr (std::is_same<Rhs, CHRhs<T>>::value) {
      const Dims d = dims_of(g);
      const int id = ch_tiled(d) ? ch_config(d, sizeof(T) == 8 ? GC_F64 : GC_F32) : -1;
      if (id >= 0 && is_tb2(d, id)) {  // 2-D: one launch per step; yA holds the other state
        T* cur = psi;
        T* other = yA;
        for (int s = 0; s < n_steps; ++s) {
          GC_TRY(tb2_step<T>(ctx, g, rhs, id, cur, other, co));
          T* tmp = cur;
          cur = other;
          other = tmp;
        }
        if (cur != psi)
          GC_HIP(hipMemcpyAsync(psi, cur, n * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
        return GC_OK;
      }
      if (id >= 0 && is_pair(d, id)) {  // two stages per pass; yA holds the other state
        T* cur = psi;
        T* other = yA;
        for (int s = 0; s < n_steps; ++s) {
          GC_TRY(pair_step<T>(ctx, g, rhs, id, cur, other, yB, a, b, co));
          T* tmp = cur;
          cur = other;
          other = tmp;
        }
        if (cur != psi)
          GC_HIP(hipMemcpyAsync(psi, cur, n * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
        return GC_OK;
      }
    } else if (g->nz >= 8) {  // diffusion: two stages per pass
      T* cur = psi;
      T* other = yA;
      for (int s = 0; s < n_steps; ++s) {
        GC_TRY(pair_step<T>(ctx, g, rhs, cur, other, yB, a, b, co));
        T* tmp = cur;
        cur = other;
        other = tmp;
      }
      if (cur != psi)
        GC_HIP(hipMemcpyAsync(psi, cur, n * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
      return GC_OK;
    }
    for (int s = 0; s < n_steps; ++s) {
      GC_TRY(launch_stage<T>(ctx, g, rhs, kS1, psi, psi, yA, a, b, co));
      GC_TRY(launch_stage<T>(ctx, g, rhs, kS2, yA, psi, yB, a, b, co));
      GC_TRY(launch_stage<T>(ctx, g, rhs, kS3, yB, psi, yA, a, b, co));
      GC_TRY(launch_stage<T>(ctx, g, rhs, kS4, yA, psi, nullptr, a, b, co));
    }
    return GC_OK;
  }
  GC_TRY(scratch(ctx, n * sizeof(T), &scr));
  T* cur = psi;
  T* other = (T*)scr;
  for (int s = 0; s < n_steps; ++s) {
    GC_TRY(launch_stage<T>(ctx, g, rhs, kEuler, cur, nullptr, other, nullptr, nullptr, co));
    T* t = cur;
    cur = other;
    other = t;
  }
  if (cur != psi) GC_HIP(hipMemcpyAsync(psi, cur, n * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
  return GC_OK;
}

gc_status check_steps(const char* who, double dt, int n_steps) {
  if (n_steps < 0)
    return set_error(GC_ERR_INVALID, "%s: n_steps must be >= 0; got %d", who, n_steps);
  if (dt < 0.0) return set_error(GC_ERR_INVALID, "%s: dt must be >= 0; got %f", who, dt);
  return GC_OK;
}

const char* integ_name(int integrator) {
  return integrator == GC_RK4 ? "solve::integrate(... RK4)" : "solve::integrate(... ExplicitEuler)";
}

gc_status base_checks(gc_context* ctx, const gc_grid* g, gc_dtype dt, const void* psi,
                      int integrator) {
  GC_TRY(check_ctx(ctx));
  GC_TRY(check_grid(g));
  GC_TRY(check_dtype(dt));
  if (!psi) return set_error(GC_ERR_INVALID, "gridcalc: null field pointer");
  if (integrator != GC_RK4 && integrator != GC_EULER)
    return set_error(GC_ERR_INVALID, "gridcalc: unknown integrator %d", integrator);
  return GC_OK;
}

template <class T>
CHRhs<T> make_ch(const gc_grid* g, const gc_ch_params* p) {
  return CHRhs<T>{make_ch_coef<T>(T(p->M), T(p->kappa), T(1.0 / (g->hx * g->hx)),
                                  T(1.0 / (g->hy * g->hy)), T(1.0 / (g->hz * g->hz)))};
}
template <class T>
LapCoef<T> lap_coef_of(const gc_grid* g) {
  return make_lap_coef<T>(T(1.0 / (g->hx * g->hx)), T(1.0 / (g->hy * g->hy)),
                          T(1.0 / (g->hz * g->hz)));
}
template <class T>
LapRhs<T> make_lap(const gc_grid* g) {
  return LapRhs<T>{lap_coef_of<T>(g)};
}
template <class T>
FvmRhs<T> make_fvm(const gc_grid* g, const void* D) {
  return FvmRhs<T>{(const T*)D, lap_coef_of<T>(g)};
}

// solve/Diffusion.hpp:45-59
gc_status check_cfl(const gc_grid* g, double D, double dt, int integrator) {
  const double limit = integrator == GC_RK4 ? 0.6963 : 0.5;
  const double inv_sum = 1.0 / (g->hx * g->hx) + 1.0 / (g->hy * g->hy) + 1.0 / (g->hz * g->hz);
  const double cfl = D * dt * inv_sum;
  if (cfl > limit)
    return set_error(GC_ERR_CFL,
                     "solve::diffuse: explicit-diffusion CFL violated for %s: D=%f, dt=%f, "
                     "h=(%f, %f, %f); D*dt*(1/hx^2 + 1/hy^2 + 1/hz^2) = %f > %f",
                     integrator == GC_RK4 ? "RK4" : "ExplicitEuler", D, dt, g->hx, g->hy, g->hz,
                     cfl, limit);
  return GC_OK;
}

}  // namespace
}  // namespace gc

using namespace gc;

extern "C" {

gc_status gc_axpy(gc_context* ctx, gc_dtype dt, size_t n, const void* a, double scale,
                  const void* b, void* out) {
  GC_TRY(check_ctx(ctx));
  GC_TRY(check_dtype(dt));
  if (n == 0) return GC_OK;
  if (!a || !b || !out) return set_error(GC_ERR_INVALID, "gridcalc: null field pointer");
  if (dt == GC_F64)
    hipLaunchKernelGGL((k_axpy<double>), dim3(pw_blocks(n)), dim3(256), 0, ctx->stream, n,
                       (const double*)a, scale, (const double*)b, (double*)out);
  else
    hipLaunchKernelGGL((k_axpy<float>), dim3(pw_blocks(n)), dim3(256), 0, ctx->stream, n,
                       (const float*)a, (float)scale, (const float*)b, (float*)out);
  GC_LAUNCHED();
  return GC_OK;
}

gc_status gc_euler_update(gc_context* ctx, gc_dtype dt, size_t n, void* psi, double step,
                          const void* r) {
  GC_TRY(check_ctx(ctx));
  GC_TRY(check_dtype(dt));
  if (n == 0) return GC_OK;
  if (!psi || !r) return set_error(GC_ERR_INVALID, "gridcalc: null field pointer");
  if (dt == GC_F64)
    hipLaunchKernelGGL((k_euler_update<double>), dim3(pw_blocks(n)), dim3(256), 0, ctx->stream, n,
                       (double*)psi, step, (const double*)r);
  else
    hipLaunchKernelGGL((k_euler_update<float>), dim3(pw_blocks(n)), dim3(256), 0, ctx->stream, n,
                       (float*)psi, (float)step, (const float*)r);
  GC_LAUNCHED();
  return GC_OK;
}

gc_status gc_rk4_combine(gc_context* ctx, gc_dtype dt, size_t n, void* psi, double step,
                         const void* k1, const void* k2, const void* k3, const void* k4) {
  GC_TRY(check_ctx(ctx));
  GC_TRY(check_dtype(dt));
  if (n == 0) return GC_OK;
  if (!psi || !k1 || !k2 || !k3 || !k4)
    return set_error(GC_ERR_INVALID, "gridcalc: null field pointer");
  const double sixth = step / 6.0, third = step / 3.0;
  if (dt == GC_F64)
    hipLaunchKernelGGL((k_rk4_combine<double>), dim3(pw_blocks(n)), dim3(256), 0, ctx->stream, n,
                       (double*)psi, sixth, third, (const double*)k1, (const double*)k2,
                       (const double*)k3, (const double*)k4);
  else
    hipLaunchKernelGGL((k_rk4_combine<float>), dim3(pw_blocks(n)), dim3(256), 0, ctx->stream, n,
                       (float*)psi, (float)sixth, (float)third, (const float*)k1,
                       (const float*)k2, (const float*)k3, (const float*)k4);
  GC_LAUNCHED();
  return GC_OK;
}

gc_status gc_ch_rhs(gc_context* ctx, const gc_grid* g, gc_dtype dt, const gc_ch_params* p,
                    const void* psi, void* out) {
  GC_TRY(base_checks(ctx, g, dt, psi, GC_RK4));
  if (!p || !out) return set_error(GC_ERR_INVALID, "gridcalc: null argument");
  if (out == psi) return set_error(GC_ERR_INVALID, "gridcalc: output must not alias the input");
  GC_TRY(check_extent(g, 1));
  if (dt == GC_F64)
    return launch_stage<double>(ctx, g, make_ch<double>(g, p), kRhsOnly, (const double*)psi,
                                nullptr, (double*)out, nullptr, nullptr, StageCoef<double>{});
  return launch_stage<float>(ctx, g, make_ch<float>(g, p), kRhsOnly, (const float*)psi, nullptr,
                             (float*)out, nullptr, nullptr, StageCoef<float>{});
}

gc_status gc_ch_integrate(gc_context* ctx, const gc_grid* g, gc_dtype dt, const gc_ch_params* p,
                          void* psi, double step, int n_steps, int integrator) {
  GC_TRY(base_checks(ctx, g, dt, psi, integrator));
  if (!p) return set_error(GC_ERR_INVALID, "gridcalc: null parameters");
  GC_TRY(check_steps(integ_name(integrator), step, n_steps));
  if (n_steps == 0) return GC_OK;
  GC_TRY(check_extent(g, 1));
  if (dt == GC_F64)
    return run_steps<double>(ctx, g, make_ch<double>(g, p), (double*)psi, step, n_steps, integrator);
  return run_steps<float>(ctx, g, make_ch<float>(g, p), (float*)psi, step, n_steps, integrator);
}

gc_status gc_diffuse_const(gc_context* ctx, const gc_grid* g, gc_dtype dt, void* psi, double D,
                           double step, int n_steps, int integrator) {
  GC_TRY(base_checks(ctx, g, dt, psi, integrator));
  // solve/Diffusion.hpp:95-105
  if (D < 0.0) return set_error(GC_ERR_INVALID, "solve::diffuse: D must be >= 0; got %f", D);
  if (step < 0.0) return set_error(GC_ERR_INVALID, "solve::diffuse: dt must be >= 0; got %f", step);
  if (n_steps < 0)
    return set_error(GC_ERR_INVALID, "solve::diffuse: n_steps must be >= 0; got %d", n_steps);
  GC_TRY(check_cfl(g, D, step, integrator));
  if (n_steps == 0) return GC_OK;
  GC_TRY(check_extent(g, 1));
  // D folded into the step (Diffusion.hpp:106-111)
  const double eff = D * step;
  if (dt == GC_F64)
    return run_steps<double>(ctx, g, make_lap<double>(g), (double*)psi, eff, n_steps, integrator);
  return run_steps<float>(ctx, g, make_lap<float>(g), (float*)psi, eff, n_steps, integrator);
}

gc_status gc_diffuse_field(gc_context* ctx, const gc_grid* g, gc_dtype dt, void* psi,
                           const void* D, double step, int n_steps, int integrator,
                           int64_t* first_bad) {
  GC_TRY(base_checks(ctx, g, dt, psi, integrator));
  if (!D) return set_error(GC_ERR_INVALID, "gridcalc: null D pointer");
  if (D == psi) return set_error(GC_ERR_INVALID, "gridcalc: D must not alias psi");
  if (first_bad) *first_bad = -1;
  // solve/Diffusion.hpp:167-174
  if (step < 0.0) return set_error(GC_ERR_INVALID, "solve::diffuse: dt must be >= 0; got %f", step);
  if (n_steps < 0)
    return set_error(GC_ERR_INVALID, "solve::diffuse: n_steps must be >= 0; got %d", n_steps);
  double dmax = 0.0;
  GC_TRY(max_face_d(ctx, g, dt, D, &dmax));
  GC_TRY(check_cfl(g, dmax, step, integrator));
  if (n_steps == 0) return GC_OK;
  // the first right-hand-side evaluation runs cellLaplacian's checks
  // (fvm/CellLaplacian.hpp:95-109); D is constant over the run, so once suffices
  GC_TRY(check_extent(g, 1));
  int64_t bad = -1;
  GC_TRY(first_nonpositive(ctx, g, dt, D, &bad));
  if (bad >= 0) {
    if (first_bad) *first_bad = bad;
    return set_error(GC_ERR_NONPOS_D,
                     "fvm::cellLaplacian: D must be strictly positive at every cell (the "
                     "harmonic-mean face-D averaging is undefined otherwise); first violation "
                     "at flat index %lld",
                     (long long)bad);
  }
  if (dt == GC_F64)
    return run_steps<double>(ctx, g, make_fvm<double>(g, D), (double*)psi, step, n_steps,
                             integrator);
  return run_steps<float>(ctx, g, make_fvm<float>(g, D), (float*)psi, step, n_steps, integrator);
}

}  // extern "C"

// ---- Cahn-Hilliard tile autotuner -------------------------------------------------------

namespace {

template <class T>
gc_status tune_ch(gc_context* ctx, const gc_grid* g, int* best, float* ms_out) {
  const Dims d = dims_of(g);
  const int ncfg = d.nz == 1 ? kN2 + kNT2 : kN3 + kNP;
  const size_t n = sites(g);
  void* scr = nullptr;
  GC_TRY(scratch(ctx, 5 * n * sizeof(T), &scr));
  GC_HIP(hipMemsetAsync(scr, 0, 5 * n * sizeof(T), ctx->stream));
  T* psi = (T*)scr;
  T* yA = psi + n;
  T* yB = yA + n;
  T* a = yB + n;
  T* b = a + n;
  const gc_ch_params p{1.0, 1.0};
  const CHRhs<T> rhs = make_ch<T>(g, &p);
  StageCoef<T> co{T(0.0005), T(0.001), T(0.001 / 6.0), T(0.001 / 3.0)};
  // steps per measurement: about 2^24 site-stages of work, at least 2
  const int steps = (int)std::max<size_t>(2, std::min<size_t>(64, (size_t(1) << 24) / n));
  float best_ms = 0.0f;
  int best_id = -1;
  T* bufs[2] = {psi, yA};
  int cur = 0;
  for (int id = 0; id < ncfg; ++id) {
    auto step = [&]() -> gc_status {
      if (is_tb2(d, id)) {
        GC_TRY(tb2_step<T>(ctx, g, rhs, id, bufs[cur], bufs[cur ^ 1], co));
        cur ^= 1;
        return GC_OK;
      }
      if (is_pair(d, id)) {
        GC_TRY(pair_step<T>(ctx, g, rhs, id, bufs[cur], bufs[cur ^ 1], yB, a, b, co));
        cur ^= 1;
        return GC_OK;
      }
      GC_TRY(launch_stage<T>(ctx, g, rhs, kS1, psi, psi, yA, a, b, co, id));
      GC_TRY(launch_stage<T>(ctx, g, rhs, kS2, yA, psi, yB, a, b, co, id));
      GC_TRY(launch_stage<T>(ctx, g, rhs, kS3, yB, psi, yA, a, b, co, id));
      GC_TRY(launch_stage<T>(ctx, g, rhs, kS4, yA, psi, nullptr, a, b, co, id));
      return GC_OK;
    };
    GC_TRY(step());  // warm-up
    float ms = 0.0f;
    GC_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
    for (int s = 0; s < steps; ++s) GC_TRY(step());
    GC_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
    GC_HIP(hipEventSynchronize(ctx->ev_stop));
    GC_HIP(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    ms /= (float)steps;
    if (ms_out) ms_out[id] = ms;
    if (best_id < 0 || ms < best_ms) {
      best_ms = ms;
      best_id = id;
    }
  }
  {
    std::lock_guard<std::mutex> lock(g_cfg_mu);
    g_cfg[CfgKey{d.nx, d.ny, d.nz, sizeof(T) == 8 ? GC_F64 : GC_F32}] = best_id;
  }
  *best = best_id;
  return GC_OK;
}

gc_status cfg_checks(const gc_grid* g) {
  GC_TRY(check_grid(g));
  if (!ch_tiled(dims_of(g)))
    return set_error(GC_ERR_UNSUPPORTED, "gridcalc: grids with 1 < Nz < 8 use the untiled kernel");
  return GC_OK;
}

}  // namespace

extern "C" {

int gc_ch_config_count(const gc_grid* g) {
  if (!g || check_grid(g) != GC_OK || !ch_tiled(dims_of(g))) return 0;
  return g->nz == 1 ? kN2 + kNT2 : kN3 + kNP;
}

int gc_ch_config_fused(const gc_grid* g, int id) {
  if (!g || id < 0 || id >= gc_ch_config_count(g)) return -1;
  if (is_tb2(dims_of(g), id)) return 3;
  if (!is_pair(dims_of(g), id)) return 0;
  return kChPairCfgs[id - kN3].q ? 2 : 1;
}

gc_status gc_ch_config_shape(const gc_grid* g, int id, int* ox, int* oy, int* zc) {
  GC_TRY(cfg_checks(g));
  const int n = gc_ch_config_count(g);
  if (id < 0 || id >= n || !ox || !oy || !zc)
    return set_error(GC_ERR_INVALID, "gridcalc: config id %d out of range [0, %d)", id, n);
  if (is_pair(dims_of(g), id)) {
    const PairTile c = kChPairCfgs[id - kN3];
    *ox = c.ox;
    *oy = c.oy;
    *zc = c.zc;
    return GC_OK;
  }
  const TileCfg c = g->nz == 1 ? (id >= kN2 ? kChTb2Cfgs[id - kN2] : kCh2Cfgs[id]) : kCh3Cfgs[id];
  *ox = c.ox;
  *oy = c.oy;
  *zc = c.zc;
  return GC_OK;
}

gc_status gc_ch_set_config(const gc_grid* g, gc_dtype dt, int id) {
  GC_TRY(cfg_checks(g));
  GC_TRY(check_dtype(dt));
  const int n = gc_ch_config_count(g);
  if (id < -1 || id >= n)
    return set_error(GC_ERR_INVALID, "gridcalc: config id %d out of range [-1, %d)", id, n);
  std::lock_guard<std::mutex> lock(g_cfg_mu);
  const CfgKey k{g->nx, g->ny, g->nz, (int)dt};
  if (id < 0) g_cfg.erase(k);
  else g_cfg[k] = id;
  return GC_OK;
}

gc_status gc_ch_get_config(const gc_grid* g, gc_dtype dt, int* id) {
  GC_TRY(cfg_checks(g));
  GC_TRY(check_dtype(dt));
  if (!id) return set_error(GC_ERR_INVALID, "gridcalc: null id");
  *id = ch_config(dims_of(g), (int)dt);
  return GC_OK;
}

gc_status gc_ch_tune(gc_context* ctx, const gc_grid* g, gc_dtype dt, int* best, float* ms) {
  GC_TRY(check_ctx(ctx));
  GC_TRY(cfg_checks(g));
  GC_TRY(check_dtype(dt));
  GC_TRY(check_extent(g, 1));
  if (!best) return set_error(GC_ERR_INVALID, "gridcalc: null best");
  return dt == GC_F64 ? tune_ch<double>(ctx, g, best, ms) : tune_ch<float>(ctx, g, best, ms);
}

}  // extern "C"
