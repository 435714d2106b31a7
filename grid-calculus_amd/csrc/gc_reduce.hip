// gc_reduce.hip -- the reduction path: integrate() (pairwise and compensated), the
// fused Cahn-Hilliard free energy and gradient energy, and the diffusivity scans.
//
// Pairwise (func/Integrate.hpp:65-74) is reproduced exactly.  The reference splits a
// range of n values at n/2 until a range holds <= 64 values and sums each such leaf
// left to right.  At depth t every range has floor(n/2^t) or ceil(n/2^t) values, so
// with E the smallest depth at which ceil(n/2^E) <= 64, every range above depth E-1
// is split and every range at depth E-1 holds 64..128 values: one leaf of 64, or two
// leaves of floor(s/2) and ceil(s/2).  One lane owns one depth-(E-1) range ("node");
// the 2^(E-1) node sums then combine in a perfect binary tree -- xor-butterfly
// shuffles inside a wavefront, then LDS, then further passes -- whose every addition
// pairs exactly the two halves the recursion pairs (IEEE addition is commutative).
//
// One lane per leaf: the two lanes 2q, 2q+1 of a wave take the two leaves of a node (the
// second one empty for a node of 64 values), so a wave covers 32 nodes.  Loads are coalesced:
// 16 lanes fetch 16 consecutive values of one leaf into the wave's own LDS rows, then every
// lane adds its leaf's values in order; the next 16 values of every leaf are already being
// fetched into registers meanwhile.  No wave reads another wave's rows, so the loop needs no
// workgroup barrier.
#include "gc_common.hpp"

#include <algorithm>
#include <climits>
#include <cstring>
#include <mutex>

namespace gc {
namespace {

constexpr int kChunk = 16;          // values per leaf per LDS round
constexpr int kPad = kChunk + 1;    // padded LDS row (conflict-free b64 reads)
constexpr int kWaves = 4;           // waves per block of the tree passes
constexpr int kLeafWaves = 8;       // waves per block of the leaf kernel
constexpr unsigned kNodesPerBlock = 32u * kLeafWaves;  // 256 nodes (512 leaves) per block

// ---- value loaders ------------------------------------------------------------------

template <class T>
struct LoadPlain {
  const T* x;
  __device__ __forceinline__ double operator()(size_t c) const { return (double)x[c]; }
};

// ---- per-site derivatives for fused densities ---------------------------------------------
//
// func::evaluate (func/Functional.hpp:88-148) materialises the Order-2 gradient, Laplacian and
// biharmonic of psi and applies the density site by site.  Here the reduction's load computes
// them at the site it reads, in double precision from the stored values, with the operators'
// accumulation order (per-axis sums in stencil-offset order from 0, then the scaled terms left
// to right), so each value equals the materialised operator's bit for bit.

struct SiteIdx {
  int i, j, k;
};

// Grids hold at most 2^31-1 sites, so site indices fit 32-bit unsigned arithmetic (a 64-bit
// division costs several times a 32-bit one).
__device__ __forceinline__ SiteIdx site_of(const Dims& d, size_t c) {
  const unsigned cc = (unsigned)c;
  const unsigned r = cc / (unsigned)d.nx;
  const unsigned k = r / (unsigned)d.ny;
  return SiteIdx{(int)(cc - r * (unsigned)d.nx), (int)(r - k * (unsigned)d.ny), (int)k};
}

template <class T>
struct Sampler {  // wrapped read of component `comp` of an NC-component AoS field
  const T* p;
  Dims d;
  __device__ __forceinline__ double operator()(int i, int j, int k, int NC = 1, int comp = 0) const {
    const unsigned s = (unsigned)wrap(k, d.nz) * (unsigned)d.plane +
                       (unsigned)wrap(j, d.ny) * (unsigned)d.nx + (unsigned)wrap(i, d.nx);
    return (double)p[(size_t)s * NC + comp];
  }
};

// diff/Gradient.hpp:59-69 (Order 2) of component comp of an NC-component field
template <class T>
__device__ __forceinline__ void grad2_site(const Sampler<T>& P, SiteIdx s, double ihx, double ihy,
                                           double ihz, double& gx, double& gy, double& gz,
                                           int NC = 1, int comp = 0) {
  double dx = 0.0, dy = 0.0, dz = 0.0;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const int o = m - 1;
    const double w = D1<2>::w(m);
    dx += w * P(s.i + o, s.j, s.k, NC, comp);
    dy += w * P(s.i, s.j + o, s.k, NC, comp);
    dz += w * P(s.i, s.j, s.k + o, NC, comp);
  }
  gx = dx * ihx;
  gy = dy * ihy;
  gz = dz * ihz;
}

// diff/Laplacian.hpp:56-71 (Order 2)
template <class T>
__device__ __forceinline__ double lap2_site(const Sampler<T>& P, SiteIdx s, double ihx2,
                                            double ihy2, double ihz2) {
  double dxx = 0.0, dyy = 0.0, dzz = 0.0;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const int o = m - 1;
    const double w = D2<2>::w(m);
    dxx += w * P(s.i + o, s.j, s.k);
    dyy += w * P(s.i, s.j + o, s.k);
    dzz += w * P(s.i, s.j, s.k + o);
  }
  return dxx * ihx2 + dyy * ihy2 + dzz * ihz2;
}

// diff/Biharmonic.hpp:67-101 (Order 2): the fourth-derivative loop, the mixed loop (mb outer,
// ma inner), then the final expression as written
template <class T>
__device__ __forceinline__ double bih2_site(const Sampler<T>& P, SiteIdx s, double ihx2,
                                            double ihy2, double ihz2, double ihx4, double ihy4,
                                            double ihz4) {
  double dx4 = 0.0, dy4 = 0.0, dz4 = 0.0;
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    const int o = m - 2;
    const double w = D4<2>::w(m);
    dx4 += w * P(s.i + o, s.j, s.k);
    dy4 += w * P(s.i, s.j + o, s.k);
    dz4 += w * P(s.i, s.j, s.k + o);
  }
  double dxy = 0.0, dxz = 0.0, dyz = 0.0;
#pragma unroll
  for (int mb = 0; mb < 3; ++mb) {
    const int ob = mb - 1;
    const double wb = D2<2>::w(mb);
#pragma unroll
    for (int ma = 0; ma < 3; ++ma) {
      const int oa = ma - 1;
      const double wab = D2<2>::w(ma) * wb;
      dxy += wab * P(s.i + oa, s.j + ob, s.k);
      dxz += wab * P(s.i + oa, s.j, s.k + ob);
      dyz += wab * P(s.i, s.j + oa, s.k + ob);
    }
  }
  return dx4 * ihx4 + dy4 * ihy4 + dz4 * ihz4 +
         2.0 * (dxy * ihx2 * ihy2 + dxz * ihx2 * ihz2 + dyz * ihy2 * ihz2);
}

struct DensityCoef {
  double ihx, ihy, ihz;     // 1/h per axis (gradient)
  double ihx2, ihy2, ihz2;  // 1/h^2 (Laplacian, biharmonic)
  double ihx4, ihy4, ihz4;  // (1/h^2)^2 (biharmonic)
  double a, b;              // the density's parameters (gc_density)
};

// The density of gc_density kind K at flat site c.  Each formula is the one its C++
// counterpart in func/Densities.hpp evaluates, operation for operation:
//  VALUE             p                                             (arity 1)
//  GRADIENT_ENERGY   |g|^2 = (gx*gx + gy*gy) + gz*gz               (arity 2)
//  CAHN_HILLIARD     -0.5*p*p + 0.25*p*p*p*p + 0.5*a*|g|^2         (arity 2, a = kappa)
//  GINZBURG_LANDAU   q = p*p - 1; 0.5*|g|^2 + a*q*q                (arity 3, a = W)
//  PFC               q2 = a*a; q4 = q2*q2; p2 = p*p;               (arity 4, a = q0, b = eps)
//                    0.5*(q4 - b)*p2 + q2*p*lap + 0.5*p*bih + 0.25*p2*p2
//  LINEAR_ELASTIC    e = 0.5*(G + G^T), G(r,c) = d u_r/d x_c;      (Vec3 field, a = lambda,
//                    tr = (e00 + e11) + e22; ee = sum of e_i*e_i in   b = mu)
//                    column-major order; 0.5*(a*tr*tr + 2.0*b*ee)
template <class T, int K>
struct LoadDensity {
  const T* p;
  Dims d;
  DensityCoef k;
  __device__ __forceinline__ double operator()(size_t c) const {
    const Sampler<T> P{p, d};
    if constexpr (K == GC_DENSITY_VALUE) {
      return (double)p[c];
    } else if constexpr (K == GC_DENSITY_LINEAR_ELASTIC) {
      const SiteIdx s = site_of(d, c);
      double G[9];  // column-major: G(r, c) at r + 3c
#pragma unroll
      for (int r = 0; r < 3; ++r) grad2_site(P, s, k.ihx, k.ihy, k.ihz, G[r], G[r + 3], G[r + 6], 3, r);
      double e[9];
#pragma unroll
      for (int cc = 0; cc < 3; ++cc)
#pragma unroll
        for (int r = 0; r < 3; ++r) e[r + 3 * cc] = 0.5 * (G[r + 3 * cc] + G[cc + 3 * r]);
      const double tr = e[0] + e[4] + e[8];
      double ee = e[0] * e[0];
#pragma unroll
      for (int q = 1; q < 9; ++q) ee += e[q] * e[q];
      return 0.5 * (k.a * tr * tr + 2.0 * k.b * ee);
    } else {
      const SiteIdx s = site_of(d, c);
      const double v = (double)p[c];
      if constexpr (K == GC_DENSITY_PFC) {
        const double lap = lap2_site(P, s, k.ihx2, k.ihy2, k.ihz2);
        const double bih = bih2_site(P, s, k.ihx2, k.ihy2, k.ihz2, k.ihx4, k.ihy4, k.ihz4);
        const double q2 = k.a * k.a;
        const double q4 = q2 * q2;
        const double p2 = v * v;
        return 0.5 * (q4 - k.b) * p2 + q2 * v * lap + 0.5 * v * bih + 0.25 * p2 * p2;
      } else {
        double gx, gy, gz;
        grad2_site(P, s, k.ihx, k.ihy, k.ihz, gx, gy, gz);
        const double sq = gx * gx + gy * gy + gz * gz;
        if constexpr (K == GC_DENSITY_GRADIENT_ENERGY) {
          return sq;
        } else if constexpr (K == GC_DENSITY_CAHN_HILLIARD) {
          return -0.5 * v * v + 0.25 * v * v * v * v + 0.5 * k.a * sq;
        } else {
          static_assert(K == GC_DENSITY_GINZBURG_LANDAU, "density kind");
          const double q = v * v - 1.0;
          return 0.5 * sq + k.a * q * q;
        }
      }
    }
  }
};

inline DensityCoef density_coef(const gc_grid* g, double a, double b) {
  DensityCoef k;
  k.ihx = 1.0 / g->hx;
  k.ihy = 1.0 / g->hy;
  k.ihz = 1.0 / g->hz;
  k.ihx2 = 1.0 / (g->hx * g->hx);
  k.ihy2 = 1.0 / (g->hy * g->hy);
  k.ihz2 = 1.0 / (g->hz * g->hz);
  k.ihx4 = k.ihx2 * k.ihx2;
  k.ihy4 = k.ihy2 * k.ihy2;
  k.ihz4 = k.ihz2 * k.ihz2;
  k.a = a;
  k.b = b;
  return k;
}

// ---- exact pairwise tree ------------------------------------------------------------

// Perfect-tree sum of `cnt` (power of two, <= 64*kWaves) consecutive lane values of the
// block; result valid in thread 0.
__device__ __forceinline__ double block_tree_sum(double s, unsigned cnt, double* sh) {
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const unsigned in_wave = cnt < 64u ? cnt : 64u;
  for (unsigned st = 1; st < in_wave; st <<= 1) s = s + __shfl_xor(s, (int)st);
  if (cnt <= 64u) return s;
  if (lane == 0) sh[wave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned nw = cnt / 64u;  // 2 or 4
    if (nw == 2) s = sh[0] + sh[1];
    else s = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  }
  return s;
}

// Orders the wave's LDS writes before its later reads (and reads before later writes) without
// a workgroup barrier: the rows are private to the wave.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Perfect-tree sum over `cnt` (power of two, <= kNodesPerBlock) consecutive nodes of the
// block whose sums are held by lane pairs (both lanes of a pair hold the node's sum); result
// valid in thread 0.
__device__ __forceinline__ double block_tree_sum_pairs(double s, unsigned cnt, double* sh) {
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const unsigned in_wave = cnt < 32u ? cnt : 32u;
  for (unsigned st = 1; st < in_wave; st <<= 1) s = s + __shfl_xor(s, (int)(2 * st));
  if (cnt <= 32u) return s;
  if (lane == 0) sh[wave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned nw = cnt / 32u;  // 2, 4 or 8 waves, combined as a perfect tree
    if (nw == 2) s = sh[0] + sh[1];
    else if (nw == 4) s = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    else s = ((sh[0] + sh[1]) + (sh[2] + sh[3])) + ((sh[4] + sh[5]) + (sh[6] + sh[7]));
  }
  return s;
}

// Depth-`depth` nodes of the recursion over [0, n), each of 64..128 values (or fewer for
// n < 128), one leaf per lane; leafmax (<= 64) bounds every leaf's size.
template <class L>
__global__ __launch_bounds__(64 * kLeafWaves) void k_pairwise_leaves(L ld, size_t n, int depth,
                                                                     unsigned P, int leafmax,
                                                                     double* __restrict__ partial) {
  __shared__ double buf[kLeafWaves][64 * kPad];
  __shared__ size_t loff[kLeafWaves][64];
  __shared__ int lsize[kLeafWaves][64];
  __shared__ double sh[kLeafWaves];
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  // logical block: each XCD takes one contiguous range of nodes, so the neighbour rows and
  // planes a stencil-reading load touches stay in that XCD's L2
  const unsigned blk = xcd_remap(blockIdx.x, gridDim.x);
  const unsigned q = blk * kNodesPerBlock + wave * 32u + (lane >> 1);
  const unsigned h = lane & 1u;

  size_t off = 0, size = 0;
  if (q < P) {
    size = n;
    for (int b = depth - 1; b >= 0; --b) {
      const size_t mid = size >> 1;
      if ((q >> b) & 1u) {
        off += mid;
        size -= mid;
      } else {
        size = mid;
      }
    }
  }
  const bool split = size > 64;  // two leaves of floor(size/2) and ceil(size/2)
  int mine = 0;
  if (split) {
    const size_t m = size >> 1;
    mine = h ? (int)(size - m) : (int)m;
    off += h ? m : 0;
  } else if (!h) {
    mine = (int)size;
  }
  loff[wave][lane] = off;
  lsize[wave][lane] = mine;
  wave_lds_sync();

  double* row = &buf[wave][0];
  const int e = (int)(lane & (kChunk - 1));
  double v[kChunk];
  auto fetch = [&](int base) {
#pragma unroll
    for (int pass = 0; pass < kChunk; ++pass) {
      const int leaf = pass * (64 / kChunk) + (int)(lane / kChunk);
      const int idx = base + e;
      v[pass] = idx < lsize[wave][leaf] ? ld(loff[wave][leaf] + (size_t)idx) : 0.0;
    }
  };
  double s = 0.0;  // func/Integrate.hpp:66-69: s = 0.0, then += in order
  fetch(0);
  for (int base = 0; base < leafmax; base += kChunk) {
#pragma unroll
    for (int pass = 0; pass < kChunk; ++pass)
      row[(pass * (64 / kChunk) + (int)(lane / kChunk)) * kPad + e] = v[pass];
    wave_lds_sync();
    if (base + kChunk < leafmax) fetch(base + kChunk);
    const int lim = min(kChunk, mine - base);
    for (int t = 0; t < lim; ++t) s += row[lane * kPad + t];
    wave_lds_sync();
  }
  // node sum: left leaf + right leaf (IEEE addition is commutative: both lanes get its bits)
  const double o = __shfl_xor(s, 1);
  s = split ? s + o : (h ? o : s);
  const unsigned cnt = P < kNodesPerBlock ? P : kNodesPerBlock;
  s = block_tree_sum_pairs(s, cnt, sh);
  if (threadIdx.x == 0) partial[blk] = s;
}

// Perfect-tree sum of consecutive groups of min(cnt, 256) values.
__global__ __launch_bounds__(64 * kWaves) void k_tree_pass(const double* __restrict__ in,
                                                            unsigned cnt,
                                                            double* __restrict__ out) {
  __shared__ double sh[kWaves];
  const unsigned g = cnt < 64u * kWaves ? cnt : 64u * kWaves;
  const unsigned idx = blockIdx.x * g + threadIdx.x;
  double s = (threadIdx.x < g) ? in[idx] : 0.0;
  s = block_tree_sum(s, g, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// Scratch of the pairwise reduction of n values: two ping-pong arrays of node sums.
inline size_t pairwise_scratch_bytes(size_t n, unsigned* nb_out, int* depth_out) {
  int E = 0;
  while (((n + ((size_t)1 << E) - 1) >> E) > 64) ++E;
  const int depth = E > 0 ? E - 1 : 0;
  const unsigned P = 1u << depth;
  const unsigned nb = (P + kNodesPerBlock - 1) / kNodesPerBlock;
  if (nb_out) *nb_out = nb;
  if (depth_out) *depth_out = depth;
  return 2 * (size_t)nb * sizeof(double) + 256;
}

// Launches the pairwise reduction on `stream` with scratch `scr` (at least
// pairwise_scratch_bytes(n) bytes); the sum lands in *final_out (device memory, or the
// device address of pinned host memory) once the launches complete.
template <class L>
gc_status pairwise_launch(hipStream_t stream, L ld, size_t n, void* scr, double* final_out) {
  unsigned nb = 0;
  int depth = 0;
  pairwise_scratch_bytes(n, &nb, &depth);
  const unsigned P = 1u << depth;
  double* a = (double*)scr;
  double* b = a + nb;
  // nodes hold floor(n / 2^depth) or ceil(n / 2^depth) values; a node of more than 64 has
  // leaves of at most ceil(size / 2), one of at most 64 is a single leaf (65: 33, but 64: 64)
  const int maxsize = (int)((n + ((size_t)1 << depth) - 1) >> depth);
  const int minsize = (int)(n >> depth);
  auto leaf_of = [](int sz) { return sz > 64 ? (sz + 1) / 2 : sz; };
  const int leafmax = std::max(leaf_of(maxsize), leaf_of(minsize));
  hipLaunchKernelGGL((k_pairwise_leaves<L>), dim3(nb), dim3(64 * kLeafWaves), 0, stream, ld, n,
                     depth, P, leafmax, nb == 1 ? final_out : a);
  GC_LAUNCHED();
  unsigned cnt = nb;
  while (cnt > 1) {
    const unsigned g = cnt < 64u * kWaves ? cnt : 64u * kWaves;
    const unsigned blocks = cnt / g;
    hipLaunchKernelGGL(k_tree_pass, dim3(blocks), dim3(64 * kWaves), 0, stream, a, cnt,
                       blocks == 1 ? final_out : b);
    GC_LAUNCHED();
    cnt = blocks;
    double* t = a;
    a = b;
    b = t;
  }
  return GC_OK;
}

template <class L>
gc_status pairwise(gc_context* ctx, L ld, size_t n, double* result) {
  void* scr = nullptr;
  GC_TRY(scratch(ctx, pairwise_scratch_bytes(n, nullptr, nullptr), &scr));
  GC_TRY(pairwise_launch(ctx->stream, ld, n, scr, ctx->host_result_dev));
  GC_HIP(hipStreamSynchronize(ctx->stream));
  *result = ctx->host_result[0];
  return GC_OK;
}

// ---- compensated (Neumaier) sum -------------------------------------------------------

struct Comp {
  double s, c;
};

// Neumaier step (func/Integrate.hpp:88-97)
__device__ __forceinline__ void neu_add(Comp& a, double x) {
  const double t = a.s + x;
  if (fabs(a.s) >= fabs(x)) a.c += (a.s - t) + x;
  else a.c += (x - t) + a.s;
  a.s = t;
}

__device__ __forceinline__ Comp neu_merge(Comp a, Comp b) {
  Comp r{a.s, a.c + b.c};
  neu_add(r, b.s);
  return r;
}

__device__ __forceinline__ Comp block_comp_sum(Comp v, Comp* sh) {
  for (int st = 1; st < 64; st <<= 1) {
    Comp o{__shfl_xor(v.s, st), __shfl_xor(v.c, st)};
    // merge in index order so every lane holds the same value
    v = ((threadIdx.x & st) == 0) ? neu_merge(v, o) : neu_merge(o, v);
  }
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    v = neu_merge(neu_merge(sh[0], sh[1]), neu_merge(sh[2], sh[3]));
  }
  return v;
}

// Logical block b sums the contiguous range [b*chunk, (b+1)*chunk) (XCD-aware order: each
// XCD takes one contiguous run of blocks, so stencil-reading loads find their neighbours in
// that XCD's L2); the partition depends on n only, so the result is deterministic.
template <class L>
__global__ __launch_bounds__(256) void k_comp_partial(L ld, size_t n, size_t chunk,
                                                      Comp* __restrict__ partial) {
  __shared__ Comp sh[4];
  const unsigned blk = xcd_remap(blockIdx.x, gridDim.x);
  const size_t begin = (size_t)blk * chunk;
  const size_t end = begin + chunk < n ? begin + chunk : n;
  Comp a{0.0, 0.0};
  for (size_t c = begin + threadIdx.x; c < end; c += 256u) neu_add(a, ld(c));
  a = block_comp_sum(a, sh);
  if (threadIdx.x == 0) partial[blk] = a;
}

__global__ __launch_bounds__(256) void k_comp_final(const Comp* __restrict__ partial, unsigned nb,
                                                    double* __restrict__ out) {
  __shared__ Comp sh[4];
  Comp a{0.0, 0.0};
  for (unsigned b = threadIdx.x; b < nb; b += 256u) a = neu_merge(a, partial[b]);
  a = block_comp_sum(a, sh);
  if (threadIdx.x == 0) out[0] = a.s + a.c;
}

constexpr unsigned kCompMaxBlocks = 4096;

inline unsigned comp_blocks(size_t n) {
  const size_t want = (n + 255) / 256;
  return (unsigned)(want < kCompMaxBlocks ? (want == 0 ? 1 : want) : kCompMaxBlocks);
}

// Scratch of the compensated reduction: one (sum, compensation) pair per block.
inline size_t comp_scratch_bytes(size_t n) { return (size_t)comp_blocks(n) * sizeof(Comp) + 256; }

template <class L>
gc_status compensated_launch(hipStream_t stream, L ld, size_t n, void* scr, double* final_out) {
  const unsigned nb = comp_blocks(n);
  const size_t per = (n + nb - 1) / nb;
  const size_t chunk = (per + 255) / 256 * 256;
  Comp* part = (Comp*)scr;
  hipLaunchKernelGGL((k_comp_partial<L>), dim3(nb), dim3(256), 0, stream, ld, n, chunk, part);
  GC_LAUNCHED();
  hipLaunchKernelGGL(k_comp_final, dim3(1), dim3(256), 0, stream, part, nb, final_out);
  GC_LAUNCHED();
  return GC_OK;
}

template <class L>
gc_status compensated(gc_context* ctx, L ld, size_t n, double* result) {
  void* scr = nullptr;
  GC_TRY(scratch(ctx, comp_scratch_bytes(n), &scr));
  GC_TRY(compensated_launch(ctx->stream, ld, n, scr, ctx->host_result_dev));
  GC_HIP(hipStreamSynchronize(ctx->stream));
  *result = ctx->host_result[0];
  return GC_OK;
}

template <class L>
gc_status reduce(gc_context* ctx, int method, L ld, size_t n, double* result) {
  if (method == GC_PAIRWISE) return pairwise(ctx, ld, n, result);
  return compensated(ctx, ld, n, result);
}

// ---- diffusivity scans ------------------------------------------------------------------

__global__ void k_init_u64(unsigned long long* p, unsigned long long v) { *p = v; }

template <class T>
__global__ __launch_bounds__(256) void k_first_nonpos(const T* __restrict__ D, size_t n,
                                                      unsigned long long* __restrict__ out) {
  const size_t G = (size_t)gridDim.x * 256u;
  unsigned long long best = ULLONG_MAX;
  for (size_t c = (size_t)blockIdx.x * 256u + threadIdx.x; c < n; c += G)
    if (!(D[c] > T(0))) {
      best = c;
      break;
    }
  for (int st = 32; st >= 1; st >>= 1) {
    const unsigned long long o = __shfl_xor(best, st);
    best = o < best ? o : best;
  }
  if ((threadIdx.x & 63u) == 0 && best != ULLONG_MAX) atomicMin(out, best);
}

template <class T>
__device__ __forceinline__ double hmean_d(double a, double b) {
  return 2.0 * a * b / (a + b);
}

// solve/Diffusion.hpp:192-206; partial maxima per block
template <class T>
__global__ __launch_bounds__(256) void k_max_face(const T* __restrict__ D, Dims d, size_t n,
                                                  double* __restrict__ partial) {
  __shared__ double sh[4];
  const size_t G = (size_t)gridDim.x * 256u;
  double m = 0.0;
  for (size_t c = (size_t)blockIdx.x * 256u + threadIdx.x; c < n; c += G) {
    const int i = (int)(c % (size_t)d.nx);
    const size_t r = c / (size_t)d.nx;
    const int j = (int)(r % (size_t)d.ny);
    const int k = (int)(r / (size_t)d.ny);
    const int ip = (i + 1) % d.nx, jp = (j + 1) % d.ny, kp = (k + 1) % d.nz;
    const size_t row = (size_t)j * d.nx, pl = (size_t)k * d.plane;
    const double Dc = (double)D[c];
    const double fx = hmean_d<T>(Dc, (double)D[pl + row + ip]);
    const double fy = hmean_d<T>(Dc, (double)D[pl + (size_t)jp * d.nx + i]);
    const double fz = hmean_d<T>(Dc, (double)D[(size_t)kp * d.plane + row + i]);
    if (fx > m) m = fx;
    if (fy > m) m = fy;
    if (fz > m) m = fz;
  }
  for (int st = 32; st >= 1; st >>= 1) {
    const double o = __shfl_xor(m, st);
    if (o > m) m = o;
  }
  if ((threadIdx.x & 63u) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = sh[0];
    for (int w = 1; w < 4; ++w)
      if (sh[w] > r) r = sh[w];
    partial[blockIdx.x] = r;
  }
}

__global__ __launch_bounds__(256) void k_max_final(const double* __restrict__ partial, unsigned nb,
                                                   double* __restrict__ out) {
  __shared__ double sh[4];
  double m = 0.0;
  for (unsigned b = threadIdx.x; b < nb; b += 256u)
    if (partial[b] > m) m = partial[b];
  for (int st = 32; st >= 1; st >>= 1) {
    const double o = __shfl_xor(m, st);
    if (o > m) m = o;
  }
  if ((threadIdx.x & 63u) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = sh[0];
    for (int w = 1; w < 4; ++w)
      if (sh[w] > r) r = sh[w];
    out[0] = r;
  }
}

inline unsigned scan_blocks(size_t n) {
  const size_t want = (n + 255) / 256;
  return (unsigned)(want < 2048 ? (want == 0 ? 1 : want) : 2048);
}

gc_status field_checks(gc_context* ctx, const gc_grid* g, gc_dtype dt, const void* p) {
  GC_TRY(check_ctx(ctx));
  GC_TRY(check_grid(g));
  GC_TRY(check_dtype(dt));
  if (!p) return set_error(GC_ERR_INVALID, "gridcalc: null field pointer");
  return GC_OK;
}

gc_status check_method(int method) {
  if (method != GC_PAIRWISE && method != GC_KAHAN)
    return set_error(GC_ERR_INVALID, "func::integrate: Tag must be func::Pairwise or func::Kahan");
  return GC_OK;
}

}  // namespace

gc_status first_nonpositive(gc_context* ctx, const gc_grid* g, gc_dtype dt, const void* D,
                            int64_t* first_bad) {
  const size_t n = sites(g);
  void* scr = nullptr;
  GC_TRY(scratch(ctx, 256, &scr));
  unsigned long long* out = (unsigned long long*)scr;
  hipLaunchKernelGGL(k_init_u64, dim3(1), dim3(1), 0, ctx->stream, out, ULLONG_MAX);
  GC_LAUNCHED();
  const unsigned nb = scan_blocks(n);
  if (dt == GC_F64)
    hipLaunchKernelGGL((k_first_nonpos<double>), dim3(nb), dim3(256), 0, ctx->stream,
                       (const double*)D, n, out);
  else
    hipLaunchKernelGGL((k_first_nonpos<float>), dim3(nb), dim3(256), 0, ctx->stream,
                       (const float*)D, n, out);
  GC_LAUNCHED();
  unsigned long long h = 0;
  GC_HIP(hipMemcpyAsync(ctx->host_result, out, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
  GC_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(&h, ctx->host_result, sizeof h);
  *first_bad = (h == ULLONG_MAX) ? -1 : (int64_t)h;
  return GC_OK;
}

gc_status max_face_d(gc_context* ctx, const gc_grid* g, gc_dtype dt, const void* D, double* dmax) {
  const size_t n = sites(g);
  const unsigned nb = scan_blocks(n);
  void* scr = nullptr;
  GC_TRY(scratch(ctx, ((size_t)nb + 1) * sizeof(double) + 256, &scr));
  double* part = (double*)scr;
  double* out = part + nb;
  const Dims d = dims_of(g);
  if (dt == GC_F64)
    hipLaunchKernelGGL((k_max_face<double>), dim3(nb), dim3(256), 0, ctx->stream,
                       (const double*)D, d, n, part);
  else
    hipLaunchKernelGGL((k_max_face<float>), dim3(nb), dim3(256), 0, ctx->stream,
                       (const float*)D, d, n, part);
  GC_LAUNCHED();
  hipLaunchKernelGGL(k_max_final, dim3(1), dim3(256), 0, ctx->stream, part, nb, out);
  GC_LAUNCHED();
  GC_HIP(hipMemcpyAsync(ctx->host_result, out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  GC_HIP(hipStreamSynchronize(ctx->stream));
  *dmax = ctx->host_result[0];
  return GC_OK;
}

// Reductions used by the time-stepping code (already validated arguments).
gc_status integrate_raw(gc_context* ctx, const gc_grid* g, gc_dtype dt, int method,
                        const void* in, double* sum_out) {
  const size_t n = sites(g);
  if (dt == GC_F64) return reduce(ctx, method, LoadPlain<double>{(const double*)in}, n, sum_out);
  return reduce(ctx, method, LoadPlain<float>{(const float*)in}, n, sum_out);
}

template <int K>
gc_status density_raw(gc_context* ctx, const gc_grid* g, gc_dtype dt, int method,
                      const void* field, double a, double b, double* sum_out) {
  const size_t n = sites(g);
  const Dims d = dims_of(g);
  const DensityCoef k = density_coef(g, a, b);
  if (dt == GC_F64)
    return reduce(ctx, method, LoadDensity<double, K>{(const double*)field, d, k}, n, sum_out);
  return reduce(ctx, method, LoadDensity<float, K>{(const float*)field, d, k}, n, sum_out);
}

gc_status free_energy_raw(gc_context* ctx, const gc_grid* g, gc_dtype dt, double kappa,
                          int method, const void* psi, double* sum_out) {
  return density_raw<GC_DENSITY_CAHN_HILLIARD>(ctx, g, dt, method, psi, kappa, 0.0, sum_out);
}

// ---- integrate() without allocation --------------------------------------------------------
//
// gc_integrate_noalloc runs gc_integrate's kernels with their partial sums in a fixed device
// array instead of the context's scratch: sized for the largest grid the library accepts
// (2^31-1 sites: 2^24 pairwise nodes -> 2 x 65536 node sums; the compensated sum needs
// 2048 x 16 B), plus the final scaled result slot.  Every kernel writes each partial before
// any kernel reads it, so nothing needs clearing between calls.  One array per device, shared
// by every call on that device.

constexpr size_t kNoAllocScratchBytes = (size_t(1) << 20) + 4096;

static __device__ __align__(256) unsigned char noalloc_scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices] = {};

// Device address of this device's noalloc_scratch (looked up once per device, with the
// device current; the host-side symbol is not a device address).
gc_status noalloc_scratch_ptr(int device, void** out) {
  if (device < 0 || device >= kMaxDevices)
    return set_error(GC_ERR_UNSUPPORTED, "gridcalc: device ordinal %d beyond %d", device,
                     kMaxDevices);
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[device]) {
    void* p = nullptr;
    GC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch)));
    g_noalloc_ptr[device] = p;
  }
  *out = g_noalloc_ptr[device];
  return GC_OK;
}

__global__ void k_store_scaled(const double* __restrict__ sum, double vol,
                               double* __restrict__ out) {
  out[0] = vol * sum[0];
}

}  // namespace

}  // namespace gc

using namespace gc;

extern "C" {

gc_status gc_integrate_noalloc(gc_context* ctx, const gc_grid* g, gc_dtype dt, int method,
                               const void* in, double* result_dev) {
  GC_TRY(field_checks(ctx, g, dt, in));
  GC_TRY(check_method(method));
  if (!result_dev) return set_error(GC_ERR_INVALID, "gridcalc: null result pointer");
  const size_t n = sites(g);
  const size_t need = (method == GC_PAIRWISE ? pairwise_scratch_bytes(n, nullptr, nullptr)
                                             : comp_scratch_bytes(n));
  if (need > kNoAllocScratchBytes)
    return set_error(GC_ERR_UNSUPPORTED,
                     "gridcalc: gc_integrate_noalloc needs %zu bytes of partial sums for %zu "
                     "sites; its fixed scratch holds %zu",
                     need, n, kNoAllocScratchBytes);
  void* scr = nullptr;
  GC_TRY(noalloc_scratch_ptr(ctx->device, &scr));
  // the sum goes to the array's last 256-byte slot, past every partial (need <= capacity)
  double* fin = (double*)((unsigned char*)scr + kNoAllocScratchBytes - 256);
  if (method == GC_PAIRWISE) {
    if (dt == GC_F64)
      GC_TRY(pairwise_launch(ctx->stream, LoadPlain<double>{(const double*)in}, n, scr, fin));
    else
      GC_TRY(pairwise_launch(ctx->stream, LoadPlain<float>{(const float*)in}, n, scr, fin));
  } else {
    if (dt == GC_F64)
      GC_TRY(compensated_launch(ctx->stream, LoadPlain<double>{(const double*)in}, n, scr, fin));
    else
      GC_TRY(compensated_launch(ctx->stream, LoadPlain<float>{(const float*)in}, n, scr, fin));
  }
  // func/Integrate.hpp:118-119: getCellVolume() * sum, cell volume (hx*hy)*hz
  hipLaunchKernelGGL(k_store_scaled, dim3(1), dim3(1), 0, ctx->stream, fin,
                     g->hx * g->hy * g->hz, result_dev);
  GC_LAUNCHED();
  return GC_OK;
}

gc_status gc_integrate(gc_context* ctx, const gc_grid* g, gc_dtype dt, int method, const void* in,
                       double* result) {
  GC_TRY(field_checks(ctx, g, dt, in));
  GC_TRY(check_method(method));
  if (!result) return set_error(GC_ERR_INVALID, "gridcalc: null result pointer");
  double s = 0.0;
  GC_TRY(integrate_raw(ctx, g, dt, method, in, &s));
  // func/Integrate.hpp:118-119: getCellVolume() * sum, cell volume (hx*hy)*hz
  *result = (g->hx * g->hy * g->hz) * s;
  return GC_OK;
}

gc_status gc_ch_free_energy(gc_context* ctx, const gc_grid* g, gc_dtype dt, double kappa,
                            int method, const void* psi, double* F) {
  GC_TRY(field_checks(ctx, g, dt, psi));
  GC_TRY(check_method(method));
  GC_TRY(check_extent(g, 1));
  if (!F) return set_error(GC_ERR_INVALID, "gridcalc: null result pointer");
  double s = 0.0;
  GC_TRY(free_energy_raw(ctx, g, dt, kappa, method, psi, &s));
  *F = (g->hx * g->hy * g->hz) * s;
  return GC_OK;
}

gc_status gc_gradient_energy(gc_context* ctx, const gc_grid* g, gc_dtype dt, int method,
                             const void* psi, double* E) {
  GC_TRY(field_checks(ctx, g, dt, psi));
  GC_TRY(check_method(method));
  GC_TRY(check_extent(g, 1));
  if (!E) return set_error(GC_ERR_INVALID, "gridcalc: null result pointer");
  double s = 0.0;
  GC_TRY(density_raw<GC_DENSITY_GRADIENT_ENERGY>(ctx, g, dt, method, psi, 0.0, 0.0, &s));
  *E = (g->hx * g->hy * g->hz) * s;
  return GC_OK;
}

gc_status gc_evaluate(gc_context* ctx, const gc_grid* g, gc_dtype dt, const gc_density* f,
                      int method, const void* field, double* result) {
  GC_TRY(field_checks(ctx, g, dt, field));
  GC_TRY(check_method(method));
  if (!f) return set_error(GC_ERR_INVALID, "gridcalc: null density");
  if (!result) return set_error(GC_ERR_INVALID, "gridcalc: null result pointer");
  // the operators func::evaluate materialises, in its order: gradient, Laplacian, biharmonic
  // (Functional.hpp:100-147), each with its own extent check
  switch (f->kind) {
    case GC_DENSITY_VALUE: break;
    case GC_DENSITY_GRADIENT_ENERGY:
    case GC_DENSITY_CAHN_HILLIARD:
    case GC_DENSITY_GINZBURG_LANDAU:
    case GC_DENSITY_LINEAR_ELASTIC: GC_TRY(check_extent(g, 1)); break;
    case GC_DENSITY_PFC:
      GC_TRY(check_extent(g, 1));
      GC_TRY(check_extent(g, 2));
      break;
    default: return set_error(GC_ERR_INVALID, "gridcalc: unknown density kind %d", (int)f->kind);
  }
  double s = 0.0;
  switch (f->kind) {
    case GC_DENSITY_VALUE:
      GC_TRY(density_raw<GC_DENSITY_VALUE>(ctx, g, dt, method, field, f->a, f->b, &s));
      break;
    case GC_DENSITY_GRADIENT_ENERGY:
      GC_TRY(density_raw<GC_DENSITY_GRADIENT_ENERGY>(ctx, g, dt, method, field, f->a, f->b, &s));
      break;
    case GC_DENSITY_CAHN_HILLIARD:
      GC_TRY(density_raw<GC_DENSITY_CAHN_HILLIARD>(ctx, g, dt, method, field, f->a, f->b, &s));
      break;
    case GC_DENSITY_GINZBURG_LANDAU:
      GC_TRY(density_raw<GC_DENSITY_GINZBURG_LANDAU>(ctx, g, dt, method, field, f->a, f->b, &s));
      break;
    case GC_DENSITY_PFC:
      GC_TRY(density_raw<GC_DENSITY_PFC>(ctx, g, dt, method, field, f->a, f->b, &s));
      break;
    default:
      GC_TRY(density_raw<GC_DENSITY_LINEAR_ELASTIC>(ctx, g, dt, method, field, f->a, f->b, &s));
      break;
  }
  // func/Integrate.hpp:118-119: getCellVolume() * sum
  *result = (g->hx * g->hy * g->hz) * s;
  return GC_OK;
}

gc_status gc_first_nonpositive(gc_context* ctx, const gc_grid* g, gc_dtype dt, const void* D,
                               int64_t* first_bad) {
  GC_TRY(field_checks(ctx, g, dt, D));
  if (!first_bad) return set_error(GC_ERR_INVALID, "gridcalc: null result pointer");
  return first_nonpositive(ctx, g, dt, D, first_bad);
}

gc_status gc_max_face_d(gc_context* ctx, const gc_grid* g, gc_dtype dt, const void* D,
                        double* dmax) {
  GC_TRY(field_checks(ctx, g, dt, D));
  if (!dmax) return set_error(GC_ERR_INVALID, "gridcalc: null result pointer");
  return max_face_d(ctx, g, dt, D, dmax);
}

}  // extern "C"
