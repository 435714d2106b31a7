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
// Loads are coalesced: 16 lanes fetch 16 consecutive values of one node into LDS,
// then every lane consumes its own node's values in order from LDS.
#include "gc_common.hpp"

#include <climits>
#include <cstring>
#include <mutex>

namespace gc {
namespace {

constexpr int kChunk = 16;          // values per node per LDS round
constexpr int kPad = kChunk + 1;    // padded LDS row (conflict-free b64 reads)
constexpr int kWaves = 4;           // waves per block in the leaf kernel

// ---- value loaders ------------------------------------------------------------------

template <class T>
struct LoadPlain {
  const T* x;
  __device__ __forceinline__ double operator()(size_t c) const { return (double)x[c]; }
};

// Gradient (Order 2, diff/Gradient.hpp:59-69) of psi at flat index c, in double.
template <class T>
__device__ __forceinline__ void grad2_at(const T* p, const Dims& d, size_t c, double ihx,
                                         double ihy, double ihz, double& gx, double& gy,
                                         double& gz) {
  const int i = (int)(c % (size_t)d.nx);
  const size_t r = c / (size_t)d.nx;
  const int j = (int)(r % (size_t)d.ny);
  const int k = (int)(r / (size_t)d.ny);
  const size_t row = (size_t)j * d.nx, pl = (size_t)k * d.plane;
  double dx = 0.0, dy = 0.0, dz = 0.0;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const int o = m - 1;
    const double w = D1<2>::w(m);
    dx += w * (double)p[pl + row + wrap(i + o, d.nx)];
    dy += w * (double)p[pl + (size_t)wrap(j + o, d.ny) * d.nx + i];
    dz += w * (double)p[(size_t)wrap(k + o, d.nz) * d.plane + row + i];
  }
  gx = dx * ihx;
  gy = dy * ihy;
  gz = dz * ihz;
}

// examples/common/cahn_hilliard.hpp:65-67:
//   -0.5*p*p + 0.25*p*p*p*p + 0.5*kappa*g.squaredNorm(),  |g|^2 = (gx^2 + gy^2) + gz^2
template <class T>
struct LoadCHDensity {
  const T* p;
  Dims d;
  double ihx, ihy, ihz, kappa;
  __device__ __forceinline__ double operator()(size_t c) const {
    double gx, gy, gz;
    grad2_at(p, d, c, ihx, ihy, ihz, gx, gy, gz);
    const double v = (double)p[c];
    const double sq = gx * gx + gy * gy + gz * gz;
    return -0.5 * v * v + 0.25 * v * v * v * v + 0.5 * kappa * sq;
  }
};

// examples/common/cahn_hilliard.hpp:74-76: g.squaredNorm()
template <class T>
struct LoadGradSq {
  const T* p;
  Dims d;
  double ihx, ihy, ihz;
  __device__ __forceinline__ double operator()(size_t c) const {
    double gx, gy, gz;
    grad2_at(p, d, c, ihx, ihy, ihz, gx, gy, gz);
    return gx * gx + gy * gy + gz * gz;
  }
};

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

// One lane per depth-`depth` node of the recursion over [0, n); every node holds at
// most `maxsize` (= ceil(n / 2^depth) <= 128) values.
template <class L>
__global__ __launch_bounds__(64 * kWaves) void k_pairwise_leaves(L ld, size_t n, int depth,
                                                                 unsigned P, int maxsize,
                                                                 double* __restrict__ partial) {
  __shared__ double buf[kWaves][64 * kPad];
  __shared__ size_t noff[kWaves][64];
  __shared__ int nsize[kWaves][64];
  __shared__ double sh[kWaves];
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const unsigned q = blockIdx.x * (64u * kWaves) + threadIdx.x;

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
  noff[wave][lane] = off;
  nsize[wave][lane] = (int)size;
  __syncthreads();

  const int split = size > 64 ? (int)(size >> 1) : INT_MAX;  // leaf boundary inside the node
  double s0 = 0.0, s1 = 0.0;
  double* row = &buf[wave][0];
  const int e = (int)(lane & (kChunk - 1));
  for (int base = 0; base < maxsize; base += kChunk) {
    // cooperative load: 16 lanes fetch 16 consecutive values of one node, 4 nodes per pass
#pragma unroll 4
    for (int pass = 0; pass < kChunk; ++pass) {
      const int node = pass * (64 / kChunk) + (int)(lane / kChunk);
      const int idx = base + e;
      if (idx < nsize[wave][node]) row[node * kPad + e] = ld(noff[wave][node] + (size_t)idx);
    }
    __syncthreads();
    const int lim = min(kChunk, (int)size - base);
    for (int t = 0; t < lim; ++t) {
      const double v = row[lane * kPad + t];
      if (base + t < split) s0 += v;
      else s1 += v;
    }
    __syncthreads();
  }
  double s = (split == INT_MAX) ? s0 : s0 + s1;
  const unsigned cnt = P < 64u * kWaves ? P : 64u * kWaves;
  s = block_tree_sum(s, cnt, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Perfect-tree sum of consecutive groups of min(cnt, 256) values.
__global__ __launch_bounds__(256) void k_tree_pass(const double* __restrict__ in, unsigned cnt,
                                                   double* __restrict__ out) {
  __shared__ double sh[4];
  const unsigned g = cnt < 256u ? cnt : 256u;
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
  const unsigned per_block = 64u * kWaves;
  const unsigned nb = (P + per_block - 1) / per_block;
  if (nb_out) *nb_out = nb;
  if (depth_out) *depth_out = depth;
  return 2 * (size_t)nb * sizeof(double) + 256;
}

// Launches the pairwise reduction on `stream` with scratch `scr` (at least
// pairwise_scratch_bytes(n) bytes); *final points at the device double holding the sum once
// the launches complete.
template <class L>
gc_status pairwise_launch(hipStream_t stream, L ld, size_t n, void* scr, const double** final) {
  unsigned nb = 0;
  int depth = 0;
  pairwise_scratch_bytes(n, &nb, &depth);
  const unsigned P = 1u << depth;
  const unsigned per_block = 64u * kWaves;
  double* a = (double*)scr;
  double* b = a + nb;
  const int maxsize = (int)((n + ((size_t)1 << depth) - 1) >> depth);
  hipLaunchKernelGGL((k_pairwise_leaves<L>), dim3(nb), dim3(per_block), 0, stream, ld, n, depth,
                     P, maxsize, a);
  GC_LAUNCHED();
  unsigned cnt = nb;
  while (cnt > 1) {
    const unsigned g = cnt < 256u ? cnt : 256u;
    const unsigned blocks = cnt / g;
    hipLaunchKernelGGL(k_tree_pass, dim3(blocks), dim3(256), 0, stream, a, cnt, b);
    GC_LAUNCHED();
    cnt = blocks;
    double* t = a;
    a = b;
    b = t;
  }
  *final = a;
  return GC_OK;
}

template <class L>
gc_status pairwise(gc_context* ctx, L ld, size_t n, double* result) {
  void* scr = nullptr;
  GC_TRY(scratch(ctx, pairwise_scratch_bytes(n, nullptr, nullptr), &scr));
  const double* fin = nullptr;
  GC_TRY(pairwise_launch(ctx->stream, ld, n, scr, &fin));
  GC_HIP(hipMemcpyAsync(ctx->host_result, fin, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
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

template <class L>
__global__ __launch_bounds__(256) void k_comp_partial(L ld, size_t n, Comp* __restrict__ partial) {
  __shared__ Comp sh[4];
  const size_t G = (size_t)gridDim.x * 256u;
  Comp a{0.0, 0.0};
  for (size_t c = (size_t)blockIdx.x * 256u + threadIdx.x; c < n; c += G) neu_add(a, ld(c));
  a = block_comp_sum(a, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = a;
}

__global__ __launch_bounds__(256) void k_comp_final(const Comp* __restrict__ partial, unsigned nb,
                                                    double* __restrict__ out) {
  __shared__ Comp sh[4];
  Comp a{0.0, 0.0};
  for (unsigned b = threadIdx.x; b < nb; b += 256u) a = neu_merge(a, partial[b]);
  a = block_comp_sum(a, sh);
  if (threadIdx.x == 0) out[0] = a.s + a.c;
}

inline unsigned comp_blocks(size_t n) {
  const size_t want = (n + 255) / 256;
  return (unsigned)(want < 2048 ? (want == 0 ? 1 : want) : 2048);
}

// Scratch of the compensated reduction: one (sum, compensation) pair per block + the result.
inline size_t comp_scratch_bytes(size_t n) { return (size_t)comp_blocks(n) * sizeof(Comp) + 256; }

template <class L>
gc_status compensated_launch(hipStream_t stream, L ld, size_t n, void* scr, const double** final) {
  const unsigned nb = comp_blocks(n);
  Comp* part = (Comp*)scr;
  double* out = (double*)(part + nb);
  hipLaunchKernelGGL((k_comp_partial<L>), dim3(nb), dim3(256), 0, stream, ld, n, part);
  GC_LAUNCHED();
  hipLaunchKernelGGL(k_comp_final, dim3(1), dim3(256), 0, stream, part, nb, out);
  GC_LAUNCHED();
  *final = out;
  return GC_OK;
}

template <class L>
gc_status compensated(gc_context* ctx, L ld, size_t n, double* result) {
  void* scr = nullptr;
  GC_TRY(scratch(ctx, comp_scratch_bytes(n), &scr));
  const double* fin = nullptr;
  GC_TRY(compensated_launch(ctx->stream, ld, n, scr, &fin));
  GC_HIP(hipMemcpyAsync(ctx->host_result, fin, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
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

gc_status free_energy_raw(gc_context* ctx, const gc_grid* g, gc_dtype dt, double kappa,
                          int method, const void* psi, double* sum_out) {
  const size_t n = sites(g);
  const Dims d = dims_of(g);
  const double ihx = 1.0 / g->hx, ihy = 1.0 / g->hy, ihz = 1.0 / g->hz;
  if (dt == GC_F64)
    return reduce(ctx, method, LoadCHDensity<double>{(const double*)psi, d, ihx, ihy, ihz, kappa},
                  n, sum_out);
  return reduce(ctx, method, LoadCHDensity<float>{(const float*)psi, d, ihx, ihy, ihz, kappa}, n,
                sum_out);
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
  const double* fin = nullptr;
  if (method == GC_PAIRWISE) {
    if (dt == GC_F64)
      GC_TRY(pairwise_launch(ctx->stream, LoadPlain<double>{(const double*)in}, n, scr, &fin));
    else
      GC_TRY(pairwise_launch(ctx->stream, LoadPlain<float>{(const float*)in}, n, scr, &fin));
  } else {
    if (dt == GC_F64)
      GC_TRY(compensated_launch(ctx->stream, LoadPlain<double>{(const double*)in}, n, scr, &fin));
    else
      GC_TRY(compensated_launch(ctx->stream, LoadPlain<float>{(const float*)in}, n, scr, &fin));
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
  const size_t n = sites(g);
  const Dims d = dims_of(g);
  const double ihx = 1.0 / g->hx, ihy = 1.0 / g->hy, ihz = 1.0 / g->hz;
  double s = 0.0;
  if (dt == GC_F64)
    GC_TRY(reduce(ctx, method, LoadGradSq<double>{(const double*)psi, d, ihx, ihy, ihz}, n, &s));
  else
    GC_TRY(reduce(ctx, method, LoadGradSq<float>{(const float*)psi, d, ihx, ihy, ihz}, n, &s));
  *E = (g->hx * g->hy * g->hz) * s;
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
