// gc_ensemble.hip -- batched 2-D Cahn-Hilliard ensembles: many independent periodic grids
// of uneven sizes and parameters advanced together, one kernel per RK4 step.
//
// Temporal blocking.  A work item is one OX x OY tile of one member grid.  The block loads
// the tile with an 8-site periodic halo into LDS and runs all four RK4 stages there: stage
// s evaluates mu on the tile + (9-2s) halo and k_s = M lap(mu) on the tile + (8-2s) halo,
// so the halo shrinks by two per stage and k4 lands exactly on the tile.  Only the new state
// is written back, to the member's second buffer (the two buffers alternate by step
// parity).  Every site value is formed with the same arithmetic as the per-stage pipeline
// (solve/RK4.hpp:122-153 with the Cahn-Hilliard right-hand side,
// examples/common/cahn_hilliard.hpp:41-57); the degenerate z terms of the 2-D Laplacians
// are exactly zero for finite data and are not formed, so results equal the reference's
// arithmetic up to the sign of an exact zero.
//
// Persistent launch: min(#work items, resident blocks) blocks stride over the work list
// (every block exits when the list is exhausted).  gc_ensemble_step replays a captured
// hipGraph of an even number of steps.
#include "gc_common.hpp"
#include "gc_arith.hpp"

#include <algorithm>
#include <map>
#include <vector>

namespace gc {
gc_status integrate_raw(gc_context* ctx, const gc_grid* g, gc_dtype dt, int method,
                        const void* in, double* sum_out);
gc_status free_energy_raw(gc_context* ctx, const gc_grid* g, gc_dtype dt, double kappa,
                          int method, const void* psi, double* sum_out);
}  // namespace gc

namespace {

using namespace gc;

struct EnsMember {
  int nx, ny;
  size_t off[2];  // element offsets of the two state buffers in the pool
  double M, kappa, ihx2, ihy2, hx, hy;
};

struct WorkItem {
  int member, x0, y0, pad;
};

template <class T>
struct Coef4 {
  T half_dt, dt, sixth_dt, third_dt;
};

__device__ __forceinline__ int wrap_mod(int x, int n) {
  x %= n;
  return x < 0 ? x + n : x;
}

// Order-2 2-D Laplacian from an LDS plane of width W (diff/Laplacian.hpp:60-70 order).
template <int W>
__device__ __forceinline__ double lap2d(const double* __restrict__ f, int o, double ihx2,
                                        double ihy2) {
  const double c = f[o];
  double dxx = 0, dyy = 0;
  dxx += f[o - 1];
  dyy += f[o - W];
  dxx += -2.0 * c;
  dyy += -2.0 * c;
  dxx += f[o + 1];
  dyy += f[o + W];
  return dxx * ihx2 + dyy * ihy2;
}

// mu and k = M lap(mu) at LDS offset o.  FP64: the reference's arithmetic without the zero
// z terms; FP32: the library's FP32 form (gc_arith.hpp).
template <int W>
__device__ __forceinline__ double ens_mu(const CHCoef<double>& cf, const double* f, int o) {
  const double l = lap2d<W>(f, o, cf.ihx2, cf.ihy2);
  const double p = f[o];
  return -p + p * p * p - cf.kappa * l;
}
template <int W>
__device__ __forceinline__ double ens_k(const CHCoef<double>& cf, const double* f, int o) {
  return lap2d<W>(f, o, cf.ihx2, cf.ihy2) * cf.M;
}
template <int W>
__device__ __forceinline__ float ens_mu(const CHCoef<float>& cf, const float* f, int o) {
  return ch_mu2(cf, f[o], f[o - 1], f[o + 1], f[o - W], f[o + W]);
}
template <int W>
__device__ __forceinline__ float ens_k(const CHCoef<float>& cf, const float* f, int o) {
  return ch_k2(cf, f[o], f[o - 1], f[o + 1], f[o - W], f[o + W]);
}

template <int OX, int OY>
struct EnsCfg {
  static constexpr int H = 8, W = OX + 2 * H, HH = OY + 2 * H, N = W * HH, NT = 256;
};

template <class T, int OX, int OY>
__global__ __launch_bounds__(256) void k_ens_step(const EnsMember* __restrict__ mem,
                                                  const WorkItem* __restrict__ work, int nwork,
                                                  T* __restrict__ pool, int parity, Coef4<T> co) {
  using C = EnsCfg<OX, OY>;
  constexpr int W = C::W, H = C::H;
  __shared__ T P[C::N];   // psi on the tile + 8 halo
  __shared__ T Y[C::N];   // stage input y_s (s >= 2)
  __shared__ T MU[C::N];  // mu of the current stage
  __shared__ T K1[OX * OY], K23[OX * OY];
  const int t = threadIdx.x;
  // The next item's tile + halo is fetched into registers while the current item runs its
  // stages, so an item starts from a tile already on chip.
  constexpr int NL = (C::N + C::NT - 1) / C::NT;
  T pre[NL];
  auto fetch = [&](int w) {
    const WorkItem wi = work[w];
    const EnsMember m = mem[wi.member];
    const T* __restrict__ src = pool + (parity ? m.off[1] : m.off[0]);
#pragma unroll
    for (int s = 0; s < NL; ++s) {
      const int idx = t + s * C::NT;
      if (idx < C::N) {
        const int px = idx % W, py = idx / W;
        pre[s] = src[(size_t)wrap_mod(wi.x0 + px - H, m.nx) +
                     (size_t)m.nx * (size_t)wrap_mod(wi.y0 + py - H, m.ny)];
      }
    }
  };
  if ((int)blockIdx.x < nwork) fetch(blockIdx.x);

  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const WorkItem wi = work[w];
    const EnsMember m = mem[wi.member];
    // selects, not m.off[parity]: a runtime index into the copied struct would place the
    // whole struct in per-lane scratch memory (80 bytes a lane, written back through L2)
    const size_t off_out = parity ? m.off[0] : m.off[1];
    T* __restrict__ dst = pool + off_out;
    const CHCoef<T> cf = make_ch_coef<T>(T(m.M), T(m.kappa), T(m.ihx2), T(m.ihy2), T(0));
    __syncthreads();  // the previous item is done with LDS
#pragma unroll
    for (int s = 0; s < NL; ++s) {
      const int idx = t + s * C::NT;
      if (idx < C::N) P[idx] = pre[s];
    }
    if (w + (int)gridDim.x < nwork) fetch(w + (int)gridDim.x);
    __syncthreads();
#pragma unroll
    for (int s = 1; s <= 4; ++s) {
      const T* __restrict__ in = (s == 1) ? P : Y;
      {  // mu on the tile + rm halo
        const int rm = 9 - 2 * s, wm = OX + 2 * rm, nm = wm * (OY + 2 * rm);
        for (int idx = t; idx < nm; idx += C::NT) {
          const int mx = idx % wm, my = idx / wm;
          const int o = (my + H - rm) * W + (mx + H - rm);
          MU[o] = ens_mu<W>(cf, in, o);
        }
      }
      __syncthreads();
      {  // k_s on the tile + rk halo, then the stage update
        const int rk = 8 - 2 * s, wk = OX + 2 * rk, nk = wk * (OY + 2 * rk);
        for (int idx = t; idx < nk; idx += C::NT) {
          const int kx = idx % wk, ky = idx / wk;
          const int o = (ky + H - rk) * W + (kx + H - rk);
          const T kv = ens_k<W>(cf, MU, o);
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
            const int gx = wi.x0 + ix, gy = wi.y0 + iy;
            if (gx < m.nx && gy < m.ny)
              dst[(size_t)gx + (size_t)m.nx * (size_t)gy] =
                  rk4_final(P[o], K1[ii], K23[ii], kv, co.sixth_dt, co.third_dt);
          }
        }
      }
      __syncthreads();
    }
  }
}

struct TileShape {
  int ox, oy;
};
constexpr TileShape kEnsF32[] = {{32, 16}, {32, 32}, {64, 16}, {64, 32}, {64, 64}};
constexpr TileShape kEnsF64[] = {{32, 16}, {32, 32}, {64, 16}, {64, 32}};
constexpr int kEnsNF32 = 5, kEnsNF64 = 4;

}  // namespace

struct gc_ensemble {
  gc_context* ctx = nullptr;
  gc_dtype dt = GC_F64;
  double step = 0.0;
  std::vector<EnsMember> members;
  std::vector<size_t> sizes;
  size_t total = 0;
  void* pool = nullptr;
  EnsMember* d_mem = nullptr;
  WorkItem* d_work = nullptr;
  int nwork = 0;
  int cfg = -1;  // tile shape id
  int parity = 0;
  int max_blocks = 0;  // resident blocks for the current shape
  std::map<int, hipGraphExec_t> graphs;  // steps per graph -> executable graph
  hipGraph_t graph_tmp = nullptr;
};

namespace {

int shape_count(gc_dtype dt) { return dt == GC_F64 ? kEnsNF64 : kEnsNF32; }
TileShape shape_of(gc_dtype dt, int id) { return dt == GC_F64 ? kEnsF64[id] : kEnsF32[id]; }

template <class T, int OX, int OY>
gc_status occupancy(int* blocks) {
  int per_cu = 0, dev = 0, cus = 0;
  GC_HIP(hipGetDevice(&dev));
  GC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  GC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_ens_step<T, OX, OY>, 256, 0));
  *blocks = std::max(1, per_cu) * cus;
  return GC_OK;
}

template <class T, int OX, int OY>
gc_status launch_ens(gc_ensemble* e, int parity) {
  const int blocks = std::min(e->nwork, e->max_blocks);
  const Coef4<T> co{T(0.5 * e->step), T(e->step), T(e->step / 6.0), T(e->step / 3.0)};
  hipLaunchKernelGGL((k_ens_step<T, OX, OY>), dim3((unsigned)blocks), dim3(256), 0,
                     e->ctx->stream, e->d_mem, e->d_work, e->nwork, (T*)e->pool, parity, co);
  GC_LAUNCHED();
  return GC_OK;
}

gc_status dispatch(gc_ensemble* e, int parity, bool occupancy_only = false) {
#define GC_ENS(T, OX, OY)                                     \
  return occupancy_only ? occupancy<T, OX, OY>(&e->max_blocks) : launch_ens<T, OX, OY>(e, parity)
  if (e->dt == GC_F32) {
    switch (e->cfg) {
      case 0: GC_ENS(float, 32, 16);
      case 1: GC_ENS(float, 32, 32);
      case 2: GC_ENS(float, 64, 16);
      case 3: GC_ENS(float, 64, 32);
      default: GC_ENS(float, 64, 64);
    }
  }
  switch (e->cfg) {
    case 0: GC_ENS(double, 32, 16);
    case 1: GC_ENS(double, 32, 32);
    case 2: GC_ENS(double, 64, 16);
    default: GC_ENS(double, 64, 32);
  }
#undef GC_ENS
}

void drop_graphs(gc_ensemble* e) {
  for (auto& kv : e->graphs) (void)hipGraphExecDestroy(kv.second);
  e->graphs.clear();
}

// Builds the work list for tile shape `id` (members in order, tiles row-major).
gc_status set_shape(gc_ensemble* e, int id) {
  const TileShape s = shape_of(e->dt, id);
  std::vector<WorkItem> work;
  for (size_t i = 0; i < e->members.size(); ++i) {
    const EnsMember& m = e->members[i];
    for (int y0 = 0; y0 < m.ny; y0 += s.oy)
      for (int x0 = 0; x0 < m.nx; x0 += s.ox) work.push_back(WorkItem{(int)i, x0, y0, 0});
  }
  if (e->d_work) GC_HIP(hipFree(e->d_work));
  e->d_work = nullptr;
  GC_HIP(hipMalloc(&e->d_work, work.size() * sizeof(WorkItem)));
  GC_HIP(hipMemcpyAsync(e->d_work, work.data(), work.size() * sizeof(WorkItem),
                        hipMemcpyHostToDevice, e->ctx->stream));
  GC_HIP(hipStreamSynchronize(e->ctx->stream));
  e->nwork = (int)work.size();
  e->cfg = id;
  drop_graphs(e);
  return dispatch(e, 0, true);
}

gc_status direct_steps(gc_ensemble* e, int n) {
  for (int s = 0; s < n; ++s) {
    GC_TRY(dispatch(e, e->parity));
    e->parity ^= 1;
  }
  return GC_OK;
}

// Replays a graph of `chunk` (even) steps captured from parity 0.
gc_status graph_steps(gc_ensemble* e, int n, int chunk) {
  if (e->parity) {
    GC_TRY(direct_steps(e, 1));
    --n;
  }
  const int reps = n / chunk, rest = n % chunk;
  if (reps > 0) {
    auto it = e->graphs.find(chunk);
    hipGraphExec_t exec = nullptr;
    if (it == e->graphs.end()) {
      GC_HIP(hipStreamBeginCapture(e->ctx->stream, hipStreamCaptureModeThreadLocal));
      gc_status st = GC_OK;
      for (int s = 0; s < chunk && st == GC_OK; ++s) st = dispatch(e, s & 1);
      hipGraph_t g = nullptr;
      const hipError_t ce = hipStreamEndCapture(e->ctx->stream, &g);
      if (st != GC_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
      }
      if (ce != hipSuccess)
        return set_error(GC_ERR_DEVICE, "hipStreamEndCapture failed: %s", hipGetErrorString(ce));
      const hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess)
        return set_error(GC_ERR_DEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(ie));
      e->graphs[chunk] = exec;
    } else {
      exec = it->second;
    }
    for (int r = 0; r < reps; ++r) GC_HIP(hipGraphLaunch(exec, e->ctx->stream));
  }
  return direct_steps(e, rest);
}

gc_status check_ens(gc_ensemble* e) {
  if (!e) return set_error(GC_ERR_INVALID, "gridcalc: null ensemble");
  return check_ctx(e->ctx);
}

gc_status check_member_index(gc_ensemble* e, int i) {
  if (i < 0 || i >= (int)e->members.size())
    return set_error(GC_ERR_INVALID, "gridcalc: ensemble member %d out of range [0, %zu)", i,
                     e->members.size());
  return GC_OK;
}

}  // namespace

extern "C" {

gc_status gc_ensemble_create(gc_context* ctx, gc_dtype dt, int n, const gc_ensemble_member* mem,
                             double step, gc_ensemble** out) {
  GC_TRY(check_ctx(ctx));
  GC_TRY(check_dtype(dt));
  if (!out || !mem || n <= 0) return set_error(GC_ERR_INVALID, "gridcalc: bad ensemble arguments");
  if (!(step >= 0.0)) return set_error(GC_ERR_INVALID, "gridcalc: ensemble dt must be >= 0");
  *out = nullptr;
  gc_ensemble* e = new gc_ensemble();
  e->ctx = ctx;
  e->dt = dt;
  e->step = step;
  const size_t w = dsize(dt);
  for (int i = 0; i < n; ++i) {
    const gc_ensemble_member& m = mem[i];
    const gc_grid g{m.nx, m.ny, 1, m.hx, m.hy, 1.0};
    gc_status st = check_grid(&g);
    if (st == GC_OK) st = check_extent(&g, 1);
    if (st != GC_OK) {
      delete e;
      return st;
    }
    EnsMember em{};
    em.nx = m.nx;
    em.ny = m.ny;
    em.M = m.M;
    em.kappa = m.kappa;
    em.ihx2 = 1.0 / (m.hx * m.hx);
    em.ihy2 = 1.0 / (m.hy * m.hy);
    em.hx = m.hx;
    em.hy = m.hy;
    const size_t sites = (size_t)m.nx * (size_t)m.ny;
    // 256-byte aligned member buffers
    const size_t pad = (256 / w);
    const size_t padded = (sites + pad - 1) / pad * pad;
    em.off[0] = e->total;
    em.off[1] = e->total + padded;
    e->total += 2 * padded;
    e->members.push_back(em);
    e->sizes.push_back(sites);
  }
  hipError_t he = hipMalloc(&e->pool, e->total * w);
  if (he == hipSuccess) he = hipMalloc(&e->d_mem, e->members.size() * sizeof(EnsMember));
  if (he != hipSuccess) {
    (void)hipGetLastError();
    gc_ensemble_destroy(e);
    return set_error(GC_ERR_NOMEM, "gridcalc: ensemble allocation failed");
  }
  gc_status st = GC_OK;
  if (hipMemcpyAsync(e->d_mem, e->members.data(), e->members.size() * sizeof(EnsMember),
                     hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
      hipMemsetAsync(e->pool, 0, e->total * w, ctx->stream) != hipSuccess)
    st = set_error(GC_ERR_DEVICE, "gridcalc: ensemble set-up copy failed");
  if (st == GC_OK) st = set_shape(e, 1);
  if (st != GC_OK) {
    gc_ensemble_destroy(e);
    return st;
  }
  *out = e;
  return GC_OK;
}

gc_status gc_ensemble_destroy(gc_ensemble* e) {
  if (!e) return GC_OK;
  if (e->ctx) {
    (void)hipSetDevice(e->ctx->device);
    (void)hipStreamSynchronize(e->ctx->stream);
  }
  drop_graphs(e);
  if (e->pool) (void)hipFree(e->pool);
  if (e->d_mem) (void)hipFree(e->d_mem);
  if (e->d_work) (void)hipFree(e->d_work);
  delete e;
  return GC_OK;
}

gc_status gc_ensemble_set_state(gc_ensemble* e, int i, const void* host) {
  GC_TRY(check_ens(e));
  GC_TRY(check_member_index(e, i));
  if (!host) return set_error(GC_ERR_INVALID, "gridcalc: null host pointer");
  const size_t w = dsize(e->dt);
  char* dev = (char*)e->pool + e->members[i].off[e->parity] * w;
  GC_HIP(hipMemcpyAsync(dev, host, e->sizes[i] * w, hipMemcpyHostToDevice, e->ctx->stream));
  GC_HIP(hipStreamSynchronize(e->ctx->stream));
  return GC_OK;
}

gc_status gc_ensemble_get_state(gc_ensemble* e, int i, void* host) {
  GC_TRY(check_ens(e));
  GC_TRY(check_member_index(e, i));
  if (!host) return set_error(GC_ERR_INVALID, "gridcalc: null host pointer");
  const size_t w = dsize(e->dt);
  const char* dev = (const char*)e->pool + e->members[i].off[e->parity] * w;
  GC_HIP(hipMemcpyAsync(host, dev, e->sizes[i] * w, hipMemcpyDeviceToHost, e->ctx->stream));
  GC_HIP(hipStreamSynchronize(e->ctx->stream));
  return GC_OK;
}

gc_status gc_ensemble_state_ptr(gc_ensemble* e, int i, void** dev) {
  GC_TRY(check_ens(e));
  GC_TRY(check_member_index(e, i));
  if (!dev) return set_error(GC_ERR_INVALID, "gridcalc: null pointer");
  *dev = (char*)e->pool + e->members[i].off[e->parity] * dsize(e->dt);
  return GC_OK;
}

gc_status gc_ensemble_step(gc_ensemble* e, int n_steps, int use_graph) {
  GC_TRY(check_ens(e));
  if (n_steps < 0)
    return set_error(GC_ERR_INVALID, "gridcalc: ensemble n_steps must be >= 0; got %d", n_steps);
  if (!use_graph) return direct_steps(e, n_steps);
  constexpr int kChunk = 20;  // steps per captured graph (even)
  return graph_steps(e, n_steps, kChunk);
}

gc_status gc_ensemble_sites(gc_ensemble* e, size_t* sites) {
  GC_TRY(check_ens(e));
  if (!sites) return set_error(GC_ERR_INVALID, "gridcalc: null pointer");
  size_t s = 0;
  for (size_t n : e->sizes) s += n;
  *sites = s;
  return GC_OK;
}

gc_status gc_ensemble_config_shape(gc_ensemble* e, int id, int* ox, int* oy) {
  GC_TRY(check_ens(e));
  if (id < 0 || id >= shape_count(e->dt) || !ox || !oy)
    return set_error(GC_ERR_INVALID, "gridcalc: ensemble tile id %d out of range", id);
  const TileShape s = shape_of(e->dt, id);
  *ox = s.ox;
  *oy = s.oy;
  return GC_OK;
}

gc_status gc_ensemble_set_config(gc_ensemble* e, int id) {
  GC_TRY(check_ens(e));
  if (id < 0 || id >= shape_count(e->dt))
    return set_error(GC_ERR_INVALID, "gridcalc: ensemble tile id %d out of range", id);
  return set_shape(e, id);
}

gc_status gc_ensemble_tune(gc_ensemble* e, int* best, float* ms_out) {
  GC_TRY(check_ens(e));
  if (!best) return set_error(GC_ERR_INVALID, "gridcalc: null best");
  // tune on a copy of the state so the members' data is unchanged
  const size_t bytes = e->total * dsize(e->dt);
  void* save = nullptr;
  GC_TRY(scratch(e->ctx, bytes, &save));
  GC_HIP(hipMemcpyAsync(save, e->pool, bytes, hipMemcpyDeviceToDevice, e->ctx->stream));
  const int parity0 = e->parity;
  int best_id = -1;
  float best_ms = 0.0f;
  for (int id = 0; id < shape_count(e->dt); ++id) {
    GC_TRY(set_shape(e, id));
    GC_TRY(direct_steps(e, 2));
    GC_HIP(hipEventRecord(e->ctx->ev_start, e->ctx->stream));
    GC_TRY(direct_steps(e, 10));
    GC_HIP(hipEventRecord(e->ctx->ev_stop, e->ctx->stream));
    GC_HIP(hipEventSynchronize(e->ctx->ev_stop));
    float ms = 0.0f;
    GC_HIP(hipEventElapsedTime(&ms, e->ctx->ev_start, e->ctx->ev_stop));
    ms /= 10.0f;
    if (ms_out) ms_out[id] = ms;
    if (best_id < 0 || ms < best_ms) {
      best_ms = ms;
      best_id = id;
    }
  }
  GC_HIP(hipMemcpyAsync(e->pool, save, bytes, hipMemcpyDeviceToDevice, e->ctx->stream));
  GC_HIP(hipStreamSynchronize(e->ctx->stream));
  e->parity = parity0;
  GC_TRY(set_shape(e, best_id));
  *best = best_id;
  return GC_OK;
}

gc_status gc_ensemble_diagnostics(gc_ensemble* e, int method, double* mass, double* F) {
  GC_TRY(check_ens(e));
  if (method != GC_PAIRWISE && method != GC_KAHAN)
    return set_error(GC_ERR_INVALID, "func::integrate: Tag must be func::Pairwise or func::Kahan");
  const size_t w = dsize(e->dt);
  for (size_t i = 0; i < e->members.size(); ++i) {
    const EnsMember& m = e->members[i];
    const gc_grid g{m.nx, m.ny, 1, m.hx, m.hy, 1.0};
    const void* p = (const char*)e->pool + m.off[e->parity] * w;
    double s = 0.0;
    if (mass) {
      GC_TRY(integrate_raw(e->ctx, &g, e->dt, method, p, &s));
      mass[i] = (g.hx * g.hy * g.hz) * s;
    }
    if (F) {
      GC_TRY(free_energy_raw(e->ctx, &g, e->dt, m.kappa, method, p, &s));
      F[i] = (g.hx * g.hy * g.hz) * s;
    }
  }
  return GC_OK;
}

}  // extern "C"
