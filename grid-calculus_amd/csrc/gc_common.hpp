// gc_common.hpp -- shared pieces of libgridcalc_hip: error state, the context
// (stream + scratch), argument validation and the device-side stencil tables.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>

#include "gridcalc_hip.h"

namespace gc {

// ---- error state ---------------------------------------------------------------
gc_status set_error(gc_status s, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

#define GC_HIP(expr)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return ::gc::set_error(GC_ERR_DEVICE, "%s failed: %s (%s:%d)", #expr,             \
                             hipGetErrorString(e_), __FILE__, __LINE__);                 \
  } while (0)

#define GC_TRY(expr)                  \
  do {                                \
    gc_status s_ = (expr);            \
    if (s_ != GC_OK) return s_;       \
  } while (0)

// Checks the launch that was just issued.
#define GC_LAUNCHED() GC_HIP(hipGetLastError())

}  // namespace gc

// ---- context -------------------------------------------------------------------
struct gc_context {
  hipStream_t stream = nullptr;
  bool owns_stream = false;
  int device = 0;
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  double* host_result = nullptr;      // pinned, small: reduction results
  double* host_result_dev = nullptr;  // its device address (final reduction kernels store here)
};

namespace gc {

// Grows the context scratch to at least `bytes` (contents are not preserved).
gc_status scratch(gc_context* ctx, size_t bytes, void** out);

gc_status check_ctx(gc_context* ctx);
// Validates a grid (extents > 0, cell sizes > 0, at most 2^31-1 sites).
gc_status check_grid(const gc_grid* g);
// requireAxisExtent for every axis (diff/detail/PreconditionAxisExtent.hpp:64-79).
gc_status check_extent(const gc_grid* g, int radius);
// requireAxisExtent with a radius per axis (x, y, z checked in that order).
gc_status check_extent3(const gc_grid* g, int rx, int ry, int rz);
gc_status check_dtype(gc_dtype dt);

inline size_t sites(const gc_grid* g) {
  return (size_t)g->nx * (size_t)g->ny * (size_t)g->nz;
}
inline size_t dsize(gc_dtype dt) { return dt == GC_F64 ? 8 : 4; }

// ---- device helpers ------------------------------------------------------------

// Periodic wrap for any index.
__device__ __forceinline__ int wrap_any(int x, int n) {
  x %= n;
  return x < 0 ? x + n : x;
}

// Periodic wrap for an index at most one extent outside [0, n) -- which is all a
// stencil of radius r reaches once n >= 2r+1 -- plus the degenerate n == 1 axis.
__device__ __forceinline__ int wrap(int i, int n) {
  if (n == 1) return 0;
  i = i < 0 ? i + n : i;
  return i >= n ? i - n : i;
}

// Stencil tables (stencil/CentralDifference.hpp:41-74, FirstDerivative.hpp:51-86,
// FourthDerivative.hpp:49-83).  Offsets run -R..R in order; weights are the exact
// rationals of the reference, evaluated in double.
template <int O> struct D2;
template <> struct D2<2> {
  static constexpr int R = 1, N = 3;
  __host__ __device__ static constexpr double w(int m) { return m == 1 ? -2.0 : 1.0; }
};
template <> struct D2<4> {
  static constexpr int R = 2, N = 5;
  __host__ __device__ static constexpr double w(int m) {
    return (m == 0 || m == 4) ? -1.0 / 12.0 : (m == 1 || m == 3) ? 4.0 / 3.0 : -5.0 / 2.0;
  }
};
template <int O> struct D1;
template <> struct D1<2> {
  static constexpr int R = 1, N = 3;
  __host__ __device__ static constexpr double w(int m) {
    return m == 0 ? -0.5 : m == 1 ? 0.0 : 0.5;
  }
};
template <> struct D1<4> {
  static constexpr int R = 2, N = 5;
  __host__ __device__ static constexpr double w(int m) {
    return m == 0 ? 1.0 / 12.0 : m == 1 ? -2.0 / 3.0 : m == 2 ? 0.0 : m == 3 ? 2.0 / 3.0
                                                                          : -1.0 / 12.0;
  }
};
template <int O> struct D4;
template <> struct D4<2> {
  static constexpr int R = 2, N = 5;
  __host__ __device__ static constexpr double w(int m) {
    return (m == 0 || m == 4) ? 1.0 : (m == 1 || m == 3) ? -4.0 : 6.0;
  }
};
template <> struct D4<4> {
  static constexpr int R = 3, N = 7;
  __host__ __device__ static constexpr double w(int m) {
    return (m == 0 || m == 6) ? -1.0 / 6.0
           : (m == 1 || m == 5) ? 2.0
           : (m == 2 || m == 4) ? -13.0 / 2.0
                                : 28.0 / 3.0;
  }
};

// stencil/ThirdDerivative.hpp: offsets -R..R
template <int O> struct D3;
template <> struct D3<2> {
  static constexpr int R = 2, N = 5;
  __host__ __device__ static constexpr double w(int m) {
    return m == 0 ? -0.5 : m == 1 ? 1.0 : m == 2 ? 0.0 : m == 3 ? -1.0 : 0.5;
  }
};
template <> struct D3<4> {
  static constexpr int R = 3, N = 7;
  __host__ __device__ static constexpr double w(int m) {
    return m == 0 ? 1.0 / 8.0 : m == 1 ? -1.0 : m == 2 ? 13.0 / 8.0 : m == 3 ? 0.0
           : m == 4 ? -13.0 / 8.0 : m == 5 ? 1.0 : -1.0 / 8.0;
  }
};

// Site geometry passed to kernels.
struct Dims {
  int nx, ny, nz;
  int plane;  // nx*ny
};
inline Dims dims_of(const gc_grid* g) { return Dims{g->nx, g->ny, g->nz, g->nx * g->ny}; }

// XCD-aware block order: the hardware hands consecutive workgroups to the 8 XCDs
// round-robin; remapping the linear id gives each XCD one contiguous range of
// tiles, so halo rows shared by neighbouring tiles stay in that XCD's L2.
__device__ __forceinline__ unsigned xcd_remap(unsigned bid, unsigned nblocks) {
  constexpr unsigned kXcd = 8;
  const unsigned per = nblocks / kXcd;      // blocks per XCD in the full rounds
  const unsigned full = per * kXcd;
  if (bid >= full) return bid;              // ragged tail keeps its id
  return (bid % kXcd) * per + bid / kXcd;
}

}  // namespace gc
