// gc_core.hip -- error state, contexts, memory and timers of libgridcalc_hip.
#include "gc_common.hpp"

#include <climits>
#include <cmath>

namespace gc {

static thread_local std::string g_last_error;

gc_status set_error(gc_status s, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return s;
}

void clear_error() { g_last_error.clear(); }

gc_status check_ctx(gc_context* ctx) {
  if (!ctx) return set_error(GC_ERR_INVALID, "gridcalc: null gc_context");
  GC_HIP(hipSetDevice(ctx->device));
  return GC_OK;
}

gc_status check_grid(const gc_grid* g) {
  if (!g) return set_error(GC_ERR_INVALID, "gridcalc: null gc_grid");
  if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0)
    return set_error(GC_ERR_INVALID, "Grid: extents (Nx, Ny, Nz) must be > 0; got (%d, %d, %d)",
                     g->nx, g->ny, g->nz);
  if (!(g->hx > 0.0) || !(g->hy > 0.0) || !(g->hz > 0.0))
    return set_error(GC_ERR_INVALID, "Grid: cell_size components must be > 0; got (%f, %f, %f)",
                     g->hx, g->hy, g->hz);
  if (sites(g) > (size_t)INT_MAX)
    return set_error(GC_ERR_UNSUPPORTED, "gridcalc: %zu sites exceed the 2^31-1 limit", sites(g));
  return GC_OK;
}

static gc_status check_axis(const char* label, int n, int radius) {
  if (radius == 0 || n == 1) return GC_OK;
  const int min_n = 2 * radius + 1;
  if (n < min_n)
    return set_error(GC_ERR_EXTENT,
                     "gridcalc: axis %s has N=%d but the requested stencil has radius %d, "
                     "which requires N>=%d on a periodic grid (or N=1 for a degenerate axis); "
                     "intermediate extents alias the [-r, r] offsets under the periodic wrap "
                     "and the stencil silently degrades. Use a larger grid or a lower-Order "
                     "stencil.",
                     label, n, radius, min_n);
  return GC_OK;
}

gc_status check_extent(const gc_grid* g, int radius) {
  GC_TRY(check_axis("x", g->nx, radius));
  GC_TRY(check_axis("y", g->ny, radius));
  GC_TRY(check_axis("z", g->nz, radius));
  return GC_OK;
}

gc_status check_extent3(const gc_grid* g, int rx, int ry, int rz) {
  GC_TRY(check_axis("x", g->nx, rx));
  GC_TRY(check_axis("y", g->ny, ry));
  GC_TRY(check_axis("z", g->nz, rz));
  return GC_OK;
}

gc_status check_dtype(gc_dtype dt) {
  if (dt != GC_F64 && dt != GC_F32) return set_error(GC_ERR_INVALID, "gridcalc: bad dtype %d", (int)dt);
  return GC_OK;
}

gc_status scratch(gc_context* ctx, size_t bytes, void** out) {
  if (bytes > ctx->scratch_bytes) {
    if (ctx->scratch) {
      GC_HIP(hipStreamSynchronize(ctx->stream));
      GC_HIP(hipFree(ctx->scratch));
      ctx->scratch = nullptr;
      ctx->scratch_bytes = 0;
    }
    // round up to 2 MiB to limit regrowth
    const size_t want = (bytes + (size_t(2) << 20) - 1) & ~((size_t(2) << 20) - 1);
    hipError_t e = hipMalloc(&ctx->scratch, want);
    if (e != hipSuccess) {
      ctx->scratch = nullptr;
      (void)hipGetLastError();
      return set_error(GC_ERR_NOMEM, "gridcalc: scratch allocation of %zu bytes failed: %s", want,
                       hipGetErrorString(e));
    }
    ctx->scratch_bytes = want;
  }
  *out = ctx->scratch;
  return GC_OK;
}

}  // namespace gc

using namespace gc;

extern "C" {

const char* gc_last_error(void) { return g_last_error.c_str(); }

int gc_version(void) { return 1 * 10000 + 0 * 100 + 0; }

gc_status gc_device_count(int* count) {
  if (!count) return set_error(GC_ERR_INVALID, "gridcalc: null count");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    n = 0;
  }
  *count = n;
  return GC_OK;
}

gc_status gc_context_create(gc_context** out, void* stream) {
  if (!out) return set_error(GC_ERR_INVALID, "gridcalc: null context pointer");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return set_error(GC_ERR_DEVICE, "gridcalc: no HIP device available");
  }
  gc_context* ctx = new gc_context();
  const hipError_t ed = hipGetDevice(&ctx->device);
  if (ed != hipSuccess) {
    delete ctx;
    (void)hipGetLastError();
    return set_error(GC_ERR_DEVICE, "hipGetDevice failed: %s", hipGetErrorString(ed));
  }
  if (stream) {
    ctx->stream = (hipStream_t)stream;
  } else {
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete ctx;
      return set_error(GC_ERR_DEVICE, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    ctx->owns_stream = true;
  }
  if (hipEventCreate(&ctx->ev_start) != hipSuccess || hipEventCreate(&ctx->ev_stop) != hipSuccess ||
      hipHostMalloc((void**)&ctx->host_result, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
      hipHostGetDevicePointer((void**)&ctx->host_result_dev, ctx->host_result, 0) != hipSuccess) {
    (void)hipGetLastError();
    gc_context_destroy(ctx);
    return set_error(GC_ERR_DEVICE, "gridcalc: context set-up failed");
  }
  *out = ctx;
  return GC_OK;
}

gc_status gc_context_destroy(gc_context* ctx) {
  if (!ctx) return GC_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
  if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
  if (ctx->host_result) (void)hipHostFree(ctx->host_result);
  if (ctx->owns_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return GC_OK;
}

gc_status gc_context_sync(gc_context* ctx) {
  GC_TRY(check_ctx(ctx));
  GC_HIP(hipStreamSynchronize(ctx->stream));
  return GC_OK;
}

void* gc_context_stream(gc_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

size_t gc_context_scratch_bytes(gc_context* ctx) { return ctx ? ctx->scratch_bytes : 0; }

gc_status gc_malloc(void** ptr, size_t bytes) {
  if (!ptr) return set_error(GC_ERR_INVALID, "gridcalc: null pointer");
  *ptr = nullptr;
  if (bytes == 0) return GC_OK;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) {
    *ptr = nullptr;
    (void)hipGetLastError();
    return set_error(GC_ERR_NOMEM, "gridcalc: hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  return GC_OK;
}

gc_status gc_free(void* ptr) {
  if (ptr) GC_HIP(hipFree(ptr));
  return GC_OK;
}

gc_status gc_host_malloc(void** ptr, size_t bytes) {
  if (!ptr) return set_error(GC_ERR_INVALID, "gridcalc: null pointer");
  *ptr = nullptr;
  if (bytes == 0) return GC_OK;
  hipError_t e = hipHostMalloc(ptr, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    *ptr = nullptr;
    (void)hipGetLastError();
    return set_error(GC_ERR_NOMEM, "gridcalc: hipHostMalloc(%zu) failed: %s", bytes,
                     hipGetErrorString(e));
  }
  return GC_OK;
}

gc_status gc_host_free(void* ptr) {
  if (ptr) GC_HIP(hipHostFree(ptr));
  return GC_OK;
}

gc_status gc_memcpy_htod(gc_context* ctx, void* dst, const void* src, size_t bytes) {
  GC_TRY(check_ctx(ctx));
  if (bytes == 0) return GC_OK;
  if (!dst || !src) return set_error(GC_ERR_INVALID, "gridcalc: null copy pointer");
  GC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  GC_HIP(hipStreamSynchronize(ctx->stream));
  return GC_OK;
}

gc_status gc_memcpy_dtoh(gc_context* ctx, void* dst, const void* src, size_t bytes) {
  GC_TRY(check_ctx(ctx));
  if (bytes == 0) return GC_OK;
  if (!dst || !src) return set_error(GC_ERR_INVALID, "gridcalc: null copy pointer");
  GC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  GC_HIP(hipStreamSynchronize(ctx->stream));
  return GC_OK;
}

gc_status gc_memcpy_dtod(gc_context* ctx, void* dst, const void* src, size_t bytes) {
  GC_TRY(check_ctx(ctx));
  if (bytes == 0) return GC_OK;
  if (!dst || !src) return set_error(GC_ERR_INVALID, "gridcalc: null copy pointer");
  GC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return GC_OK;
}

gc_status gc_timer_start(gc_context* ctx) {
  GC_TRY(check_ctx(ctx));
  GC_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
  return GC_OK;
}

gc_status gc_timer_stop(gc_context* ctx, float* ms) {
  GC_TRY(check_ctx(ctx));
  if (!ms) return set_error(GC_ERR_INVALID, "gridcalc: null ms");
  GC_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
  GC_HIP(hipEventSynchronize(ctx->ev_stop));
  GC_HIP(hipEventElapsedTime(ms, ctx->ev_start, ctx->ev_stop));
  return GC_OK;
}

// ---- asynchronous snapshots ------------------------------------------------------------

struct gc_snapshot {
  gc_context* ctx = nullptr;
  hipStream_t copy = nullptr;
  hipEvent_t ready = nullptr, done = nullptr;
  void* dev = nullptr;
  void* host = nullptr;
  size_t bytes = 0;
  bool pending = false;
};

gc_status gc_snapshot_destroy(gc_snapshot* s) {
  if (!s) return GC_OK;
  (void)hipSetDevice(s->ctx->device);
  if (s->copy) (void)hipStreamSynchronize(s->copy);
  if (s->ready) (void)hipEventDestroy(s->ready);
  if (s->done) (void)hipEventDestroy(s->done);
  if (s->copy) (void)hipStreamDestroy(s->copy);
  if (s->dev) (void)hipFree(s->dev);
  if (s->host) (void)hipHostFree(s->host);
  delete s;
  return GC_OK;
}

gc_status gc_snapshot_create(gc_context* ctx, size_t bytes, gc_snapshot** out) {
  if (!out) return set_error(GC_ERR_INVALID, "gridcalc: null snapshot pointer");
  *out = nullptr;
  GC_TRY(check_ctx(ctx));
  if (bytes == 0) return set_error(GC_ERR_INVALID, "gridcalc: empty snapshot");
  gc_snapshot* s = new gc_snapshot();
  s->ctx = ctx;
  s->bytes = bytes;
  if (hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&s->ready, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&s->done, hipEventDisableTiming) != hipSuccess ||
      hipMalloc(&s->dev, bytes) != hipSuccess ||
      hipHostMalloc(&s->host, bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    gc_snapshot_destroy(s);
    return set_error(GC_ERR_NOMEM, "gridcalc: snapshot set-up of %zu bytes failed", bytes);
  }
  *out = s;
  return GC_OK;
}

gc_status gc_snapshot_capture(gc_snapshot* s, const void* src_dev, size_t bytes) {
  if (!s) return set_error(GC_ERR_INVALID, "gridcalc: null snapshot");
  GC_TRY(check_ctx(s->ctx));
  if (!src_dev) return set_error(GC_ERR_INVALID, "gridcalc: null snapshot source");
  if (bytes != s->bytes)
    return set_error(GC_ERR_INVALID, "gridcalc: snapshot of %zu bytes into a slot of %zu", bytes,
                     s->bytes);
  if (s->pending) GC_HIP(hipEventSynchronize(s->done));  // the staging buffer is free again
  GC_HIP(hipMemcpyAsync(s->dev, src_dev, bytes, hipMemcpyDeviceToDevice, s->ctx->stream));
  GC_HIP(hipEventRecord(s->ready, s->ctx->stream));
  GC_HIP(hipStreamWaitEvent(s->copy, s->ready, 0));
  GC_HIP(hipMemcpyAsync(s->host, s->dev, bytes, hipMemcpyDeviceToHost, s->copy));
  GC_HIP(hipEventRecord(s->done, s->copy));
  s->pending = true;
  return GC_OK;
}

gc_status gc_snapshot_wait(gc_snapshot* s, const void** host) {
  if (!s || !host) return set_error(GC_ERR_INVALID, "gridcalc: null snapshot argument");
  GC_TRY(check_ctx(s->ctx));
  if (s->pending) GC_HIP(hipEventSynchronize(s->done));
  s->pending = false;
  *host = s->host;
  return GC_OK;
}

}  // extern "C"
