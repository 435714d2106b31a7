// Host-side glue between the C++ header API and libgridcalc_hip.so (include/gridcalc_hip.h):
// the process-wide execution context, translation of gc_status codes into the exceptions
// the reference throws, and a caching device-memory pool for Field storage.
//
// Every Field operation of this library runs on one gc_context (one HIP stream) at a time,
// so device buffers returned to the pool can be handed out again at once: later work on the
// same stream is ordered after the work that last used them.  Switching to another context
// (setContext) first waits for the outgoing context's stream, so a block released by work
// still queued there is never reused by work on the new stream.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <map>
#include <mutex>
#include <new>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>

#include <gridcalc/core/Grid.hpp>
#include <gridcalc_hip.h>

namespace gridcalc::hip {

// Throws the exception the reference throws for the same condition: std::invalid_argument
// for precondition failures, std::runtime_error for device failures.
inline void check(gc_status s) {
  if (s == GC_OK) return;
  const std::string msg = gc_last_error();
  if (s == GC_ERR_INVALID || s == GC_ERR_EXTENT || s == GC_ERR_CFL || s == GC_ERR_NONPOS_D)
    throw std::invalid_argument(msg);
  throw std::runtime_error("gridcalc (device): " + msg);
}

namespace detail {

struct State {
  std::mutex mu;
  gc_context* ctx = nullptr;
  gc_context* user_ctx = nullptr;
  std::multimap<std::size_t, void*> free_blocks;  // cached device blocks by size
  std::size_t cached_bytes = 0;
  std::size_t uploads = 0, downloads = 0, upload_bytes = 0, download_bytes = 0;
  std::multimap<std::size_t, void*> free_host;  // cached page-locked host blocks by size
  std::size_t cached_host_bytes = 0;
  std::set<void*> pageable;  // host blocks that fell back to malloc (pinning refused)
};

// Never destroyed: device memory and the context are reclaimed with the process, so no
// HIP call runs during static destruction (after the HIP runtime itself may be gone).
inline State& state() {
  static State* s = new State();
  return *s;
}

}  // namespace detail

// The context every Field operation uses: a caller-installed one (setContext) or a
// private context created on first use.
inline gc_context* context() {
  auto& s = detail::state();
  std::lock_guard<std::mutex> lock(s.mu);
  if (s.user_ctx) return s.user_ctx;
  if (!s.ctx) check(gc_context_create(&s.ctx, nullptr));
  return s.ctx;
}

// Installs a caller-owned context (e.g. bound to the caller's HIP stream); nullptr
// restores the private one.  The caller keeps ownership.  Waits for all work queued on the
// outgoing context first (pooled blocks are reused without per-block events, so the old
// stream must be idle before another stream can receive them).
inline void setContext(gc_context* ctx) {
  auto& s = detail::state();
  std::lock_guard<std::mutex> lock(s.mu);
  gc_context* outgoing = s.user_ctx ? s.user_ctx : s.ctx;
  gc_context* incoming = ctx ? ctx : s.ctx;
  if (outgoing && outgoing != incoming) check(gc_context_sync(outgoing));
  s.user_ctx = ctx;
}

inline void synchronize() { check(gc_context_sync(context())); }

// Caching pool: exact-size reuse; blocks are returned to the pool, not to HIP.
inline void* allocate(std::size_t bytes) {
  if (bytes == 0) return nullptr;
  auto& s = detail::state();
  {
    std::lock_guard<std::mutex> lock(s.mu);
    auto it = s.free_blocks.find(bytes);
    if (it != s.free_blocks.end()) {
      void* p = it->second;
      s.free_blocks.erase(it);
      s.cached_bytes -= bytes;
      return p;
    }
  }
  void* p = nullptr;
  gc_status st = gc_malloc(&p, bytes);
  if (st == GC_ERR_NOMEM) {  // release the cache and retry once
    auto& st2 = detail::state();
    std::lock_guard<std::mutex> lock(st2.mu);
    for (auto& kv : st2.free_blocks) gc_free(kv.second);
    st2.free_blocks.clear();
    st2.cached_bytes = 0;
    st = gc_malloc(&p, bytes);
  }
  check(st);
  return p;
}

inline void release(void* p, std::size_t bytes) noexcept {
  if (!p) return;
  auto& s = detail::state();
  std::lock_guard<std::mutex> lock(s.mu);
  s.free_blocks.emplace(bytes, p);
  s.cached_bytes += bytes;
}

// Host <-> device field transfers made on behalf of Field (counted: transferStats()).
// Host storage of Field: page-locked blocks (gc_host_malloc) cached by exact size, so the
// transfers a caller's host loops cause run as DMA from pinned memory; up to kHostCacheBytes of
// released blocks stay cached.  When the system refuses to pin more memory, blocks come from
// malloc instead (still correct, transfers staged).
constexpr std::size_t kHostCacheBytes = std::size_t(1) << 30;

inline void* host_allocate(std::size_t bytes) {
  if (bytes == 0) return nullptr;
  auto& s = detail::state();
  {
    std::lock_guard<std::mutex> lock(s.mu);
    auto it = s.free_host.find(bytes);
    if (it != s.free_host.end()) {
      void* p = it->second;
      s.free_host.erase(it);
      s.cached_host_bytes -= bytes;
      return p;
    }
  }
  void* p = nullptr;
  if (gc_host_malloc(&p, bytes) == GC_OK && p) return p;
  p = std::malloc(bytes);
  if (!p) throw std::bad_alloc();
  std::lock_guard<std::mutex> lock(s.mu);
  s.pageable.insert(p);
  return p;
}

inline void host_release(void* p, std::size_t bytes) noexcept {
  if (!p) return;
  auto& s = detail::state();
  std::lock_guard<std::mutex> lock(s.mu);
  auto pg = s.pageable.find(p);
  if (pg != s.pageable.end()) {
    s.pageable.erase(pg);
    std::free(p);
    return;
  }
  if (s.cached_host_bytes + bytes <= kHostCacheBytes) {
    s.free_host.emplace(bytes, p);
    s.cached_host_bytes += bytes;
  } else {
    (void)gc_host_free(p);
  }
}

// std::allocator replacement for Field's host vector.
template <class T>
struct HostAllocator {
  using value_type = T;
  HostAllocator() noexcept = default;
  template <class U>
  HostAllocator(const HostAllocator<U>&) noexcept {}
  T* allocate(std::size_t n) { return static_cast<T*>(host_allocate(n * sizeof(T))); }
  void deallocate(T* p, std::size_t n) noexcept { host_release(p, n * sizeof(T)); }
  template <class U>
  bool operator==(const HostAllocator<U>&) const noexcept { return true; }
  template <class U>
  bool operator!=(const HostAllocator<U>&) const noexcept { return false; }
};

inline void upload(void* dev, const void* host, std::size_t bytes) {
  check(gc_memcpy_htod(context(), dev, host, bytes));
  auto& s = detail::state();
  std::lock_guard<std::mutex> lock(s.mu);
  ++s.uploads;
  s.upload_bytes += bytes;
}
inline void download(void* host, const void* dev, std::size_t bytes) {
  check(gc_memcpy_dtoh(context(), host, dev, bytes));
  auto& s = detail::state();
  std::lock_guard<std::mutex> lock(s.mu);
  ++s.downloads;
  s.download_bytes += bytes;
}

struct TransferStats {
  std::size_t uploads, downloads, upload_bytes, download_bytes;
};

// Field transfers since the process started (or the last resetTransferStats()): how often a
// caller's host loops made the library move a field between host and device.
inline TransferStats transferStats() {
  auto& s = detail::state();
  std::lock_guard<std::mutex> lock(s.mu);
  return TransferStats{s.uploads, s.downloads, s.upload_bytes, s.download_bytes};
}
inline void resetTransferStats() {
  auto& s = detail::state();
  std::lock_guard<std::mutex> lock(s.mu);
  s.uploads = s.downloads = s.upload_bytes = s.download_bytes = 0;
}
inline void copyDevice(void* dst, const void* src, std::size_t bytes) {
  check(gc_memcpy_dtod(context(), dst, src, bytes));
}

inline gc_grid toGrid(const core::Grid& g) {
  const auto& h = g.getCellSize();
  return gc_grid{g.getNx(), g.getNy(), g.getNz(), h(0), h(1), h(2)};
}

template <class T>
struct dtype_of;
template <>
struct dtype_of<double> {
  static constexpr gc_dtype value = GC_F64;
};
template <>
struct dtype_of<float> {
  static constexpr gc_dtype value = GC_F32;
};

}  // namespace gridcalc::hip
