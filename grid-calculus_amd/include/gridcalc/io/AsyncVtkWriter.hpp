// io::AsyncVtkWriter: snapshots of Field<double> written while the time stepping continues,
// in the legacy ASCII VTK STRUCTURED_POINTS format of examples/common/vtk_writer.hpp:34-67
// (the same bytes: header lines, 17 significant digits per value, one value per line).
//
// write() takes the field's current contents -- on the device: one device-to-device copy on
// the library's stream into a snapshot slot, then a device-to-host copy on the slot's own
// stream (gc_snapshot_*), so later kernels on the library's stream run while it proceeds; a
// host-current field is copied on the host -- and returns; a worker thread waits for the copy,
// formats the file and writes it.  Up to `slots` snapshots are in flight; write() blocks when
// all are busy.  flush() waits for every queued file and rethrows the first I/O error; the
// destructor flushes (printing, not throwing, an error).
#pragma once

#include <charconv>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <exception>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include <gridcalc/core/Field.hpp>
#include <gridcalc/core/Grid.hpp>
#include <gridcalc/hip/Runtime.hpp>

namespace gridcalc::io {

namespace detail {

// Appends v as printf("%.17g") (what std::ostream with precision(17) writes for a double).
inline void appendG17(std::string& out, double v) {
  char buf[40];
  const auto r = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::general, 17);
  out.append(buf, r.ptr);
}

inline void writeVtk(const std::string& path, const std::string& name, const core::Grid& g,
                     const double* data) {
  std::FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("AsyncVtkWriter: cannot open '" + path + "' for writing");
  const auto& h = g.getCellSize();
  std::string head = "# vtk DataFile Version 3.0\ngridcalc field: " + name +
                     "\nASCII\nDATASET STRUCTURED_POINTS\nDIMENSIONS " +
                     std::to_string(g.getNx()) + ' ' + std::to_string(g.getNy()) + ' ' +
                     std::to_string(g.getNz()) + "\nORIGIN 0 0 0\nSPACING ";
  appendG17(head, h(0));
  head += ' ';
  appendG17(head, h(1));
  head += ' ';
  appendG17(head, h(2));
  head += "\nPOINT_DATA " + std::to_string(g.getSize()) + "\nSCALARS " + name +
          " double 1\nLOOKUP_TABLE default\n";
  bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size();
  std::string chunk;
  chunk.reserve(1 << 20);
  const std::size_t n = g.getSize();
  for (std::size_t i = 0; i < n && ok; ++i) {
    appendG17(chunk, data[i]);
    chunk += '\n';
    if (chunk.size() > (1 << 20) - 64 || i + 1 == n) {
      ok = std::fwrite(chunk.data(), 1, chunk.size(), f) == chunk.size();
      chunk.clear();
    }
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) throw std::runtime_error("AsyncVtkWriter: write to '" + path + "' failed");
}

}  // namespace detail

class AsyncVtkWriter {
 public:
  explicit AsyncVtkWriter(int slots = 2) : _slots(slots > 0 ? slots : 1) {
    _worker = std::thread([this] { run(); });
  }
  AsyncVtkWriter(const AsyncVtkWriter&) = delete;
  AsyncVtkWriter& operator=(const AsyncVtkWriter&) = delete;

  ~AsyncVtkWriter() {
    try {
      flush();
    } catch (const std::exception& e) {
      std::fprintf(stderr, "%s\n", e.what());
    }
    {
      std::lock_guard<std::mutex> lock(_mu);
      _stop = true;
    }
    _cv.notify_all();
    _worker.join();
    for (auto& s : _slots) gc_snapshot_destroy(s.snap);
  }

  void write(const core::Field<double>& f, const std::string& path, const std::string& name) {
    std::unique_lock<std::mutex> lock(_mu);
    _cv.wait(lock, [&] { return freeSlot() >= 0; });
    const int i = freeSlot();
    Slot& s = _slots[i];
    s.busy = true;
    lock.unlock();
    Job job{i, f.getGrid(), path, name, !f.deviceCurrent()};
    const std::size_t bytes = f.getSize() * sizeof(double);
    try {
      if (job.from_host) {
        s.host_copy.resize(f.getSize());
        std::memcpy(s.host_copy.data(), f.data(), bytes);
      } else {
        gc_context* ctx = hip::context();
        if (!s.snap || s.bytes != bytes || s.ctx != ctx) {
          gc_snapshot_destroy(s.snap);
          s.snap = nullptr;
          hip::check(gc_snapshot_create(ctx, bytes, &s.snap));
          s.bytes = bytes;
          s.ctx = ctx;
        }
        hip::check(gc_snapshot_capture(s.snap, f.deviceData(), bytes));
      }
    } catch (...) {
      std::lock_guard<std::mutex> l(_mu);
      s.busy = false;
      _cv.notify_all();
      throw;
    }
    lock.lock();
    _jobs.push_back(std::move(job));
    lock.unlock();
    _cv.notify_all();
  }

  void flush() {
    std::unique_lock<std::mutex> lock(_mu);
    _cv.wait(lock, [&] { return _jobs.empty() && _running == 0; });
    if (_error) {
      auto e = _error;
      _error = nullptr;
      std::rethrow_exception(e);
    }
  }

 private:
  struct Slot {
    gc_snapshot* snap = nullptr;
    gc_context* ctx = nullptr;
    std::size_t bytes = 0;
    std::vector<double> host_copy;
    bool busy = false;
  };
  struct Job {
    int slot;
    core::Grid grid;
    std::string path, name;
    bool from_host;
  };

  int freeSlot() const {
    for (std::size_t i = 0; i < _slots.size(); ++i)
      if (!_slots[i].busy) return (int)i;
    return -1;
  }

  void run() {
    std::unique_lock<std::mutex> lock(_mu);
    for (;;) {
      _cv.wait(lock, [&] { return _stop || !_jobs.empty(); });
      if (_jobs.empty()) return;  // stop requested and nothing left
      Job job = std::move(_jobs.front());
      _jobs.pop_front();
      ++_running;
      Slot& s = _slots[job.slot];
      lock.unlock();
      try {
        const double* data = s.host_copy.data();
        if (!job.from_host) {
          const void* host = nullptr;
          hip::check(gc_snapshot_wait(s.snap, &host));
          data = static_cast<const double*>(host);
        }
        detail::writeVtk(job.path, job.name, job.grid, data);
      } catch (...) {
        std::lock_guard<std::mutex> l(_mu);
        if (!_error) _error = std::current_exception();
      }
      lock.lock();
      s.busy = false;
      --_running;
      _cv.notify_all();
    }
  }

  std::vector<Slot> _slots;
  std::deque<Job> _jobs;
  std::mutex _mu;
  std::condition_variable _cv;
  int _running = 0;
  bool _stop = false;
  std::exception_ptr _error;
  std::thread _worker;
};

}  // namespace gridcalc::io
