// core::Field<T, Policy>: an owning, i-fastest field on a Grid (core/Field.hpp:36-145 of
// the reference), with the same constructors and host accessors.
//
// Storage lives on the host and/or on the device.  Library operators read and write the
// device copy; the host accessors (operator(), data(), begin()/end()) return references into
// the host copy.  A state flag records which copy is current:
//   - a non-const host accessor makes the host copy current and the device copy stale;
//   - a const host accessor downloads a newer device copy first and keeps both current;
//   - deviceData() uploads a newer host copy first; deviceDataForWrite()/
//     deviceDataDiscard() make the device copy current and the host copy stale.
// So callers that mix host loops and library operators (as the reference's examples do)
// see the reference's semantics, and chains of library operators never leave the device.
// Const host access mutates the cache, so a Field must not be read from several threads
// at once while its device copy is newer.
// The host accessors that may download (operator(), data(), begin(), end()) are not
// noexcept, unlike the reference's host-only ones: a device failure while downloading
// throws std::runtime_error with the library's message instead of ending the process.
#pragma once

#include <cstddef>
#include <cstring>
#include <type_traits>
#include <utility>
#include <vector>

#include <gridcalc/core/Grid.hpp>
#include <gridcalc/core/IndexPolicy.hpp>
#include <gridcalc/hip/Runtime.hpp>

namespace gridcalc::core {

namespace detail {
// Tag for library-internal fields whose contents are produced on the device.
struct device_only_t {};
inline constexpr device_only_t device_only{};
}  // namespace detail

template <class T, class Policy = IndexPolicy::Periodic>
class Field {
  static_assert(std::is_trivially_copyable_v<T>, "Field<T>: T must be trivially copyable");

 public:
  using value_type = T;
  using policy_type = Policy;

  // Zero-initialised field (Field.hpp:49).
  explicit Field(const Grid& grid) : _grid(grid), _data(grid.getSize()), _state(kHost) {}

  // Broadcast constructor (Field.hpp:55).
  Field(const Grid& grid, const T& value)
      : _grid(grid), _data(grid.getSize(), value), _state(kHost) {}

  // Sampling constructor: f(i*hx, j*hy, k*hz) at every site (Field.hpp:66-80).
  template <class F,
            std::enable_if_t<std::is_invocable_r_v<T, F&, double, double, double>, int> = 0>
  Field(const Grid& grid, F&& f) : _grid(grid), _data(grid.getSize()), _state(kHost) {
    const auto& h = _grid.getCellSize();
    for (int k = 0; k < _grid.getNz(); ++k)
      for (int j = 0; j < _grid.getNy(); ++j)
        for (int i = 0; i < _grid.getNx(); ++i)
          _data[_grid.getLinearIndex(i, j, k)] =
              static_cast<T>(f(static_cast<double>(i) * h(0), static_cast<double>(j) * h(1),
                               static_cast<double>(k) * h(2)));
  }

  // Library-internal: storage on the device only, contents written by a kernel next.
  Field(const Grid& grid, detail::device_only_t) : _grid(grid), _state(kDevice) {
    _dev = hip::allocate(bytes());
  }

  Field(const Field& o) : _grid(o._grid), _state(o._state) {
    if (o._state != kDevice) _data = o._data;
    if (o._state != kHost && o._dev) {
      _dev = hip::allocate(bytes());
      hip::copyDevice(_dev, o._dev, bytes());
    } else if (_state == kBoth) {
      _state = kHost;
    }
  }
  Field(Field&& o) noexcept
      : _grid(o._grid), _data(std::move(o._data)), _dev(o._dev), _state(o._state) {
    o._dev = nullptr;
    o._state = kHost;
  }
  Field& operator=(const Field& o) {
    if (this != &o) {
      Field tmp(o);
      swap(tmp);
    }
    return *this;
  }
  Field& operator=(Field&& o) noexcept {
    if (this != &o) {
      Field tmp(std::move(o));
      swap(tmp);
    }
    return *this;
  }
  ~Field() { hip::release(_dev, bytes()); }

  T& operator()(int i, int j, int k) {
    toHost(true);
    return _data[index(i, j, k)];
  }
  const T& operator()(int i, int j, int k) const {
    toHost(false);
    return _data[index(i, j, k)];
  }

  const Grid& getGrid() const noexcept { return _grid; }
  std::size_t getSize() const noexcept { return _grid.getSize(); }
  std::size_t size() const noexcept { return getSize(); }

  T* data() {
    toHost(true);
    return _data.data();
  }
  const T* data() const {
    toHost(false);
    return _data.data();
  }
  T* begin() { return data(); }
  const T* begin() const { return data(); }
  T* end() { return data() + getSize(); }
  const T* end() const { return data() + getSize(); }

  // ---- device residency -----------------------------------------------------------
  // Device pointer to current contents (uploads a newer host copy).
  const T* deviceData() const {
    if (_state == kHost) {
      if (!_dev) _dev = hip::allocate(bytes());
      hip::upload(_dev, _data.data(), bytes());
      _state = kBoth;
    }
    return static_cast<const T*>(_dev);
  }
  // Device pointer to current contents that the caller will modify.
  T* deviceDataForWrite() {
    deviceData();
    _state = kDevice;
    return static_cast<T*>(_dev);
  }
  // Device pointer the caller will overwrite entirely (no upload).
  T* deviceDataDiscard() {
    if (!_dev) _dev = hip::allocate(bytes());
    _state = kDevice;
    return static_cast<T*>(_dev);
  }
  bool hostCurrent() const noexcept { return _state != kDevice; }
  bool deviceCurrent() const noexcept { return _state != kHost; }

  void swap(Field& o) noexcept {
    std::swap(_grid, o._grid);
    _data.swap(o._data);
    std::swap(_dev, o._dev);
    std::swap(_state, o._state);
  }

 private:
  enum State { kHost, kDevice, kBoth };

  std::size_t bytes() const noexcept { return _grid.getSize() * sizeof(T); }
  std::size_t index(int i, int j, int k) const noexcept {
    return _grid.getLinearIndex(Policy::wrap(i, _grid.getNx()), Policy::wrap(j, _grid.getNy()),
                                Policy::wrap(k, _grid.getNz()));
  }
  void toHost(bool will_write) const {
    if (_state == kDevice) {
      if (_data.size() != getSize()) _data.resize(getSize());
      hip::download(_data.data(), _dev, bytes());
      _state = kBoth;
    }
    if (will_write) _state = kHost;
  }

  Grid _grid;
  mutable std::vector<T, hip::HostAllocator<T>> _data;  // page-locked host copy
  mutable void* _dev = nullptr;
  mutable State _state;
};

}  // namespace gridcalc::core
