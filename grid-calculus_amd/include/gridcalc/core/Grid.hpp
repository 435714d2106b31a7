// core::Grid: extents, cell size, cell volume and the i-fastest linear index
// (core/Grid.hpp:28-148 of the reference; same interface, validation and messages).
#pragma once

#include <cstddef>
#include <stdexcept>
#include <string>

#include <gridcalc/core/EigenAliases.hpp>

namespace gridcalc::core {

class Grid {
 public:
  Grid(int Nx, int Ny, int Nz, const Vec3d& cell_size)
      : _Nx(Nx), _Ny(Ny), _Nz(Nz), _cell_size(cell_size) {
    if (Nx <= 0 || Ny <= 0 || Nz <= 0) {
      throw std::invalid_argument("Grid: extents (Nx, Ny, Nz) must be > 0; got (" +
                                  std::to_string(Nx) + ", " + std::to_string(Ny) + ", " +
                                  std::to_string(Nz) + ")");
    }
    if (!(cell_size(0) > 0.0) || !(cell_size(1) > 0.0) || !(cell_size(2) > 0.0)) {
      throw std::invalid_argument("Grid: cell_size components must be > 0; got (" +
                                  std::to_string(cell_size(0)) + ", " +
                                  std::to_string(cell_size(1)) + ", " +
                                  std::to_string(cell_size(2)) + ")");
    }
  }

  int getNx() const noexcept { return _Nx; }
  int getNy() const noexcept { return _Ny; }
  int getNz() const noexcept { return _Nz; }
  const Vec3d& getCellSize() const noexcept { return _cell_size; }

  // bit-exact comparison of extents and cell sizes
  bool operator==(const Grid& o) const noexcept {
    return _Nx == o._Nx && _Ny == o._Ny && _Nz == o._Nz && _cell_size(0) == o._cell_size(0) &&
           _cell_size(1) == o._cell_size(1) && _cell_size(2) == o._cell_size(2);
  }
  bool operator!=(const Grid& o) const noexcept { return !(*this == o); }

  double getCellVolume() const noexcept { return _cell_size(0) * _cell_size(1) * _cell_size(2); }

  std::size_t getSize() const noexcept {
    return static_cast<std::size_t>(_Nx) * static_cast<std::size_t>(_Ny) *
           static_cast<std::size_t>(_Nz);
  }
  std::size_t size() const noexcept { return getSize(); }

  std::size_t getLinearIndex(int i, int j, int k) const noexcept {
    return static_cast<std::size_t>(i) +
           static_cast<std::size_t>(_Nx) *
               (static_cast<std::size_t>(j) +
                static_cast<std::size_t>(_Ny) * static_cast<std::size_t>(k));
  }

 private:
  int _Nx;
  int _Ny;
  int _Nz;
  Vec3d _cell_size;
};

}  // namespace gridcalc::core
