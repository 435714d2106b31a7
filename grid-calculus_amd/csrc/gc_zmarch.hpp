// gc_zmarch.hpp -- the z-marching LDS engine shared by the scalar-input stencil operators
// (gc_stencil.hip) and the mixed partial derivatives (gc_mixed.hip).
#pragma once

#include "gc_common.hpp"

#include <type_traits>

namespace gc {
namespace {

// ---- z-marching LDS engine for the scalar-input operators --------------------------------
//
// A block owns an OX x OY tile of xy-columns and a chunk of consecutive planes.  It keeps the
// 2R+1 planes z-R..z+R of the input field (tile + R-site halo, periodic wrap applied while
// filling) in an LDS ring, so each input value is read from memory once per block plus the
// halo; the next plane is fetched into registers while the current output plane is formed.
// The operators read neighbours through an accessor and keep the reference's per-axis
// accumulation order, so results equal the one-site-at-a-time evaluation bit for bit.

template <int R, int OX, int OY>
struct ZTile {
  static constexpr int NT = 256;
  static constexpr int PW = OX + 2 * R, PH = OY + 2 * R, PN = PW * PH, NP = 2 * R + 1;
  static constexpr int NS = NP + 1;  // ring slots: one more than a site reads (k_zmarch)
  static constexpr int PL = (PN + NT - 1) / NT;  // plane slots filled per thread
  static constexpr int ROWS = NT / OX;
  static constexpr int OR = OX * OY / NT;        // output sites per thread per plane
  static_assert(NT % OX == 0 && (OX * OY) % NT == 0, "tile must match the block");
};

// Neighbour access into the ring: F(0, dx, dy, dz) at the current output site.
template <class T, int R, int PN, int NS, int PW>
struct RingAt {
  const T* ring;     // [NS][PN]
  const int* slot;   // slot of plane z + dz at slot[dz + R]
  int base;          // (oy + R) * PW + (ox + R)
  __device__ __forceinline__ T operator()(int, int dx, int dy, int dz) const {
    return ring[slot[dz + R] * PN + base + dy * PW + dx];
  }
};

// Values per site of an operator that declares NOUT (and provides values(F, v)), else 0.
template <class Op, class = void>
struct ZOut {
  static constexpr int N = 0;
};
template <class Op>
struct ZOut<Op, std::void_t<decltype(Op::NOUT)>> {
  static constexpr int N = Op::NOUT;
};

template <class T, int OX, int OY, class Op>
__global__ __launch_bounds__(256) void k_zmarch(Op op, const T* __restrict__ f,
                                                T* __restrict__ out, Dims d, int ntx, int nty,
                                                int zlen) {
  constexpr int R = Op::R;
  using Z = ZTile<R, OX, OY>;
  __shared__ T ring[Z::NS * Z::PN];
  const int t = threadIdx.x;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int tx = (int)(bid % (unsigned)ntx);
  const unsigned rr = bid / (unsigned)ntx;
  const int ty = (int)(rr % (unsigned)nty);
  const int zc = (int)(rr / (unsigned)nty);
  const int x0 = tx * OX, y0 = ty * OY, z0 = zc * zlen;
  const int nout = min(zlen, d.nz - z0);

  unsigned poff[Z::PL];
  bool pin[Z::PL];
#pragma unroll
  for (int s = 0; s < Z::PL; ++s) {
    const int idx = t + s * Z::NT;
    pin[s] = idx < Z::PN;
    const int px = idx % Z::PW, py = idx / Z::PW;
    poff[s] = pin[s] ? (unsigned)(wrap_any(x0 + px - R, d.nx) + d.nx * wrap_any(y0 + py - R, d.ny))
                     : 0u;
  }
  T reg[Z::PL];
  auto load = [&](int plane) {
    const size_t zo = (size_t)d.plane * (size_t)wrap_any(plane, d.nz);
#pragma unroll
    for (int s = 0; s < Z::PL; ++s)
      if (pin[s]) reg[s] = f[zo + poff[s]];
  };
  auto store = [&](int slot) {
#pragma unroll
    for (int s = 0; s < Z::PL; ++s)
      if (pin[s]) ring[slot * Z::PN + t + s * Z::NT] = reg[s];
  };

  // Plane z0 - R + i lives in slot i % NS.  With NS = 2R + 2 slots the plane stored in
  // iteration k goes to the slot that iteration k - 1 did not read, so one barrier per plane
  // suffices: a thread stores in iteration k only after every thread has passed the barrier
  // of iteration k - 1, hence finished iteration k - 2, the last reader of that slot.
  for (int i = 0; i < 2 * R; ++i) {
    load(z0 - R + i);
    store(i);
  }
  load(z0 + R);
  const int ox = t % OX, oy0 = t / OX;
  // More than 16 bytes of output per site: a wave is one tile row (OX = 64 = the wave size) for
  // each q, so it stages the row's values in LDS and writes them 64 consecutive words per store
  // instruction instead of one word of each site at the site stride.  Only the wave itself
  // touches its row buffer: wave-level ordering suffices.
  constexpr int NO = ZOut<Op>::N;
  constexpr bool kRow = NO > 1 && NO * sizeof(T) > 16 && OX == 64;
  __shared__ T orow[kRow ? Z::ROWS : 1][kRow ? OX * NO : 1];
  auto wave_sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  for (int k = 0; k < nout; ++k) {
    store((k + 2 * R) % Z::NS);
    __syncthreads();
    if (k + 1 < nout) load(z0 + k + 1 + R);
    int slot[Z::NP];
#pragma unroll
    for (int j = 0; j < Z::NP; ++j) slot[j] = (k + j) % Z::NS;
    const size_t zo = (size_t)d.plane * (size_t)(z0 + k);
#pragma unroll
    for (int q = 0; q < Z::OR; ++q) {
      const int oy = oy0 + q * Z::ROWS;
      if constexpr (kRow) {
        // every lane reads its ring position (in range for the whole tile), the row's sites
        // inside the grid are written
        const RingAt<T, R, Z::PN, Z::NS, Z::PW> F{ring, slot, (oy + R) * Z::PW + (ox + R)};
        T v[NO];
        op.values(F, v);
        T* rw = orow[oy0];
#pragma unroll
        for (int j = 0; j < NO; ++j) rw[ox * NO + j] = v[j];
        wave_sync();
        const int rowlen = y0 + oy < d.ny ? min(OX, d.nx - x0) * NO : 0;
        T* gp = out + (size_t)NO * (zo + (size_t)d.nx * (size_t)(y0 + oy) + (size_t)x0);
#pragma unroll
        for (int j = 0; j < NO; ++j) {
          const int e = ox + OX * j;
          if (e < rowlen) gp[e] = rw[e];
        }
        wave_sync();
      } else if (x0 + ox < d.nx && y0 + oy < d.ny) {
        const RingAt<T, R, Z::PN, Z::NS, Z::PW> F{ring, slot, (oy + R) * Z::PW + (ox + R)};
        op(F, out, zo + (size_t)(x0 + ox) + (size_t)d.nx * (size_t)(y0 + oy));
      }
    }
  }
}

constexpr int kZOX = 64, kZOY = 8, kZLen = 32;

// Tile height of an operator on the z-march engines: Op::kTileRows if the operator declares
// one, else kZOY.  A taller tile halves the y-halo share of every plane a block loads; it pays
// for the light operators and costs occupancy for those with a larger ring or more work.
template <class Op, class = void>
struct ZRows {
  static constexpr int OY = kZOY;
};
template <class Op>
struct ZRows<Op, std::void_t<decltype(Op::kTileRows)>> {
  static constexpr int OY = Op::kTileRows;
};

template <class T, class Op>
void zmarch_launch(gc_context* ctx, const Dims& d, const Op& op, const T* f, T* out) {
  constexpr int OY = ZRows<Op>::OY;
  const int ntx = (d.nx + kZOX - 1) / kZOX, nty = (d.ny + OY - 1) / OY;
  const int nzc = (d.nz + kZLen - 1) / kZLen;
  hipLaunchKernelGGL((k_zmarch<T, kZOX, OY, Op>), dim3((unsigned)(ntx * nty * nzc)), dim3(256),
                     0, ctx->stream, op, f, out, d, ntx, nty, kZLen);
}

}  // namespace
}  // namespace gc
