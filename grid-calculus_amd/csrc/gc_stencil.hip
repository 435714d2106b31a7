// gc_stencil.hip -- the stencil engine: Laplacian, gradient, divergence, biharmonic
// and the finite-volume cell Laplacian on periodic grids.
//
// Every FP64 operator keeps the reference's arithmetic: per-axis partial sums accumulated
// in stencil-offset order starting from 0, then the scaled axis terms summed left to
// right; this file is compiled with -ffp-contract=off, so double-precision results
// equal the CPU path bit for bit.  FP32 is this library's own form, compared with the FP64
// oracle within stated bounds: the same order for the first-derivative and Laplacian
// operators, fused multiply-adds for the cell Laplacian, and the separable form for the
// biharmonic (k_bih_sep).
//
// Two engines, both marching through chunks of z-planes per block with block ids remapped so
// each XCD owns a contiguous range of tiles (neighbouring tiles meet in the same L2):
// scalar-input operators keep a ring of input planes in LDS; vector- and tensor-field
// operators keep the components differentiated along z in per-thread register queues and
// stage only the current plane (x or y halo) in LDS.  The operators move 2-12 words per site
// and are bounded by HBM.
#include "gc_common.hpp"
#include "gc_zmarch.hpp"
#include "gc_arith.hpp"

#include <type_traits>
#include <utility>

#include <cstring>
#include <type_traits>

namespace gc {
namespace {

// diff/Laplacian.hpp:56-71
template <class T, int O>
struct OpLap {
  static constexpr int R = D2<O>::R;
  static constexpr int kTileRows = 16;  // 512^3 FP64: O2 0.444 -> 0.428 ms, O4 0.562 -> 0.524
  T ihx2, ihy2, ihz2;
  template <class A>
  __device__ __forceinline__ void operator()(const A& F, T* __restrict__ out, size_t c) const {
    using S = D2<O>;
    T dxx = 0, dyy = 0, dzz = 0;
#pragma unroll
    for (int m = 0; m < S::N; ++m) {
      const int o = m - S::R;
      const T w = T(S::w(m));
      dxx += w * F(0, o, 0, 0);
      dyy += w * F(0, 0, o, 0);
      dzz += w * F(0, 0, 0, o);
    }
    out[c] = dxx * ihx2 + dyy * ihy2 + dzz * ihz2;
  }
};

// diff/Gradient.hpp:56-70 (scalar -> Vec3, AoS)
template <class T, int O>
struct OpGrad {
  static constexpr int R = D1<O>::R;
  static constexpr int kTileRows = 16;  // 512^3 FP64: O2 1.082 -> 0.978 ms, O4 1.124 -> 1.007
  static constexpr int NOUT = 3;        // values per site (the z-march engine's row staging)
  T ihx, ihy, ihz;
  template <class A>
  __device__ __forceinline__ void values(const A& F, T (&v)[NOUT]) const {
    using S = D1<O>;
    T dx = 0, dy = 0, dz = 0;
#pragma unroll
    for (int m = 0; m < S::N; ++m) {
      const int o = m - S::R;
      const T w = T(S::w(m));
      dx += w * F(0, o, 0, 0);
      dy += w * F(0, 0, o, 0);
      dz += w * F(0, 0, 0, o);
    }
    v[0] = dx * ihx;
    v[1] = dy * ihy;
    v[2] = dz * ihz;
  }
  template <class A>
  __device__ __forceinline__ void operator()(const A& F, T* __restrict__ out, size_t c) const {
    T v[NOUT];
    values(F, v);
    T* o3 = out + 3 * c;
    o3[0] = v[0];
    o3[1] = v[1];
    o3[2] = v[2];
  }
};

// diff/Biharmonic.hpp:67-101: S4 loop, mixed loop (mb outer, ma inner), final expression in
// the reference's order.
template <class T, int O>
struct OpBih {
  static constexpr int R = D4<O>::R;
  static_assert(D4<O>::R >= D2<O>::R, "ring radius");
  T ihx2, ihy2, ihz2, ihx4, ihy4, ihz4;
  template <class A>
  __device__ __forceinline__ void operator()(const A& F, T* __restrict__ out, size_t c) const {
    using S2 = D2<O>;
    using S4 = D4<O>;
    T dx4 = 0, dy4 = 0, dz4 = 0;
#pragma unroll
    for (int m = 0; m < S4::N; ++m) {
      const int o = m - S4::R;
      const T w = T(S4::w(m));
      dx4 += w * F(0, o, 0, 0);
      dy4 += w * F(0, 0, o, 0);
      dz4 += w * F(0, 0, 0, o);
    }
    T dxy = 0, dxz = 0, dyz = 0;
#pragma unroll
    for (int mb = 0; mb < S2::N; ++mb) {
      const int ob = mb - S2::R;
      const double wb = S2::w(mb);
#pragma unroll
      for (int ma = 0; ma < S2::N; ++ma) {
        const int oa = ma - S2::R;
        const T wab = T(S2::w(ma) * wb);
        dxy += wab * F(0, oa, ob, 0);
        dxz += wab * F(0, oa, 0, ob);
        dyz += wab * F(0, 0, oa, ob);
      }
    }
    out[c] = dx4 * ihx4 + dy4 * ihy4 + dz4 * ihz4 +
             T(2) * (dxy * ihx2 * ihy2 + dxz * ihx2 * ihz2 + dyz * ihy2 * ihz2);
  }
};

// fvm::cellLaplacian (fvm/CellLaplacian.hpp:115-141) with register queues along z.  A block
// of 64 OY / OR threads owns a 64 x OY tile of columns and a chunk of planes; each thread
// owns OR consecutive rows of one column and keeps psi at z - 1, z, z + 1, D at z, z + 1 and the face
// below each of its sites in registers.  LDS holds the psi and D planes z (tile + 1-site
// halo), double-buffered, so one barrier per plane suffices: iteration k stages plane z0 + k
// (tile sites from the owners' registers, halo sites requested PF planes ahead) into
// buffer k & 1, whose previous readers (iteration k - 2) have all passed the barrier of
// iteration k - 1.  Face coefficients: x faces of both sides, the y face above each row (the
// face below it is the row beneath's, held by the same thread), the z face above (the face
// below is the previous plane's), each by face_coef(D_lower, D_upper) and the fluxes by fvm_k
// (gc_arith.hpp): FP64 is the reference's harmonic mean and order (the per-cell
// evaluation's bits), FP32 the hardware reciprocal and fused multiply-adds.  Every D value
// the block loads is tested (!(D > 0)) and the smallest failing flat index goes to *bad by
// atomicMin; on a failure the output values are meaningless and the caller reports the error.
template <class T, int OY, int OR, int PF>
__global__ __launch_bounds__(64 * OY / OR) void k_cell_zq(T ihx2, T ihy2, T ihz2,
                                                          const T* __restrict__ psi,
                                                          const T* __restrict__ D,
                                                          T* __restrict__ out,
                                                          unsigned long long* __restrict__ bad,
                                                          Dims d, int ntx, int nty, int zlen) {
  constexpr int OX = 64, NT = OX * OY / OR;
  constexpr int PW = OX + 2, PH = OY + 2, PN = PW * PH, NR = PN - OX * OY;
  static_assert(OY % OR == 0 && NR <= NT, "one halo site per thread");
  __shared__ T pb[2][PN];
  __shared__ T db[2][PN];
  const int t = threadIdx.x;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int tx = (int)(bid % (unsigned)ntx);
  const unsigned rr = bid / (unsigned)ntx;
  const int ty = (int)(rr % (unsigned)nty);
  const int zc = (int)(rr / (unsigned)nty);
  const int x0 = tx * OX, y0 = ty * OY, z0 = zc * zlen;
  const int nout = min(zlen, d.nz - z0);
  const int ox = t % OX, oyb = (t / OX) * OR;  // this thread's rows oyb .. oyb + OR - 1
  auto goff = [&](int px, int py) {  // plane coordinates (halo 1) -> wrapped plane offset
    return (unsigned)(wrap_any(x0 + px - 1, d.nx) + d.nx * wrap_any(y0 + py - 1, d.ny));
  };
  unsigned coff[OR];
  bool cin[OR];
#pragma unroll
  for (int q = 0; q < OR; ++q) {
    coff[q] = goff(ox + 1, oyb + q + 1);
    cin[q] = x0 + ox < d.nx && y0 + oyb + q < d.ny;
  }
  // the halo site this thread stages: top row, bottom row, then left and right ends of rows
  int hs = -1;
  unsigned hoff = 0u;
  if (t < NR) {
    int px, py;
    if (t < PW) {
      px = t;
      py = 0;
    } else if (t < 2 * PW) {
      px = t - PW;
      py = PH - 1;
    } else {
      const int r = t - 2 * PW;
      py = 1 + (r >> 1);
      px = (r & 1) ? PW - 1 : 0;
    }
    hs = py * PW + px;
    hoff = goff(px, py);
  }
  auto zoff = [&](int z) { return (size_t)d.plane * (size_t)wrap_any(z, d.nz); };
  unsigned long long first = ~0ull;
  auto check = [&](T v, size_t idx) {
    if (!(v > T(0))) first = min(first, (unsigned long long)idx);
  };
  // Register queues, index j of a column queue = plane z - 1 + j (psi) or z + j (D); the last
  // entry receives the plane requested PF iterations before its use.  Halo queue entry j =
  // plane z + j.
  T ps[OR][3 + PF] = {}, ds[OR][2 + PF] = {}, fzm[OR];
  T hps[PF + 1] = {}, hds[PF + 1] = {};
  auto fetch_cols = [&](int z, int j, int jd) {
    const size_t zo = zoff(z);
#pragma unroll
    for (int q = 0; q < OR; ++q) {
      ps[q][j] = psi[zo + coff[q]];
      if (jd >= 0) {
        ds[q][jd] = D[zo + coff[q]];
        check(ds[q][jd], zo + coff[q]);
      }
    }
  };
  auto fetch_halo = [&](int z, int j) {
    if (hs >= 0) {
      const size_t zo = zoff(z);
      hps[j] = psi[zo + hoff];
      hds[j] = D[zo + hoff];
      check(hds[j], zo + hoff);
    }
  };
  {
    T dm[OR];
    const size_t zo = zoff(z0 - 1);
#pragma unroll
    for (int q = 0; q < OR; ++q) {
      dm[q] = D[zo + coff[q]];
      check(dm[q], zo + coff[q]);
    }
    fetch_cols(z0 - 1, 0, -1);
#pragma unroll
    for (int j = 0; j <= PF; ++j) fetch_cols(z0 + j, j + 1, j);
#pragma unroll
    for (int j = 0; j < PF; ++j) fetch_halo(z0 + j, j);
#pragma unroll
    for (int q = 0; q < OR; ++q) fzm[q] = face_coef(dm[q], ds[q][0], ihz2);
  }
  for (int k = 0; k < nout; ++k) {
    const int b = k & 1;
    T* P = pb[b];
    T* Q = db[b];
#pragma unroll
    for (int q = 0; q < OR; ++q) {
      const int o = (oyb + q + 1) * PW + ox + 1;
      P[o] = ps[q][1];
      Q[o] = ds[q][0];
    }
    if (hs >= 0) {
      P[hs] = hps[0];
      Q[hs] = hds[0];
    }
    __syncthreads();
    if (k + PF < nout) {
      fetch_halo(z0 + k + PF, PF);
      fetch_cols(z0 + k + 1 + PF, 2 + PF, 1 + PF);
    }
    const size_t zo = (size_t)d.plane * (size_t)(z0 + k);
    T fyb = face_coef(Q[oyb * PW + ox + 1], ds[0][0], ihy2);  // face below row oyb
#pragma unroll
    for (int q = 0; q < OR; ++q) {
      const int o = (oyb + q + 1) * PW + ox + 1;
      const T dcq = ds[q][0];
      const T fxm = face_coef(Q[o - 1], dcq, ihx2);
      const T fxp = face_coef(dcq, Q[o + 1], ihx2);
      const T dup = q + 1 < OR ? ds[q + 1][0] : Q[o + PW];
      const T fyp = face_coef(dcq, dup, ihy2);
      const T fzp = face_coef(dcq, ds[q][1], ihz2);
      const T v = fvm_k(ps[q][1], P[o - 1], P[o + 1], P[o - PW], P[o + PW], ps[q][0], ps[q][2],
                        fxm, fxp, fyb, fyp, fzm[q], fzp, ihx2, ihy2, ihz2);
      if (cin[q]) out[zo + (size_t)(x0 + ox) + (size_t)d.nx * (size_t)(y0 + oyb + q)] = v;
      fyb = fyp;
      fzm[q] = fzp;
    }
#pragma unroll
    for (int q = 0; q < OR; ++q) {
#pragma unroll
      for (int j = 0; j < 2 + PF; ++j) ps[q][j] = ps[q][j + 1];
#pragma unroll
      for (int j = 0; j < 1 + PF; ++j) ds[q][j] = ds[q][j + 1];
    }
#pragma unroll
    for (int j = 0; j < PF; ++j) {
      hps[j] = hps[j + 1];
      hds[j] = hds[j + 1];
    }
  }
  for (int st = 32; st >= 1; st >>= 1) {
    const unsigned long long o = __shfl_xor(first, st);
    first = o < first ? o : first;
  }
  if ((t & 63) == 0 && first != ~0ull) atomicMin(bad, first);
}

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v) { *p = v; }


// ---- axis stencils: z in register queues (used for the order-4 Laplacian) --------------
//
// Operators that read only along the axes through a site take their z neighbours from a
// per-column register queue of the 2R+1 planes around the output plane; LDS holds one plane
// with an x/y halo of R for the in-plane neighbours.  The tile sites of that plane come from
// the owners' queues (no second load), the halo sites from a prefetch.  Per site 4R LDS
// reads instead of 6R+1 from a (2R+1)-plane ring, and one LDS plane instead of 2R+1 (more
// workgroups per CU).  The operator code and its accumulation order are the ring engine's.
template <class T, int R, int PW>
struct AxisAt {
  const T* pl;  // the plane, tile + R halo, row width PW
  int base;     // (oy + R) * PW + (ox + R)
  const T* q;   // the column over z - R .. z + R
  __device__ __forceinline__ T operator()(int, int dx, int dy, int dz) const {
    if (dz != 0 || (dx == 0 && dy == 0)) return q[dz + R];  // axis stencils: dx = dy = 0 here
    return pl[base + dy * PW + dx];
  }
};

template <class T, int OX, int OY, class Op>
__global__ __launch_bounds__(256) void k_axis_zq(Op op, const T* __restrict__ f,
                                                 T* __restrict__ out, Dims d, int ntx, int nty,
                                                 int zlen) {
  constexpr int R = Op::R, NT = 256, PW = OX + 2 * R, PH = OY + 2 * R, PN = PW * PH;
  constexpr int ROWS = NT / OX, OR = OX * OY / NT, PL = (PN + NT - 1) / NT;
  static_assert(NT % OX == 0 && (OX * OY) % NT == 0, "tile must match the block");
  __shared__ T pl[PN];
  const int t = threadIdx.x;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int tx = (int)(bid % (unsigned)ntx);
  const unsigned rr = bid / (unsigned)ntx;
  const int ty = (int)(rr % (unsigned)nty);
  const int zc = (int)(rr / (unsigned)nty);
  const int x0 = tx * OX, y0 = ty * OY, z0 = zc * zlen;
  const int nout = min(zlen, d.nz - z0);
  const int ox = t % OX, oy0 = t / OX;

  // halo sites of the plane this thread stages (tile sites come from the owners' queues)
  unsigned hoff[PL];
  int hslot[PL];
#pragma unroll
  for (int s = 0; s < PL; ++s) {
    const int idx = t + s * NT;
    const int px = idx % PW, py = idx / PW;
    const bool tile = px >= R && px < R + OX && py >= R && py < R + OY;
    hslot[s] = (idx < PN && !tile) ? idx : -1;
    hoff[s] = hslot[s] >= 0
                  ? (unsigned)(wrap_any(x0 + px - R, d.nx) + d.nx * wrap_any(y0 + py - R, d.ny))
                  : 0u;
  }
  unsigned coff[OR];
  bool cin[OR];
#pragma unroll
  for (int q = 0; q < OR; ++q) {
    const int gy = y0 + oy0 + q * ROWS;
    cin[q] = x0 + ox < d.nx && gy < d.ny;
    coff[q] = (unsigned)(wrap_any(x0 + ox, d.nx) + d.nx * wrap_any(gy, d.ny));
  }
  auto plane_ptr = [&](int z) { return f + (size_t)d.plane * (size_t)wrap_any(z, d.nz); };
  T qv[OR][2 * R + 1];  // own columns over z - R .. z + R
#pragma unroll
  for (int j = 0; j < 2 * R + 1; ++j) {
    const T* fp = plane_ptr(z0 - R + j);
#pragma unroll
    for (int q = 0; q < OR; ++q) qv[q][j] = fp[coff[q]];
  }
  T hv[PL], qn[OR];
  {
    const T* fp = plane_ptr(z0);
#pragma unroll
    for (int s = 0; s < PL; ++s)
      if (hslot[s] >= 0) hv[s] = fp[hoff[s]];
  }
  for (int k = 0; k < nout; ++k) {
#pragma unroll
    for (int s = 0; s < PL; ++s)
      if (hslot[s] >= 0) pl[hslot[s]] = hv[s];
#pragma unroll
    for (int q = 0; q < OR; ++q) pl[(oy0 + q * ROWS + R) * PW + ox + R] = qv[q][R];
    __syncthreads();
    if (k + 1 < nout) {
      const T* fp = plane_ptr(z0 + k + 1);
#pragma unroll
      for (int s = 0; s < PL; ++s)
        if (hslot[s] >= 0) hv[s] = fp[hoff[s]];
      const T* fq = plane_ptr(z0 + k + 1 + R);
#pragma unroll
      for (int q = 0; q < OR; ++q) qn[q] = fq[coff[q]];
    }
    const size_t zo = (size_t)d.plane * (size_t)(z0 + k);
#pragma unroll
    for (int q = 0; q < OR; ++q) {
      const int oy = oy0 + q * ROWS;
      if (cin[q]) {
        const AxisAt<T, R, PW> F{pl, (oy + R) * PW + (ox + R), qv[q]};
        op(F, out, zo + (size_t)(x0 + ox) + (size_t)d.nx * (size_t)(y0 + oy));
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < OR; ++q) {
#pragma unroll
      for (int j = 0; j < 2 * R; ++j) qv[q][j] = qv[q][j + 1];
      qv[q][2 * R] = qn[q];
    }
  }
}

template <class T, class Op>
void axis_launch(gc_context* ctx, const Dims& d, const Op& op, const T* in, T* out) {
  constexpr int OY = ZRows<Op>::OY;
  const int ntx = (d.nx + kZOX - 1) / kZOX, nty = (d.ny + OY - 1) / OY;
  const int nzc = (d.nz + kZLen - 1) / kZLen;
  hipLaunchKernelGGL((k_axis_zq<T, kZOX, OY, Op>), dim3((unsigned)(ntx * nty * nzc)),
                     dim3(256), 0, ctx->stream, op, in, out, d, ntx, nty, kZLen);
}

// Calls f(integral_constant<int, P>) for each P of the sequence, in order.
template <int... Ps, class F>
__device__ __forceinline__ void unroll_seq(std::integer_sequence<int, Ps...>, F&& f) {
  (f(std::integral_constant<int, Ps>{}), ...);
}

// ---- FP64 biharmonic: z-accumulating sums with shared products -----------------------------
//
// Biharmonic.hpp:69-109 sums each of dz4, dxz and dyz with the z offset in the outer position,
// so a site's terms arrive plane by plane in the order the z-march reads the planes.  Here
// every input plane is staged in LDS once and its contributions are added to the running sums
// of every site it belongs to: dz4 of the 2 R4 + 1 sites z - R4 .. z + R4, dxz and dyz of the
// 2 R2 + 1 sites z - R2 .. z + R2 (in the reference's inner order), while dx4, dy4 and dxy of
// the site in the plane itself are formed in one go.  A value times a weight is the same
// rounded product whichever site it serves, so each distinct (weight, value) product is formed
// once per thread and plane (the unrolled code names each product once per value and weight;
// the symmetric tables make many of them equal): for order 4, 63 multiplications per site
// instead of 96, for the same 96 additions, same operands, same order, same bits.  Weights
// that are powers of two (every order-2 mixed weight, the order-2 -4 and the order-4 2) take
// fused multiply-adds, which round the exact product once, like the reference's separate
// product and sum (acc_w).  No value queues: each thread owns NY rows of one column and
// keeps, per row, the running sums of the sites in flight (dz4: 2 R4, dxz and dyz: 2 R2) and,
// for the R4 sites between their own plane and their last dz4 term, A = dx4 ihx4 + dy4 ihy4
// and X = dxy ihx2 ihy2 (turned into T = 2 (X + dxz ihx2 ihz2 + dyz ihy2 ihz2) when dxz and
// dyz complete); the plane loop is unrolled by the number of site slots L so these live in
// fixed registers.  The output of site z - R4 is (A + dz4 ihz4) + T, the reference's
// expression order.  PB planes per barrier: the planes of step q are staged in buffer pair
// q & 1, whose previous readers (step q - 2) have passed the barrier of step q - 1; the next
// PB planes are requested right after the barrier.  A chunk of nout output planes reads
// nout + 2 R4 planes; the R4 planes at each end feed only sums of sites outside the chunk,
// which are formed but never stored (the in-plane sums are skipped there).  32-wide tiles keep
// a thread's whole 7 x 7 window within the immediate-offset reach of one LDS address.

// acc + fl(w v) for a compile-time weight w.  When |w| is a power of two, w v is exact (short of
// overflow, |v| > DBL_MAX / |w|), so the fused multiply-add rounds once to the same value;
// w = +-1 is a plain addition or subtraction.
__host__ __device__ constexpr bool pow2_weight(double a) {
  a = a < 0 ? -a : a;
  if (a == 0.0) return false;
  while (a >= 2.0) a *= 0.5;
  while (a < 1.0) a *= 2.0;
  return a == 1.0;
}
template <class T, class W>
__device__ __forceinline__ T acc_w(T acc, W wc, T v) {
  constexpr double w = W::value();
  if constexpr (w == 1.0) return acc + v;
  else if constexpr (w == -1.0) return acc - v;
  else if constexpr (pow2_weight(w)) return __builtin_fma(T(w), v, acc);
  else return acc + T(w) * v;
}
template <class S, int M>
struct W1 {  // a one-axis weight
  static constexpr double value() { return S::w(M); }
};
template <class S, int MA, int MB>
struct W2 {  // a mixed weight w(ma) * w(mb), formed in double as the reference does
  static constexpr double value() { return S::w(MA) * S::w(MB); }
};

template <class T, int O, int OX, int NY, int PB, int WPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE, 8))) void k_bih_za(
    OpBih<T, O> op, const T* __restrict__ f, T* __restrict__ out, Dims d, int ntx, int nty,
    int zlen) {
  using S2 = D2<O>;
  using S4 = D4<O>;
  constexpr int R4 = S4::R, R2 = S2::R, N2 = S2::N, N4 = S4::N;
  static_assert(R4 > R2, "the mixed sums complete before dz4");
  static_assert(PB == 1 || PB == 2, "planes per barrier");
  // site slots = unroll period: 2 R4 + 1 sites are in flight; a multiple of 2 PB, so the LDS
  // buffer pair alternates at compile time
  constexpr int L = ((2 * R4 + 1 + 2 * PB - 1) / (2 * PB)) * (2 * PB);
  constexpr int NT = 256, OY = (NT / OX) * NY;
  constexpr int PW = OX + 2 * R4, PH = OY + 2 * R4, PN = PW * PH;
  constexpr int PL = (PN + NT - 1) / NT;
  __shared__ T pb[2 * PB * PN];
  const int t = threadIdx.x;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int tx = (int)(bid % (unsigned)ntx);
  const unsigned rr = bid / (unsigned)ntx;
  const int ty = (int)(rr % (unsigned)nty);
  const int zc = (int)(rr / (unsigned)nty);
  const int x0 = tx * OX, y0 = ty * OY, z0 = zc * zlen;
  const int nout = min(zlen, d.nz - z0);
  const int np = nout + 2 * R4;

  unsigned poff[PL];
  bool pin[PL];
#pragma unroll
  for (int s = 0; s < PL; ++s) {
    const int idx = t + s * NT;
    pin[s] = idx < PN;
    const int px = idx % PW, py = idx / PW;
    poff[s] = pin[s] ? (unsigned)(wrap_any(x0 + px - R4, d.nx) +
                                  d.nx * wrap_any(y0 + py - R4, d.ny))
                     : 0u;
  }
  T reg[PB][PL];
  auto load = [&](T* r, int plane) {
    const size_t zo = (size_t)d.plane * (size_t)wrap_any(plane, d.nz);
#pragma unroll
    for (int s = 0; s < PL; ++s)
      if (pin[s]) r[s] = f[zo + poff[s]];
  };
  const int ox = t % OX, oyt = (t / OX) * NY;
  // the thread's window origin (x - R4, y - R4): every LDS offset below is non-negative, so
  // the reads take immediate offsets from one base register
  const int base = oyt * PW + ox;
  // that origin in each of the 2 PB staged planes, kept in registers as they are: folded into
  // every address, the buffer offsets push the immediates out of the ds_read2 range and each
  // read pair would get an address addition of its own
  int wofs[2 * PB];
#pragma unroll
  for (int i = 0; i < 2 * PB; ++i) {
    wofs[i] = base + i * PN;
    asm volatile("" : "+v"(wofs[i]));
  }
  bool vxy[NY];
  size_t ocol[NY];
#pragma unroll
  for (int j = 0; j < NY; ++j) {
    const int gy = y0 + oyt + j;
    vxy[j] = x0 + ox < d.nx && gy < d.ny;
    ocol[j] = (size_t)(x0 + ox) + (size_t)d.nx * (size_t)gy;
  }
  // per row and site slot (site plane - (z0 - R4)) mod L: running sums and finished parts
  T dz[NY][L], xz[NY][L], yz[NY][L], A[NY][L], X[NY][L];
#pragma unroll
  for (int j = 0; j < NY; ++j)
#pragma unroll
    for (int s = 0; s < L; ++s) dz[j][s] = xz[j][s] = yz[j][s] = A[j][s] = X[j][s] = T(0);

#pragma unroll
  for (int i = 0; i < PB; ++i)
    if (i < np) load(reg[i], z0 - R4 + i);
  // plane z0 - R4 + p from staged plane `st` (P = p mod L)
  auto plane = [&](auto Pc, int p, int st) {
      constexpr int P = decltype(Pc)::value;
      const int w0 = wofs[st];
      auto V = [&](int dx, int dy) { return pb[w0 + (dy + R4) * PW + (dx + R4)]; };
      // slot of the site o planes below this one
      auto sl = [](int o) { return ((P - o) % L + L) % L; };
      const bool inplane = p >= R4 && p - R4 < nout;
#pragma unroll
      for (int j = 0; j < NY; ++j) {
        const T c0 = V(0, j);
        using MS4 = std::make_integer_sequence<int, N4>;
        using MS2 = std::make_integer_sequence<int, N2>;
        // dz4 of the sites z - R4 .. z + R4 (this plane at offset o = m - R4 from the site)
        unroll_seq(MS4{}, [&](auto Mc) {
          constexpr int m = decltype(Mc)::value, o = m - R4;
          T& acc = dz[j][sl(o)];
          acc = acc_w(o == -R4 ? T(0) : acc, W1<S4, m>{}, c0);
        });
        // dxz and dyz of the sites z - R2 .. z + R2 (z offset outer, x or y offset inner)
        unroll_seq(MS2{}, [&](auto Bc) {
          constexpr int mb = decltype(Bc)::value, o = mb - R2;
          T& ax = xz[j][sl(o)];
          T& ay = yz[j][sl(o)];
          unroll_seq(MS2{}, [&](auto Ac) {
            constexpr int ma = decltype(Ac)::value;
            const bool first = o == -R2 && ma == 0;
            ax = acc_w(first ? T(0) : ax, W2<S2, ma, mb>{}, V(ma - R2, j));
          });
          unroll_seq(MS2{}, [&](auto Ac) {
            constexpr int ma = decltype(Ac)::value;
            const bool first = o == -R2 && ma == 0;
            ay = acc_w(first ? T(0) : ay, W2<S2, ma, mb>{}, V(0, j + ma - R2));
          });
        });
        // the mixed sums of the site R2 planes below are complete
        {
          constexpr int s = (((P - R2) % L) + L) % L;
          X[j][s] = T(2) * ((X[j][s] + xz[j][s] * op.ihx2 * op.ihz2) +
                            yz[j][s] * op.ihy2 * op.ihz2);
        }
        if (inplane) {
          T dx4 = 0, dy4 = 0;
          unroll_seq(MS4{}, [&](auto Mc) {
            constexpr int m = decltype(Mc)::value;
            dx4 = acc_w(dx4, W1<S4, m>{}, V(m - R4, j));
            dy4 = acc_w(dy4, W1<S4, m>{}, V(0, j + m - R4));
          });
          T dxy = 0;
          unroll_seq(MS2{}, [&](auto Bc) {
            constexpr int mb = decltype(Bc)::value;
            unroll_seq(MS2{}, [&](auto Ac) {
              constexpr int ma = decltype(Ac)::value;
              dxy = acc_w(dxy, W2<S2, ma, mb>{}, V(ma - R2, j + mb - R2));
            });
          });
          A[j][P] = dx4 * op.ihx4 + dy4 * op.ihy4;
          X[j][P] = dxy * op.ihx2 * op.ihy2;
        }
      }
      // the site R4 planes below is complete
      const int so = p - 2 * R4;
      if (so >= 0 && so < nout) {
        constexpr int s = (((P - R4) % L) + L) % L;
        const size_t zo = (size_t)d.plane * (size_t)(z0 + so);
#pragma unroll
        for (int j = 0; j < NY; ++j)
          if (vxy[j]) out[zo + ocol[j]] = (A[j][s] + dz[j][s] * op.ihz4) + X[j][s];
      }
  };
  for (int kb = 0; kb < np; kb += L) {
    // PB planes per barrier: stage them, request the next PB, then form each
    unroll_seq(std::make_integer_sequence<int, L / PB>{}, [&](auto Qc) {
      constexpr int Q = decltype(Qc)::value;
      const int p0 = kb + Q * PB;
      if (p0 >= np) return;
      T* b = pb + (Q & 1) * PB * PN;
#pragma unroll
      for (int i = 0; i < PB; ++i)
        if (p0 + i < np) {
#pragma unroll
          for (int s = 0; s < PL; ++s)
            if (pin[s]) b[i * PN + t + s * NT] = reg[i][s];
        }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < PB; ++i)
        if (p0 + PB + i < np) load(reg[i], z0 - R4 + p0 + PB + i);
      unroll_seq(std::make_integer_sequence<int, PB>{}, [&](auto Ic) {
        constexpr int i = decltype(Ic)::value;
        if (p0 + i < np) plane(std::integral_constant<int, Q * PB + i>{}, p0 + i, (Q & 1) * PB + i);
      });
    });
  }
}

// ---- FP32 biharmonic: separable form on a z-march ------------------------------------------
//
// The FP32 form (this library's own arithmetic, compared with the FP64 oracle within stated
// bounds) uses that the mixed fourth derivatives of the tensor-product stencil factor:
//   Σ_a Σ_b w2(a) w2(b) f(x+a, y+b) = D2y(D2x f),  likewise D2z(D2x f) and D2z(D2y f),
// so per plane a = D2x f and b = D2y f are formed once per site (a also on the R2 rows above
// and below the tile, staged in LDS for D2y), and every z sum comes from per-column register
// queues.  Per site about 21 LDS reads and 60 fused multiply-adds, against 96 LDS reads and
// 100 multiply-adds for the per-site mixed sums.  Iteration k stages plane p = z0 - R4 + k
// (tile sites from the owners' queues, the halo from a prefetch), forms a, b, D4x f, D4y f and
// D2y a on it, and finishes output plane p - R4:
//   out = D4x ihx4 + D4y ihy4 + D4z ihz4 + 2 (D2y a ihx2 ihy2 + D2z a ihx2 ihz2 + D2z b ihy2 ihz2).
struct BihSepCoef {
  float ihx4, ihy4, ihz4, cxy, cxz, cyz;  // cxy = 2 ihx2 ihy2, ...
};

template <int O, int OY>
__global__ __launch_bounds__(256) void k_bih_sep(BihSepCoef c, const float* __restrict__ f,
                                                 float* __restrict__ out, Dims d, int ntx,
                                                 int nty, int zlen) {
  using S2 = D2<O>;
  using S4 = D4<O>;
  constexpr int R4 = S4::R, R2 = S2::R, OX = 64, NT = 256;
  constexpr int PW = OX + 2 * R4, PH = OY + 2 * R4, PN = PW * PH;  // f plane, halo R4
  constexpr int AH = OY + 2 * R2, AN = OX * AH;                     // a plane, y halo R2
  constexpr int ROWS = NT / OX, OR = OX * OY / NT, PL = (PN + NT - 1) / NT;
  constexpr int NAH = 2 * R2 * OX, LA = (NAH + NT - 1) / NT;       // a halo sites
  constexpr int QF = 2 * R4 + 1, QA = R4 + R2 + 1, QH = R4 + 1;     // queue lengths
  static_assert((OX * OY) % NT == 0 && R4 >= R2, "tile must match the block");
  __shared__ float fp[PN];
  __shared__ float ap[AN];
  const int t = threadIdx.x;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int tx = (int)(bid % (unsigned)ntx);
  const unsigned rr = bid / (unsigned)ntx;
  const int ty = (int)(rr % (unsigned)nty);
  const int zc = (int)(rr / (unsigned)nty);
  const int x0 = tx * OX, y0 = ty * OY, z0 = zc * zlen;
  const int nout = min(zlen, d.nz - z0);
  const int ox = t % OX, oy0 = t / OX;

  // halo sites of the f plane this thread stages (tile sites come from the owners' queues)
  unsigned hoff[PL];
  int hslot[PL];
#pragma unroll
  for (int s = 0; s < PL; ++s) {
    const int idx = t + s * NT;
    const int px = idx % PW, py = idx / PW;
    const bool tile = px >= R4 && px < R4 + OX && py >= R4 && py < R4 + OY;
    hslot[s] = (idx < PN && !tile) ? idx : -1;
    hoff[s] = hslot[s] >= 0
                  ? (unsigned)(wrap_any(x0 + px - R4, d.nx) + d.nx * wrap_any(y0 + py - R4, d.ny))
                  : 0u;
  }
  // a-plane halo rows this thread forms: a-plane row ar, column ac
  int ahr[LA], ahc[LA];
#pragma unroll
  for (int s = 0; s < LA; ++s) {
    const int idx = t + s * NT;
    if (idx < NAH) {
      const int hr = idx / OX;  // 0 .. 2 R2 - 1
      ahr[s] = hr < R2 ? hr : OY + hr;
      ahc[s] = idx % OX;
    } else {
      ahr[s] = -1;
      ahc[s] = 0;
    }
  }
  unsigned coff[OR];
  bool cin[OR];
#pragma unroll
  for (int q = 0; q < OR; ++q) {
    const int gy = y0 + oy0 + q * ROWS;
    cin[q] = x0 + ox < d.nx && gy < d.ny;
    coff[q] = (unsigned)(wrap_any(x0 + ox, d.nx) + d.nx * wrap_any(gy, d.ny));
  }
  auto plane_ptr = [&](int z) { return f + (size_t)d.plane * (size_t)wrap_any(z, d.nz); };
  // register rings of period QP = 2 R4 + 1: plane p_k = z0 - R4 + k sits in slot k mod QP of
  // each ring (f over the last QP planes, a and b over the last R4 + R2 + 1, the in-plane part
  // over the last R4 + 1); the plane loop is unrolled by QP so every slot is a compile-time
  // register and nothing is moved between planes
  constexpr int QP = QF;
  static_assert(QA <= QP && QH <= QP, "ring period");
  float fq[OR][QP] = {}, aq[OR][QP] = {}, bq[OR][QP] = {}, hq[OR][QP] = {};
  float hv[PL], fn[OR];
  auto prefetch = [&](int z) {
    const float* src = plane_ptr(z);
#pragma unroll
    for (int s = 0; s < PL; ++s)
      if (hslot[s] >= 0) hv[s] = src[hoff[s]];
#pragma unroll
    for (int q = 0; q < OR; ++q) fn[q] = src[coff[q]];
  };
  prefetch(z0 - R4);
  const int iters = nout + 2 * R4;
  for (int kb = 0; kb < iters; kb += QP) {
    unroll_seq(std::make_integer_sequence<int, QP>{}, [&](auto Pc) {
      constexpr int P = decltype(Pc)::value;  // k mod QP
      const int k = kb + P;
      if (k >= iters) return;  // uniform over the block
      auto slot = [](int back) { return (P - back + 2 * QP) % QP; };  // plane p_k - back
#pragma unroll
      for (int q = 0; q < OR; ++q) fq[q][P] = fn[q];
#pragma unroll
      for (int s = 0; s < PL; ++s)
        if (hslot[s] >= 0) fp[hslot[s]] = hv[s];
#pragma unroll
      for (int q = 0; q < OR; ++q) fp[(oy0 + q * ROWS + R4) * PW + ox + R4] = fq[q][P];
      __syncthreads();
      if (k + 1 < iters) prefetch(z0 - R4 + k + 1);
      // a = D2x f on the a-plane halo rows
#pragma unroll
      for (int s = 0; s < LA; ++s) {
        if (ahr[s] >= 0) {
          const float* row = fp + (ahr[s] - R2 + R4) * PW + ahc[s] + R4;
          float a = 0.f;
#pragma unroll
          for (int m = 0; m < S2::N; ++m) a = fmaf(float(S2::w(m)), row[m - R2], a);
          ap[ahr[s] * OX + ahc[s]] = a;
        }
      }
      float d4x[OR], d4y[OR];
#pragma unroll
      for (int q = 0; q < OR; ++q) {
        const int oy = oy0 + q * ROWS;
        const float* cen = fp + (oy + R4) * PW + ox + R4;
        float a = 0.f, b = 0.f, x4 = 0.f, y4 = 0.f;
#pragma unroll
        for (int m = 0; m < S2::N; ++m) {
          const int o = m - R2;
          a = fmaf(float(S2::w(m)), cen[o], a);
          b = fmaf(float(S2::w(m)), cen[o * PW], b);
        }
#pragma unroll
        for (int m = 0; m < S4::N; ++m) {
          const int o = m - R4;
          x4 = fmaf(float(S4::w(m)), cen[o], x4);
          y4 = fmaf(float(S4::w(m)), cen[o * PW], y4);
        }
        ap[(oy + R2) * OX + ox] = a;
        aq[q][P] = a;
        bq[q][P] = b;
        d4x[q] = x4;
        d4y[q] = y4;
      }
      __syncthreads();
      const int z = z0 - 2 * R4 + k;  // the output plane this iteration completes
#pragma unroll
      for (int q = 0; q < OR; ++q) {
        const int oy = oy0 + q * ROWS;
        float axy = 0.f;
#pragma unroll
        for (int m = 0; m < S2::N; ++m) axy = fmaf(float(S2::w(m)), ap[(oy + m) * OX + ox], axy);
        hq[q][P] = fmaf(d4x[q], c.ihx4, fmaf(d4y[q], c.ihy4, axy * c.cxy));
        if (k >= 2 * R4 && cin[q]) {
          // plane z + o is p_k - (R4 - o)
          float z4 = 0.f, az = 0.f, bz = 0.f;
#pragma unroll
          for (int m = 0; m < S4::N; ++m) z4 = fmaf(float(S4::w(m)), fq[q][slot(2 * R4 - m)], z4);
#pragma unroll
          for (int m = 0; m < S2::N; ++m) {
            az = fmaf(float(S2::w(m)), aq[q][slot(R4 + R2 - m)], az);
            bz = fmaf(float(S2::w(m)), bq[q][slot(R4 + R2 - m)], bz);
          }
          out[(size_t)d.plane * (size_t)z + (size_t)(x0 + ox) + (size_t)d.nx * (size_t)(y0 + oy)] =
              fmaf(z4, c.ihz4, fmaf(az, c.cxz, fmaf(bz, c.cyz, hq[q][slot(R4)])));
        }
      }
    });
  }
}

// ---- FP32 order-2 biharmonic: the Laplacian of the Laplacian -----------------------------
//
// With the order-2 stencils the fourth-derivative weights (1, -4, 6, -4, 1) are the second-
// derivative weights (1, -2, 1) applied twice and every mixed term is a product of two second
// differences, so the order-2 biharmonic equals L(L f) for the 7-point Laplacian L with
// weights 1/h^2 (ihx4 = ihx2^2).  The FP32 form evaluates it that way: per plane l = L f on
// the tile + 1 ring (LDS plane of f with a 2-site halo), then L l on the tile (LDS plane of l
// with a 1-site halo); z neighbours of both come from per-column register queues.  Each f
// and l plane is double-buffered in LDS, so one barrier per plane suffices: iteration k
// stages f of plane q = z0 - 3 + k (tile and ring-1 columns from the owners' queues, the
// outer ring from a prefetch), forms l(q), and forms the output plane q - 1 from l(q - 1)
// (staged in the previous iteration) and the l queue.
template <int OY>
__global__ __launch_bounds__(256) void k_bih_ll(LapCoef<float> lc, const float* __restrict__ f,
                                                float* __restrict__ out, Dims d, int ntx,
                                                int nty, int zlen) {
  constexpr int OX = 64, NT = 256, ROWS = NT / OX, OR = OX * OY / NT;
  constexpr int FW = OX + 4, FH = OY + 4, FN = FW * FH;  // f plane, halo 2
  constexpr int LW = OX + 2, LH = OY + 2, LN = LW * LH;  // l plane, halo 1
  constexpr int NR1 = LN - OX * OY, NR2 = FN - LN;       // ring-1 columns, outer-ring sites
  static_assert(NR1 <= NT && NR2 <= NT && (OX * OY) % NT == 0, "one ring site per thread");
  __shared__ float fb[2][FN];
  __shared__ float lb[2][LN];
  const int t = threadIdx.x;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int tx = (int)(bid % (unsigned)ntx);
  const unsigned rr = bid / (unsigned)ntx;
  const int ty = (int)(rr % (unsigned)nty);
  const int zc = (int)(rr / (unsigned)nty);
  const int x0 = tx * OX, y0 = ty * OY, z0 = zc * zlen;
  const int nout = min(zlen, d.nz - z0);
  const int ox = t % OX, oy0 = t / OX;
  auto goff = [&](int px, int py) {  // f-plane coordinates (halo 2) -> wrapped plane offset
    return (unsigned)(wrap_any(x0 + px - 2, d.nx) + d.nx * wrap_any(y0 + py - 2, d.ny));
  };
  // tile columns
  unsigned coff[OR];
  bool cin[OR];
#pragma unroll
  for (int q = 0; q < OR; ++q) {
    const int oy = oy0 + q * ROWS;
    coff[q] = goff(ox + 2, oy + 2);
    cin[q] = x0 + ox < d.nx && y0 + oy < d.ny;
  }
  // the ring-1 column this thread owns (l on it), in l-plane coordinates
  int r1x = -1, r1y = 0;
  if (t < NR1) {
    if (t < LW) {
      r1x = t;
      r1y = 0;
    } else if (t < 2 * LW) {
      r1x = t - LW;
      r1y = LH - 1;
    } else {
      const int r = t - 2 * LW;
      r1y = 1 + (r >> 1);
      r1x = (r & 1) ? LW - 1 : 0;
    }
  }
  const unsigned r1off = r1x >= 0 ? goff(r1x + 1, r1y + 1) : 0u;
  // the outer-ring site this thread stages, in f-plane coordinates
  int r2s = -1;
  unsigned r2off = 0u;
  if (t < NR2) {
    int px, py;
    if (t < FW) {
      px = t;
      py = 0;
    } else if (t < 2 * FW) {
      px = t - FW;
      py = FH - 1;
    } else {
      const int r = t - 2 * FW;
      py = 1 + (r >> 1);
      px = (r & 1) ? FW - 1 : 0;
    }
    r2s = py * FW + px;
    r2off = goff(px, py);
  }
  auto plane_ptr = [&](int z) { return f + (size_t)d.plane * (size_t)wrap_any(z, d.nz); };
  float fq[OR][3] = {}, f1q[3] = {}, lq[OR][3] = {};
  float fn[OR], f1n = 0.f, r2v = 0.f;
  auto fetch = [&](int z_cols, int z_ring2) {  // columns of plane z_cols, outer ring of z_ring2
    const float* pc = plane_ptr(z_cols);
#pragma unroll
    for (int q = 0; q < OR; ++q) fn[q] = pc[coff[q]];
    if (r1x >= 0) f1n = pc[r1off];
    if (r2s >= 0) r2v = plane_ptr(z_ring2)[r2off];
  };
  fetch(z0 - 2, z0 - 3);
  const int iters = nout + 4;
  for (int k = 0; k < iters; ++k) {
    const int b = k & 1;
    // plane p = z0 - 2 + k arrives; the queues then hold p - 2, p - 1, p
#pragma unroll
    for (int q = 0; q < OR; ++q) {
      fq[q][0] = fq[q][1];
      fq[q][1] = fq[q][2];
      fq[q][2] = fn[q];
    }
    f1q[0] = f1q[1];
    f1q[1] = f1q[2];
    f1q[2] = f1n;
    // stage f of plane q = p - 1
    float* F = fb[b];
#pragma unroll
    for (int q = 0; q < OR; ++q) F[(oy0 + q * ROWS + 2) * FW + ox + 2] = fq[q][1];
    if (r1x >= 0) F[(r1y + 1) * FW + r1x + 1] = f1q[1];
    if (r2s >= 0) F[r2s] = r2v;
    __syncthreads();
    if (k + 1 < iters) fetch(z0 - 1 + k, z0 - 2 + k);
    // l(q) on the tile and the ring-1 column
    float* Lp = lb[b];
#pragma unroll
    for (int q = 0; q < OR; ++q) {
      const int oy = oy0 + q * ROWS;
      const int o = (oy + 2) * FW + ox + 2;
      const float l = lap7(lc, fq[q][1], F[o - 1], F[o + 1], F[o - FW], F[o + FW], fq[q][0], fq[q][2]);
      Lp[(oy + 1) * LW + ox + 1] = l;
      lq[q][0] = lq[q][1];
      lq[q][1] = lq[q][2];
      lq[q][2] = l;
    }
    if (r1x >= 0) {
      const int o = (r1y + 1) * FW + r1x + 1;
      Lp[r1y * LW + r1x] = lap7(lc, f1q[1], F[o - 1], F[o + 1], F[o - FW], F[o + FW], f1q[0], f1q[2]);
    }
    // output plane q - 1 = z0 - 4 + k from l(q - 1) (staged in the previous iteration)
    if (k >= 4) {
      const float* Lm = lb[b ^ 1];
      const size_t zo = (size_t)d.plane * (size_t)(z0 - 4 + k);
#pragma unroll
      for (int q = 0; q < OR; ++q) {
        const int oy = oy0 + q * ROWS;
        const int o = (oy + 1) * LW + ox + 1;
        if (cin[q])
          out[zo + (size_t)(x0 + ox) + (size_t)d.nx * (size_t)(y0 + oy)] =
              lap7(lc, lq[q][1], Lm[o - 1], Lm[o + 1], Lm[o - LW], Lm[o + LW], lq[q][0], lq[q][2]);
      }
    }
  }
}

// ---- vector- and tensor-field operators: z-accumulating engine ----------------------------
//
// The inputs are AoS (CIN values per site).  A block owns a 64 x 8 tile (one column per
// thread) and marches a chunk of planes.  Each plane's own-column sites are loaded once, all
// components together: a kernel that fetched a plane's z-differentiated components R planes
// before its x- and y-differentiated ones found the lines gone from L2 the second time and
// read the input 2-2.5 times from HBM.  Components differentiated along z are not queued:
// plane p adds its term w[p-z'+R] v(p) to the running z sum of every output plane
// z' = p-R..p+R on arrival, so each output's terms come in increasing offset, the reference's
// order from 0.  Components differentiated along x and y are staged in LDS with an x or y halo
// when the plane is itself an output plane, and the in-plane part of the result is kept in
// registers until that output's z sum completes R planes later.  The z sum and in-plane part
// of output plane z0+j sit in register j mod (2R+1); the plane loop is unrolled by 2R+1 so
// every such index is a compile-time constant.

template <int R, int TOY>
struct VecTile {
  static constexpr int OX = 64, OY = TOY, NT = 64 * TOY;  // one column per thread
  static constexpr int PXW = OX + 2 * R;            // x-staged plane width
  static constexpr int PYH = OY + 2 * R;            // y-staged plane height
  static constexpr int NXH = 2 * R * OY, NYH = 2 * R * OX;  // halo sites
  static constexpr int LXH = (NXH + NT - 1) / NT, LYH = (NYH + NT - 1) / NT;
};

template <class T, int O, class Op, int TOY>
__global__ __launch_bounds__(64 * TOY) void k_vec_za(Op op, const T* __restrict__ in,
                                                T* __restrict__ out, Dims d, int ntx, int nty,
                                                int zlen) {
  using S = D1<O>;
  constexpr int R = S::R, QP = 2 * R + 1;
  constexpr int NXc = Op::NXc, NYc = Op::NYc, NZc = Op::NZc, CIN = Op::CIN, NTQ = Op::NTQ;
  using V = VecTile<R, TOY>;
  __shared__ T px[NXc][V::OY * V::PXW];
  __shared__ T py[NYc][V::PYH * V::OX];
  const int t = threadIdx.x;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int tx = (int)(bid % (unsigned)ntx);
  const unsigned rr = bid / (unsigned)ntx;
  const int ty = (int)(rr % (unsigned)nty);
  const int zc = (int)(rr / (unsigned)nty);
  const int x0 = tx * V::OX, y0 = ty * V::OY, z0 = zc * zlen;
  const int nout = min(zlen, d.nz - z0);
  const int ox = t % V::OX, oy = t / V::OX;
  // own column, wrapped (a column past the grid edge is another tile's halo)
  const bool cin = x0 + ox < d.nx && y0 + oy < d.ny;
  const unsigned coff = (unsigned)(wrap_any(x0 + ox, d.nx) + d.nx * wrap_any(y0 + oy, d.ny));
  // x halo: columns -R..-1 and OX..OX+R-1 of each row; y halo: rows -R..-1, OY..OY+R-1
  unsigned xhoff[V::LXH], yhoff[V::LYH];
  int xhslot[V::LXH], yhslot[V::LYH];
#pragma unroll
  for (int s = 0; s < V::LXH; ++s) {
    const int h = t + s * V::NT;
    const int row = h / (2 * R), k = h % (2 * R);
    const int lx = k < R ? k : V::OX + k;  // plane column (0..PXW-1)
    xhslot[s] = h < V::NXH ? row * V::PXW + lx : -1;
    xhoff[s] = h < V::NXH ? (unsigned)(wrap_any(x0 + lx - R, d.nx) + d.nx * wrap_any(y0 + row, d.ny)) : 0u;
  }
#pragma unroll
  for (int s = 0; s < V::LYH; ++s) {
    const int h = t + s * V::NT;
    const int col = h % V::OX, k = h / V::OX;
    const int ly = k < R ? k : V::OY + k;  // plane row (0..PYH-1)
    yhslot[s] = h < V::NYH ? ly * V::OX + col : -1;
    yhoff[s] = h < V::NYH ? (unsigned)(wrap_any(x0 + col, d.nx) + d.nx * wrap_any(y0 + ly - R, d.ny)) : 0u;
  }

  // More than 16 bytes of output per site: a wave is one tile row (OX = 64 = the wave size), so
  // it stages its row's outputs in LDS and writes them as runs of 64 consecutive words -- each
  // store instruction covers whole lines instead of one word of each of 64 sites NOUT words
  // apart.  Only the wave itself touches its row buffer: wave-level ordering suffices.  512^3:
  // gradient of a vector field FP64 3.98 -> 2.60 ms (order 2), 4.32 -> 2.99 ms (order 4), FP32
  // 1.56 -> 1.34 ms (order 2); 12-byte sites (FP32 divergence of a tensor) lost 3 % staged.
  constexpr int NO = Op::NOUT;
  constexpr bool kRow = NO > 1 && NO * sizeof(T) > 16;
  __shared__ T orow[kRow ? TOY : 1][kRow ? V::OX * NO : 1];
  const int rowlen = y0 + oy < d.ny ? min(V::OX, d.nx - x0) * NO : 0;  // output words of the row
  auto wave_sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };

  T own[CIN];                    // the own site of the plane being processed, every component
  T xh[V::LXH][NXc], yh[V::LYH][NYc];
  T acc[NZc][QP];                // z sums of output planes z0 + j at register j mod QP
  T tq[NTQ][QP];                 // their in-plane parts
  const int steps = nout + 2 * R;  // planes z0-R .. z0+nout-1+R
  auto is_out = [&](int st) { return st >= R && st < R + nout; };
  auto load = [&](int st) {
    const size_t po = (size_t)d.plane * (size_t)wrap_any(z0 - R + st, d.nz);
    const T* sp = in + (size_t)CIN * (po + coff);
#pragma unroll
    for (int i = 0; i < CIN; ++i) own[i] = sp[i];
    if (is_out(st)) {
#pragma unroll
      for (int s = 0; s < V::LXH; ++s)
        if (xhslot[s] >= 0)
#pragma unroll
          for (int i = 0; i < NXc; ++i) xh[s][i] = in[(size_t)CIN * (po + xhoff[s]) + Op::xc(i)];
#pragma unroll
      for (int s = 0; s < V::LYH; ++s)
        if (yhslot[s] >= 0)
#pragma unroll
          for (int i = 0; i < NYc; ++i) yh[s][i] = in[(size_t)CIN * (po + yhoff[s]) + Op::yc(i)];
    }
  };
  load(0);
  const size_t obase = (size_t)d.plane * (size_t)z0 + coff;
  for (int sb = 0; sb < steps; sb += QP) {
    unroll_seq(std::make_integer_sequence<int, QP>{}, [&](auto Pc) {
      constexpr int P = decltype(Pc)::value;  // step st = sb + P, st mod QP = P
      const int st = sb + P;
      if (st >= steps) return;
      // z sums: this plane's term for output j = st - R - o (offset o)
#pragma unroll
      for (int m = 0; m < S::N; ++m) {
        const int o = m - S::R;
        const int j = st - R - o;
        const int slot = ((P - R - o) % QP + QP) % QP;
        if (j >= 0 && j < nout) {
          const T w = T(S::w(m));
#pragma unroll
          for (int i = 0; i < NZc; ++i) {
            if (m == 0) acc[i][slot] = T(0);
            acc[i][slot] += w * own[Op::zc(i)];
          }
        }
      }
      if (is_out(st)) {  // in-plane part of output j = st - R
#pragma unroll
        for (int i = 0; i < NXc; ++i) px[i][oy * V::PXW + ox + R] = own[Op::xc(i)];
#pragma unroll
        for (int i = 0; i < NYc; ++i) py[i][(oy + R) * V::OX + ox] = own[Op::yc(i)];
#pragma unroll
        for (int s = 0; s < V::LXH; ++s)
          if (xhslot[s] >= 0)
#pragma unroll
            for (int i = 0; i < NXc; ++i) px[i][xhslot[s]] = xh[s][i];
#pragma unroll
        for (int s = 0; s < V::LYH; ++s)
          if (yhslot[s] >= 0)
#pragma unroll
            for (int i = 0; i < NYc; ++i) py[i][yhslot[s]] = yh[s][i];
        __syncthreads();
        if (st + 1 < steps) load(st + 1);
        constexpr int sin = ((P - R) % QP + QP) % QP;
        auto X = [&](int i, int o) { return px[i][oy * V::PXW + ox + R + o]; };
        auto Y = [&](int i, int o) { return py[i][(oy + R + o) * V::OX + ox]; };
        T tv[NTQ];
        op.template inplane<O>(X, Y, tv);
#pragma unroll
        for (int q = 0; q < NTQ; ++q) tq[q][sin] = tv[q];
        __syncthreads();
      } else if (st + 1 < steps) {
        load(st + 1);
      }
      const int jd = st - 2 * R;  // output whose z sum this plane completed
      if (jd >= 0 && jd < nout) {  // the same for the whole block
        constexpr int sd = ((P - 2 * R) % QP + QP) % QP;
        T tv[NTQ], dz[NZc], res[NO];
#pragma unroll
        for (int q = 0; q < NTQ; ++q) tv[q] = tq[q][sd];
#pragma unroll
        for (int i = 0; i < NZc; ++i) dz[i] = acc[i][sd];
        op.finish(tv, dz, res);
        if constexpr (kRow) {
          T* rw = orow[oy];
#pragma unroll
          for (int q = 0; q < NO; ++q) rw[ox * NO + q] = res[q];
          wave_sync();
          T* gp = out + (size_t)NO * ((size_t)d.plane * (size_t)(z0 + jd) +
                                      (size_t)d.nx * (size_t)(y0 + oy) + (size_t)x0);
#pragma unroll
          for (int k = 0; k < NO; ++k) {
            const int e = ox + V::OX * k;
            if (e < rowlen) gp[e] = rw[e];
          }
          wave_sync();
        } else if (cin) {
          T* o = out + (size_t)NO * (obase + (size_t)d.plane * (size_t)jd);
#pragma unroll
          for (int q = 0; q < NO; ++q) o[q] = res[q];
        }
      }
    });
  }
}

// diff/Divergence.hpp:56-69 (Vec3 -> scalar): V_x along x, V_y along y, V_z along z,
// ((ax ihx + ay ihy) + az ihz)
template <class T>
struct OpDivVec {
  static constexpr int CIN = 3, NXc = 1, NYc = 1, NZc = 1, NTQ = 1, NOUT = 1;
  __host__ __device__ static constexpr int xc(int) { return 0; }
  __host__ __device__ static constexpr int yc(int) { return 1; }
  __host__ __device__ static constexpr int zc(int) { return 2; }
  T ihx, ihy, ihz;
  template <int O, class FX, class FY>
  __device__ __forceinline__ void inplane(FX X, FY Y, T (&tq)[NTQ]) const {
    using S = D1<O>;
    T ax = 0, ay = 0;
#pragma unroll
    for (int m = 0; m < S::N; ++m) {
      const int o = m - S::R;
      const T w = T(S::w(m));
      ax += w * X(0, o);
      ay += w * Y(0, o);
    }
    tq[0] = ax * ihx + ay * ihy;
  }
  __device__ __forceinline__ void finish(const T (&tq)[NTQ], const T (&dz)[NZc], T (&o)[NOUT]) const {
    o[0] = tq[0] + dz[0] * ihz;
  }
};

// diff/Divergence.hpp:113-131 (Mat3 -> Vec3): column a of M along axis a,
// out_i = (d_x M(i,0) ihx + d_y M(i,1) ihy) + d_z M(i,2) ihz
template <class T>
struct OpDivMat {
  static constexpr int CIN = 9, NXc = 3, NYc = 3, NZc = 3, NTQ = 3, NOUT = 3;
  __host__ __device__ static constexpr int xc(int i) { return i; }
  __host__ __device__ static constexpr int yc(int i) { return 3 + i; }
  __host__ __device__ static constexpr int zc(int i) { return 6 + i; }
  T ihx, ihy, ihz;
  template <int O, class FX, class FY>
  __device__ __forceinline__ void inplane(FX X, FY Y, T (&tq)[NTQ]) const {
    using S = D1<O>;
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
      T dx = 0, dy = 0;
#pragma unroll
      for (int m = 0; m < S::N; ++m) {
        const int o = m - S::R;
        const T w = T(S::w(m));
        dx += w * X(comp, o);
        dy += w * Y(comp, o);
      }
      tq[comp] = dx * ihx + dy * ihy;
    }
  }
  __device__ __forceinline__ void finish(const T (&tq)[NTQ], const T (&dz)[NZc], T (&o)[NOUT]) const {
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) o[comp] = tq[comp] + dz[comp] * ihz;
  }
};

// diff/Gradient.hpp:113-134 (Vec3 -> Mat3, M(comp, axis) at comp + 3*axis): every component
// along every axis
template <class T>
struct OpGradVec {
  // The in-plane entries wait for the z entries in registers: stored at once, each site's
  // line would be written twice, R planes apart, and the partial writes do not merge in L2
  // (5.5 ms against 3.96 ms at 512^3 in FP64).
  static constexpr int CIN = 3, NXc = 3, NYc = 3, NZc = 3, NTQ = 6, NOUT = 9;
  __host__ __device__ static constexpr int xc(int i) { return i; }
  __host__ __device__ static constexpr int yc(int i) { return i; }
  __host__ __device__ static constexpr int zc(int i) { return i; }
  T ihx, ihy, ihz;
  template <int O, class FX, class FY>
  __device__ __forceinline__ void inplane(FX X, FY Y, T (&tq)[NTQ]) const {
    using S = D1<O>;
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
      T dx = 0, dy = 0;
#pragma unroll
      for (int m = 0; m < S::N; ++m) {
        const int o = m - S::R;
        const T w = T(S::w(m));
        dx += w * X(comp, o);
        dy += w * Y(comp, o);
      }
      tq[comp] = dx * ihx;
      tq[3 + comp] = dy * ihy;
    }
  }
  __device__ __forceinline__ void finish(const T (&tq)[NTQ], const T (&dz)[NZc], T (&o)[NOUT]) const {
#pragma unroll
    for (int q = 0; q < 6; ++q) o[q] = tq[q];
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) o[6 + comp] = dz[comp] * ihz;
  }
};

constexpr int kVLen = 32;  // planes per block

constexpr int kVOY = 8;  // tile height of the vector / tensor kernels (64 x kVOY sites)

template <class T, int O, class Op>
void vec_launch(gc_context* ctx, const Dims& d, const Op& op, const T* in, T* out) {
  const int ntx = (d.nx + 63) / 64, nty = (d.ny + kVOY - 1) / kVOY;
  const int nzc = (d.nz + kVLen - 1) / kVLen;
  hipLaunchKernelGGL((k_vec_za<T, O, Op, kVOY>), dim3((unsigned)(ntx * nty * nzc)),
                     dim3(64 * kVOY), 0, ctx->stream, op, in, out, d, ntx, nty, kVLen);
}

template <class T>
struct Inv {
  T x, y, z;
};
inline Inv<double> inv_h(const gc_grid* g) { return {1.0 / g->hx, 1.0 / g->hy, 1.0 / g->hz}; }
inline Inv<double> inv_h2(const gc_grid* g) {
  return {1.0 / (g->hx * g->hx), 1.0 / (g->hy * g->hy), 1.0 / (g->hz * g->hz)};
}

}  // namespace

// ---- typed launchers (shared with the time-stepping kernels) -----------------------

template <class T>
static gc_status run_laplacian(gc_context* ctx, const gc_grid* g, int order, const T* in, T* out) {
  const Dims d = dims_of(g);
  const Inv<double> ih = inv_h2(g);
  // order 4 reads 4 z neighbours: the register queue pays (0.49 against 0.55 ms FP64 on
  // 512^3); order 2 runs as fast on the plane ring
  if (order == 2)
    zmarch_launch<T>(ctx, d, OpLap<T, 2>{T(ih.x), T(ih.y), T(ih.z)}, in, out);
  else
    axis_launch<T>(ctx, d, OpLap<T, 4>{T(ih.x), T(ih.y), T(ih.z)}, in, out);
  GC_LAUNCHED();
  return GC_OK;
}

gc_status launch_laplacian(gc_context* ctx, const gc_grid* g, gc_dtype dt, int order,
                           const void* in, void* out) {
  return dt == GC_F64 ? run_laplacian<double>(ctx, g, order, (const double*)in, (double*)out)
                      : run_laplacian<float>(ctx, g, order, (const float*)in, (float*)out);
}

}  // namespace gc

using namespace gc;

namespace {

gc_status common_checks(gc_context* ctx, const gc_grid* g, gc_dtype dt, int order,
                        const void* in, const void* out) {
  GC_TRY(check_ctx(ctx));
  GC_TRY(check_grid(g));
  GC_TRY(check_dtype(dt));
  if (order != 2 && order != 4)
    return set_error(GC_ERR_INVALID, "gridcalc: Order must be 2 or 4; got %d", order);
  if (!in || !out) return set_error(GC_ERR_INVALID, "gridcalc: null field pointer");
  if (in == out) return set_error(GC_ERR_INVALID, "gridcalc: output must not alias the input");
  return GC_OK;
}

template <class T, int O>
void first_deriv_launch(int which, gc_context* ctx, const gc_grid* g, const void* in, void* out) {
  const Dims d = dims_of(g);
  const Inv<double> ih = inv_h(g);
  const T x = T(ih.x), y = T(ih.y), z = T(ih.z);
  switch (which) {
    case 0:
      // the plane ring: the register-queue engine measured 7 % slower for the gradient
      zmarch_launch<T>(ctx, d, OpGrad<T, O>{x, y, z}, (const T*)in, (T*)out);
      break;
    case 1:
      vec_launch<T, O>(ctx, d, OpGradVec<T>{x, y, z}, (const T*)in, (T*)out);
      break;
    case 2:
      vec_launch<T, O>(ctx, d, OpDivVec<T>{x, y, z}, (const T*)in, (T*)out);
      break;
    default:
      vec_launch<T, O>(ctx, d, OpDivMat<T>{x, y, z}, (const T*)in, (T*)out);
      break;
  }
}

gc_status first_deriv(int which, gc_context* ctx, const gc_grid* g, gc_dtype dt, int order,
                      const void* in, void* out) {
  GC_TRY(common_checks(ctx, g, dt, order, in, out));
  GC_TRY(check_extent(g, order == 2 ? D1<2>::R : D1<4>::R));
  if (dt == GC_F64) {
    if (order == 2) first_deriv_launch<double, 2>(which, ctx, g, in, out);
    else first_deriv_launch<double, 4>(which, ctx, g, in, out);
  } else {
    if (order == 2) first_deriv_launch<float, 2>(which, ctx, g, in, out);
    else first_deriv_launch<float, 4>(which, ctx, g, in, out);
  }
  GC_LAUNCHED();
  return GC_OK;
}

template <class T, int O>
void bih_launch(gc_context* ctx, const gc_grid* g, const void* in, void* out) {
  const Dims d = dims_of(g);
  const Inv<double> i2 = inv_h2(g);
  // inv_h4 = inv_h2 * inv_h2 in double (Biharmonic.hpp:62-64), then cast
  const OpBih<T, O> op{T(i2.x), T(i2.y), T(i2.z), T(i2.x * i2.x), T(i2.y * i2.y), T(i2.z * i2.z)};
  if constexpr (std::is_same<T, float>::value && O == 2) {
    // FP32 order 2: the Laplacian of the Laplacian (k_bih_ll), 64 x 8 tiles
    const LapCoef<float> lc = make_lap_coef<float>(float(i2.x), float(i2.y), float(i2.z));
    const int ntx = (d.nx + 63) / 64, nty = (d.ny + 7) / 8, nzc = (d.nz + kZLen - 1) / kZLen;
    hipLaunchKernelGGL((k_bih_ll<8>), dim3((unsigned)(ntx * nty * nzc)), dim3(256), 0,
                       ctx->stream, lc, (const float*)in, (float*)out, d, ntx, nty, kZLen);
  } else if constexpr (std::is_same<T, float>::value) {
    // FP32 order 4: the separable form (k_bih_sep), 64 x 8 tiles
    const BihSepCoef c{float(i2.x * i2.x), float(i2.y * i2.y), float(i2.z * i2.z),
                       float(2.0 * i2.x * i2.y), float(2.0 * i2.x * i2.z),
                       float(2.0 * i2.y * i2.z)};
    const int ntx = (d.nx + 63) / 64, nty = (d.ny + 7) / 8, nzc = (d.nz + kZLen - 1) / kZLen;
    hipLaunchKernelGGL((k_bih_sep<O, 8>), dim3((unsigned)(ntx * nty * nzc)), dim3(256), 0,
                       ctx->stream, c, (const float*)in, (float*)out, d, ntx, nty, kZLen);
  } else {
    // FP64 (both orders): the z-accumulating kernel on 32 x 8 tiles, one row per thread; order
    // 4 stages two planes per barrier (1.19 ms on 512^3 against 1.22 with one; the register-
    // queue kernel it replaced took 1.59 ms), order 2 one (0.67 ms; the plane ring 0.80 ms)
    constexpr int PB = O == 4 ? 2 : 1, WPE = O == 4 ? 3 : 1;
    const int ntx = (d.nx + 31) / 32, nty = (d.ny + 7) / 8;
    // 64-plane chunks (each chunk re-reads 2 R4 planes), shorter where the grid would not fill
    // the GPU with workgroups
    int zl = 64;
    while (zl > 8 && (long)ntx * nty * ((d.nz + zl - 1) / zl) < 2048) zl /= 2;
    const int nzc = (d.nz + zl - 1) / zl;
    hipLaunchKernelGGL((k_bih_za<T, O, 32, 1, PB, WPE>), dim3((unsigned)(ntx * nty * nzc)),
                       dim3(256), 0, ctx->stream, op, (const T*)in, (T*)out, d, ntx, nty, zl);
  }
}

}  // namespace

namespace gc {
// The checked operator: k_cell_zq, then one 8-byte copy for the D > 0 verdict.
gc_status cell_laplacian_checked(gc_context* ctx, const gc_grid* g, gc_dtype dt, const void* psi,
                                 const void* D, void* out, int64_t* first_bad) {
  const Dims d = dims_of(g);
  const Inv<double> ih = inv_h2(g);
  void* scr = nullptr;
  GC_TRY(scratch(ctx, 256, &scr));
  unsigned long long* bad = (unsigned long long*)scr;
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(1), 0, ctx->stream, bad, ~0ull);
  GC_LAUNCHED();
  // 64 x 16 tiles of 4-row columns (256 threads) for FP64, 64 x 32 of 4-row columns (512
  // threads) for FP32: the fastest of 64 x {4, 8, 16, 32} tiles with 1, 2 or 4 rows per
  // thread on 512^3 (FP64 0.92 ms, FP32 0.41 ms)
  const int ntx = (d.nx + 63) / 64;
  const int nzc = (d.nz + kZLen - 1) / kZLen;
  if (dt == GC_F64) {
    const int nty = (d.ny + 15) / 16;
    hipLaunchKernelGGL((k_cell_zq<double, 16, 4, 1>), dim3((unsigned)(ntx * nty * nzc)),
                       dim3(256), 0, ctx->stream, ih.x, ih.y, ih.z, (const double*)psi,
                       (const double*)D, (double*)out, bad, d, ntx, nty, kZLen);
  } else {
    const int nty = (d.ny + 31) / 32;
    hipLaunchKernelGGL((k_cell_zq<float, 32, 4, 1>), dim3((unsigned)(ntx * nty * nzc)),
                       dim3(512), 0, ctx->stream, float(ih.x), float(ih.y), float(ih.z),
                       (const float*)psi, (const float*)D, (float*)out, bad, d, ntx, nty,
                       kZLen);
  }
  GC_LAUNCHED();
  unsigned long long h = 0;
  GC_HIP(hipMemcpyAsync(ctx->host_result, bad, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
  GC_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(&h, ctx->host_result, sizeof h);
  *first_bad = (h == ~0ull) ? -1 : (int64_t)h;
  return GC_OK;
}
// defined in gc_reduce.hip
gc_status first_nonpositive(gc_context* ctx, const gc_grid* g, gc_dtype dt, const void* D,
                            int64_t* first_bad);
}  // namespace gc

extern "C" {

gc_status gc_laplacian(gc_context* ctx, const gc_grid* g, gc_dtype dt, int order, const void* in,
                       void* out) {
  GC_TRY(common_checks(ctx, g, dt, order, in, out));
  GC_TRY(check_extent(g, order == 2 ? D2<2>::R : D2<4>::R));
  return launch_laplacian(ctx, g, dt, order, in, out);
}

gc_status gc_gradient(gc_context* ctx, const gc_grid* g, gc_dtype dt, int order, const void* in,
                      void* out_vec3) {
  return first_deriv(0, ctx, g, dt, order, in, out_vec3);
}

gc_status gc_gradient_vec(gc_context* ctx, const gc_grid* g, gc_dtype dt, int order,
                          const void* in_vec3, void* out_mat3) {
  return first_deriv(1, ctx, g, dt, order, in_vec3, out_mat3);
}

gc_status gc_divergence(gc_context* ctx, const gc_grid* g, gc_dtype dt, int order,
                        const void* in_vec3, void* out) {
  return first_deriv(2, ctx, g, dt, order, in_vec3, out);
}

gc_status gc_divergence_mat(gc_context* ctx, const gc_grid* g, gc_dtype dt, int order,
                            const void* in_mat3, void* out_vec3) {
  return first_deriv(3, ctx, g, dt, order, in_mat3, out_vec3);
}

gc_status gc_biharmonic(gc_context* ctx, const gc_grid* g, gc_dtype dt, int order, const void* in,
                        void* out) {
  GC_TRY(common_checks(ctx, g, dt, order, in, out));
  // max(S4 radius, S2 radius) (Biharmonic.hpp:52-57)
  GC_TRY(check_extent(g, order == 2 ? D4<2>::R : D4<4>::R));
  if (dt == GC_F64) {
    if (order == 2) bih_launch<double, 2>(ctx, g, in, out);
    else bih_launch<double, 4>(ctx, g, in, out);
  } else {
    if (order == 2) bih_launch<float, 2>(ctx, g, in, out);
    else bih_launch<float, 4>(ctx, g, in, out);
  }
  GC_LAUNCHED();
  return GC_OK;
}

gc_status gc_cell_laplacian(gc_context* ctx, const gc_grid* g, gc_dtype dt, const void* psi,
                            const void* D, void* out, int64_t* first_bad) {
  GC_TRY(common_checks(ctx, g, dt, 2, psi, out));
  if (!D) return set_error(GC_ERR_INVALID, "gridcalc: null D pointer");
  if (D == out) return set_error(GC_ERR_INVALID, "gridcalc: output must not alias D");
  GC_TRY(check_extent(g, 1));
  int64_t bad = -1;
  GC_TRY(cell_laplacian_checked(ctx, g, dt, psi, D, out, &bad));
  if (first_bad) *first_bad = bad;
  if (bad >= 0)
    return set_error(GC_ERR_NONPOS_D,
                     "fvm::cellLaplacian: D must be strictly positive at every cell (the "
                     "harmonic-mean face-D averaging is undefined otherwise); first violation "
                     "at flat index %lld",
                     (long long)bad);
  return GC_OK;
}

}  // extern "C"
