// gc_cnn.hip -- the learned corrector for 2-D Cahn-Hilliard: a 6-layer 3x3 periodic CNN
// (1 -> 32 -> 32 -> 32 -> 32 -> 32 -> 1 channels, GELU, residual output) trained and run
// on the device.
//
//   a1 = gelu(conv(x, W1) + b1),  a_l = gelu(conv(a_{l-1}, W_l) + b_l) (l = 2..5),
//   out = conv(a5, W6) + b6,      prediction = c + scale * (out - mean(out))
//
// conv(x, W)[y, x] = sum_tap W[tap] x[(y+dy) mod H, (x+dx) mod W], tap = 3(dy+1) + (dx+1).
// In the forward pass the 32 -> 32 layers take FP16 weights and activations (10-bit
// significands: with BF16's 7 the roundings alone left the trained network's one-step error
// several times its exact-arithmetic error); the 1 -> 32 and 32 -> 1 layers take the state
// and their weights in FP32; every accumulation is FP32.  The four 32 -> 32 layers carry 98 %
// of the FLOPs and run as implicit GEMMs on the matrix cores (v_mfma_f32_32x32x16_f16):
// M = 32 output channels, N = 32 sites, K = 9 taps x 32 input channels (18 MFMAs per 32 x 32
// output tile; a lane ends with 16 channels of one site).  The backward pass runs in BF16
// (v_mfma_f32_32x32x16_bf16), whose FP32 exponent range holds the small loss gradients that
// FP16 would flush: backward-data is the same convolution with the taps flipped and the
// weight matrix transposed (BF16 weights, BF16 pre-activation gradients); backward-weights is
// a GEMM over sites whose operands are staged channel-major in LDS, the FP16 activations
// rounded to BF16 there.  The 1 -> 32 and 32 -> 1 layers (288 MACs per site) run on the
// vector ALUs.
//
// Parameters are one FP32 vector (37601 values): W1[co][tap], b1, {W_l[co][tap][ci], b_l}
// for l = 2..5, W6[tap][ci], b6.  Activations are NHWC (32 channels contiguous per site).
#include "gc_common.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace gc {
gc_status integrate_raw(gc_context* ctx, const gc_grid* g, gc_dtype dt, int method,
                        const void* in, double* sum_out);
gc_status free_energy_raw(gc_context* ctx, const gc_grid* g, gc_dtype dt, double kappa,
                          int method, const void* psi, double* sum_out);
}  // namespace gc

namespace {

using namespace gc;

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));

// One 32 x 32 x 16 matrix-core step in either 16-bit format.
__device__ __forceinline__ f32x16 mfma16(const f16x8& a, const f16x8& b, const f32x16& c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x16 mfma16(const bf16x8& a, const bf16x8& b, const f32x16& c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

constexpr int kC = 32;            // channels
constexpr int kMidW = 9 * 32 * 32;  // weights of a 32 -> 32 layer
constexpr int kOffW1 = 0, kOffB1 = 288, kOffMid = 320, kMidStride = kMidW + 32;
constexpr int kOffW6 = kOffMid + 4 * kMidStride, kOffB6 = kOffW6 + 288;
constexpr int kNParams = kOffB6 + 1;  // 37601
static_assert(kNParams == 37601, "parameter count");

constexpr int kTR = 8, kTC = 32, kSP = 40;  // conv tile rows, cols, LDS site stride (bf16)

// GELU(z) = z Phi(z) ~= z sigma(z P(z^2)), sigma(u) = 1 / (1 + e^-u), with
// P(t) = 1.5950206 + 0.074006644 t - 7.0221731e-4 t^2 fitted (least maximum error on
// |z| <= 9) to the exact z Phi(z): |error| <= 2.6e-5, a tenth of a bfloat16 step at 0.5.
// Inside P, z is clamped to [-9, 9], where sigma is already 1 or 0 to FP32 precision.  Two
// hardware transcendentals (exp2 with log2(e) folded into P, reciprocal) and seven other
// instructions.
__device__ __forceinline__ float gelu_f(float z) {
  constexpr float L = -1.44269504088896341f;  // -log2(e)
  const float zc = __builtin_amdgcn_fmed3f(z, -9.0f, 9.0f);
  const float t = zc * zc;
  const float P = fmaf(t, fmaf(t, L * -7.0221731e-4f, L * 0.074006644f), L * 1.5950206f);
  const float e = __builtin_amdgcn_exp2f(zc * P);
  return z * __builtin_amdgcn_rcpf(1.0f + e);
}
__device__ __forceinline__ float gelu_grad_f(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
         z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

__device__ __forceinline__ int wrap1(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// ---- layer 1 (1 -> 32), matrix cores in FP32 ----------------------------------------------
// The state and W1 enter in FP32 (a BF16-rounded input would carry noise of 2^-9 |c| that the
// network answers as if it were part of the state).  v_mfma_f32_32x32x2f32: M = 32 output
// channels, N = 32 sites, K = the 9 taps and a zero tenth, 2 per instruction; lane (h, j)
// (h = lane / 32) supplies W1[j][2 kk + h] and tap 2 kk + h of site j.  The accumulator
// layout is that of the 32 -> 32 layers: a lane ends with channels 8 g + 4 h + i of site j
// in acc[4 g + i].
__device__ __forceinline__ void conv1_weights(const float* __restrict__ w1, int lane, float (&wa)[5]) {
  const int j = lane & 31, h = lane >> 5;
#pragma unroll
  for (int kk = 0; kk < 5; ++kk) {
    const int tap = 2 * kk + h;
    wa[kk] = tap < 9 ? w1[j * 9 + tap] : 0.0f;
  }
}
template <class X>
__device__ __forceinline__ f32x16 conv1_mfma(const float (&wa)[5], int lane, const X& x) {
  const int h = lane >> 5;
  f32x16 acc = {};
#pragma unroll
  for (int kk = 0; kk < 5; ++kk) {
    const int tap = 2 * kk + h;
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(wa[kk], tap < 9 ? x(tap) : 0.0f, acc, 0, 0, 0);
  }
  return acc;
}

// Block: 4 waves, each 32 consecutive sites of one row (W is a multiple of 32).
template <class TS>
__global__ __launch_bounds__(256) void k_conv1(const TS* __restrict__ c, const int* __restrict__ idx,
                                               const float* __restrict__ w1,
                                               const float* __restrict__ b1, float* __restrict__ z1,
                                               _Float16* __restrict__ a1, int H, int W, int B) {
  const int HW = H * W;
  const int lane = threadIdx.x & 63, j = lane & 31, cq = (lane >> 5) * 4;
  const long s0 = ((long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 32;
  if (s0 >= (long)B * HW) return;
  const long s = s0 + j;
  const int b = (int)(s / HW), p = (int)(s % HW), y = p / W, x = p % W;
  const TS* img = c + (size_t)(idx ? idx[b] : b) * HW;
  float wa[5];
  conv1_weights(w1, lane, wa);
  const f32x16 acc = conv1_mfma(wa, lane, [&](int tap) {
    return (float)img[wrap1(y + tap / 3 - 1, H) * W + wrap1(x + tap % 3 - 1, W)];
  });
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    float v[4];
    f16x4 av;
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = acc[4 * g + i] + b1[g * 8 + cq + i];
    if (z1) *(float4*)(z1 + (size_t)s * kC + g * 8 + cq) = make_float4(v[0], v[1], v[2], v[3]);
#pragma unroll
    for (int i = 0; i < 4; ++i) av[i] = (_Float16)gelu_f(v[i]);
    *(f16x4*)(a1 + (size_t)s * kC + g * 8 + cq) = av;
  }
}

// ---- 32 -> 32 layers, matrix cores --------------------------------------------------------
// MODE 0 (forward, FP16): out = fp16(gelu(acc + bias)), zout = acc + bias (optional).
// MODE 1 (backward-data, BF16): out = bf16(acc * gelu'(zprev)).
// Block: 256 threads, one kTR x kTC site tile of one image; wave w computes rows 2w, 2w+1.
template <int MODE>
struct MidTypes {
  typedef _Float16 T;
  typedef f16x8 T8;
  typedef f16x4 T4;
};
template <>
struct MidTypes<1> {
  typedef __bf16 T;
  typedef bf16x8 T8;
  typedef bf16x4 T4;
};
template <int MODE>
__global__ __launch_bounds__(256) void k_conv_mid(const typename MidTypes<MODE>::T* __restrict__ in,
                                                  const typename MidTypes<MODE>::T* __restrict__ wt,
                                                  const float* __restrict__ bias,
                                                  const float* __restrict__ zprev,
                                                  float* __restrict__ zout,
                                                  typename MidTypes<MODE>::T* __restrict__ out,
                                                  int H, int W, int tiles_x, int tiles_y) {
  typedef typename MidTypes<MODE>::T T;
  typedef typename MidTypes<MODE>::T8 T8;
  typedef typename MidTypes<MODE>::T4 T4;
  __shared__ __attribute__((aligned(16))) T xs[(kTR + 2) * (kTC + 2) * kSP];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int tx = blockIdx.x % tiles_x;
  const int r1 = blockIdx.x / tiles_x;
  const int ty = r1 % tiles_y, b = r1 / tiles_y;
  const int x0 = tx * kTC, y0 = ty * kTR;
  const size_t img = (size_t)b * H * W;
  constexpr int NS = (kTR + 2) * (kTC + 2);
  for (int q = t; q < NS * 4; q += 256) {
    const int site = q >> 2, ch = (q & 3) * 8;
    const int ly = site / (kTC + 2), lx = site % (kTC + 2);
    const int gy = wrap1(y0 + ly - 1, H), gx = wrap1(x0 + lx - 1, W);
    *(T8*)(xs + site * kSP + ch) = *(const T8*)(in + (img + (size_t)gy * W + gx) * kC + ch);
  }
  const int j = lane & 31, kh = (lane >> 5) * 8;
  T8 wreg[18];
#pragma unroll
  for (int kc = 0; kc < 18; ++kc) wreg[kc] = *(const T8*)(wt + (size_t)(kc * 64 + lane) * 8);
  __syncthreads();
#pragma unroll
  for (int rr = 0; rr < 2; ++rr) {
    const int r = wave * 2 + rr;
    f32x16 acc = {};
#pragma unroll
    for (int kc = 0; kc < 18; ++kc) {
      const int tap = kc >> 1, dy = tap / 3 - 1, dx = tap % 3 - 1;
      const int site = (r + 1 + dy) * (kTC + 2) + (j + 1 + dx);
      const T8 a = *(const T8*)(xs + site * kSP + (kc & 1) * 16 + kh);
      acc = mfma16(wreg[kc], a, acc);
    }
    // lane: site x0 + j of the row, channels 8 g + 4 (lane / 32) + i in acc[4 g + i]
    const size_t o0 = (img + (size_t)(y0 + r) * W + x0 + j) * kC + (lane >> 5) * 4;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const size_t o = o0 + g * 8;
      float v[4];
      T4 ov;
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = acc[4 * g + i];
      if (MODE == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] += bias[g * 8 + (lane >> 5) * 4 + i];
        if (zout) *(float4*)(zout + o) = make_float4(v[0], v[1], v[2], v[3]);
#pragma unroll
        for (int i = 0; i < 4; ++i) ov[i] = (T)gelu_f(v[i]);
      } else {
        const float4 zp = *(const float4*)(zprev + o);
        ov[0] = (T)(v[0] * gelu_grad_f(zp.x));
        ov[1] = (T)(v[1] * gelu_grad_f(zp.y));
        ov[2] = (T)(v[2] * gelu_grad_f(zp.z));
        ov[3] = (T)(v[3] * gelu_grad_f(zp.w));
      }
      *(T4*)(out + o) = ov;
    }
  }
}

// ---- compensated reductions shared by the learned-integration kernels ------------------

// Sum of the per-tile partial sums part[0..nparts): wave 0 of the block, lane l summing
// part[l], part[l + 64], ... in order, then a fixed xor tree (every block gets the same bits);
// the result is stored to *dst (LDS) before the barrier.
__device__ __forceinline__ void block_sum_parts(const double* __restrict__ part, int nparts,
                                                double* dst) {
  if (threadIdx.x < 64) {
    double v = 0.0;
    for (int i = threadIdx.x; i < nparts; i += 64) v += part[i];
    for (int st = 32; st >= 1; st >>= 1) v += __shfl_xor(v, st);
    if (threadIdx.x == 0) *dst = v;
  }
  __syncthreads();
}

struct Comp2 {
  double ms, mc, fs, fc;
};

__device__ __forceinline__ void neu(double& s, double& c, double x) {
  const double t = s + x;
  if (fabs(s) >= fabs(x)) c += (s - t) + x;
  else c += (x - t) + s;
  s = t;
}

// One wave's xor-tree merge of per-lane partials (sums with TwoSum, compensations added);
// every lane ends with the wave total.
__device__ __forceinline__ Comp2 wave_merge(Comp2 v) {
  for (int st = 1; st < 64; st <<= 1) {
    Comp2 o{__shfl_xor(v.ms, st), __shfl_xor(v.mc, st), __shfl_xor(v.fs, st), __shfl_xor(v.fc, st)};
    Comp2 lo = (threadIdx.x & st) ? o : v, hi = (threadIdx.x & st) ? v : o;
    double s1 = lo.ms, c1 = lo.mc + hi.mc;
    neu(s1, c1, hi.ms);
    double s2 = lo.fs, c2 = lo.fc + hi.fc;
    neu(s2, c2, hi.fs);
    v = Comp2{s1, c1, s2, c2};
  }
  return v;
}

// The first nw waves' totals (sh[w], written by lane 0 of each) merged in wave order; for
// thread 0.
__device__ __forceinline__ Comp2 waves_merge(const Comp2* sh, int nw) {
  Comp2 r = sh[0];
  for (int w = 1; w < nw; ++w) {
    double c1 = r.mc + sh[w].mc, c2 = r.fc + sh[w].fc;
    neu(r.ms, c1, sh[w].ms);
    neu(r.fs, c2, sh[w].fs);
    r.mc = c1;
    r.fc = c2;
  }
  return r;
}

// Mass and Cahn-Hilliard density (examples/common/cahn_hilliard.hpp:65-67) of a site of a 2-D
// periodic grid with value p and neighbours xm, ym, xp, yp (x - 1, y - 1, x + 1, y + 1),
// accumulated (Neumaier) into ms/mc and fs/fc.
__device__ __forceinline__ void site_mass_fe5(double p, double xm, double ym, double xp,
                                              double yp, double ihx, double ihy, double kappa,
                                              double& ms, double& mc, double& fs, double& fc) {
  double dx = 0.0, dy = 0.0;
  dx += -0.5 * xm;
  dy += -0.5 * ym;
  dx += 0.0 * p;
  dy += 0.0 * p;
  dx += 0.5 * xp;
  dy += 0.5 * yp;
  const double gx = dx * ihx, gy = dy * ihy, gz = 0.0;
  const double sq = gx * gx + gy * gy + gz * gz;
  neu(ms, mc, p);
  neu(fs, fc, -0.5 * p * p + 0.25 * p * p * p * p + 0.5 * kappa * sq);
}

// The same for site (i, j) whose values come from val(flat index).
template <class V>
__device__ __forceinline__ void site_mass_fe(const V& val, int i, int j, int nx, int ny,
                                             double ihx, double ihy, double kappa, double& ms,
                                             double& mc, double& fs, double& fc) {
  site_mass_fe5(val((size_t)j * nx + i), val((size_t)j * nx + wrap1(i - 1, nx)),
                val((size_t)wrap1(j - 1, ny) * nx + i), val((size_t)j * nx + wrap1(i + 1, nx)),
                val((size_t)wrap1(j + 1, ny) * nx + i), ihx, ihy, kappa, ms, mc, fs, fc);
}

// One wave: lane l merges the block partials l, l + 64, ... in order, then the lanes merge in
// a fixed xor tree; lane 0 writes the two results.
__device__ __forceinline__ void mass_fe_reduce(const Comp2* __restrict__ part, int nb,
                                               double* __restrict__ out) {
  double ms = 0.0, mc = 0.0, fs = 0.0, fc = 0.0;
  for (int b = threadIdx.x; b < nb; b += 64) {
    mc += part[b].mc;
    fc += part[b].fc;
    neu(ms, mc, part[b].ms);
    neu(fs, fc, part[b].fs);
  }
  const Comp2 v = wave_merge(Comp2{ms, mc, fs, fc});
  if (threadIdx.x == 0) {
    out[0] = v.ms + v.mc;
    out[1] = v.fs + v.fc;
  }
}

// ---- whole-network inference in two kernels ------------------------------------------------
// A block owns an 8 x 32 output tile.  Kernel A evaluates layers 1..3 on it, each on the tile
// plus the halo the later layers of the kernel need (layer 1 on a 2-site halo, layer 2 on a
// 1-site halo, layer 3 on the tile), from the state region (halo 3) staged in LDS as FP32, and
// writes layer 3's activations (FP16, site-major, like the per-layer path) to global memory;
// kernel B reads them on a 3-site halo and evaluates layers 4..6 the same way.  Activations
// ping-pong between two LDS buffers (site stride kSP).  Every value is formed with the
// per-layer kernels' arithmetic (same FMA order, same MFMA operand fragments and K order), so
// the output is bit-identical to theirs.  Recomputed halos cost 1.34x the tile's matrix work
// in layers 2..5 (1368 sites in 32-site chunks for 1024) and 1.7x its layer-1 work; one
// kernel for all six layers needs halos up to 6 sites (1984 sites, 3x layer-1 work), which
// costs more than the extra launch and the 64 bytes per site of layer-3 activations.
constexpr int kFR = kTR, kFC = kTC;                // output tile
constexpr int kAH0 = kFR + 6, kAW0 = kFC + 6;      // A: state region (halo 3)
constexpr int kAN1 = (kFR + 4) * (kFC + 4);        // A: layer-1 region (halo 2)
constexpr int kAN2 = (kFR + 2) * (kFC + 2);        // A: layer-2 region (halo 1)
constexpr int kBN3 = (kFR + 6) * (kFC + 6);        // B: layer-3 region (halo 3)
constexpr int kBN4 = (kFR + 4) * (kFC + 4);        // B: layer-4 region (halo 2)

constexpr int kFNT = 1024;  // threads of the network kernels (16 waves)
constexpr int kFNW = kFNT / 64;
static_assert(kFNT >= 2 * kFR * kFC, "layer 6 maps two lanes to each tile site");

// Chained mode (learned integration, one image): the state kernel A advances is formed on
// the fly from the previous macro step, c_k = c_{k-1} + scale (out_{k-1} - mean(out_{k-1}))
// (k_apply's expression, so the same values); the block writes it for its own tile, writes
// its tile's mass and free-energy partial sums (merged by a later kernel) and evaluates the
// network on it.  Two launches per macro step instead of three.
template <class TS>
struct Chain {
  const TS* cprev;       // state of the previous macro step
  const float* oprev;    // the network's output for it
  const double* pprev;   // its per-tile sums
  int nparts;
  double scale, ihx, ihy, kappa;
  TS* cdst;              // this macro step's state (the block's tile)
  Comp2* mpart;          // per-tile (mass, density) partial sums
};

// One 32 -> 32 layer (index l of the four) on the matrix cores, output region = tile + halo
// HO, input (in, LDS, site stride kSP) one site wider.  The output region's sites, row-major,
// are cut into chunks of 32, one per wave (16 waves; every region here has at most 14);
// store(site, g, v) receives channels 8 g + cq .. 8 g + cq + 3 of an output site as FP16.
// wt: the layer's weight fragments (fragment-major, one contiguous 1 KiB per read), staged
// in LDS by the kernel before its first 32 -> 32 layer.
template <int HO, class Store>
__device__ __forceinline__ void mid_layer(const _Float16* __restrict__ in, int l,
                                          const _Float16* __restrict__ wt,
                                          const float* __restrict__ p, int wave, int lane,
                                          Store store) {
  constexpr int Wo = kFC + 2 * HO, Ho = kFR + 2 * HO, Wi = Wo + 2, NS = Wo * Ho;
  constexpr int nch = (NS + 31) / 32;
  const int j = lane & 31, kh = (lane >> 5) * 8, cq = (lane >> 5) * 4;
  float bj[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) bj[i] = p[kOffMid + l * kMidStride + kMidW + (i / 4) * 8 + cq + i % 4];
  for (int c = wave; c < nch; c += kFNW) {
    const int s0 = min(c * 32 + j, NS - 1);
    const int i0 = (s0 / Wo + 1) * Wi + s0 % Wo + 1;
    // operand fragments of step kc; requested kAhead steps before their matrix instruction
    auto frag_b = [&](int kc) {
      const int tap = kc >> 1, off = (tap / 3 - 1) * Wi + tap % 3 - 1;
      return *(const f16x8*)(in + (i0 + off) * kSP + (kc & 1) * 16 + kh);
    };
    auto frag_a = [&](int kc) { return *(const f16x8*)(wt + (size_t)(kc * 64 + lane) * 8); };
    constexpr int kAhead = 4;
    f16x8 fa[kAhead], fb[kAhead];
#pragma unroll
    for (int u = 0; u < kAhead; ++u) {
      fa[u] = frag_a(u);
      fb[u] = frag_b(u);
    }
    f32x16 acc = {};
#pragma unroll
    for (int kc = 0; kc < 18; ++kc) {
      acc = mfma16(fa[kc % kAhead], fb[kc % kAhead], acc);
      if (kc + kAhead < 18) {
        fa[kc % kAhead] = frag_a(kc + kAhead);
        fb[kc % kAhead] = frag_b(kc + kAhead);
      }
      // keep the reads where they are: the scheduler would otherwise sink them to just
      // before their matrix instruction and expose the LDS latency at every step
      __builtin_amdgcn_sched_barrier(0);
    }
    if (c * 32 + j < NS) {
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        f16x4 ov;
#pragma unroll
        for (int i = 0; i < 4; ++i) ov[i] = (_Float16)gelu_f(acc[4 * g + i] + bj[4 * g + i]);
        store(c * 32 + j, g * 8 + cq, ov);
      }
    }
  }
}

// The fragments of two consecutive 32 -> 32 layers (36 KiB of FP16), fetched into registers
// at kernel start and stored to LDS once the registers are free of the loads' latency.
constexpr int kWQ = (2 * kMidW / 8 + kFNT - 1) / kFNT;  // 16-byte pieces per thread
__device__ __forceinline__ void fetch_w2(const _Float16* __restrict__ w, int t, f16x8* r) {
#pragma unroll
  for (int i = 0; i < kWQ; ++i) {
    const int q = t + i * kFNT;
    if (q < 2 * kMidW / 8) r[i] = ((const f16x8*)w)[q];
  }
}
__device__ __forceinline__ void store_w2(_Float16* ws, int t, const f16x8* r) {
#pragma unroll
  for (int i = 0; i < kWQ; ++i) {
    const int q = t + i * kFNT;
    if (q < 2 * kMidW / 8) ((f16x8*)ws)[q] = r[i];
  }
}

// Kernel A: layers 1..3 of the network on the block's tile; layer 3's activations to a3.
template <class TS, bool APPLY>
__global__ __launch_bounds__(kFNT) void k_fwd_a(const TS* __restrict__ c,
                                               const float* __restrict__ w1,
                                               const float* __restrict__ b1,
                                               const _Float16* __restrict__ wmid,
                                               const float* __restrict__ p,
                                               _Float16* __restrict__ a3, int H, int W,
                                               int tiles_x, int tiles_y, Chain<TS> ch) {
  __shared__ float cs[kAH0 * kAW0];
  __shared__ TS csx[APPLY ? kAH0 * kAW0 : 1];  // chained: the new state on the region
  __shared__ Comp2 msh[kFR * kFC / 64];
  __shared__ float w1s[288 + 32];
  __shared__ __attribute__((aligned(16))) _Float16 bufA[kAN1 * kSP];
  __shared__ __attribute__((aligned(16))) _Float16 bufB[kAN2 * kSP];
  __shared__ __attribute__((aligned(16))) _Float16 wls[2 * kMidW];  // layers 2 and 3
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int tx = blockIdx.x % tiles_x;
  const int r1 = blockIdx.x / tiles_x;
  const int ty = r1 % tiles_y, b = r1 / tiles_y;
  const int x0 = tx * kFC, y0 = ty * kFR;
  constexpr int NR = kAH0 * kAW0;
  f16x8 wreg[kWQ];  // layers 2 and 3's weights, requested after the state's loads
  // layer 1's weights and biases, requested first (staged in LDS with the state)
  const float w1v = t < 288 ? w1[t] : (t < 288 + 32 ? b1[t - 288] : 0.0f);
  if constexpr (APPLY) {
    // Every load is issued before the first use: the previous tile sums first (the mean
    // waits only for them), then the state region.
    // The sums are fetched eight per lane at a time and added in index order, as a plain
    // lane-strided loop would (which waits for each load before issuing the next).
    double sv = 0.0;
    for (int k0 = lane; k0 < ch.nparts; k0 += 8 * 64) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = k0 + 64 * u < ch.nparts ? ch.pprev[k0 + 64 * u] : 0.0;
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (k0 + 64 * u < ch.nparts) sv += v[u];
    }
    constexpr int kNH = (NR + kFNT - 1) / kFNT;
    double hc[kNH];
    float ho[kNH];
#pragma unroll
    for (int r = 0; r < kNH; ++r) {
      const int q = t + r * kFNT;
      if (q < NR) {
        const size_t g = (size_t)wrap1(y0 + q / kAW0 - 3, H) * W + wrap1(x0 + q % kAW0 - 3, W);
        hc[r] = (double)ch.cprev[g];
        ho[r] = ch.oprev[g];
      }
    }
    fetch_w2(wmid, t, wreg);
    // mean of the previous output, in every wave: lane 0's value of the lane-strided sums
    // and xor tree of block_sum_parts, so the value k_apply uses
    for (int st = 32; st >= 1; st >>= 1) sv += __shfl_xor(sv, st);
    const double mean = __shfl(sv, 0) / (double)((size_t)H * W);
    auto nv = [&](double cp, float op) { return (TS)(cp + ch.scale * ((double)op - mean)); };
#pragma unroll
    for (int r = 0; r < kNH; ++r) {
      const int q = t + r * kFNT;
      if (q < NR) {
        const TS x = nv(hc[r], ho[r]);
        csx[q] = x;
        cs[q] = (float)x;
      }
    }
  } else {
    fetch_w2(wmid, t, wreg);
    const TS* img = c + (size_t)b * H * W;
    constexpr int kNH = (NR + kFNT - 1) / kFNT;
    TS hv[kNH];
#pragma unroll
    for (int r = 0; r < kNH; ++r) {
      const int q = t + r * kFNT;
      if (q < NR) hv[r] = img[(size_t)wrap1(y0 + q / kAW0 - 3, H) * W + wrap1(x0 + q % kAW0 - 3, W)];
    }
#pragma unroll
    for (int r = 0; r < kNH; ++r) {
      const int q = t + r * kFNT;
      if (q < NR) cs[q] = (float)hv[r];
    }
  }
  static_assert(kFNT >= 288 + 32, "one layer-1 weight or bias per thread");
  if (t < 288 + 32) w1s[t] = w1v;
  __syncthreads();
  // layer 1 on the halo-2 region (kFR + 4) x (kFC + 4), in 32-site chunks like layers 2..5;
  // chained: the two waves without a layer-1 chunk write the tile's new state and reduce its
  // mass and density meanwhile, tile site v * 64 + lane in pass v of the four (one wave's
  // worth of sites per pass, so the sums merge exactly as with one site per thread)
  if constexpr (APPLY) {
    static_assert((kAN1 + 31) / 32 <= kFNW - 2 && kFR * kFC == 4 * 64, "idle waves in layer 1");
    if (wave >= kFNW - 2) {
#pragma unroll
      for (int pv = 0; pv < 2; ++pv) {
        const int v = (wave - (kFNW - 2)) * 2 + pv, st = v * 64 + lane;
        const int q0 = (st / kFC + 3) * kAW0 + st % kFC + 3;
        const TS x0v = csx[q0], xm = csx[q0 - 1], ym = csx[q0 - kAW0], xp = csx[q0 + 1],
                 yp = csx[q0 + kAW0];
        ch.cdst[(size_t)(y0 + st / kFC) * W + x0 + st % kFC] = x0v;
        double ms = 0.0, mc = 0.0, fs = 0.0, fc = 0.0;
        site_mass_fe5((double)x0v, (double)xm, (double)ym, (double)xp, (double)yp, ch.ihx,
                      ch.ihy, ch.kappa, ms, mc, fs, fc);
        const Comp2 m2 = wave_merge(Comp2{ms, mc, fs, fc});
        if (lane == 0) msh[v] = m2;
      }
    }
  }
  {
    constexpr int RW = kFC + 4, RN = kAN1, nch = (RN + 31) / 32;
    const int j = lane & 31, cq = (lane >> 5) * 4;
    float wa[5];
    conv1_weights(w1s, lane, wa);
    auto chunk = [&](int cc) {
      const int sr = min(cc * 32 + j, RN - 1), ry = sr / RW, rx = sr % RW;
      return conv1_mfma(wa, lane, [&](int tap) { return cs[(ry + tap / 3) * kAW0 + rx + tap % 3]; });
    };
    auto emit1 = [&](const f32x16& acc, int sr) {
      if (sr < RN) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          f16x4 av;
#pragma unroll
          for (int i = 0; i < 4; ++i) av[i] = (_Float16)gelu_f(acc[4 * g + i] + w1s[288 + g * 8 + cq + i]);
          *(f16x4*)(bufA + sr * kSP + g * 8 + cq) = av;
        }
      }
    };
    for (int cc = wave; cc < nch; cc += kFNW) emit1(chunk(cc), cc * 32 + j);
  }
  store_w2(wls, t, wreg);
  __syncthreads();
  if constexpr (APPLY)
    if (t == 0) ch.mpart[blockIdx.x] = waves_merge(msh, kFR * kFC / 64);
  mid_layer<1>(bufA, 0, wls, p, wave, lane, [&](int s, int ch0, const f16x4& v) {
    *(f16x4*)(bufB + s * kSP + ch0) = v;
  });
  __syncthreads();
  const size_t img = (size_t)b * H * W;
  mid_layer<0>(bufB, 1, wls + kMidW, p, wave, lane, [&](int s, int ch0, const f16x4& v) {
    *(f16x4*)(a3 + (img + (size_t)(y0 + s / kFC) * W + x0 + s % kFC) * kC + ch0) = v;
  });
}

// Kernel B: layers 4..6 from layer 3's activations; output to out, the tile's output sum to
// part[block].
__global__ __launch_bounds__(kFNT) void k_fwd_b(const _Float16* __restrict__ a3,
                                               const _Float16* __restrict__ wmid,
                                               const float* __restrict__ p,
                                               const float* __restrict__ w6,
                                               const float* __restrict__ b6,
                                               float* __restrict__ out,
                                               double* __restrict__ part, int H, int W,
                                               int tiles_x, int tiles_y) {
  __shared__ double psum[kFR * kFC / 32];
  __shared__ float w6s[288];
  __shared__ __attribute__((aligned(16))) _Float16 bufA[kBN3 * kSP];
  __shared__ __attribute__((aligned(16))) _Float16 bufB[kBN4 * kSP];
  __shared__ __attribute__((aligned(16))) _Float16 wls[2 * kMidW];  // layers 4 and 5
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int tx = blockIdx.x % tiles_x;
  const int r1 = blockIdx.x / tiles_x;
  const int ty = r1 % tiles_y, b = r1 / tiles_y;
  const int x0 = tx * kFC, y0 = ty * kFR;
  const size_t img = (size_t)b * H * W;
  const float w6v = t < 288 ? w6[t] : 0.0f;  // layer 6's weights, staged in LDS below
  f16x8 wreg[kWQ];
  fetch_w2(wmid + 2 * kMidW, t, wreg);
  // layer 3's activations on the halo-3 region: every load issued before the first store
  constexpr int NQ = (kBN3 * 4 + kFNT - 1) / kFNT;
  f16x8 rv[NQ];
#pragma unroll
  for (int r = 0; r < NQ; ++r) {
    const int q = t + r * kFNT;
    if (q < kBN3 * 4) {
      const int site = q >> 2, ch = (q & 3) * 8;
      const int gy = wrap1(y0 + site / (kFC + 6) - 3, H), gx = wrap1(x0 + site % (kFC + 6) - 3, W);
      rv[r] = *(const f16x8*)(a3 + (img + (size_t)gy * W + gx) * kC + ch);
    }
  }
#pragma unroll
  for (int r = 0; r < NQ; ++r) {
    const int q = t + r * kFNT;
    if (q < kBN3 * 4) *(f16x8*)(bufA + (q >> 2) * kSP + (q & 3) * 8) = rv[r];
  }
  if (t < 288) w6s[t] = w6v;
  store_w2(wls, t, wreg);
  __syncthreads();
  mid_layer<2>(bufA, 2, wls, p, wave, lane, [&](int s, int ch0, const f16x4& v) {
    *(f16x4*)(bufB + s * kSP + ch0) = v;
  });
  __syncthreads();
  mid_layer<1>(bufB, 3, wls + kMidW, p, wave, lane, [&](int s, int ch0, const f16x4& v) {
    *(f16x4*)(bufA + s * kSP + ch0) = v;
  });
  __syncthreads();
  // layer 6 on the tile from the layer-5 region (halo 1, in bufA): lanes 0..31 of wave w
  // take tile sites 32 w .. 32 w + 31 over channels 0..15, lanes 32..63 the same sites over
  // channels 16..31 (k_conv6's two accumulators)
  if (wave < kFR * kFC / 32) {
    constexpr int Wi = kFC + 2;
    const int ts = wave * 32 + (lane & 31), h = lane >> 5;
    const int ry = ts / kFC, rx = ts % kFC;
    float acc = 0.0f;
    for (int tp = 0; tp < 9; ++tp) {
      const int site = (ry + tp / 3) * Wi + rx + tp % 3;
#pragma unroll
      for (int ch = 16 * h; ch < 16 * h + 16; ch += 8) {
        const f16x8 v = *(const f16x8*)(bufA + site * kSP + ch);
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = fmaf(w6s[tp * kC + ch + u], (float)v[u], acc);
      }
    }
    const float o = (acc + __shfl_xor(acc, 32)) + b6[0];
    if (h == 0) out[img + (size_t)(y0 + ry) * W + x0 + rx] = o;
    // the tile's sum of the output (for the mean removal), in a fixed order
    double sd = h == 0 ? (double)o : 0.0;
    for (int st = 32; st >= 1; st >>= 1) sd += __shfl_xor(sd, st);
    if (lane == 0) psum[wave] = sd;
  }
  __syncthreads();
  if (t == 0) {
    double sd = psum[0];
    for (int w = 1; w < kFR * kFC / 32; ++w) sd += psum[w];
    part[blockIdx.x] = sd;
  }
}

// Backward-weights of a 32 -> 32 layer: dW[co][tap][ci] = sum_s dz[s][co] x[s+off][ci],
// db[co] = sum_s dz[s][co].  Block: one tile; operands staged channel-major in LDS (x, FP16,
// rounded to BF16 there, in three copies shifted by dx so every 16-byte read is aligned).  Per-block partial sums go
// to part[block][9248] (reduced later in a fixed order).
__global__ __launch_bounds__(256) void k_conv_mid_bwdw(const __bf16* __restrict__ dz,
                                                       const _Float16* __restrict__ x,
                                                       float* __restrict__ part, int H, int W,
                                                       int tiles_x, int tiles_y, int ntiles) {
  constexpr int CS = kTC + 8;  // column stride of the channel-major planes (bf16)
  __shared__ __attribute__((aligned(16))) __bf16 dzT[kC * kTR * CS];
  __shared__ __attribute__((aligned(16))) __bf16 xT[3][kC * (kTR + 2) * CS];
  __shared__ float red[4][kC * 32];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int j = lane & 31, kh = (lane >> 5) * 8;
  f32x16 acc[9];
#pragma unroll
  for (int tp = 0; tp < 9; ++tp) acc[tp] = f32x16{};
  float dbacc = 0.0f;
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int tx = tile % tiles_x;
    const int r1 = tile / tiles_x;
    const int ty = r1 % tiles_y, b = r1 / tiles_y;
    const int x0 = tx * kTC, y0 = ty * kTR;
    const size_t img = (size_t)b * H * W;
    __syncthreads();
    for (int q = t; q < kTR * kTC * 4; q += 256) {  // dz tile, transposed
      const int site = q >> 2, ch = (q & 3) * 8;
      const int ly = site / kTC, lx = site % kTC;
      const bf16x8 v = *(const bf16x8*)(dz + (img + (size_t)(y0 + ly) * W + x0 + lx) * kC + ch);
#pragma unroll
      for (int u = 0; u < 8; ++u) dzT[(ch + u) * kTR * CS + ly * CS + lx] = v[u];
    }
    for (int q = t; q < (kTR + 2) * (kTC + 2) * 4; q += 256) {  // x tile + halo, 3 shifts
      const int site = q >> 2, ch = (q & 3) * 8;
      const int ly = site / (kTC + 2), lx = site % (kTC + 2);  // lx: column x0 + lx - 1
      const int gy = wrap1(y0 + ly - 1, H), gx = wrap1(x0 + lx - 1, W);
      const f16x8 v16 = *(const f16x8*)(x + (img + (size_t)gy * W + gx) * kC + ch);
      bf16x8 v;
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = (__bf16)(float)v16[u];
#pragma unroll
      for (int sh = 0; sh < 3; ++sh) {  // copy sh holds column (x0 + c + sh - 1) at c
        const int c = lx - sh;
        if (c >= 0 && c < kTC) {
#pragma unroll
          for (int u = 0; u < 8; ++u) xT[sh][(ch + u) * (kTR + 2) * CS + ly * CS + c] = v[u];
        }
      }
    }
    __syncthreads();
    for (int rr = 0; rr < 2; ++rr) {
      const int r = wave * 2 + rr;
#pragma unroll
      for (int h = 0; h < 2; ++h) {  // 16-site chunks of the row
        const bf16x8 a = *(const bf16x8*)(dzT + j * kTR * CS + r * CS + h * 16 + kh);
#pragma unroll
        for (int tp = 0; tp < 9; ++tp) {
          const int dy = tp / 3 - 1, sh = tp % 3;  // dx = sh - 1
          const bf16x8 bb =
              *(const bf16x8*)(xT[sh] + j * (kTR + 2) * CS + (r + 1 + dy) * CS + h * 16 + kh);
          acc[tp] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bb, acc[tp], 0, 0, 0);
        }
      }
    }
    // bias gradient: thread (co = t % 32, row = t / 32)
    {
      const int co = t & 31, ly = t >> 5;
      float s = 0.0f;
      for (int lx = 0; lx < kTC; ++lx) s += (float)dzT[co * kTR * CS + ly * CS + lx];
      dbacc += s;
    }
  }
  // reduce the four waves' accumulators tap by tap: acc[tp][q] -> D[co = row i][ci = j]
  float* out = part + (size_t)blockIdx.x * (kMidW + kC);
  for (int tp = 0; tp < 9; ++tp) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int co = (q >> 2) * 8 + (lane >> 5) * 4 + (q & 3);
      red[wave][co * 32 + j] = acc[tp][q];
    }
    __syncthreads();
    for (int e = t; e < kC * 32; e += 256) {
      const int co = e >> 5, ci = e & 31;
      out[co * 288 + tp * 32 + ci] = (red[0][e] + red[1][e]) + (red[2][e] + red[3][e]);
    }
  }
  __syncthreads();
  red[0][t] = dbacc;
  __syncthreads();
  if (t < kC) {
    float s = 0.0f;
    for (int ly = 0; ly < kTR; ++ly) s += red[0][ly * 32 + t];
    out[kMidW + t] = s;
  }
}

// ---- layer 6 (32 -> 1), vector ALU, FP32 weights ------------------------------------------
__global__ __launch_bounds__(256) void k_conv6(const _Float16* __restrict__ a5,
                                               const float* __restrict__ w6,
                                               const float* __restrict__ b6,
                                               float* __restrict__ out, int H, int W, int B) {
  const int HW = H * W;
  const long s = (long)blockIdx.x * 256 + threadIdx.x;
  if (s >= (long)B * HW) return;
  const int b = (int)(s / HW), p = (int)(s % HW), y = p / W, x = p % W;
  const size_t img = (size_t)b * HW;
  float acc[2] = {0.0f, 0.0f};  // channels 0..15 and 16..31, each over taps, then channels
  for (int tp = 0; tp < 9; ++tp) {
    const size_t n = img + (size_t)wrap1(y + tp / 3 - 1, H) * W + wrap1(x + tp % 3 - 1, W);
#pragma unroll
    for (int ch = 0; ch < kC; ch += 8) {
      const f16x8 v = *(const f16x8*)(a5 + n * kC + ch);
#pragma unroll
      for (int u = 0; u < 8; ++u) acc[ch / 16] = fmaf(w6[tp * kC + ch + u], (float)v[u], acc[ch / 16]);
    }
  }
  out[s] = (acc[0] + acc[1]) + b6[0];
}

// Loss on the normalised increment: d = out - (target - c)/scale; L = mean(d^2);
// dout = 2 d / N.  Per-block partial sums of d^2 in double.
template <class TS>
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ out, const TS* __restrict__ c,
                                              const TS* __restrict__ target,
                                              const int* __restrict__ idx, float inv_scale,
                                              float* __restrict__ dout, double* __restrict__ part,
                                              int HW, int B) {
  __shared__ double sh[4];
  const long s = (long)blockIdx.x * 256 + threadIdx.x;
  double d2 = 0.0;
  if (s < (long)B * HW) {
    const int b = (int)(s / HW), p = (int)(s % HW);
    const size_t src = (size_t)(idx ? idx[b] : b) * HW + p;
    const float delta = (float)((double)target[src] - (double)c[src]) * inv_scale;
    const float d = out[s] - delta;
    dout[s] = 2.0f * d / (float)((long)B * HW);
    d2 = (double)d * (double)d;
  }
  for (int st = 32; st >= 1; st >>= 1) d2 += __shfl_xor(d2, st);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = d2;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// Backward-data of layer 6 fused with gelu': dz5[s][ci] = gelu'(z5) sum_tap W6[tap][ci]
// dout[s - off(tap)].
__global__ __launch_bounds__(256) void k_conv6_bwdd(const float* __restrict__ dout,
                                                    const float* __restrict__ w6,
                                                    const float* __restrict__ z5,
                                                    __bf16* __restrict__ dz5, int H, int W, int B) {
  const int HW = H * W;
  const long s = (long)blockIdx.x * 256 + threadIdx.x;
  if (s >= (long)B * HW) return;
  const int b = (int)(s / HW), p = (int)(s % HW), y = p / W, x = p % W;
  const size_t img = (size_t)b * HW;
  float dv[9];
#pragma unroll
  for (int tp = 0; tp < 9; ++tp)
    dv[tp] = dout[img + (size_t)wrap1(y - (tp / 3 - 1), H) * W + wrap1(x - (tp % 3 - 1), W)];
#pragma unroll
  for (int ch = 0; ch < kC; ch += 8) {
    bf16x8 o;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      float acc = 0.0f;
#pragma unroll
      for (int tp = 0; tp < 9; ++tp) acc = fmaf(w6[tp * kC + ch + u], dv[tp], acc);
      o[u] = (__bf16)(acc * gelu_grad_f(z5[(size_t)s * kC + ch + u]));
    }
    *(bf16x8*)(dz5 + (size_t)s * kC + ch) = o;
  }
}

// Backward-weights of layer 6: thread e < 288 owns W6[tap = e/32][ci = e%32], e = 288 owns
// b6.  A block sums over a contiguous range of sites.
__global__ __launch_bounds__(320) void k_conv6_bwdw(const float* __restrict__ dout,
                                                    const _Float16* __restrict__ a5,
                                                    float* __restrict__ part, int H, int W,
                                                    long nsites, long per_block) {
  const int e = threadIdx.x;
  const int tp = e / 32, ci = e % 32;
  const int HW = H * W;
  float acc = 0.0f;
  const long s0 = (long)blockIdx.x * per_block;
  const long s1 = min(nsites, s0 + per_block);
  for (long s = s0; s < s1; ++s) {
    const float d = dout[s];
    if (e < 288) {
      const int b = (int)(s / HW), p = (int)(s % HW), y = p / W, x = p % W;
      const size_t n = (size_t)b * HW + (size_t)wrap1(y + tp / 3 - 1, H) * W + wrap1(x + tp % 3 - 1, W);
      acc = fmaf(d, (float)a5[n * kC + ci], acc);
    } else if (e == 288) {
      acc += d;
    }
  }
  if (e < 289) part[(size_t)blockIdx.x * 289 + e] = acc;
}

// Backward-weights of layer 1: thread e < 288 owns W1[co = e/9][tap = e%9], 288 <= e < 320
// owns b1[e-288].
template <class TS>
__global__ __launch_bounds__(320) void k_conv1_bwdw(const __bf16* __restrict__ dz1,
                                                    const TS* __restrict__ c,
                                                    const int* __restrict__ idx,
                                                    float* __restrict__ part, int H, int W,
                                                    long nsites, long per_block) {
  const int e = threadIdx.x;
  const int co = e < 288 ? e / 9 : e - 288, tp = e % 9;
  const int HW = H * W;
  float acc = 0.0f;
  const long s0 = (long)blockIdx.x * per_block;
  const long s1 = min(nsites, s0 + per_block);
  for (long s = s0; s < s1; ++s) {
    const int b = (int)(s / HW), p = (int)(s % HW), y = p / W, x = p % W;
    const float d = (float)dz1[(size_t)s * kC + co];
    if (e < 288) {
      const TS* img = c + (size_t)(idx ? idx[b] : b) * HW;
      const float xv = (float)img[wrap1(y + tp / 3 - 1, H) * W + wrap1(x + tp % 3 - 1, W)];
      acc = fmaf(d, xv, acc);
    } else {
      acc += d;
    }
  }
  part[(size_t)blockIdx.x * 320 + e] = acc;
}

// out[i] = sum over p of part[p][i], p in order (deterministic); with accumulate, added to
// out[i] instead of replacing it.
__global__ __launch_bounds__(256) void k_reduce_parts(const float* __restrict__ part, int nparts,
                                                      int width, float* __restrict__ out,
                                                      int accumulate) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= width) return;
  float s = 0.0f;
  for (int p = 0; p < nparts; ++p) s += part[(size_t)p * width + i];
  out[i] = accumulate ? out[i] + s : s;
}

// ---- unrolled-chain training (several residual applications, gradient through all) -----

// Input gradient of layer 1, added to dx: dx[s] += sum_tap sum_co W1[co][tap] dz1[s - off(tap)][co]
// (the forward pass reads x[s + off(tap)]).
__global__ __launch_bounds__(256) void k_conv1_bwdd(const __bf16* __restrict__ dz1,
                                                    const float* __restrict__ w1,
                                                    float* __restrict__ dx, int H, int W, int B) {
  const int HW = H * W;
  const long s = (long)blockIdx.x * 256 + threadIdx.x;
  if (s >= (long)B * HW) return;
  const int b = (int)(s / HW), p = (int)(s % HW), y = p / W, x = p % W;
  const size_t img = (size_t)b * HW;
  float acc = 0.0f;
#pragma unroll
  for (int tp = 0; tp < 9; ++tp) {
    const size_t n = img + (size_t)wrap1(y - (tp / 3 - 1), H) * W + wrap1(x - (tp % 3 - 1), W);
#pragma unroll
    for (int ch = 0; ch < kC; ch += 8) {
      const bf16x8 v = *(const bf16x8*)(dz1 + n * kC + ch);
#pragma unroll
      for (int u = 0; u < 8; ++u) acc = fmaf(w1[(ch + u) * 9 + tp], (float)v[u], acc);
    }
  }
  dx[s] += acc;
}

// sums[b] = sum of image b's HW values in double (block b; fixed order).
__global__ __launch_bounds__(1024) void k_img_sum(const float* __restrict__ v, long HW,
                                                  double* __restrict__ sums) {
  __shared__ double sh[16];
  const float* img = v + (size_t)blockIdx.x * HW;
  double s = 0.0;
  for (long i = threadIdx.x; i < HW; i += 1024) s += (double)img[i];
  for (int st = 32; st >= 1; st >>= 1) s += __shfl_xor(s, st);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < 16; ++w) t += sh[w];
    sums[blockIdx.x] = t;
  }
}

// c_new = c + scale (out - mean(out)) with image b's mean sums[b] / HW (FP32 state).
__global__ __launch_bounds__(256) void k_chain_step(const float* __restrict__ c,
                                                    const float* __restrict__ out,
                                                    const double* __restrict__ sums,
                                                    double scale, long HW, long n,
                                                    float* __restrict__ cn) {
  const long s = (long)blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  const double mean = sums[s / HW] / (double)HW;
  cn[s] = (float)((double)c[s] + scale * ((double)out[s] - mean));
}

// Output gradient of one application from the state gradient G: dout = scale (G - mean(G)).
__global__ __launch_bounds__(256) void k_chain_dout(const float* __restrict__ G,
                                                    const double* __restrict__ sums,
                                                    double scale, long HW, long n,
                                                    float* __restrict__ dout) {
  const long s = (long)blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  const double mean = sums[s / HW] / (double)HW;
  dout[s] = (float)(scale * ((double)G[s] - mean));
}

// Loss term of the state after k applications: e = (c_k - t) / scale with t = snapshot
// idx[b] + k of data; G = (accumulate ? G : 0) + coef e; part[block] = sum of w e^2.
template <class TS>
__global__ __launch_bounds__(256) void k_chain_err(const float* __restrict__ ck,
                                                   const TS* __restrict__ data,
                                                   const int* __restrict__ idx, int k,
                                                   double inv_scale, double w, double coef,
                                                   int accumulate, float* __restrict__ G,
                                                   double* __restrict__ part, long HW, long n) {
  __shared__ double sh[4];
  const long s = (long)blockIdx.x * 256 + threadIdx.x;
  double e2 = 0.0;
  if (s < n) {
    const long b = s / HW, p = s % HW;
    const double t = (double)data[(size_t)(idx[b] + k) * HW + p];
    const double e = ((double)ck[s] - t) * inv_scale;
    const float ge = (float)(coef * e);
    G[s] = accumulate ? G[s] + ge : ge;
    e2 = w * e * e;
  }
  for (int st = 32; st >= 1; st >>= 1) e2 += __shfl_xor(e2, st);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = e2;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// c0[b] = data[idx[b]] in FP32.
template <class TS>
__global__ __launch_bounds__(256) void k_gather(const TS* __restrict__ data,
                                                const int* __restrict__ idx, long HW, long n,
                                                float* __restrict__ c0) {
  const long s = (long)blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  c0[s] = (float)data[(size_t)idx[s / HW] * HW + s % HW];
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, int n,
                                              float lr, float b1, float b2, float eps, float bc1,
                                              float bc2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// FP32 master parameters -> 16-bit working copies of the 32 -> 32 matrices, as the
// matrix-core A operands: forward in FP16 (wmid, rows co, columns tap * 32 + ci) and
// backward-data in BF16 (wmidT, rows ci, columns tap' * 32 + co with tap' = 8 - tap, i.e.
// [ci][tap'*32 + co] = W[co][(8 - tap')*32 + ci]).  W1 and W6 are used in FP32 directly.  Stored fragment-major: for K block kc (16 columns),
// lane (h, j) of a wave holds row j, columns kc * 16 + 8 h .. + 7, at
// ((kc * 64 + 32 h + j) * 8), so each of a wave's 18 fragment loads reads 1 KiB contiguously.
__device__ __forceinline__ int frag_index(int row, int col) {
  return ((col / 16) * 64 + ((col % 16) / 8) * 32 + row) * 8 + col % 8;
}
__global__ __launch_bounds__(256) void k_pack(const float* __restrict__ p, _Float16* __restrict__ wmid,
                                              __bf16* __restrict__ wmidT) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < 4 * kMidW) {
    const int l = i / kMidW, e = i % kMidW;
    const float* W = p + kOffMid + l * kMidStride;
    const int r = e / 288, k = e % 288;
    wmid[l * kMidW + frag_index(r, k)] = (_Float16)W[e];
    const int tpp = k / 32, co = k % 32;  // wmidT row r = ci, column k = tap' * 32 + co
    wmidT[l * kMidW + frag_index(r, k)] = (__bf16)W[co * 288 + (8 - tpp) * 32 + r];
  }
}


// Residual update with the mean of the increment removed (mass conservation):
// c_new = c + scale * (out - sum(out)/N).
template <class TS>
__global__ __launch_bounds__(256) void k_apply(const TS* __restrict__ c, TS* __restrict__ cn,
                                               const float* __restrict__ out,
                                               const double* __restrict__ part, int nparts,
                                               double scale, long n) {
  __shared__ double sum;
  block_sum_parts(part, nparts, &sum);
  const long s = (long)blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  const double mean = sum / (double)n;
  cn[s] = (TS)((double)c[s] + scale * ((double)out[s] - mean));
}



template <class TS, class TD>
__global__ __launch_bounds__(256) void k_cast(const TS* __restrict__ a, TD* __restrict__ b, long n) {
  const long s = (long)blockIdx.x * 256 + threadIdx.x;
  if (s < n) b[s] = (TD)a[s];
}

}  // namespace

struct gc_cnn {
  gc_context* ctx = nullptr;
  int H = 0, W = 0, Bmax = 0;
  double scale = 1.0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  _Float16* wmid = nullptr;  // forward weights of the 32 -> 32 layers (FP16)
  __bf16* wmidT = nullptr;   // backward-data weights (BF16)
  float* z[5] = {};
  _Float16* a[5] = {};       // activations (FP16)
  __bf16 *dzA = nullptr, *dzB = nullptr;
  float *out = nullptr, *dout = nullptr;
  float* part = nullptr;
  size_t part_floats = 0;
  double* dpart = nullptr;
  int* idx = nullptr;
  unsigned* done = nullptr;  // learned integration: blocks finished (last-block reduction)
  void* hist = nullptr;     // learned integration: proposals of one chunk
  double* diag = nullptr;   // learned integration: (mass, free energy) per proposal
  float* out2 = nullptr;    // learned integration: second network output (chained steps)
  double* hdiag = nullptr;  // learned integration: pinned copies of two chunks' diagnostics
  hipEvent_t ev[2] = {};    // learned integration: each chunk's diagnostics copied
  double* cpart = nullptr;  // learned integration: second tile sums, then reduction partials
  float* chain_c = nullptr;    // chain training: states c_0 .. c_K of the batch (FP32)
  float* chain_g = nullptr;    // chain training: gradient of the loss with respect to c_k
  double* chain_sum = nullptr; // chain training: per-image sums
  double* chain_part = nullptr;// chain training: loss partial sums, one row per step
  int chain_k = 0;             // chain training: steps the buffers hold
  int adam_t = 0;
  bool packed = false;
};

namespace {

constexpr int kPartBlocks = 512;  // blocks of the weight-gradient kernels

gc_status cnn_alloc(void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return set_error(GC_ERR_NOMEM, "gridcalc: CNN allocation of %zu bytes failed", bytes);
  }
  return GC_OK;
}

gc_status pack(gc_cnn* n) {
  hipLaunchKernelGGL(k_pack, dim3((4 * kMidW + 255) / 256), dim3(256), 0, n->ctx->stream, n->p,
                     n->wmid, n->wmidT);
  GC_LAUNCHED();
  n->packed = true;
  return GC_OK;
}

gc_status check_cnn(gc_cnn* n, int B) {
  if (!n) return set_error(GC_ERR_INVALID, "gridcalc: null CNN");
  GC_TRY(check_ctx(n->ctx));
  if (B < 1 || B > n->Bmax)
    return set_error(GC_ERR_INVALID, "gridcalc: batch %d outside [1, %d]", B, n->Bmax);
  return GC_OK;
}

gc_status upload_idx(gc_cnn* n, const int* idx, int B) {
  if (!idx) return GC_OK;
  GC_HIP(hipMemcpyAsync(n->idx, idx, B * sizeof(int), hipMemcpyHostToDevice, n->ctx->stream));
  return GC_OK;
}

// Forward pass on B images (c[idx[b]] or c[b]); train keeps the pre-activations.
template <class TS>
gc_status forward(gc_cnn* n, const TS* c, const int* d_idx, int B, bool train) {
  if (!n->packed) GC_TRY(pack(n));
  const long ns = (long)B * n->H * n->W;
  const unsigned g1 = (unsigned)((ns + 127) / 128);  // 128 sites per block
  hipStream_t st = n->ctx->stream;
  hipLaunchKernelGGL((k_conv1<TS>), dim3(g1), dim3(256), 0, st, c, d_idx, n->p + kOffW1, n->p + kOffB1,
                     train ? n->z[0] : nullptr, n->a[0], n->H, n->W, B);
  GC_LAUNCHED();
  const int tx = n->W / kTC, ty = n->H / kTR;
  for (int l = 0; l < 4; ++l) {
    hipLaunchKernelGGL((k_conv_mid<0>), dim3((unsigned)(tx * ty * B)), dim3(256), 0, st, n->a[l],
                       n->wmid + (size_t)l * kMidW, n->p + kOffMid + l * kMidStride + kMidW,
                       nullptr, train ? n->z[l + 1] : nullptr, n->a[l + 1], n->H, n->W, tx, ty);
    GC_LAUNCHED();
  }
  hipLaunchKernelGGL(k_conv6, dim3(g1), dim3(256), 0, st, n->a[4], n->p + kOffW6, n->p + kOffB6, n->out,
                     n->H, n->W, B);
  GC_LAUNCHED();
  return GC_OK;
}

// Inference: the whole network in two kernels per 8 x 32 tile (k_fwd_a, k_fwd_b), layer 3's
// activations passing through n->a[2]; output to n->out.
template <class TS>
gc_status forward_fused(gc_cnn* n, const TS* c, int B, float* out = nullptr,
                        double* part = nullptr) {
  if (!n->packed) GC_TRY(pack(n));
  const int tx = n->W / kTC, ty = n->H / kTR;
  const dim3 grid((unsigned)(tx * ty * B));
  hipLaunchKernelGGL((k_fwd_a<TS, false>), grid, dim3(kFNT), 0, n->ctx->stream, c,
                     n->p + kOffW1, n->p + kOffB1, n->wmid, n->p, n->a[2], n->H, n->W, tx, ty,
                     Chain<TS>{});
  GC_LAUNCHED();
  hipLaunchKernelGGL(k_fwd_b, grid, dim3(kFNT), 0, n->ctx->stream, n->a[2], n->wmid, n->p,
                     n->p + kOffW6, n->p + kOffB6, out ? out : n->out, part ? part : n->dpart,
                     n->H, n->W, tx, ty);
  GC_LAUNCHED();
  return GC_OK;
}

// One chained macro step on one image (see Chain): the network's output to out, its tile
// sums to part.
template <class TS>
gc_status forward_chain(gc_cnn* n, const Chain<TS>& ch, float* out, double* part) {
  if (!n->packed) GC_TRY(pack(n));
  const int tx = n->W / kTC, ty = n->H / kTR;
  const dim3 grid((unsigned)(tx * ty));
  hipLaunchKernelGGL((k_fwd_a<TS, true>), grid, dim3(kFNT), 0, n->ctx->stream,
                     (const TS*)nullptr, n->p + kOffW1, n->p + kOffB1, n->wmid, n->p, n->a[2],
                     n->H, n->W, tx, ty, ch);
  GC_LAUNCHED();
  hipLaunchKernelGGL(k_fwd_b, grid, dim3(kFNT), 0, n->ctx->stream, n->a[2], n->wmid, n->p,
                     n->p + kOffW6, n->p + kOffB6, out, part, n->H, n->W, tx, ty);
  GC_LAUNCHED();
  return GC_OK;
}

// Backward pass from dout; gradients into n->g (added to it with accumulate); with dx, the
// gradient with respect to the network input is added to dx.
template <class TS>
gc_status backward(gc_cnn* n, const TS* c, const int* d_idx, int B, bool accumulate = false,
                   float* dx = nullptr) {
  const int acc = accumulate ? 1 : 0;
  hipStream_t st = n->ctx->stream;
  const long ns = (long)B * n->H * n->W;
  const unsigned g1 = (unsigned)((ns + 255) / 256);
  const long per = (ns + kPartBlocks - 1) / kPartBlocks;
  // layer 6
  hipLaunchKernelGGL(k_conv6_bwdw, dim3(kPartBlocks), dim3(320), 0, st, n->dout, n->a[4], n->part,
                     n->H, n->W, ns, per);
  GC_LAUNCHED();
  hipLaunchKernelGGL(k_reduce_parts, dim3((289 + 255) / 256), dim3(256), 0, st, n->part,
                     kPartBlocks, 289, n->g + kOffW6, acc);
  GC_LAUNCHED();
  hipLaunchKernelGGL(k_conv6_bwdd, dim3(g1), dim3(256), 0, st, n->dout, n->p + kOffW6, n->z[4], n->dzA,
                     n->H, n->W, B);
  GC_LAUNCHED();
  const int tx = n->W / kTC, ty = n->H / kTR, ntiles = tx * ty * B;
  __bf16* dz = n->dzA;
  __bf16* dzn = n->dzB;
  for (int l = 3; l >= 0; --l) {  // layer l+2 in 1-based numbering: a[l] -> z[l+1]
    hipLaunchKernelGGL(k_conv_mid_bwdw, dim3(kPartBlocks), dim3(256), 0, st, dz, n->a[l],
                       n->part, n->H, n->W, tx, ty, ntiles);
    GC_LAUNCHED();
    hipLaunchKernelGGL(k_reduce_parts, dim3((kMidStride + 255) / 256), dim3(256), 0, st, n->part,
                       kPartBlocks, kMidStride, n->g + kOffMid + l * kMidStride, acc);
    GC_LAUNCHED();
    hipLaunchKernelGGL((k_conv_mid<1>), dim3((unsigned)ntiles), dim3(256), 0, st, dz,
                       n->wmidT + (size_t)l * kMidW, nullptr, n->z[l], nullptr, dzn, n->H, n->W,
                       tx, ty);
    GC_LAUNCHED();
    __bf16* tmp = dz;
    dz = dzn;
    dzn = tmp;
  }
  hipLaunchKernelGGL((k_conv1_bwdw<TS>), dim3(kPartBlocks), dim3(320), 0, st, dz, c, d_idx,
                     n->part, n->H, n->W, ns, per);
  GC_LAUNCHED();
  hipLaunchKernelGGL(k_reduce_parts, dim3((320 + 255) / 256), dim3(256), 0, st, n->part,
                     kPartBlocks, 320, n->g + kOffW1, acc);
  GC_LAUNCHED();
  if (dx) {
    hipLaunchKernelGGL(k_conv1_bwdd, dim3(g1), dim3(256), 0, st, dz, n->p + kOffW1, dx, n->H,
                       n->W, B);
    GC_LAUNCHED();
  }
  return GC_OK;
}

template <class TS>
gc_status loss_and_dout(gc_cnn* n, const TS* c, const TS* target, const int* d_idx, int B,
                        double* loss) {
  const long ns = (long)B * n->H * n->W;
  const unsigned g1 = (unsigned)((ns + 255) / 256);
  hipLaunchKernelGGL((k_loss<TS>), dim3(g1), dim3(256), 0, n->ctx->stream, n->out, c, target,
                     d_idx, (float)(1.0 / n->scale), n->dout, n->dpart, n->H * n->W, B);
  GC_LAUNCHED();
  if (loss) {
    std::vector<double> h(g1);
    GC_HIP(hipMemcpyAsync(h.data(), n->dpart, g1 * sizeof(double), hipMemcpyDeviceToHost,
                          n->ctx->stream));
    GC_HIP(hipStreamSynchronize(n->ctx->stream));
    double s = 0.0;
    for (double x : h) s += x;
    *loss = s / (double)ns;
  }
  return GC_OK;
}

constexpr int kChainMax = 8;  // residual applications one chain evaluation may unroll

gc_status ensure_chain(gc_cnn* n, int K) {
  if (n->chain_k >= K) return GC_OK;
  GC_HIP(hipStreamSynchronize(n->ctx->stream));
  void* old[] = {n->chain_c, n->chain_g, n->chain_sum, n->chain_part};
  for (void* p : old)
    if (p) (void)hipFree(p);
  n->chain_c = n->chain_g = nullptr;
  n->chain_sum = n->chain_part = nullptr;
  n->chain_k = 0;
  const size_t ns = (size_t)n->Bmax * n->H * n->W, g1 = (ns + 255) / 256;
  GC_TRY(cnn_alloc((void**)&n->chain_c, (size_t)(K + 1) * ns * sizeof(float)));
  GC_TRY(cnn_alloc((void**)&n->chain_g, ns * sizeof(float)));
  GC_TRY(cnn_alloc((void**)&n->chain_sum, (size_t)n->Bmax * sizeof(double)));
  GC_TRY(cnn_alloc((void**)&n->chain_part, (size_t)K * g1 * sizeof(double)));
  n->chain_k = K;
  return GC_OK;
}

// The chain's states c_0..c_K by the per-layer kernels (so that the recomputed training
// forward pass of each application sees the same bits), then the applications in reverse:
// output gradient scale (G - mean G), activations recomputed, weight gradients accumulated,
// G carried to the previous state through the residual path and the input gradient.
template <class TS>
gc_status chain_loss_grad(gc_cnn* n, const TS* data, const int* d_idx, int B, int K,
                          const double* w, double* loss) {
  hipStream_t st = n->ctx->stream;
  const long HW = (long)n->H * n->W, ns = (long)B * HW;
  const unsigned g1 = (unsigned)((ns + 255) / 256);
  const size_t stride = (size_t)n->Bmax * HW;
  auto state = [&](int k) { return n->chain_c + (size_t)k * stride; };
  const double inv_s = 1.0 / n->scale, N = (double)ns;
  hipLaunchKernelGGL((k_gather<TS>), dim3(g1), dim3(256), 0, st, data, d_idx, HW, ns, state(0));
  GC_LAUNCHED();
  for (int k = 1; k <= K; ++k) {
    GC_TRY(forward<float>(n, state(k - 1), nullptr, B, false));
    hipLaunchKernelGGL(k_img_sum, dim3(B), dim3(1024), 0, st, n->out, HW, n->chain_sum);
    GC_LAUNCHED();
    hipLaunchKernelGGL(k_chain_step, dim3(g1), dim3(256), 0, st, state(k - 1), n->out,
                       n->chain_sum, n->scale, HW, ns, state(k));
    GC_LAUNCHED();
  }
  hipLaunchKernelGGL((k_chain_err<TS>), dim3(g1), dim3(256), 0, st, state(K), data, d_idx, K,
                     inv_s, w[K - 1], 2.0 * w[K - 1] * inv_s / N, 0, n->chain_g,
                     n->chain_part + (size_t)(K - 1) * g1, HW, ns);
  GC_LAUNCHED();
  for (int k = K; k >= 1; --k) {
    hipLaunchKernelGGL(k_img_sum, dim3(B), dim3(1024), 0, st, n->chain_g, HW, n->chain_sum);
    GC_LAUNCHED();
    hipLaunchKernelGGL(k_chain_dout, dim3(g1), dim3(256), 0, st, n->chain_g, n->chain_sum,
                       n->scale, HW, ns, n->dout);
    GC_LAUNCHED();
    GC_TRY(forward<float>(n, state(k - 1), nullptr, B, true));
    GC_TRY(backward<float>(n, state(k - 1), nullptr, B, k < K, k > 1 ? n->chain_g : nullptr));
    if (k > 1) {
      hipLaunchKernelGGL((k_chain_err<TS>), dim3(g1), dim3(256), 0, st, state(k - 1), data, d_idx,
                         k - 1, inv_s, w[k - 2], 2.0 * w[k - 2] * inv_s / N, 1, n->chain_g,
                         n->chain_part + (size_t)(k - 2) * g1, HW, ns);
      GC_LAUNCHED();
    }
  }
  if (loss) {
    std::vector<double> h((size_t)K * g1);
    GC_HIP(hipMemcpyAsync(h.data(), n->chain_part, h.size() * sizeof(double),
                          hipMemcpyDeviceToHost, st));
    GC_HIP(hipStreamSynchronize(st));
    double s = 0.0;
    for (double x : h) s += x;
    *loss = s / N;
  }
  return GC_OK;
}

}  // namespace

extern "C" {

gc_status gc_cnn_create(gc_context* ctx, int H, int W, int batch_max, gc_cnn** out) {
  GC_TRY(check_ctx(ctx));
  if (!out) return set_error(GC_ERR_INVALID, "gridcalc: null CNN pointer");
  if (H < kTR || W < kTC || H % kTR || W % kTC || batch_max < 1)
    return set_error(GC_ERR_UNSUPPORTED,
                     "gridcalc: the corrector needs H a multiple of %d and W a multiple of %d",
                     kTR, kTC);
  *out = nullptr;
  gc_cnn* n = new gc_cnn();
  n->ctx = ctx;
  n->H = H;
  n->W = W;
  n->Bmax = batch_max;
  const size_t ns = (size_t)batch_max * H * W;
  const size_t np4 = kNParams * sizeof(float);
  gc_status st = GC_OK;
  auto A = [&](void** p, size_t b) {
    if (st == GC_OK) st = cnn_alloc(p, b);
  };
  A((void**)&n->p, np4);
  A((void**)&n->g, np4);
  A((void**)&n->m, np4);
  A((void**)&n->v, np4);
  A((void**)&n->wmid, 4 * kMidW * 2);
  A((void**)&n->wmidT, 4 * kMidW * 2);
  for (int l = 0; l < 5; ++l) {
    A((void**)&n->z[l], ns * kC * sizeof(float));
    A((void**)&n->a[l], ns * kC * 2);
  }
  A((void**)&n->dzA, ns * kC * 2);
  A((void**)&n->dzB, ns * kC * 2);
  A((void**)&n->out, ns * sizeof(float));
  A((void**)&n->dout, ns * sizeof(float));
  n->part_floats = (size_t)kPartBlocks * kMidStride;
  A((void**)&n->part, n->part_floats * sizeof(float));
  A((void**)&n->dpart, ((ns + 255) / 256 + 1024) * sizeof(double));
  A((void**)&n->idx, batch_max * sizeof(int));
  A((void**)&n->done, sizeof(unsigned));
  if (st == GC_OK && (hipMemsetAsync(n->p, 0, np4, ctx->stream) != hipSuccess ||
                      hipMemsetAsync(n->done, 0, sizeof(unsigned), ctx->stream) != hipSuccess ||
                      hipMemsetAsync(n->m, 0, np4, ctx->stream) != hipSuccess ||
                      hipMemsetAsync(n->v, 0, np4, ctx->stream) != hipSuccess))
    st = set_error(GC_ERR_DEVICE, "gridcalc: CNN initialisation failed");
  if (st != GC_OK) {
    gc_cnn_destroy(n);
    return st;
  }
  *out = n;
  return GC_OK;
}

gc_status gc_cnn_destroy(gc_cnn* n) {
  if (!n) return GC_OK;
  if (n->ctx) (void)hipStreamSynchronize(n->ctx->stream);
  void* ptrs[] = {n->p, n->g, n->m, n->v, n->wmid, n->wmidT, n->dzA, n->dzB,
                  n->out, n->dout, n->part, n->dpart, n->idx, n->done, n->hist, n->diag,
                  n->out2, n->cpart, n->chain_c, n->chain_g, n->chain_sum, n->chain_part};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (n->hdiag) (void)hipHostFree(n->hdiag);
  for (hipEvent_t e : n->ev)
    if (e) (void)hipEventDestroy(e);
  for (int l = 0; l < 5; ++l) {
    if (n->z[l]) (void)hipFree(n->z[l]);
    if (n->a[l]) (void)hipFree(n->a[l]);
  }
  delete n;
  return GC_OK;
}

int gc_cnn_param_count(void) { return kNParams; }

gc_status gc_cnn_set_params(gc_cnn* n, const float* host, double scale) {
  GC_TRY(check_cnn(n, 1));
  if (!host) return set_error(GC_ERR_INVALID, "gridcalc: null parameters");
  if (!(scale > 0.0)) return set_error(GC_ERR_INVALID, "gridcalc: output scale must be > 0");
  GC_HIP(hipMemcpyAsync(n->p, host, kNParams * sizeof(float), hipMemcpyHostToDevice,
                        n->ctx->stream));
  GC_HIP(hipMemsetAsync(n->m, 0, kNParams * sizeof(float), n->ctx->stream));
  GC_HIP(hipMemsetAsync(n->v, 0, kNParams * sizeof(float), n->ctx->stream));
  GC_HIP(hipStreamSynchronize(n->ctx->stream));
  n->scale = scale;
  n->adam_t = 0;
  n->packed = false;
  return GC_OK;
}

gc_status gc_cnn_get_params(gc_cnn* n, float* host) {
  GC_TRY(check_cnn(n, 1));
  if (!host) return set_error(GC_ERR_INVALID, "gridcalc: null parameters");
  GC_HIP(hipMemcpyAsync(host, n->p, kNParams * sizeof(float), hipMemcpyDeviceToHost,
                        n->ctx->stream));
  GC_HIP(hipStreamSynchronize(n->ctx->stream));
  return GC_OK;
}

gc_status gc_cnn_get_grads(gc_cnn* n, float* host) {
  GC_TRY(check_cnn(n, 1));
  if (!host) return set_error(GC_ERR_INVALID, "gridcalc: null gradient buffer");
  GC_HIP(hipMemcpyAsync(host, n->g, kNParams * sizeof(float), hipMemcpyDeviceToHost,
                        n->ctx->stream));
  GC_HIP(hipStreamSynchronize(n->ctx->stream));
  return GC_OK;
}

gc_status gc_cnn_forward(gc_cnn* n, gc_dtype dt, const void* c, int B, float* out_dev) {
  GC_TRY(check_cnn(n, B));
  GC_TRY(check_dtype(dt));
  if (!c) return set_error(GC_ERR_INVALID, "gridcalc: null input");
  if (dt == GC_F64) GC_TRY(forward_fused<double>(n, (const double*)c, B));
  else GC_TRY(forward_fused<float>(n, (const float*)c, B));
  if (out_dev)
    GC_HIP(hipMemcpyAsync(out_dev, n->out, (size_t)B * n->H * n->W * sizeof(float),
                          hipMemcpyDeviceToDevice, n->ctx->stream));
  return GC_OK;
}

gc_status gc_cnn_forward_layers(gc_cnn* n, gc_dtype dt, const void* c, int B, float* out_dev) {
  GC_TRY(check_cnn(n, B));
  GC_TRY(check_dtype(dt));
  if (!c) return set_error(GC_ERR_INVALID, "gridcalc: null input");
  if (dt == GC_F64) GC_TRY(forward<double>(n, (const double*)c, nullptr, B, false));
  else GC_TRY(forward<float>(n, (const float*)c, nullptr, B, false));
  if (out_dev)
    GC_HIP(hipMemcpyAsync(out_dev, n->out, (size_t)B * n->H * n->W * sizeof(float),
                          hipMemcpyDeviceToDevice, n->ctx->stream));
  return GC_OK;
}

gc_status gc_cnn_loss_grad(gc_cnn* n, gc_dtype dt, const void* c, const void* target,
                           const int* idx, int B, double* loss) {
  GC_TRY(check_cnn(n, B));
  GC_TRY(check_dtype(dt));
  if (!c || !target) return set_error(GC_ERR_INVALID, "gridcalc: null data");
  GC_TRY(upload_idx(n, idx, B));
  const int* d_idx = idx ? n->idx : nullptr;
  if (dt == GC_F64) {
    GC_TRY(forward<double>(n, (const double*)c, d_idx, B, true));
    GC_TRY(loss_and_dout<double>(n, (const double*)c, (const double*)target, d_idx, B, loss));
    return backward<double>(n, (const double*)c, d_idx, B);
  }
  GC_TRY(forward<float>(n, (const float*)c, d_idx, B, true));
  GC_TRY(loss_and_dout<float>(n, (const float*)c, (const float*)target, d_idx, B, loss));
  return backward<float>(n, (const float*)c, d_idx, B);
}

gc_status gc_cnn_loss_grad_chain(gc_cnn* n, gc_dtype dt, const void* data, size_t n_images,
                                 const int* idx, int B, int K, const double* weights,
                                 double* loss) {
  GC_TRY(check_cnn(n, B));
  GC_TRY(check_dtype(dt));
  if (!data || !idx) return set_error(GC_ERR_INVALID, "gridcalc: null data or chain indices");
  if (K < 1 || K > kChainMax)
    return set_error(GC_ERR_INVALID, "gridcalc: chain of %d applications (1 to %d)", K, kChainMax);
  for (int b = 0; b < B; ++b)
    if (idx[b] < 0 || (size_t)idx[b] + (size_t)K >= n_images)
      return set_error(GC_ERR_INVALID,
                       "gridcalc: chain %d starts at image %d; its %d targets pass image %zu", b,
                       idx[b], K, n_images);
  double w[kChainMax];
  for (int k = 0; k < K; ++k) {
    w[k] = weights ? weights[k] : 1.0 / K;
    if (!(w[k] >= 0.0) || !std::isfinite(w[k]))
      return set_error(GC_ERR_INVALID, "gridcalc: chain step weight %d is not finite and >= 0", k);
  }
  GC_TRY(ensure_chain(n, K));
  GC_TRY(upload_idx(n, idx, B));
  if (dt == GC_F64) return chain_loss_grad<double>(n, (const double*)data, n->idx, B, K, w, loss);
  return chain_loss_grad<float>(n, (const float*)data, n->idx, B, K, w, loss);
}

gc_status gc_cnn_adam(gc_cnn* n, double lr, double beta1, double beta2, double eps) {
  GC_TRY(check_cnn(n, 1));
  n->adam_t += 1;
  const float bc1 = (float)(1.0 - std::pow(beta1, n->adam_t));
  const float bc2 = (float)(1.0 - std::pow(beta2, n->adam_t));
  hipLaunchKernelGGL(k_adam, dim3((kNParams + 255) / 256), dim3(256), 0, n->ctx->stream, n->p,
                     n->g, n->m, n->v, kNParams, (float)lr, (float)beta1, (float)beta2, (float)eps,
                     bc1, bc2);
  GC_LAUNCHED();
  return pack(n);
}

gc_status gc_cnn_predict(gc_cnn* n, gc_dtype dt, const void* c, void* c_new, int B) {
  GC_TRY(check_cnn(n, B));
  GC_TRY(check_dtype(dt));
  if (!c || !c_new) return set_error(GC_ERR_INVALID, "gridcalc: null state");
  const long HW = (long)n->H * n->W;
  hipStream_t st = n->ctx->stream;
  if (dt == GC_F64) GC_TRY(forward_fused<double>(n, (const double*)c, B));
  else GC_TRY(forward_fused<float>(n, (const float*)c, B));
  const int T = (n->W / kTC) * (n->H / kTR);  // tiles (partial sums) per image
  for (int b = 0; b < B; ++b) {
    const unsigned g1 = (unsigned)((HW + 255) / 256);
    if (dt == GC_F64)
      hipLaunchKernelGGL((k_apply<double>), dim3(g1), dim3(256), 0, st, (const double*)c + b * HW,
                         (double*)c_new + b * HW, n->out + b * HW, n->dpart + (size_t)b * T, T,
                         n->scale, HW);
    else
      hipLaunchKernelGGL((k_apply<float>), dim3(g1), dim3(256), 0, st, (const float*)c + b * HW,
                         (float*)c_new + b * HW, n->out + b * HW, n->dpart + (size_t)b * T, T,
                         n->scale, HW);
    GC_LAUNCHED();
  }
  return GC_OK;
}

gc_status gc_cast(gc_context* ctx, gc_dtype from, gc_dtype to, size_t n, const void* src,
                  void* dst) {
  GC_TRY(check_ctx(ctx));
  GC_TRY(check_dtype(from));
  GC_TRY(check_dtype(to));
  if (n == 0) return GC_OK;
  if (!src || !dst) return set_error(GC_ERR_INVALID, "gridcalc: null pointer");
  const unsigned g1 = (unsigned)((n + 255) / 256);
  if (from == GC_F64 && to == GC_F32)
    hipLaunchKernelGGL((k_cast<double, float>), dim3(g1), dim3(256), 0, ctx->stream,
                       (const double*)src, (float*)dst, (long)n);
  else if (from == GC_F32 && to == GC_F64)
    hipLaunchKernelGGL((k_cast<float, double>), dim3(g1), dim3(256), 0, ctx->stream,
                       (const float*)src, (double*)dst, (long)n);
  else
    GC_HIP(hipMemcpyAsync(dst, src, n * dsize(from), hipMemcpyDeviceToDevice, ctx->stream));
  GC_LAUNCHED();
  return GC_OK;
}

}  // extern "C"

// ---- learned time stepping with the plain-RK4 fallback -------------------------------------
//
// One macro step advances a 2-D state by `substeps` RK4 steps of size `step`: the network
// proposes c_new = c + scale*(out - mean(out)); the proposal is accepted unless the mean of
// the state drifts by more than mass_tol from its initial value or the free energy
// increases (Cahn-Hilliard is a gradient flow: F never increases along exact solutions);
// otherwise the macro step is recomputed with the plain fused RK4 pipeline.  Mass and free
// energy come from one fused FP64 compensated reduction per macro step.

namespace {


// Sum of c and of the Cahn-Hilliard density (examples/common/cahn_hilliard.hpp:65-67) over a
// 2-D periodic grid, compensated, in one pass.
template <class T>
__global__ __launch_bounds__(256) void k_mass_fe(const T* __restrict__ c, int nx, int ny,
                                                 double ihx, double ihy, double kappa,
                                                 Comp2* __restrict__ part) {
  __shared__ Comp2 sh[4];
  const long n = (long)nx * ny;
  auto val = [&](size_t q) { return (double)c[q]; };
  double ms = 0.0, mc = 0.0, fs = 0.0, fc = 0.0;
  for (long s = (long)blockIdx.x * 256 + threadIdx.x; s < n; s += (long)gridDim.x * 256)
    site_mass_fe(val, (int)(s % nx), (int)(s / nx), nx, ny, ihx, ihy, kappa, ms, mc, fs, fc);
  const Comp2 v = wave_merge(Comp2{ms, mc, fs, fc});
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = waves_merge(sh, 4);
}


// The residual update of k_apply fused with k_mass_fe: each thread forms the new state at its
// site and at the four neighbours (the same expression as k_apply, so the same values),
// writes its own site and accumulates the mass and the density of the new state.
template <class T>
__global__ __launch_bounds__(256) void k_apply_mass_fe(const T* __restrict__ c, T* __restrict__ cn,
                                                       const float* __restrict__ out,
                                                       const double* __restrict__ osum,
                                                       int nparts, double scale, int nx, int ny,
                                                       double ihx, double ihy, double kappa,
                                                       Comp2* __restrict__ part,
                                                       unsigned* __restrict__ done,
                                                       double* __restrict__ result) {
  __shared__ Comp2 sh[4];
  __shared__ double sum;
  __shared__ bool last;
  block_sum_parts(osum, nparts, &sum);
  const long n = (long)nx * ny;
  const double mean = sum / (double)n;
  auto nv = [&](size_t q) { return (T)((double)c[q] + scale * ((double)out[q] - mean)); };
  auto val = [&](size_t q) { return (double)nv(q); };
  double ms = 0.0, mc = 0.0, fs = 0.0, fc = 0.0;
  for (long s = (long)blockIdx.x * 256 + threadIdx.x; s < n; s += (long)gridDim.x * 256) {
    cn[s] = nv((size_t)s);
    site_mass_fe(val, (int)(s % nx), (int)(s / nx), nx, ny, ihx, ihy, kappa, ms, mc, fs, fc);
  }
  const Comp2 v = wave_merge(Comp2{ms, mc, fs, fc});
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    part[blockIdx.x] = waves_merge(sh, 4);
    if (done) {
      __threadfence();
      last = atomicAdd(done, 1u) == gridDim.x - 1;
    }
  }
  if (!done) return;  // the partials are merged by a later kernel
  __syncthreads();
  if (last) {  // every partial is written: reduce them here instead of in another launch
    __threadfence();
    if (threadIdx.x < 64) mass_fe_reduce(part, (int)gridDim.x, result);
    if (threadIdx.x == 0) *done = 0u;  // ready for the next launch
  }
}

__global__ __launch_bounds__(64) void k_mass_fe_final(const Comp2* __restrict__ part, int nb,
                                                      double* __restrict__ out) {
  mass_fe_reduce(part, nb, out);
}

// Block k - 1 merges the partials of proposal k = 1..K (slot k of parts, stride partials per
// slot; nb partials, nb_last for proposal K) into out[2k..2k+1].
__global__ __launch_bounds__(64) void k_mass_fe_slots(const Comp2* __restrict__ parts,
                                                      int stride, int nb, int nb_last, int K,
                                                      double* __restrict__ out) {
  const int k = blockIdx.x + 1;
  mass_fe_reduce(parts + (size_t)k * stride, k == K ? nb_last : nb, out + 2 * k);
}

constexpr int kMFBlocks = 128;
constexpr int kChunk = 16;  // macro steps proposed between two host decisions

// Enqueues (sum of c, sum of the density) of one 2-D state into out_dev[0..1].
gc_status mass_fe_async(gc_cnn* n, const gc_grid* g, gc_dtype dt, const void* c, double kappa,
                        double* out_dev) {
  hipStream_t st = n->ctx->stream;
  Comp2* part = (Comp2*)n->dpart;
  if (dt == GC_F64)
    hipLaunchKernelGGL((k_mass_fe<double>), dim3(kMFBlocks), dim3(256), 0, st, (const double*)c,
                       g->nx, g->ny, 1.0 / g->hx, 1.0 / g->hy, kappa, part);
  else
    hipLaunchKernelGGL((k_mass_fe<float>), dim3(kMFBlocks), dim3(256), 0, st, (const float*)c,
                       g->nx, g->ny, 1.0 / g->hx, 1.0 / g->hy, kappa, part);
  GC_LAUNCHED();
  hipLaunchKernelGGL(k_mass_fe_final, dim3(1), dim3(64), 0, st, part, kMFBlocks, out_dev);
  GC_LAUNCHED();
  return GC_OK;
}

// Learned proposals of one chunk: slot k of hist receives proposal k (k = 1..K) of the chain
// starting at c, c_k = c_{k-1} + scale*(out(c_{k-1}) - mean), and diag[2k..2k+1] its (sum of
// c_k, sum of its density).  Step 1 evaluates the network on c; step k = 2..K forms c_{k-1}
// from the previous step's output, writes it, reduces its diagnostics and evaluates the
// network on it (the two network outputs and tile sums alternate); a last kernel forms c_K
// with its diagnostics, and one more merges the diagnostics of all K.  2K + 2 launches for K
// proposals.
template <class TS>
gc_status propose_chunk(gc_cnn* n, const gc_grid* g, const TS* c, TS* hist, int K,
                        double kappa, double* diag) {
  hipStream_t st = n->ctx->stream;
  const size_t ns = sites(g);
  const int T = (n->W / kTC) * (n->H / kTR);
  float* outs[2] = {n->out, n->out2};
  double* parts[2] = {n->dpart, n->cpart};
  // diagnostics partials: slot k (S workgroup partials) for proposal k
  const int S = std::max(T, kMFBlocks);
  Comp2* red = (Comp2*)(n->cpart + T);
  const double ihx = 1.0 / g->hx, ihy = 1.0 / g->hy;
  auto state = [&](int k) { return k == 0 ? c : hist + (size_t)k * ns; };
  GC_TRY(forward_fused<TS>(n, c, 1, outs[0], parts[0]));
  for (int k = 2; k <= K; ++k) {
    const int a = (k - 2) & 1, b = (k - 1) & 1;
    Chain<TS> ch{state(k - 2), outs[a], parts[a], T, n->scale, ihx, ihy, kappa,
                 hist + (size_t)(k - 1) * ns, red + (size_t)(k - 1) * S};
    GC_TRY(forward_chain<TS>(n, ch, outs[b], parts[b]));
  }
  const int a = (K - 1) & 1;
  hipLaunchKernelGGL((k_apply_mass_fe<TS>), dim3(kMFBlocks), dim3(256), 0, st, state(K - 1),
                     hist + (size_t)K * ns, outs[a], parts[a], T, n->scale, g->nx, g->ny, ihx,
                     ihy, kappa, red + (size_t)K * S, (unsigned*)nullptr, (double*)nullptr);
  GC_LAUNCHED();
  hipLaunchKernelGGL(k_mass_fe_slots, dim3((unsigned)K), dim3(64), 0, st, red, S, T, kMFBlocks,
                     K, diag);
  GC_LAUNCHED();
  return GC_OK;
}

gc_status fetch(gc_cnn* n, double* host, const double* dev, int count) {
  GC_HIP(hipMemcpyAsync(host, dev, count * sizeof(double), hipMemcpyDeviceToHost, n->ctx->stream));
  GC_HIP(hipStreamSynchronize(n->ctx->stream));
  return GC_OK;
}

// Macro steps are proposed in chunks of up to kChunk without waiting for the host: proposal k
// of a chunk is written to history slot k and its mass and free energy to diag[k].  One copy
// of diag per chunk then lets the host apply the acceptance rule in order; the state after the
// last accepted proposal becomes psi, and the first rejected macro step (if any) is recomputed
// by plain RK4 before the next chunk starts.  While the host decides on one chunk, the GPU
// already proposes the next one from the chunk's last proposal (two proposal buffers); it is
// discarded if the chunk has a rejection.  The decisions, and so the result, are those of
// checking every proposal as it is made.
//
// Macro step m is proposed by the first network i with first[i] <= m < end[i]; steps no
// network covers run as plain RK4.  A chunk never spans two segments; the chunk after the
// last one of a segment is already proposed by the next segment's network.  The proposal
// buffers, diagnostics and events belong to nets[0]; each network uses its own scratch.
gc_status learned_run(gc_cnn* const* nets, const int* first, const int* end, int n_nets,
                      const gc_grid* g, gc_dtype dt, const gc_ch_params* p, void* psi,
                      double step, int substeps, int n_macro, double mass_tol, int* n_fallback) {
  if (!nets || !first || !end || n_nets < 1)
    return set_error(GC_ERR_INVALID, "gridcalc: no networks");
  for (int i = 0; i < n_nets; ++i) GC_TRY(check_cnn(nets[i], 1));
  GC_TRY(check_grid(g));
  GC_TRY(check_dtype(dt));
  if (!p || !psi) return set_error(GC_ERR_INVALID, "gridcalc: null argument");
  gc_cnn* n = nets[0];
  for (int i = 0; i < n_nets; ++i) {
    if (g->nz != 1 || g->nx != nets[i]->W || g->ny != nets[i]->H)
      return set_error(GC_ERR_INVALID, "gridcalc: the corrector was built for %d x %d x 1 grids",
                       nets[i]->W, nets[i]->H);
    if (nets[i]->ctx != n->ctx)
      return set_error(GC_ERR_INVALID, "gridcalc: the networks must share one context");
  }
  if (n_macro < 0 || substeps < 1 || !(step >= 0.0))
    return set_error(GC_ERR_INVALID, "gridcalc: bad learned-integration arguments");
  const size_t bytes = sites(g) * dsize(dt);
  constexpr int kSlots = kChunk + 1;  // per proposal buffer: slot 0 unused, proposals 1..K
  if (!n->hist) GC_TRY(cnn_alloc(&n->hist, 2 * (size_t)kSlots * sites(g) * sizeof(double)));
  if (!n->diag) GC_TRY(cnn_alloc((void**)&n->diag, 2 * 2 * kSlots * sizeof(double)));
  if (!n->hdiag &&
      hipHostMalloc((void**)&n->hdiag, 2 * 2 * kSlots * sizeof(double), hipHostMallocDefault) !=
          hipSuccess) {
    n->hdiag = nullptr;
    return set_error(GC_ERR_NOMEM, "gridcalc: pinned host buffer allocation failed");
  }
  for (hipEvent_t& e : n->ev)
    if (!e) GC_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  const int T = (n->W / kTC) * (n->H / kTR);
  for (int i = 0; i < n_nets; ++i) {
    gc_cnn* q = nets[i];
    if (!q->out2) GC_TRY(cnn_alloc((void**)&q->out2, sites(g) * sizeof(float)));
    if (!q->cpart)  // second tile sums, then per proposal one Comp2 (4 doubles) per workgroup
      GC_TRY(cnn_alloc((void**)&q->cpart,
                       (T + 4 * (size_t)kSlots * std::max(T, kMFBlocks)) * sizeof(double)));
  }
  hipStream_t st = n->ctx->stream;
  auto slot = [&](int buf, int k) { return (char*)n->hist + ((size_t)buf * kSlots + k) * bytes; };
  double* hd[2] = {n->hdiag, n->hdiag + 2 * kSlots};
  // The network proposing macro step m (-1: plain RK4) and the end of its segment.
  auto segment = [&](int m, int& net, int& stop) {
    for (int i = 0; i < n_nets; ++i)
      if (first[i] <= m && m < end[i]) {
        net = i;
        stop = std::min(end[i], n_macro);
        for (int j = 0; j < i; ++j)  // an earlier network takes over where its range starts
          if (first[j] > m && first[j] < end[j]) stop = std::min(stop, first[j]);
        return;
      }
    net = -1;
    stop = n_macro;
    for (int i = 0; i < n_nets; ++i)
      if (first[i] > m) stop = std::min(stop, first[i]);
  };
  // Proposals 1..K of network q from src into buffer buf; their diagnostics reach hd[buf]
  // when ev[buf] fires.
  auto launch = [&](gc_cnn* q, const void* src, int buf, int K) -> gc_status {
    double* dd = n->diag + (size_t)buf * 2 * kSlots;
    if (dt == GC_F64)
      GC_TRY(propose_chunk<double>(q, g, (const double*)src, (double*)slot(buf, 0), K, p->kappa,
                                   dd));
    else
      GC_TRY(propose_chunk<float>(q, g, (const float*)src, (float*)slot(buf, 0), K, p->kappa,
                                  dd));
    GC_HIP(hipMemcpyAsync(hd[buf] + 2, dd + 2, 2 * K * sizeof(double), hipMemcpyDeviceToHost,
                          st));
    GC_HIP(hipEventRecord(n->ev[buf], st));
    return GC_OK;
  };
  const double vol = (g->hx * g->hy) * g->hz;
  const double vol_total = (double)sites(g) * vol;
  double h[2];
  GC_TRY(mass_fe_async(n, g, dt, psi, p->kappa, n->diag));
  GC_TRY(fetch(n, h, n->diag, 2));
  const double Minit = vol * h[0];
  double Fprev = vol * h[1];
  int fallbacks = 0;
  int m = 0;       // macro steps decided; psi holds the state after them
  int cur = 0;     // buffer of the chunk being decided
  int K = 0;       // its proposals (0: none in flight)
  int ahead = 0;   // macro steps covered by the chunks in flight
  // Starts a chunk at macro step m from psi if a network covers m; otherwise runs the plain
  // segment at m and refreshes Fprev.
  auto restart = [&]() -> gc_status {
    while (m < n_macro) {
      int net, stop;
      segment(m, net, stop);
      if (net >= 0) {
        K = std::min(kChunk, stop - m);
        GC_TRY(launch(nets[net], psi, cur, K));
        ahead = m + K;
        return GC_OK;
      }
      GC_TRY(gc_ch_integrate(n->ctx, g, dt, p, psi, step, (stop - m) * substeps, GC_RK4));
      GC_TRY(mass_fe_async(n, g, dt, psi, p->kappa, n->diag));
      GC_TRY(fetch(n, h, n->diag, 2));
      Fprev = vol * h[1];
      m = stop;
    }
    K = 0;
    return GC_OK;
  };
  GC_TRY(restart());
  while (m < n_macro && K > 0) {
    // While the host waits for chunk `cur`, the GPU proposes the next chunk from its last
    // proposal; that chunk stands only if every proposal of `cur` is accepted.
    int Kn = 0;
    if (ahead < n_macro) {
      int net, stop;
      segment(ahead, net, stop);
      if (net >= 0) {
        Kn = std::min(kChunk, stop - ahead);
        GC_TRY(launch(nets[net], slot(cur, K), 1 - cur, Kn));
      }
    }
    GC_HIP(hipEventSynchronize(n->ev[cur]));
    int good = 0;
    for (int k = 1; k <= K; ++k) {
      const double Mk = vol * hd[cur][2 * k], Fk = vol * hd[cur][2 * k + 1];
      if (!(std::isfinite(Fk) && Fk <= Fprev && std::fabs(Mk - Minit) / vol_total <= mass_tol))
        break;
      Fprev = Fk;
      good = k;
    }
    if (good > 0)  // ordered after the next chunk's reads of slot(cur, K), before its reuse
      GC_HIP(hipMemcpyAsync(psi, slot(cur, good), bytes, hipMemcpyDeviceToDevice, st));
    m += good;
    if (good == K) {
      if (Kn > 0) {
        cur = 1 - cur;
        K = Kn;
        ahead += Kn;
      } else {
        GC_TRY(restart());  // a plain segment (or the end) follows
      }
      continue;
    }
    // a rejected proposal: the chunk in flight (if any) started from a discarded state
    ++fallbacks;
    GC_TRY(gc_ch_integrate(n->ctx, g, dt, p, psi, step, substeps, GC_RK4));
    GC_TRY(mass_fe_async(n, g, dt, psi, p->kappa, n->diag));
    GC_TRY(fetch(n, h, n->diag, 2));
    Fprev = vol * h[1];
    m += 1;
    GC_TRY(restart());
  }
  GC_HIP(hipStreamSynchronize(st));
  if (n_fallback) *n_fallback = fallbacks;
  return GC_OK;
}

}  // namespace

extern "C" {

gc_status gc_learned_integrate(gc_cnn* n, const gc_grid* g, gc_dtype dt, const gc_ch_params* p,
                               void* psi, double step, int substeps, int n_macro,
                               double mass_tol, int* n_fallback) {
  const int first = 0, end = n_macro;
  return learned_run(&n, &first, &end, 1, g, dt, p, psi, step, substeps, n_macro, mass_tol,
                     n_fallback);
}

gc_status gc_learned_integrate_phases(gc_cnn* const* nets, const int* first, const int* end,
                                      int n_nets, const gc_grid* g, gc_dtype dt,
                                      const gc_ch_params* p, void* psi, double step,
                                      int substeps, int n_macro, double mass_tol,
                                      int* n_fallback) {
  return learned_run(nets, first, end, n_nets, g, dt, p, psi, step, substeps, n_macro, mass_tol,
                     n_fallback);
}

}  // extern "C"
