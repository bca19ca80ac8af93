// Narrow-band network kernels for gfx950 (MI355X): MFMA forward, backward
// and Adam for the Fourier-feature MLP of mlp_kernels.h.
//
// MFMA tiling (v_mfma_f32_16x16x16_bf16, D = A B + C, 16x16 blocks):
//   A operand: lane l holds A[l % 16][4 (l / 16) + j], j = 0..3
//   B operand: lane l holds B[4 (l / 16) + j][l % 16]
//   D result:  lane l holds D[4 (l / 16) + j][l % 16]
// The inference forward uses gfx950's double-K form
// (v_mfma_f32_16x16x32_bf16: lane l holds K values 8 (l / 16) + j,
// j = 0..7) on two adjacent 16-wide K blocks at once: lane l passes its
// four values of block 2i then its four of block 2i+1, for A and B alike,
// which the instruction reads as one 32-wide block with its K index
// permuted the same way in both operands — the same sum of products.
// Its features come in another row order than training's (F1I, see
// mlp_kernels.h): in each K block every lane evaluates the same function
// (sin or cos of one coordinate) at its own frequencies, c_k = 2^(k / 2) / 2
// for k = 4 j + g, g = l / 16, i.e. c_g 4^j with c_g a power-of-two multiple
// of 1/2 or of sqrt(2)/2 — the same f32 multipliers, so the same features.
// A wavefront pushes 16 samples (the N dimension) through the network:
// layer outputs land in exactly the layout the next layer's B operand
// needs, so activations stay in registers from the features to the output.
// The backward-data pass uses the same layout with W^T fragments; weight
// gradients are products over the batch of hidden-major bf16 activations
// (split over the batch, accumulated with f32 atomics).
#include <hip/hip_runtime.h>

#include "mlp_kernels.h"

namespace iso {
namespace mlp {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// c_k = 2^(k / 2) / 2 revolutions per unit (2 pi c_k = pi 2^(k / 2)),
// rounded to f32: powers of two for even k, sqrt(2) scaled for odd k.
__constant__ float c_rev[kFreq] = {
    0x1p-1f, 0x1.6a09e6p-1f, 0x1p+0f, 0x1.6a09e6p+0f,
    0x1p+1f, 0x1.6a09e6p+1f, 0x1p+2f, 0x1.6a09e6p+2f,
    0x1p+3f, 0x1.6a09e6p+3f, 0x1p+4f, 0x1.6a09e6p+4f,
    0x1p+5f, 0x1.6a09e6p+5f, 0x1p+6f, 0x1.6a09e6p+6f};

typedef float f32x4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4v __attribute__((ext_vector_type(4)));

// f32 -> bf16, nearest with ties to even (gfx950's v_cvt_pk_bf16_f32; a NaN
// stays a quiet NaN).
__device__ __forceinline__ uint16_t to_bf16(float f) {
  return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f));
}

__device__ __forceinline__ float from_bf16(uint16_t b) {
  return __uint_as_float(static_cast<uint32_t>(b) << 16);
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));

// Two adjacent K blocks in one double-K MFMA (see the tiling note above).
__device__ __forceinline__ f32x4 mfma2(bf16x4 a0, bf16x4 a1, bf16x4 b0, bf16x4 b1, f32x4 c) {
  const s16x8 a = __builtin_shufflevector(a0, a1, 0, 1, 2, 3, 4, 5, 6, 7);
  const s16x8 b = __builtin_shufflevector(b0, b1, 0, 1, 2, 3, 4, 5, 6, 7);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                 __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// Four f32 to bf16 (two v_cvt_pk_bf16_f32).
__device__ __forceinline__ bf16x4 pack4f(float a, float b, float c, float d) {
  return __builtin_bit_cast(bf16x4, __builtin_convertvector(f32x4v{a, b, c, d}, bf16x4v));
}

__device__ __forceinline__ bf16x4 pack4(uint16_t a, uint16_t b, uint16_t c, uint16_t d) {
  bf16x4 r;
  r[0] = static_cast<short>(a);
  r[1] = static_cast<short>(b);
  r[2] = static_cast<short>(c);
  r[3] = static_cast<short>(d);
  return r;
}

// Feature fi (0..111) of point p: p_d, sin(2 pi r), cos(2 pi r) with
// r = c_k p_d in revolutions, zero padding.  v_sin_f32 / v_cos_f32 take
// revolutions; the fractional part of an f32 r is exact, so reducing r to
// it first changes nothing and keeps every argument in their input range.
__device__ __forceinline__ float feature(int fi, float px, float py, float pz) {
  if (fi < 3) return fi == 0 ? px : (fi == 1 ? py : pz);
  if (fi >= kFeat) return 0.0f;
  const int q = fi - 3;
  const int k = q / 6, r = q - 6 * k;
  const int d = r < 3 ? r : r - 3;
  // the fractional part (exact) keeps the argument in the instructions' range
  const float rev = __builtin_amdgcn_fractf(c_rev[k] * (d == 0 ? px : (d == 1 ? py : pz)));
  return r < 3 ? __builtin_amdgcn_sinf(rev) : __builtin_amdgcn_cosf(rev);
}

// Fragments of one 16-row block set, read from LDS: lane's 4 bf16 of block (mb, kb).
__device__ __forceinline__ bf16x4 frag(const uint16_t* F, int n_kb, int mb, int kb, int lane) {
  const uint16_t* p = F + ((mb * n_kb + kb) * 64 + lane) * 4;
  return *reinterpret_cast<const bf16x4*>(p);
}

struct Tile {
  bf16x4 h1[4], h2[4], h3[4];  // activations (B layout of the next layer), bf16
  float f;                     // output of this lane's sample
};

// Training forward pass of 16 samples; lane l's sample is l % 16 at
// (px, py, pz).  sF: fragments F1 F2 F3 in LDS; sB: b1 b2 b3 w4 b4 in LDS.
// Layer 1 consumes the features one K-block at a time (computed just in
// time, and stored hidden-major to xdst[row * n + s] when xdst is given).
__device__ __forceinline__ void forward_tile(const uint16_t* sF, const float* sB, float px,
                                             float py, float pz, int lane, Tile& T,
                                             uint16_t* xdst, int64_t n, int64_t s) {
  const int r0 = 4 * (lane >> 4);
  const uint16_t* F1 = sF + kFragF1;
  const uint16_t* F2 = sF + kFragF2;
  const uint16_t* F3 = sF + kFragF3;
  const float* b1 = sB;
  const float* b2 = sB + kHidden;
  const float* b3 = sB + 2 * kHidden;
  const float* w4 = sB + 3 * kHidden;
  const float b4 = sB[4 * kHidden];
  f32x4 acc1[4];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) acc1[mb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  auto features = [&](int kb) {
    const int f0 = 16 * kb + r0;
    uint16_t xb[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) xb[j] = to_bf16(feature(f0 + j, px, py, pz));
    if (xdst) {
#pragma unroll
      for (int j = 0; j < 4; ++j) xdst[(f0 + j) * n + s] = xb[j];
    }
    return pack4(xb[0], xb[1], xb[2], xb[3]);
  };
#pragma unroll 1
  for (int kb = 0; kb < kKB1; ++kb) {
    const bf16x4 x = features(kb);
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) acc1[mb] = mfma(frag(F1, kKB1, mb, kb, lane), x, acc1[mb]);
  }
  // A 64-wide K product: four MFMAs.
  auto product64 = [&](const uint16_t* F, int mb, const bf16x4* h) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) acc = mfma(frag(F, 4, mb, kb, lane), h[kb], acc);
    return acc;
  };
  // Bias, ReLU and the bf16 rounding of a layer's 16 x 16 output block.
  auto activate = [&](const f32x4& acc, const float* b) {
    return pack4f(fmaxf(acc[0] + b[r0 + 0], 0.0f), fmaxf(acc[1] + b[r0 + 1], 0.0f),
                  fmaxf(acc[2] + b[r0 + 2], 0.0f), fmaxf(acc[3] + b[r0 + 3], 0.0f));
  };
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) T.h1[mb] = activate(acc1[mb], b1 + 16 * mb);
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) T.h2[mb] = activate(product64(F2, mb, T.h1), b2 + 16 * mb);
  float part = 0.0f;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    T.h3[mb] = activate(product64(F3, mb, T.h2), b3 + 16 * mb);
#pragma unroll
    for (int j = 0; j < 4; ++j)
      part += w4[16 * mb + r0 + j] * from_bf16(static_cast<uint16_t>(T.h3[mb][j]));
  }
  // The four lanes of a sample (l, l^16, l^32, l^48) hold its 64 terms.
  part += __shfl_xor(part, 16, 64);
  part += __shfl_xor(part, 32, 64);
  T.f = part + b4;
}

constexpr int kThreads = 256;
constexpr int kLdsFrags = kFragT3;  // forward fragments only (F1I F2 F3)

// Loads fragments [0, n_frag) and the bias block into LDS.
__device__ __forceinline__ void load_model(const Model& M, uint16_t* sF, int n_frag, float* sB) {
  const uint32_t* src = reinterpret_cast<const uint32_t*>(M.frags);
  uint32_t* dst = reinterpret_cast<uint32_t*>(sF);
  for (int i = threadIdx.x; i < n_frag / 2; i += blockDim.x) dst[i] = src[i];
  for (int i = threadIdx.x; i < 4 * kHidden + 1; i += blockDim.x) {
    // b1 b2 b3 w4 b4
    const int which = i / kHidden, k = i - which * kHidden;
    const int off = which == 0 ? kOffB1 : which == 1 ? kOffB2 : which == 2 ? kOffB3 : kOffW4;
    sB[i] = i == 4 * kHidden ? M.theta[kOffB4] : M.theta[off + k];
  }
}

// Inference forward of kT 16-sample tiles at once: every weight fragment
// read from LDS feeds all kT tiles' MFMAs, 1/kT of the LDS traffic per
// sample of one tile at a time.  Pairs of K blocks go through double-K
// MFMAs.  The features come in the inference order of F1I: block kb < 6
// holds function kb (sin of x, y, z, then cos) at the lane's four
// frequencies, block 6 the coordinates (rows 0-2 of lanes 0-15) and
// padding.  p[t]: this lane's sample of tile t (lane l's sample is l % 16);
// f[t]: its output.
template <int kT>
__device__ __forceinline__ void forward_tiles(const uint16_t* sF, const float* sB,
                                              const float (*p)[3], int lane, float* f) {
  const int r0 = 4 * (lane >> 4);
  const int g = lane >> 4;
  const float cg = __builtin_amdgcn_ldexpf((g & 1) ? 0x1.6a09e6p-1f : 0x1p-1f, g >> 1);
  const uint16_t* F1 = sF + kFragF1;
  const uint16_t* F2 = sF + kFragF2;
  const uint16_t* F3 = sF + kFragF3;
  const float* b1 = sB;
  const float* b2 = sB + kHidden;
  const float* b3 = sB + 2 * kHidden;
  const float* w4 = sB + 3 * kHidden;
  const float b4 = sB[4 * kHidden];
  auto feats = [&](const float* q, int kb) {
    float xf[4];
    if (kb < 6) {
      const float m = cg * q[kb % 3];  // c_k q for j = 0
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        // c_g 4^j q: a power-of-two scaling of m, exact
        const float rev = __builtin_amdgcn_fractf(m * static_cast<float>(1 << (2 * j)));
        xf[j] = kb < 3 ? __builtin_amdgcn_sinf(rev) : __builtin_amdgcn_cosf(rev);
      }
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) xf[j] = g == 0 && j < 3 ? q[j] : 0.0f;
    }
    return pack4f(xf[0], xf[1], xf[2], xf[3]);
  };
  // Bias, ReLU and the bf16 rounding of a layer's 16 x 16 output block.
  auto activate = [&](const f32x4& acc, const float* b) {
    return pack4f(fmaxf(acc[0] + b[r0 + 0], 0.0f), fmaxf(acc[1] + b[r0 + 1], 0.0f),
                  fmaxf(acc[2] + b[r0 + 2], 0.0f), fmaxf(acc[3] + b[r0 + 3], 0.0f));
  };
  f32x4 z[kT][4];
#pragma unroll
  for (int t = 0; t < kT; ++t)
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) z[t][mb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 1
  for (int kb = 0; kb + 1 < kKB1; kb += 2) {
    bf16x4 x0[kT], x1[kT];
#pragma unroll
    for (int t = 0; t < kT; ++t) {
      x0[t] = feats(p[t], kb);
      x1[t] = feats(p[t], kb + 1);
    }
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      const bf16x4 w0 = frag(F1, kKB1, mb, kb, lane), w1 = frag(F1, kKB1, mb, kb + 1, lane);
#pragma unroll
      for (int t = 0; t < kT; ++t) z[t][mb] = mfma2(w0, w1, x0[t], x1[t], z[t][mb]);
    }
  }
  if (kKB1 & 1) {
    bf16x4 x[kT];
#pragma unroll
    for (int t = 0; t < kT; ++t) x[t] = feats(p[t], kKB1 - 1);
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      const bf16x4 w = frag(F1, kKB1, mb, kKB1 - 1, lane);
#pragma unroll
      for (int t = 0; t < kT; ++t) z[t][mb] = mfma(w, x[t], z[t][mb]);
    }
  }
  bf16x4 h[kT][4];
#pragma unroll
  for (int t = 0; t < kT; ++t)
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) h[t][mb] = activate(z[t][mb], b1 + 16 * mb);
  // Hidden layers 2 and 3 (W2, W3): 64-wide K, two double-K MFMAs a block.
  auto layer = [&](const uint16_t* F, const float* b) {
    bf16x4 y[kT][4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      f32x4 a[kT];
#pragma unroll
      for (int t = 0; t < kT; ++t) a[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kp = 0; kp < 4; kp += 2) {
        const bf16x4 w0 = frag(F, 4, mb, kp, lane), w1 = frag(F, 4, mb, kp + 1, lane);
#pragma unroll
        for (int t = 0; t < kT; ++t) a[t] = mfma2(w0, w1, h[t][kp], h[t][kp + 1], a[t]);
      }
#pragma unroll
      for (int t = 0; t < kT; ++t) y[t][mb] = activate(a[t], b + 16 * mb);
    }
#pragma unroll
    for (int t = 0; t < kT; ++t)
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) h[t][mb] = y[t][mb];
  };
  layer(F2, b2);
  layer(F3, b3);
#pragma unroll
  for (int t = 0; t < kT; ++t) {
    float q = 0.0f;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        q += w4[16 * mb + r0 + j] * from_bf16(static_cast<uint16_t>(h[t][mb][j]));
    // The four lanes of a sample (l, l^16, l^32, l^48) hold its 64 terms.
    q += __shfl_xor(q, 16, 64);
    q += __shfl_xor(q, 32, 64);
    f[t] = q + b4;
  }
}

// Tiles of 16 samples per fragment read in the inference forward.
constexpr int kFwdTiles = 4;
// 512 threads: the fragments' 40 KB of LDS then allow as many waves per SIMD
// as the registers do.
constexpr int kFwdThreads = 512;

__global__ __launch_bounds__(kFwdThreads) void k_forward(Model M, const float* __restrict__ points,
                                                         int64_t n, float* __restrict__ out) {
  __shared__ __align__(16) uint16_t sF[kLdsFrags];
  __shared__ float sB[4 * kHidden + 1];
  // F1I in F1's place, then F2 F3, and the bias block
  {
    const uint32_t* f1 = reinterpret_cast<const uint32_t*>(M.frags + kFragF1I);
    const uint32_t* f23 = reinterpret_cast<const uint32_t*>(M.frags + kFragF2);
    uint32_t* dst = reinterpret_cast<uint32_t*>(sF);
    constexpr int n1 = (kFragF2 - kFragF1) / 2, n = kLdsFrags / 2;
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = i < n1 ? f1[i] : f23[i - n1];
  }
  load_model(M, sF, 0, sB);
  __syncthreads();
  const int lane = __lane_id();
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * (kFwdThreads / 64) + (threadIdx.x >> 6);
  const int64_t n_waves = static_cast<int64_t>(gridDim.x) * (kFwdThreads / 64);
  constexpr int64_t kStep = 16 * kFwdTiles;
  for (int64_t t0 = wave * kStep; t0 < n; t0 += n_waves * kStep) {
    // The fragments are re-read from LDS for every kFwdTiles tiles rather
    // than held in registers across them: fewer registers, more waves per
    // SIMD.
    asm volatile("" ::: "memory");
    float p[kFwdTiles][3], f[kFwdTiles];
#pragma unroll
    for (int t = 0; t < kFwdTiles; ++t) {
      const int64_t si = t0 + 16 * t + (lane & 15);
#pragma unroll
      for (int a = 0; a < 3; ++a) p[t][a] = si < n ? points[3 * si + a] : 0.0f;
    }
    forward_tiles<kFwdTiles>(sF, sB, p, lane, f);
    if (lane < 16) {
#pragma unroll
      for (int t = 0; t < kFwdTiles; ++t) {
        const int64_t si = t0 + 16 * t + lane;
        if (si < n) out[si] = f[t];
      }
    }
  }
}

__device__ __forceinline__ uint32_t hash32(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t h = a * 0x9E3779B1u ^ (b + 0x85EBCA77u) * 0xC2B2AE3Du ^ (c + 0x27D4EB2Fu) * 0x165667B1u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  return h;
}

// Forward + backward-data of a batch; writes hidden-major activations and
// per-sample dL/df, accumulates the loss.
__global__ __launch_bounds__(kThreads) void k_train_fb(Model M, Batch B, Samples S,
                                                       uint32_t step) {
  __shared__ __align__(16) uint16_t sF[kTrainFrags];
  __shared__ float sB[4 * kHidden + 1];
  load_model(M, sF, kTrainFrags, sB);
  __syncthreads();
  const int lane = __lane_id();
  const int r0 = 4 * (lane >> 4);
  const int col = lane & 15;
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * (kThreads / 64) + (threadIdx.x >> 6);
  const int64_t n_waves = static_cast<int64_t>(gridDim.x) * (kThreads / 64);
  const float* w4 = sB + 3 * kHidden;
  const float inv_n = 1.0f / static_cast<float>(B.n);
  float loss = 0.0f;
  for (int64_t t0 = wave * 16; t0 < B.n; t0 += n_waves * 16) {
    const int64_t s = t0 + col;  // batch sample (B.n is a multiple of 16)
    const uint64_t src =
        S.gather ? hash32(S.seed, step, static_cast<uint32_t>(s)) % S.count
                 : static_cast<uint64_t>(s);
    const float px = S.points[3 * src], py = S.points[3 * src + 1], pz = S.points[3 * src + 2];
    const float t = fminf(fmaxf(S.sdf[src], -S.clamp), S.clamp);
    const float wgt = S.weight ? S.weight[src] : 1.0f;
    Tile T;
    forward_tile(sF, sB, px, py, pz, lane, T, B.x, B.n, s);
    // Least squares plus (overshoot - 1) e^2, e the part of the error that
    // lies outside the interval between 0 and the target.
    const float r = T.f - t;
    const float e = T.f - fminf(fmaxf(T.f, fminf(t, 0.0f)), fmaxf(t, 0.0f));
    const float ow = S.overshoot - 1.0f;
    const float g = wgt * (r + ow * e) * inv_n;
    if (lane < 16) {
      loss += wgt * (r * r + ow * e * e);
      B.g[s] = g;
    }
    // dz3 = bf16(g w4 [h3 > 0]), dz2 = bf16((W3^T dz3) [h2 > 0]), dz1 likewise.
    bf16x4 d3[4], d2[4], d1[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      uint16_t d[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float h = from_bf16(static_cast<uint16_t>(T.h3[mb][j]));
        d[j] = to_bf16(h > 0.0f ? g * w4[16 * mb + r0 + j] : 0.0f);
      }
      d3[mb] = pack4(d[0], d[1], d[2], d[3]);
    }
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) acc = mfma(frag(sF + kFragT3, 4, mb, kb, lane), d3[kb], acc);
      uint16_t d[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float h = from_bf16(static_cast<uint16_t>(T.h2[mb][j]));
        d[j] = to_bf16(h > 0.0f ? acc[j] : 0.0f);
      }
      d2[mb] = pack4(d[0], d[1], d[2], d[3]);
    }
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) acc = mfma(frag(sF + kFragT2, 4, mb, kb, lane), d2[kb], acc);
      uint16_t d[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float h = from_bf16(static_cast<uint16_t>(T.h1[mb][j]));
        d[j] = to_bf16(h > 0.0f ? acc[j] : 0.0f);
      }
      d1[mb] = pack4(d[0], d[1], d[2], d[3]);
    }
    // Hidden-major stores: row 16 mb + r0 + j, column s.
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t o = (16 * mb + r0 + j) * B.n + s;
        B.h1[o] = static_cast<uint16_t>(T.h1[mb][j]);
        B.h2[o] = static_cast<uint16_t>(T.h2[mb][j]);
        B.h3[o] = static_cast<uint16_t>(T.h3[mb][j]);
        B.dz1[o] = static_cast<uint16_t>(d1[mb][j]);
        B.dz2[o] = static_cast<uint16_t>(d2[mb][j]);
        B.dz3[o] = static_cast<uint16_t>(d3[mb][j]);
      }
  }
  // wave sum of the loss terms (lanes 0..15 hold them)
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) loss += __shfl_xor(loss, d, 64);
  if (lane == 0 && loss != 0.0f) atomicAdd(M.loss, loss);
}

// Weight gradients: block (layer, mb, nb) of dW_l = DZ_l Hprev^T over one
// chunk of the batch, accumulated into M.grad.  One wavefront per (block,
// chunk).  Output blocks: layer 1 4x7, layer 2 4x4, layer 3 4x4.
constexpr int kWgradBlocks = 4 * kKB1 + 16 + 16;
constexpr int kWgradChunk = 1024;

__global__ __launch_bounds__(kThreads) void k_wgrad(Model M, Batch B) {
  const int lane = __lane_id();
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * (kThreads / 64) + (threadIdx.x >> 6);
  const int64_t n_chunks = (B.n + kWgradChunk - 1) / kWgradChunk;
  if (wave >= kWgradBlocks * n_chunks) return;
  const int blk = static_cast<int>(wave % kWgradBlocks);
  const int64_t chunk = wave / kWgradBlocks;
  int layer, mb, nb;
  if (blk < 4 * kKB1) {
    layer = 1;
    mb = blk / kKB1;
    nb = blk % kKB1;
  } else if (blk < 4 * kKB1 + 16) {
    layer = 2;
    mb = (blk - 4 * kKB1) / 4;
    nb = (blk - 4 * kKB1) % 4;
  } else {
    layer = 3;
    mb = (blk - 4 * kKB1 - 16) / 4;
    nb = (blk - 4 * kKB1 - 16) % 4;
  }
  const uint16_t* dz = layer == 1 ? B.dz1 : layer == 2 ? B.dz2 : B.dz3;
  const uint16_t* hp = layer == 1 ? B.x : layer == 2 ? B.h1 : B.h2;
  const int cols = layer == 1 ? kFeatPad : kHidden;
  const int off = layer == 1 ? kOffW1 : layer == 2 ? kOffW2 : kOffW3;
  const int r0 = 4 * (lane >> 4), col = lane & 15;
  const int64_t s_end = (chunk + 1) * kWgradChunk < B.n ? (chunk + 1) * kWgradChunk : B.n;
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int64_t s0 = chunk * kWgradChunk; s0 < s_end; s0 += 16) {
    const bf16x4 a = *reinterpret_cast<const bf16x4*>(dz + (16 * mb + col) * B.n + s0 + r0);
    const bf16x4 b = *reinterpret_cast<const bf16x4*>(hp + (16 * nb + col) * B.n + s0 + r0);
    acc = mfma(a, b, acc);
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) atomicAdd(&M.grad[off + (16 * mb + r0 + j) * cols + 16 * nb + col], acc[j]);
}

// Bias gradients and the output layer: db_l[i] = sum_s dz_l[i][s],
// dw4[i] = sum_s g[s] h3[i][s], db4 = sum_s g[s].  One wavefront per
// (row, chunk of samples): lanes stride over the chunk (coalesced reads of
// the hidden-major rows), then a wave reduction and one atomic per wave.
// Rows 0..191 biases, 192..255 w4, 256 b4.
__global__ __launch_bounds__(kThreads) void k_bgrad(Model M, Batch B) {
  const int64_t n_chunks = (B.n + kWgradChunk - 1) / kWgradChunk;
  const int64_t t = (static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x) >> 6;
  if (t >= 257 * n_chunks) return;  // uniform per wave
  const int lane = __lane_id();
  const int row = static_cast<int>(t % 257);
  const int64_t chunk = t / 257;
  const int64_t s0 = chunk * kWgradChunk;
  const int64_t s1 = s0 + kWgradChunk < B.n ? s0 + kWgradChunk : B.n;
  float acc = 0.0f;
  int dst;
  if (row < 192) {
    const int layer = row / 64, i = row % 64;
    const uint16_t* dz = (layer == 0 ? B.dz1 : layer == 1 ? B.dz2 : B.dz3) + i * B.n;
    for (int64_t s = s0 + lane; s < s1; s += 64) acc += from_bf16(dz[s]);
    dst = (layer == 0 ? kOffB1 : layer == 1 ? kOffB2 : kOffB3) + i;
  } else if (row < 256) {
    const uint16_t* h3 = B.h3 + (row - 192) * B.n;
    for (int64_t s = s0 + lane; s < s1; s += 64) acc += B.g[s] * from_bf16(h3[s]);
    dst = kOffW4 + row - 192;
  } else {
    for (int64_t s = s0 + lane; s < s1; s += 64) acc += B.g[s];
    dst = kOffB4;
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, 64);
  if (lane == 0) atomicAdd(&M.grad[dst], acc);
}

// Inference feature row of training feature column c (see kFragF1I).
__device__ __forceinline__ int inference_row(int c) {
  if (c < 3) return 96 + c;       // coordinates
  if (c >= kFeat) return c;       // padding
  const int q = c - 3, k = q / 6, t = q - 6 * k;  // frequency k, function t
  return 16 * t + 4 * (k & 3) + (k >> 2);
}

// Fragment slot of master weight index p (layers 1-3), and of its transpose
// (layers 2-3) or its inference-order copy (layer 1).
__device__ __forceinline__ void frag_slots(int p, int* f, int* ft) {
  *f = -1;
  *ft = -1;
  int base, cols, kbn, fbase, tbase;
  if (p < kOffB1) {
    base = kOffW1, cols = kFeatPad, kbn = kKB1, fbase = kFragF1, tbase = -1;
  } else if (p >= kOffW2 && p < kOffB2) {
    base = kOffW2, cols = kHidden, kbn = 4, fbase = kFragF2, tbase = kFragT2;
  } else if (p >= kOffW3 && p < kOffB3) {
    base = kOffW3, cols = kHidden, kbn = 4, fbase = kFragF3, tbase = kFragT3;
  } else {
    return;
  }
  const int q = p - base;
  const int row = q / cols, c = q - row * cols;
  // W[row][c] sits in block (row / 16, c / 16) at lane (row % 16) + 16 ((c % 16) / 4), slot c % 4
  *f = fbase + (((row >> 4) * kbn + (c >> 4)) * 64 + (row & 15) + 16 * ((c & 15) >> 2)) * 4 + (c & 3);
  if (base == kOffW1) {
    const int ci = inference_row(c);
    *ft = kFragF1I + (((row >> 4) * kbn + (ci >> 4)) * 64 + (row & 15) + 16 * ((ci & 15) >> 2)) * 4 +
          (ci & 3);
  }
  if (tbase >= 0) {
    // W^T[c][row]: block (c / 16, row / 16), lane (c % 16) + 16 ((row % 16) / 4), slot row % 4
    *ft = tbase + (((c >> 4) * 4 + (row >> 4)) * 64 + (c & 15) + 16 * ((row & 15) >> 2)) * 4 +
          (row & 3);
  }
}

__global__ __launch_bounds__(kThreads) void k_pack(Model M) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= kParams) return;
  int f, ft;
  frag_slots(p, &f, &ft);
  const uint16_t b = to_bf16(M.theta[p]);
  if (f >= 0) M.frags[f] = b;
  if (ft >= 0) M.frags[ft] = b;
}

__global__ __launch_bounds__(kThreads) void k_adam(Model M, float lr, float beta1, float beta2,
                                                   float eps, float c1, float c2) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= kParams) return;
  const float g = M.grad[p];
  const float m = beta1 * M.m[p] + (1.0f - beta1) * g;
  const float v = beta2 * M.v[p] + (1.0f - beta2) * g * g;
  M.m[p] = m;
  M.v[p] = v;
  const float w = M.theta[p] - lr * (m / c1) / (sqrtf(v / c2) + eps);
  M.theta[p] = w;
  M.grad[p] = 0.0f;
  if (p == 0) M.loss[0] = 0.0f;
  int f, ft;
  frag_slots(p, &f, &ft);
  const uint16_t b = to_bf16(w);
  if (f >= 0) M.frags[f] = b;
  if (ft >= 0) M.frags[ft] = b;
}

__global__ __launch_bounds__(kThreads) void k_zero_grad(Model M) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p < kParams) M.grad[p] = 0.0f;
  if (p == 0) M.loss[0] = 0.0f;
}

unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + kThreads - 1) / kThreads); }

}  // namespace

hipError_t launch_pack(const Model& M, hipStream_t s) {
  hipLaunchKernelGGL(k_pack, dim3(blocks_for(kParams)), dim3(kThreads), 0, s, M);
  return hipGetLastError();
}

hipError_t launch_forward(const Model& M, const float* points, int64_t n, float* out,
                          hipStream_t s) {
  if (n <= 0) return hipSuccess;
  constexpr int64_t per_block = 16 * kFwdTiles * (kFwdThreads / 64);  // kFwdTiles tiles per wavefront
  int64_t blocks = (n + per_block - 1) / per_block;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_forward, dim3(static_cast<unsigned>(blocks)), dim3(kFwdThreads), 0, s, M,
                     points, n, out);
  return hipGetLastError();
}

hipError_t launch_train_batch(const Model& M, const Batch& B, const Samples& S, uint32_t step,
                              hipStream_t s) {
  if (B.n <= 0 || (B.n & 15) || S.count == 0) return hipErrorInvalidValue;
  int64_t blocks = (B.n + 16 * 4 - 1) / (16 * 4);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_train_fb, dim3(static_cast<unsigned>(blocks)), dim3(kThreads), 0, s, M, B,
                     S, step);
  const int64_t n_chunks = (B.n + kWgradChunk - 1) / kWgradChunk;
  hipLaunchKernelGGL(k_wgrad, dim3(blocks_for(64 * kWgradBlocks * n_chunks)), dim3(kThreads), 0, s,
                     M, B);
  hipLaunchKernelGGL(k_bgrad, dim3(blocks_for(64 * 257 * n_chunks)), dim3(kThreads), 0, s, M, B);
  return hipGetLastError();
}

hipError_t launch_adam(const Model& M, uint32_t step, float lr, float beta1, float beta2,
                       float eps, hipStream_t s) {
  const double c1 = 1.0 - __builtin_pow(static_cast<double>(beta1), static_cast<double>(step));
  const double c2 = 1.0 - __builtin_pow(static_cast<double>(beta2), static_cast<double>(step));
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(kParams)), dim3(kThreads), 0, s, M, lr, beta1, beta2,
                     eps, static_cast<float>(c1), static_cast<float>(c2));
  return hipGetLastError();
}

hipError_t launch_zero_grad(const Model& M, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_grad, dim3(blocks_for(kParams)), dim3(kThreads), 0, s, M);
  return hipGetLastError();
}

}  // namespace mlp
}  // namespace iso
