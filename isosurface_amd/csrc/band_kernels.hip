// Learned narrow-band extraction kernels for gfx950 (MI355X); see
// band_kernels.h for the method.  The network itself is in mlp_kernels.hip.
#include <hip/hip_runtime.h>

#include "band_kernels.h"
#include "mc_device.h"
#include "sdf_eval.h"

namespace iso {

namespace {

using namespace mcd;

struct BrickBox {
  int64_t lo[3], hi[3];  // sample range [lo, hi] per axis
};

__device__ __forceinline__ BrickBox brick_box(const Lattice& L, const BrickGrid& G, int64_t b) {
  const int64_t bx = b % G.nbx;
  const int64_t r = b / G.nbx;
  const int64_t by = r % G.nby;
  const int64_t bz = r / G.nby;
  const int64_t n[3] = {L.nx, L.ny, L.nz};
  const int64_t c[3] = {bx, by, bz};
  BrickBox B;
  for (int a = 0; a < 3; ++a) {
    B.lo[a] = kBrick * c[a];
    B.hi[a] = B.lo[a] + kBrick < n[a] - 1 ? B.lo[a] + kBrick : n[a] - 1;
  }
  return B;
}

__global__ __launch_bounds__(256) void k_brick_centres(Lattice L, BrickGrid G,
                                                       float* __restrict__ centres,
                                                       float* __restrict__ half_diag) {
  const int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= G.n) return;
  const BrickBox B = brick_box(L, G, b);
  float d2 = 0.0f;
  for (int a = 0; a < 3; ++a) {
    const float mid = 0.5f * static_cast<float>(B.lo[a] + B.hi[a]);
    centres[3 * b + a] = L.origin[a] + mid * L.spacing[a];
    const float ext = static_cast<float>(B.hi[a] - B.lo[a]) * L.spacing[a];
    d2 += ext * ext;
  }
  half_diag[b] = 0.5f * sqrtf(d2);
}

struct FlagLoad {
  const float* pred;
  const float* half_diag;
  float margin;
  uint8_t* flags;
  __device__ unsigned long long operator()(uint64_t i) const {
    const float p = pred[i];
    // A NaN prediction is never trusted: the brick stays in the band.
    const bool band = !(fabsf(p) > half_diag[i] + margin);
    flags[i] = static_cast<uint8_t>((band ? kBrickBand : 0) | (!(p > 0.0f) ? kBrickInside : 0));
    return band ? 1ull : 0ull;
  }
};

struct FlagSink {
  uint32_t* list;
  __device__ void operator()(uint64_t i, unsigned long long v, unsigned long long ex) const {
    if (v) list[ex] = static_cast<uint32_t>(i);
  }
};

__device__ __forceinline__ int64_t brick_of(int64_t c, int64_t nb) {
  const int64_t b = c / kBrick;
  return b < nb ? b : nb - 1;
}

// Background signs: the predicted sign of each sample's brick (the brick
// holding the sample's cell-space position, clamped at the far faces).
__global__ __launch_bounds__(256) void k_band_background(Lattice L, BrickGrid G,
                                                         const uint8_t* __restrict__ flags,
                                                         int64_t n_points, int64_t n_words,
                                                         uint64_t* __restrict__ S,
                                                         ResetArgs R) {
  reset_state(R);
  const int lane = __lane_id();
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int64_t n_waves = static_cast<int64_t>(gridDim.x) * 4;
  const uint32_t nx = static_cast<uint32_t>(L.nx);
  const uint32_t nxy = static_cast<uint32_t>(L.nx * L.ny);
  for (int64_t w = wave; w < n_words; w += n_waves) {
    const int64_t f = w * 64 + lane;
    bool inside = false;
    if (f < n_points) {
      const uint32_t ff = static_cast<uint32_t>(f);
      const uint32_t z = ff / nxy;
      const uint32_t r = ff - z * nxy;
      const uint32_t y = r / nx;
      const uint32_t x = r - y * nx;
      const int64_t b = (brick_of(z, G.nbz) * G.nby + brick_of(y, G.nby)) * G.nbx +
                        brick_of(x, G.nbx);
      inside = (flags[b] & kBrickInside) != 0;
    }
    const uint64_t m = __ballot(inside);
    if (lane == 0) S[w] = m;
  }
}

__device__ __forceinline__ bool in_band(const uint8_t* flags, const BrickGrid& G, int64_t bx,
                                        int64_t by, int64_t bz) {
  if (bx >= G.nbx || by >= G.nby || bz >= G.nbz) return false;
  return (flags[(bz * G.nby + by) * G.nbx + bx] & kBrickBand) != 0;
}

// One lane per (band brick, sample row): row (ly, lz) of the brick's
// 9 x 9 rows, x across the brick.  A brick evaluates its far face along an
// axis only where the neighbour along that axis is not in the band, so every
// sample of the band's closed bricks is evaluated (the band brick with the
// largest coordinate sum among those holding it always takes it).
constexpr int kRows = (kBrick + 1) * (kBrick + 1);

__global__ __launch_bounds__(256) void k_band_exact(const DevOp* __restrict__ ops, int n_ops,
                                                    Lattice L, BrickGrid G,
                                                    const uint8_t* __restrict__ flags,
                                                    const uint32_t* __restrict__ band_list,
                                                    uint64_t n_band, uint64_t* __restrict__ S,
                                                    unsigned long long* evaluated) {
  const uint64_t n_items = n_band * kRows;
  uint32_t count = 0;
  for (uint64_t it = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; it < n_items;
       it += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const int64_t b = band_list[it / kRows];
    const int row = static_cast<int>(it % kRows);
    const int ly = row % (kBrick + 1), lz = row / (kBrick + 1);
    const int64_t bx = b % G.nbx, by = (b / G.nbx) % G.nby, bz = b / (G.nbx * G.nby);
    const int64_t y = kBrick * by + ly, z = kBrick * bz + lz;
    if (y > L.ny - 1 || z > L.nz - 1) continue;
    if (ly == kBrick && in_band(flags, G, bx, by + 1, bz)) continue;
    if (lz == kBrick && in_band(flags, G, bx, by, bz + 1)) continue;
    const int64_t x0 = kBrick * bx;
    int64_t x1 = x0 + kBrick;  // inclusive far face unless the +x neighbour takes it
    if (in_band(flags, G, bx + 1, by, bz)) x1 -= 1;
    if (x1 > L.nx - 1) x1 = L.nx - 1;
    const float py = L.origin[1] + static_cast<float>(y) * L.spacing[1];
    const float pz = L.origin[2] + static_cast<float>(z) * L.spacing[2];
    const int64_t f0 = (z * L.ny + y) * L.nx + x0;
    const int64_t w0 = f0 >> 6;
    uint64_t set0 = 0, clr0 = 0, set1 = 0, clr1 = 0;
    for (int64_t x = x0; x <= x1; ++x) {
      const float px = L.origin[0] + static_cast<float>(x) * L.spacing[0];
      const float v = sdf_eval(ops, n_ops, px, py, pz);
      const int64_t f = f0 + (x - x0);
      const uint64_t bit = 1ull << (f & 63);
      const bool inside = !(v > 0.0f);  // classify_corners, distance.rs:52-54
      if ((f >> 6) == w0) {
        if (inside) set0 |= bit;
        else clr0 |= bit;
      } else {
        if (inside) set1 |= bit;
        else clr1 |= bit;
      }
      ++count;
    }
    unsigned long long* W = reinterpret_cast<unsigned long long*>(S);
    if (set0) atomicOr(&W[w0], set0);
    if (clr0) atomicAnd(&W[w0], ~clr0);
    if (set1) atomicOr(&W[w0 + 1], set1);
    if (clr1) atomicAnd(&W[w0 + 1], ~clr1);
  }
  unsigned long long c = count;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
  if (__lane_id() == 0 && c) atomicAdd(evaluated, c);
}

__global__ __launch_bounds__(256) void k_band_missed(Lattice L, McDims D, BrickGrid G,
                                                     const uint8_t* __restrict__ flags,
                                                     const uint32_t* __restrict__ active,
                                                     uint64_t capacity,
                                                     const unsigned long long* __restrict__ totals,
                                                     unsigned long long* missed) {
  uint64_t n = totals[kTotActive];
  if (n > capacity) n = capacity;
  const uint32_t spr = static_cast<uint32_t>(D.segs_per_row);
  const uint32_t cyn = static_cast<uint32_t>(L.ny - 1);
  uint32_t count = 0;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t rec = active[i];
    const uint32_t g = rec >> 6;
    const uint32_t row = g / spr;
    const uint32_t w = g - row * spr;
    const int64_t cz = row / cyn, cy = row - (row / cyn) * cyn;
    const int64_t cx = 64 * static_cast<int64_t>(w) + (rec & 63u);
    const int64_t b = ((cz / kBrick) * G.nby + cy / kBrick) * G.nbx + cx / kBrick;
    if (!(flags[b] & kBrickBand)) ++count;
  }
  unsigned long long c = count;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
  if (__lane_id() == 0 && c) atomicAdd(missed, c);
}

unsigned grid_of(uint64_t n, uint64_t cap) {
  uint64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return static_cast<unsigned>(b);
}

}  // namespace

BrickGrid brick_grid(const Lattice& L) {
  BrickGrid G;
  G.nbx = (L.nx - 1 + kBrick - 1) / kBrick;
  G.nby = (L.ny - 1 + kBrick - 1) / kBrick;
  G.nbz = (L.nz - 1 + kBrick - 1) / kBrick;
  G.n = G.nbx * G.nby * G.nbz;
  return G;
}

hipError_t launch_brick_centres(const Lattice& L, const BrickGrid& G, float* centres,
                                float* half_diag, hipStream_t s) {
  hipLaunchKernelGGL(k_brick_centres, dim3(grid_of(G.n, 1u << 30)), dim3(256), 0, s, L, G,
                     centres, half_diag);
  return hipGetLastError();
}

hipError_t launch_brick_flags(const BrickGrid& G, const float* pred, const float* half_diag,
                              float margin, uint8_t* flags, uint32_t* band_list,
                              const scan::State& st, hipStream_t s) {
  FlagLoad load{pred, half_diag, margin, flags};
  FlagSink sink{band_list};
  return scan::launch<scan::kItems>(static_cast<uint64_t>(G.n), load, sink, st, s);
}

hipError_t launch_band_background(const Lattice& L, const McDims& D, const BrickGrid& G,
                                  const uint8_t* flags, const McBuffers& B, hipStream_t s) {
  ResetArgs R;
  R.signs = B.signs;
  R.n_words = D.n_sign_words;
  R.agg = B.agg;
  R.incl = B.incl;
  R.n_blocks = D.n_blocks;
  R.totals = B.totals;
  int64_t blocks = (D.n_sign_words + 3) / 4;
  const int64_t min_blocks = (D.n_blocks + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  if (blocks < min_blocks) blocks = min_blocks;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_band_background, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, L,
                     G, flags, D.n_points, D.n_sign_words, B.signs, R);
  return hipGetLastError();
}

hipError_t launch_band_exact(const DevOp* ops, int n_ops, const Lattice& L, const BrickGrid& G,
                             const uint8_t* flags, const uint32_t* band_list, uint64_t n_band,
                             const McBuffers& B, unsigned long long* evaluated, hipStream_t s) {
  if (n_band == 0) return hipSuccess;
  hipLaunchKernelGGL(k_band_exact, dim3(grid_of(n_band * kRows, 1u << 20)), dim3(256), 0, s, ops,
                     n_ops, L, G, flags, band_list, n_band, B.signs, evaluated);
  return hipGetLastError();
}

hipError_t launch_band_missed(const Lattice& L, const McDims& D, const BrickGrid& G,
                              const uint8_t* flags, const McBuffers& B,
                              unsigned long long* missed, hipStream_t s) {
  hipLaunchKernelGGL(k_band_missed, dim3(grid_of(B.active_capacity, 16384)), dim3(256), 0, s, L,
                     D, G, flags, B.active, B.active_capacity, B.totals, missed);
  return hipGetLastError();
}

}  // namespace iso
