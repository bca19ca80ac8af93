// Learned narrow-band extraction kernels for gfx950 (MI355X); see
// band_kernels.h for the method.  The network itself is in mlp_kernels.hip.
#include <hip/hip_runtime.h>

#include "band_kernels.h"
#include "mc_device.h"
#include <cmath>
#include <type_traits>

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
    B.lo[a] = static_cast<int64_t>(G.brick) * c[a];
    B.hi[a] = B.lo[a] + G.brick < n[a] - 1 ? B.lo[a] + G.brick : n[a] - 1;
  }
  return B;
}

__global__ __launch_bounds__(256) void k_brick_centres(Lattice L, BrickGrid G,
                                                       float* __restrict__ centres,
                                                       float* __restrict__ half_diag) {
  const int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= G.n) return;
  const BrickBox B = brick_box(L, G, b);
  // Every brick is tested with the half-diagonal of a full brick (bricks
  // cut by the lattice's far faces are smaller), so one calibration holds
  // for all of them.
  float d2 = 0.0f;
  for (int a = 0; a < 3; ++a) {
    const float mid = 0.5f * static_cast<float>(B.lo[a] + B.hi[a]);
    centres[3 * b + a] = L.origin[a] + mid * L.spacing[a];
    const float ext = static_cast<float>(G.brick) * L.spacing[a];
    d2 += ext * ext;
  }
  half_diag[b] = 0.5f * sqrtf(d2);
}

// Adds the number of lanes with `on` set to *counter (one atomic per wave).
__device__ __forceinline__ void count_lanes(bool on, unsigned long long* counter) {
  const uint64_t m = __ballot(on);
  if (on && __lane_id() == __builtin_ctzll(m))
    atomicAdd(counter, static_cast<unsigned long long>(__popcll(m)));
}

// Every program is 1-Lipschitz (exact primitive distances, min, max,
// negation, translation and the polynomial smooth minimum, whose gradient is
// a convex combination of its operands'), so a centre value farther from
// zero than the distance to any sample of the box, plus the f32 evaluation
// error at both, proves that every sample of the box has the centre's sign.
// `widen` bounds the distance between the computed centre and any computed
// sample position (the half-diagonal, with the f32 rounding of both
// positions, pos_slack); the evaluation's own bound E holds at every point
// within `widen` of the centre, so the sample's f32 value lies within
// v +- (widen + 2E) and the sign is proven when |v| exceeds that.
__device__ __forceinline__ uint8_t certify_at(const DevOp* __restrict__ ops, int n_ops, float x,
                                              float y, float z, float half_diag,
                                              float pos_slack) {
  const float widen = half_diag * (1.0f + kCertifyRelSlack) + pos_slack;
  float e;
  const float v = sdf_eval_bound(ops, n_ops, x, y, z, widen, &e);
  return fabsf(v) > widen + 2.0f * e + kCertifyAbsSlack
             ? static_cast<uint8_t>(kBrickProven | (!(v > 0.0f) ? kBrickInside : 0))
             : static_cast<uint8_t>(0);
}

__global__ __launch_bounds__(256) void k_coarse_certify(const DevOp* __restrict__ ops, int n_ops,
                                                        Lattice L, BrickGrid G,
                                                        uint8_t* __restrict__ coarse,
                                                        unsigned long long* certified,
                                                        float pos_slack) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  count_lanes(s < G.nc, certified);
  if (s >= G.nc) return;
  const int64_t sx = s % G.ncx;
  const int64_t r = s / G.ncx;
  const int64_t c[3] = {sx, r % G.ncy, r / G.ncy};
  const int64_t n[3] = {L.nx, L.ny, L.nz};
  const int64_t kSpan = static_cast<int64_t>(G.brick) * kCoarse;  // cells per edge
  float p[3];
  float d2 = 0.0f;
  for (int a = 0; a < 3; ++a) {
    // The closed sample box [lo, hi] holds every brick box of the super-brick.
    const int64_t lo = kSpan * c[a];
    const int64_t hi = lo + kSpan < n[a] - 1 ? lo + kSpan : n[a] - 1;
    p[a] = L.origin[a] + 0.5f * static_cast<float>(lo + hi) * L.spacing[a];
    const float ext = static_cast<float>(kSpan) * L.spacing[a];
    d2 += ext * ext;
  }
  coarse[s] = certify_at(ops, n_ops, p[0], p[1], p[2], 0.5f * sqrtf(d2), pos_slack);
}

// Brick proofs: one wavefront per super-brick, one lane per brick (4 x 4 x 4
// bricks, x fastest), so a wavefront either copies its super-brick's proof
// or evaluates all of its bricks' centres together.
__global__ __launch_bounds__(256) void k_brick_certify(const DevOp* __restrict__ ops, int n_ops,
                                                       BrickGrid G,
                                                       const uint8_t* __restrict__ coarse,
                                                       const float* __restrict__ centres,
                                                       const float* __restrict__ half_diag,
                                                       uint8_t* __restrict__ proof,
                                                       unsigned long long* certified,
                                                       float pos_slack) {
  static_assert(kCoarse * kCoarse * kCoarse == 64, "one lane per brick of a super-brick");
  const int lane = __lane_id();
  const int64_t s = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (s >= G.nc) return;  // uniform per wave
  const int64_t sx = s % G.ncx;
  const int64_t r = s / G.ncx;
  const int64_t bx = kCoarse * sx + lane % kCoarse;
  const int64_t by = kCoarse * (r % G.ncy) + (lane / kCoarse) % kCoarse;
  const int64_t bz = kCoarse * (r / G.ncy) + lane / (kCoarse * kCoarse);
  const bool live = bx < G.nbx && by < G.nby && bz < G.nbz;
  const int64_t b = (bz * G.nby + by) * G.nbx + bx;
  const uint8_t cf = coarse[s];
  const bool evaluate = live && !(cf & kBrickProven);
  count_lanes(evaluate, certified);
  uint8_t out = cf;
  if (evaluate) {
    out = certify_at(ops, n_ops, centres[3 * b], centres[3 * b + 1], centres[3 * b + 2],
                     half_diag[b], pos_slack);
  }
  if (live) proof[b] = out;
}

struct FlagLoad {
  const float* pred;       // network predictions; null: no network
  const float* half_diag;
  const uint8_t* proof;    // non-null: a brick leaves the band only by its proof
  float margin;
  uint8_t* flags;
  unsigned long long* overturned;  // network culls no proof allows
  __device__ unsigned long long operator()(uint64_t i) const {
    bool band = true, inside = false;
    if (pred) {
      const float p = pred[i];
      // A NaN prediction is never trusted: the brick stays in the band.
      band = !(fabsf(p) > half_diag[i] + margin);
      inside = !(p > 0.0f);
    }
    if (proof) {
      const uint8_t pf = proof[i];
      const bool proven = (pf & kBrickProven) != 0;
      count_lanes(pred != nullptr && !band && !proven, overturned);
      band = !proven;
      inside = proven && (pf & kBrickInside) != 0;
    }
    flags[i] = static_cast<uint8_t>((band ? kBrickBand : 0) | (inside ? kBrickInside : 0));
    return band ? 1ull : 0ull;
  }
};

struct FlagSink {
  uint32_t* list;
  __device__ void operator()(uint64_t i, unsigned long long v, unsigned long long ex) const {
    if (v) list[ex] = static_cast<uint32_t>(i);
  }
};

__device__ __forceinline__ int64_t brick_of(int64_t c, int shift, int64_t nb) {
  const int64_t b = c >> shift;
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
      const int64_t b =
          (brick_of(z, G.shift, G.nbz) * G.nby + brick_of(y, G.shift, G.nby)) * G.nbx +
          brick_of(x, G.shift, G.nbx);
      inside = (flags[b] & kBrickInside) != 0;
    }
    const uint64_t m = __ballot(inside);
    if (lane == 0) S[w] = m;
  }
}

// Neighbour bricks (bit (dz+1)*9 + (dy+1)*3 + dx+1) that also hold a sample
// of class (cx, cy, cz) — per axis 0: on the brick's low face, 1: inside,
// 2: on its high face — and come later in brick order.  The band brick
// latest in brick order among those holding a sample evaluates it, so each
// exact sample is evaluated once.
__device__ __forceinline__ uint32_t later_holders(int cx, int cy, int cz) {
  uint32_t m = 0;
  for (int dz = cz == 0 ? -1 : 0; dz <= (cz == 2 ? 1 : 0); ++dz)
    for (int dy = cy == 0 ? -1 : 0; dy <= (cy == 2 ? 1 : 0); ++dy)
      for (int dx = cx == 0 ? -1 : 0; dx <= (cx == 2 ? 1 : 0); ++dx)
        if (dz > 0 || (dz == 0 && (dy > 0 || (dy == 0 && dx > 0))))
          m |= 1u << ((dz + 1) * 9 + (dy + 1) * 3 + dx + 1);
  return m;
}

constexpr int kExactThreads = 256;

// Sizes of the exact-sample kernel for bricks of 2^kShift cells per edge.
// Eight-cell bricks take fewer bricks per workgroup pass, so a workgroup's
// LDS stays near 17 KB and eight workgroups still fit a CU.
template <int kShift>
struct ExactShape {
  static constexpr int kBrick = 1 << kShift;              // cells per brick edge
  static constexpr int kSide = kBrick + 1;                 // samples per edge of a brick box
  static constexpr int kBoxRows = kSide * kSide;           // x rows of a brick's sample box
  static constexpr int kBoxSamples = kBoxRows * kSide;
  static constexpr int kGroupBricks = kShift == 2 ? 32 : 8;  // band bricks per workgroup pass
  static constexpr int kBricksPerWave = kGroupBricks / (kExactThreads / 64);
  // box position lx + kSide (ly + kSide lz)
  using Pos = typename std::conditional<(kBoxSamples <= 256), uint8_t, uint16_t>::type;
  static_assert(kSide <= 32, "a box row is one 32-bit mask");
};

// Thirty-two band bricks (of 4^3 cells; eight of 8^3) per workgroup pass.
// Each wavefront takes a quarter of them: it reads a brick's and its 26 neighbours' band flags, walks the brick's
// closed sample box (x fastest) and appends the samples the brick owns to
// the brick's list in LDS.  Then all 256 lanes evaluate the 32 lists
// as one queue (a lane's brick by binary search of the lists' offsets), so
// lanes stay busy however few samples each brick owns, store the values,
// and gather sign bits per x row in LDS; last, one lane per row merges the
// row into the sign words (at most two words, OR for inside, AND-NOT for
// outside, over the background the bricks' flags wrote).
template <int kShift>
__global__ __launch_bounds__(kExactThreads) void
k_band_exact(
    const DevOp* __restrict__ ops, int n_ops, Lattice L, BrickGrid G,
    const uint8_t* __restrict__ flags, const uint32_t* __restrict__ band_list, uint64_t n_band,
    uint64_t* __restrict__ S, float* __restrict__ values, unsigned long long* evaluated) {
  using X = ExactShape<kShift>;
  constexpr int kBrick = X::kBrick, kSide = X::kSide, kBoxRows = X::kBoxRows;
  constexpr int kGroupBricks = X::kGroupBricks, kBricksPerWave = X::kBricksPerWave;
  __shared__ uint32_t s_in[kGroupBricks][kBoxRows];
  __shared__ uint32_t s_own[kGroupBricks][kBoxRows];
  __shared__ typename X::Pos s_pos[kGroupBricks][X::kBoxSamples];
  __shared__ uint32_t s_lo[kGroupBricks][3];
  __shared__ uint32_t s_n[kGroupBricks];
  __shared__ uint32_t s_pre[kGroupBricks + 1];
  const int lane = __lane_id();
  const int wv = threadIdx.x >> 6;
  uint32_t count = 0;
  for (uint64_t base = static_cast<uint64_t>(blockIdx.x) * kGroupBricks; base < n_band;
       base += static_cast<uint64_t>(gridDim.x) * kGroupBricks) {
    for (int r = threadIdx.x; r < kGroupBricks * kBoxRows; r += kExactThreads) {
      (&s_in[0][0])[r] = 0;
      (&s_own[0][0])[r] = 0;
    }
    __syncthreads();
#pragma unroll 1
    for (int q = 0; q < kBricksPerWave; ++q) {
      const int j = wv * kBricksPerWave + q;
      const uint64_t item = base + j;
      uint32_t n_own = 0;
      if (item < n_band) {  // uniform per wave
        const int64_t b = band_list[item];
        const int64_t bc[3] = {b % G.nbx, (b / G.nbx) % G.nby, b / (G.nbx * G.nby)};
        const int64_t nbk[3] = {G.nbx, G.nby, G.nbz};
        const int64_t n[3] = {L.nx, L.ny, L.nz};
        // Band flags of the 26 neighbours (and the brick itself, bit 13).
        bool band = false;
        if (lane < 27) {
          const int64_t nq[3] = {bc[0] + lane % 3 - 1, bc[1] + (lane / 3) % 3 - 1,
                                 bc[2] + lane / 9 - 1};
          if (nq[0] >= 0 && nq[1] >= 0 && nq[2] >= 0 && nq[0] < nbk[0] && nq[1] < nbk[1] &&
              nq[2] < nbk[2])
            band = (flags[(nq[2] * G.nby + nq[1]) * G.nbx + nq[0]] & kBrickBand) != 0;
        }
        const uint32_t nb = static_cast<uint32_t>(__ballot(band));
        int64_t lo[3];
        int ext[3];
        for (int a = 0; a < 3; ++a) {
          lo[a] = kBrick * bc[a];
          int64_t hi = lo[a] + kBrick < n[a] - 1 ? lo[a] + kBrick : n[a] - 1;
          // the high face goes to the next brick along the axis when it is in the band
          const int up = a == 0 ? 14 : a == 1 ? 16 : 22;
          if (hi == lo[a] + kBrick && (nb >> up) & 1u) hi -= 1;
          ext[a] = static_cast<int>(hi - lo[a] + 1);
        }
        const int exy = ext[0] * ext[1];
        const int total = exy * ext[2];
#pragma unroll 1
        for (int t0 = 0; t0 < total; t0 += 64) {  // uniform trip count
          const int t = t0 + lane;
          bool own = false;
          int lx = 0, ly = 0, lz = 0;
          if (t < total) {
            lz = t / exy;
            const int rem = t - lz * exy;
            ly = rem / ext[0];
            lx = rem - ly * ext[0];
            const int cx = lx == 0 ? 0 : lx == kBrick ? 2 : 1;
            const int cy = ly == 0 ? 0 : ly == kBrick ? 2 : 1;
            const int cz = lz == 0 ? 0 : lz == kBrick ? 2 : 1;
            own = !(later_holders(cx, cy, cz) & nb);
          }
          const uint64_t m = __ballot(own);
          if (own) {
            const uint32_t k = n_own + static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1)));
            s_pos[j][k] = static_cast<typename X::Pos>(lx + kSide * (ly + kSide * lz));
            atomicOr(&s_own[j][ly + kSide * lz], 1u << lx);
          }
          n_own += static_cast<uint32_t>(__popcll(m));
        }
        if (lane == 0) {
          s_lo[j][0] = static_cast<uint32_t>(lo[0]);
          s_lo[j][1] = static_cast<uint32_t>(lo[1]);
          s_lo[j][2] = static_cast<uint32_t>(lo[2]);
        }
      }
      if (lane == 0) s_n[j] = n_own;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t a = 0;
      for (int j = 0; j < kGroupBricks; ++j) {
        s_pre[j] = a;
        a += s_n[j];
      }
      s_pre[kGroupBricks] = a;
    }
    __syncthreads();
    const uint32_t T = s_pre[kGroupBricks];
    // Whole wavefronts step through the queue (the culled evaluator works on
    // all 64 lanes); lanes past its end repeat its last entry and drop it.
#pragma unroll 1
    for (uint32_t t0 = threadIdx.x & ~63u; t0 < T; t0 += kExactThreads) {
      const bool valid = t0 + (threadIdx.x & 63u) < T;
      const uint32_t t = valid ? t0 + (threadIdx.x & 63u) : T - 1;
      int j = 0;  // the brick whose list holds queue entry t: s_pre[j] <= t < s_pre[j + 1]
#pragma unroll
      for (int step = kGroupBricks / 2; step >= 1; step >>= 1)
        if (s_pre[j + step] <= t) j += step;
      const int pos = s_pos[j][t - s_pre[j]];
      const int lz = pos / kBoxRows;
      const int ly = (pos - lz * kBoxRows) / kSide;
      const int lx = pos - lz * kBoxRows - ly * kSide;
      const int64_t x = static_cast<int64_t>(s_lo[j][0]) + lx;
      const int64_t y = static_cast<int64_t>(s_lo[j][1]) + ly;
      const int64_t z = static_cast<int64_t>(s_lo[j][2]) + lz;
      const float v =
          sdf_eval_culled(ops, n_ops, L.origin[0] + static_cast<float>(x) * L.spacing[0],
                          L.origin[1] + static_cast<float>(y) * L.spacing[1],
                          L.origin[2] + static_cast<float>(z) * L.spacing[2]);
      if (!valid) continue;
      if (values) values[(z * L.ny + y) * L.nx + x] = v;
      // classify_corners, distance.rs:52-54
      if (!(v > 0.0f)) atomicOr(&s_in[j][ly + kSide * lz], 1u << lx);
      ++count;
    }
    __syncthreads();
    for (int r = threadIdx.x; r < kGroupBricks * kBoxRows; r += kExactThreads) {
      const int j = r / kBoxRows;
      const int row = r - j * kBoxRows;
      const uint64_t own = s_own[j][row];
      if (!own) continue;
      const uint64_t in = s_in[j][row] & own;
      const uint64_t out = own & ~in;
      const int ly = row % kSide, lz = row / kSide;
      const int64_t f0 = ((static_cast<int64_t>(s_lo[j][2]) + lz) * L.ny +
                          static_cast<int64_t>(s_lo[j][1]) + ly) * L.nx +
                         static_cast<int64_t>(s_lo[j][0]);
      const int64_t w0 = f0 >> 6;
      const int sh = static_cast<int>(f0 & 63);
      unsigned long long* W = reinterpret_cast<unsigned long long*>(S);
      if (in << sh) atomicOr(&W[w0], in << sh);
      if (out << sh) atomicAnd(&W[w0], ~(out << sh));
      if (sh) {
        const uint64_t in1 = in >> (64 - sh), out1 = out >> (64 - sh);
        if (in1) atomicOr(&W[w0 + 1], in1);
        if (out1) atomicAnd(&W[w0 + 1], ~out1);
      }
    }
    __syncthreads();
  }
  unsigned long long c = count;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
  if (__lane_id() == 0 && c) atomicAdd(evaluated, c);
}

// Cells whose marching-cubes case differs between two sign arrays of the
// same lattice (a band extraction's and the dense one's): a cell differs
// when any of its eight corner bits does.  One thread per 64-cell segment.
__global__ __launch_bounds__(256) void k_band_diff(Lattice L, McDims D,
                                                   const uint64_t* __restrict__ Sa,
                                                   const uint64_t* __restrict__ Sb,
                                                   unsigned long long* differing) {
  const int64_t nx = L.nx, nxy = L.nx * L.ny;
  const int64_t spr = D.segs_per_row;
  const int64_t cyn = L.ny - 1;
  uint32_t count = 0;
  for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < D.n_segs;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t row = g / spr, w = g - (g / spr) * spr;
    const int64_t cz = row / cyn, cy = row - (row / cyn) * cyn;
    const int64_t f00 = cz * nxy + cy * nx + 64 * w;
    uint64_t m = 0, n = 0;
    const int64_t offs[4] = {0, nx, nxy, nxy + nx};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      uint64_t na, nb;
      const uint64_t a = row_word(Sa, f00 + offs[k], &na);
      const uint64_t b = row_word(Sb, f00 + offs[k], &nb);
      m |= a ^ b;
      n |= na ^ nb;
    }
    // cell i has corners at x = i and i + 1
    count += __popcll((m | (m >> 1) | (n << 63)) & seg_valid(nx, w));
  }
  unsigned long long c = count;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
  if (__lane_id() == 0 && c) atomicAdd(differing, c);
}

unsigned grid_of(uint64_t n, uint64_t cap) {
  uint64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return static_cast<unsigned>(b);
}

}  // namespace

BrickGrid brick_grid(const Lattice& L, int shift) {
  BrickGrid G;
  G.shift = shift;
  G.brick = 1 << shift;
  G.nbx = (L.nx - 1 + G.brick - 1) / G.brick;
  G.nby = (L.ny - 1 + G.brick - 1) / G.brick;
  G.nbz = (L.nz - 1 + G.brick - 1) / G.brick;
  G.n = G.nbx * G.nby * G.nbz;
  G.ncx = (G.nbx + kCoarse - 1) / kCoarse;
  G.ncy = (G.nby + kCoarse - 1) / kCoarse;
  G.ncz = (G.nbz + kCoarse - 1) / kCoarse;
  G.nc = G.ncx * G.ncy * G.ncz;
  return G;
}

hipError_t launch_coarse_certify(const DevOp* ops, int n_ops, const Lattice& L,
                                 const BrickGrid& G, uint8_t* coarse,
                                 unsigned long long* certified, hipStream_t s) {
  if (!ops || !coarse || !certified) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_coarse_certify, dim3(grid_of(G.nc, 1u << 30)), dim3(256), 0, s, ops, n_ops,
                     L, G, coarse, certified, certify_position_slack(L));
  return hipGetLastError();
}

hipError_t launch_brick_centres(const Lattice& L, const BrickGrid& G, float* centres,
                                float* half_diag, hipStream_t s) {
  hipLaunchKernelGGL(k_brick_centres, dim3(grid_of(G.n, 1u << 30)), dim3(256), 0, s, L, G,
                     centres, half_diag);
  return hipGetLastError();
}

float certify_position_slack(const Lattice& L) {
  // A lattice position origin + (float)i * spacing rounds twice (the
  // product and the sum), each by at most 2^-24 of the magnitude, so a
  // computed sample or centre coordinate is within 2^-23 M of the exact
  // lattice position, M bounding |origin| + i |spacing| over the lattice.
  // Two such points (centre and sample) add at most 2 sqrt(3) 2^-23 M to
  // their Euclidean distance.
  double m = 0.0;
  const int64_t n[3] = {L.nx, L.ny, L.nz};
  for (int a = 0; a < 3; ++a) {
    const double ma = std::fabs(static_cast<double>(L.origin[a])) +
                      static_cast<double>(n[a] > 0 ? n[a] - 1 : 0) *
                          std::fabs(static_cast<double>(L.spacing[a]));
    if (ma > m) m = ma;
  }
  return static_cast<float>(2.0 * std::sqrt(3.0) * std::ldexp(m, -23) * 1.01);
}

hipError_t launch_brick_certify(const DevOp* ops, int n_ops, const Lattice& L,
                                const BrickGrid& G, const uint8_t* coarse, const float* centres,
                                const float* half_diag, uint8_t* proof,
                                unsigned long long* certified, hipStream_t s) {
  if (!ops || !coarse || !centres || !half_diag || !proof || !certified)
    return hipErrorInvalidValue;
  const uint64_t blocks = (static_cast<uint64_t>(G.nc) + 3) / 4;
  if (blocks > 0x7fffffffu) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_brick_certify, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, ops,
                     n_ops, G, coarse, centres, half_diag, proof, certified,
                     certify_position_slack(L));
  return hipGetLastError();
}

hipError_t launch_brick_flags(const BrickGrid& G, const float* pred, const float* half_diag,
                              const uint8_t* proof, float margin, uint8_t* flags,
                              uint32_t* band_list, unsigned long long* overturned,
                              const scan::State& st, hipStream_t s) {
  if (!pred && !proof) return hipErrorInvalidValue;
  if (proof && !overturned) return hipErrorInvalidValue;
  FlagLoad load{pred, half_diag, proof, margin, flags, overturned};
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
                             const McBuffers& B, float* values, unsigned long long* evaluated,
                             hipStream_t s) {
  if (n_band == 0) return hipSuccess;
  if (!ops || !flags || !band_list || !B.signs || !evaluated) return hipErrorInvalidValue;
  auto launch = [&](auto kernel, uint64_t group) {
    uint64_t blocks = (n_band + group - 1) / group;
    if (blocks > (1u << 20)) blocks = 1u << 20;
    hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(blocks)), dim3(kExactThreads), 0, s, ops,
                       n_ops, L, G, flags, band_list, n_band, B.signs, values, evaluated);
    return hipGetLastError();
  };
  if (G.shift == 2) return launch(k_band_exact<2>, ExactShape<2>::kGroupBricks);
  if (G.shift == 3) return launch(k_band_exact<3>, ExactShape<3>::kGroupBricks);
  return hipErrorInvalidValue;
}

hipError_t launch_band_diff(const Lattice& L, const McDims& D, const uint64_t* signs_a,
                            const uint64_t* signs_b, unsigned long long* differing,
                            hipStream_t s) {
  if (!signs_a || !signs_b || !differing) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_band_diff, dim3(grid_of(static_cast<uint64_t>(D.n_segs), 16384)),
                     dim3(256), 0, s, L, D, signs_a, signs_b, differing);
  return hipGetLastError();
}

}  // namespace iso
