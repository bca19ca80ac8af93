// Linear-hashed octree marching cubes for gfx950 (MI355X).
//
// LinearHashedMarchingCubes::extract (linear_hashed_marching_cubes.rs:62-84)
// over ImplicitOctree::traverse (traversal/implicit_octree.rs:39-99):
//   1. level    per depth: sample each node at its Morton centre, decide
//               refinement (implicit_octree.rs:47-59), scan the decisions
//               and write the refined nodes' children in key order;
//   2. primal   every leaf marks its primal vertices (morton.rs:120-138) on a
//               dense lattice with the deepest leaf level (atomic max); a
//               compaction lists them in ascending Morton order;
//   3. cubes    per primal vertex the eight dual corners (morton.rs:142-146),
//               each the deepest existing node on its parent chain (binary
//               search in the sorted key array); cube index; triangle count;
//   4. own      each dual-cube edge (MortonKey, index_cache.rs:63-73) is owned
//               by the first cube, in order, whose triangles use it; owned
//               counts and triangle counts are scanned into first ids;
//   5. emit     owners write positions (their first local edge's crossing);
//               every cube writes its index triples.
// The output is the reference's for the primal-vertex order "ascending
// Morton key" (the reference's own order is a std HashMap's, random per
// process).  Morton::level() keeps the reference's quirk: the truncated
// constant 0.48089834696 (morton.rs:35, 54-59) puts the first keys of each
// depth one level low (exactly the power of 8 below 8^11, then more: a
// table of split points found with the host's libm log).  Deep trees
// (beyond kLhmcDenseDepth) list their primal vertices by a radix sort of
// the leaves' (vertex, level) pairs instead of a dense lattice.  A node is refined while its level is
// below M = max(2, max_depth); the origin-corner node of depth M has level
// M - 1, so nodes reach depth M + 1, the "dense depth" of the primal lattice.
#include <hip/hip_runtime.h>

#include <cmath>
#include <rocprim/device/device_radix_sort.hpp>

#include "lhmc_kernels.h"
#include "mc_tables.h"
#include "sdf_eval.h"

namespace iso {

namespace {

// morton.rs:17-32
constexpr uint64_t kMask2 = 0x7FC0000FF80001FFull;
constexpr uint64_t kMask1 = 0x01C0E070381C0E07ull;
constexpr uint64_t kMask0 = 0x9249249249249249ull;
constexpr uint64_t kTz = 0x4924924924924924ull;
constexpr uint64_t kTy = 0x2492492492492492ull;
constexpr uint64_t kTx = 0x9249249249249249ull;
constexpr uint64_t kT1 = 0xB6DB6DB6DB6DB6DBull;
constexpr uint64_t kT2 = 0xDB6DB6DB6DB6DB6Dull;
constexpr uint64_t kT3 = 0x6DB6DB6DB6DB6DB6ull;
constexpr int kMaxLevel = 21;
constexpr float kSqrtOf3 = 1.73205080757f;  // distance.rs:18
constexpr unsigned long long kEmpty = ~0ull;

__constant__ tables::McTables c_lt = tables::kTables;
__constant__ uint64_t c_level_split[kLhmcMaxLevels];  // lhmc_level_splits
__constant__ int c_remap[8] = {2, 3, 1, 0, 6, 7, 5, 4};  // REMAP_CUBE, marching_cubes_tables.rs:29

__device__ __forceinline__ int depth_of(uint64_t k) { return (63 - __clzll(k)) / 3; }

// Morton::level() (morton.rs:53-59): the depth, one less for the first
// c_level_split[d] keys of depth d.
__device__ __forceinline__ int level_of(uint64_t k) {
  if (k == 0) return 0;
  const int d = depth_of(k);
  return (d >= 1 && k - (1ull << (3 * d)) < c_level_split[d]) ? d - 1 : d;
}

__device__ __forceinline__ float size_of(uint64_t k) {  // morton.rs:73-75
  return 1.0f / static_cast<float>(2ull << level_of(k));
}

struct F3 {
  float x, y, z;
};

__device__ __forceinline__ F3 center_of(uint64_t k) {  // morton.rs:78-116
  uint64_t bz = (k >> 2) & kMask0;
  uint64_t by = (k >> 1) & kMask0;
  uint64_t bx = k & kMask0;
  const int lv = level_of(k);
  if (lv > 1) {
    bz = (bz | (bz >> 2) | (bz >> 4)) & kMask1;
    by = (by | (by >> 2) | (by >> 4)) & kMask1;
    bx = (bx | (bx >> 2) | (bx >> 4)) & kMask1;
    if (lv > 3) {
      bz = (bz | (bz >> 6) | (bz >> 12)) & kMask2;
      by = (by | (by >> 6) | (by >> 12)) & kMask2;
      bx = (bx | (bx >> 6) | (bx >> 12)) & kMask2;
      if (lv > 9) {
        bz = bz | (bz >> 18) | (bz >> 36);
        by = by | (by >> 18) | (by >> 36);
        bx = bx | (bx >> 18) | (bx >> 36);
      }
    }
  }
  const uint64_t length_mask = (1ull << lv) - 1;
  bz &= length_mask;
  by &= length_mask;
  bx &= length_mask;
  const float s = size_of(k);
  const float s2 = s * 2.0f;
  return {static_cast<float>(bx) * s2 + s, static_cast<float>(by) * s2 + s,
          static_cast<float>(bz) * s2 + s};
}

__device__ __forceinline__ uint64_t m_add(uint64_t a, uint64_t b) {  // morton.rs:153-163
  return (((a | kT1) + (b & kTz)) & kTz) | (((a | kT2) + (b & kTy)) & kTy) |
         (((a | kT3) + (b & kTx)) & kTx);
}

__device__ __forceinline__ uint64_t m_sub(uint64_t a, uint64_t b) {  // morton.rs:165-175
  return (((a & kTz) - (b & kTz)) & kTz) | (((a & kTy) - (b & kTy)) & kTy) |
         (((a & kTx) - (b & kTx)) & kTx);
}

__device__ __forceinline__ uint64_t primal_vertex(uint64_t key, int lv, unsigned which) {
  const uint64_t k = 1ull << (3 * lv);  // morton.rs:120-138
  const uint64_t k_plus_one = k << 1;
  const uint64_t vk = m_add(key, which);
  const uint64_t dk = m_sub(vk, k);
  if (vk >= k_plus_one || (dk & kTx) == 0 || (dk & kTy) == 0 || (dk & kTz) == 0) return 0;
  return vk << (3 * (kMaxLevel - lv));
}

__device__ __forceinline__ uint64_t dual_vertex(uint64_t key, int lv, unsigned which) {
  return m_sub(key >> (3 * (kMaxLevel - lv)), which);  // morton.rs:142-146
}

// Index of node `m` in the sorted key array, or -1.
__device__ __forceinline__ int64_t find_node(const LhmcNodes& N, uint64_t m) {
  const int d = depth_of(m);
  if (d >= N.n_levels) return -1;
  int64_t lo = N.level_off[d], hi = N.level_off[d + 1];
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (N.keys[mid] < m) lo = mid + 1;
    else hi = mid;
  }
  return (lo < N.level_off[d + 1] && N.keys[lo] == m) ? lo : -1;
}

__device__ __forceinline__ uint64_t hash_edge(unsigned long long k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

// Slot of edge key `k`, inserting it if absent; ~0 if the table is full.
__device__ __forceinline__ uint64_t table_insert(const LhmcEdgeTable& T, unsigned long long k) {
  uint64_t h = hash_edge(k) & T.mask;
  for (uint64_t probe = 0; probe <= T.mask; ++probe) {
    const unsigned long long cur = atomicCAS(&T.keys[h], kEmpty, k);
    if (cur == kEmpty || cur == k) return h;
    h = (h + 1) & T.mask;
  }
  return ~0ull;
}

__device__ __forceinline__ uint64_t table_find(const LhmcEdgeTable& T, unsigned long long k) {
  uint64_t h = hash_edge(k) & T.mask;
  for (uint64_t probe = 0; probe <= T.mask; ++probe) {
    const unsigned long long cur = T.keys[h];
    if (cur == k) return h;
    if (cur == kEmpty) return ~0ull;
    h = (h + 1) & T.mask;
  }
  return ~0ull;
}

__device__ __forceinline__ unsigned long long edge_key(const uint32_t* corner, int e) {
  uint32_t a = corner[tables::kEdgeCorners[e][0]], b = corner[tables::kEdgeCorners[e][1]];
  if (a > b) {
    const uint32_t t = a;
    a = b;
    b = t;
  }
  return (static_cast<unsigned long long>(a) << 32) | b;
}

// ---- 1. one octree depth ----------------------------------------------------
struct LevelLoad {
  const DevOp* ops;
  int n_ops;
  int max_depth;
  const uint64_t* keys;
  float* values;
  uint8_t* leaf;
  int64_t off;
  __device__ unsigned long long operator()(uint64_t i) const {
    const uint64_t key = keys[off + i];
    const F3 c = center_of(key);
    const float v = sdf_eval(ops, n_ops, c.x, c.y, c.z);
    const int lv = level_of(key);
    const float sz = size_of(key);
    // Signed::within_extent (distance.rs:60-62)
    const bool refine = lv < 2 || (lv < max_depth && __builtin_fabsf(v) < sz * kSqrtOf3);
    values[off + i] = v;
    leaf[off + i] = refine ? 0 : 1;
    return refine ? 1ull : 0ull;
  }
};

struct LevelSink {
  uint64_t* keys;
  int64_t off, n, next_capacity;
  __device__ void operator()(uint64_t i, unsigned long long v, unsigned long long ex) const {
    if (!v || static_cast<int64_t>(8 * ex + 8) > next_capacity) return;
    const uint64_t key = keys[off + i];
    uint64_t* dst = keys + off + n + 8 * ex;
#pragma unroll
    for (unsigned c = 0; c < 8; ++c) dst[c] = (key << 3) | c;
  }
};

// ---- 2. primal vertices -----------------------------------------------------
__global__ __launch_bounds__(256) void k_primal_mark(LhmcNodes N, int64_t n_nodes, int d1,
                                                     uint32_t* __restrict__ dense) {
  const uint64_t lmask = (1ull << (3 * d1)) - 1;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n_nodes;
       i += stride) {
    if (!N.leaf[i]) continue;
    const uint64_t key = N.keys[i];
    const int lv = level_of(key);
    for (unsigned w = 0; w < 8; ++w) {
      const uint64_t v = primal_vertex(key, lv, w);
      if (v != 0) atomicMax(&dense[(v >> (3 * (kMaxLevel - d1))) & lmask], static_cast<uint32_t>(lv + 1));
    }
  }
}

struct PrimalLoad {
  const uint32_t* dense;
  __device__ unsigned long long operator()(uint64_t i) const { return dense[i] ? 1ull : 0ull; }
};

struct PrimalSink {
  const uint32_t* dense;
  uint64_t* primal;
  uint8_t* level;
  uint64_t capacity;
  int d1;
  __device__ void operator()(uint64_t i, unsigned long long v, unsigned long long ex) const {
    if (!v || ex >= capacity) return;
    // the lattice index back to the vertex's Morton key at MAX_LEVEL
    primal[ex] = (i | (1ull << (3 * d1))) << (3 * (kMaxLevel - d1));
    level[ex] = static_cast<uint8_t>(dense[i] - 1);
  }
};

// Sorted path: every node's eight primal vertices as (key, level) pairs;
// a node that is not a leaf, or a vertex outside the domain, writes key 0
// (every primal key has bit 63 set, so 0 sorts first and marks none).
__global__ __launch_bounds__(256) void k_primal_pairs(LhmcNodes N, int64_t n_nodes,
                                                      uint64_t* __restrict__ keys,
                                                      uint8_t* __restrict__ levels) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n_nodes;
       i += stride) {
    const bool leaf = N.leaf[i] != 0;
    const uint64_t key = N.keys[i];
    const int lv = level_of(key);
    for (unsigned w = 0; w < 8; ++w) {
      const uint64_t v = leaf ? primal_vertex(key, lv, w) : 0ull;
      keys[8 * i + w] = v;
      levels[8 * i + w] = static_cast<uint8_t>(lv);
    }
  }
}

// The last pair of each run of equal nonzero keys (a vertex is shared by at
// most eight leaves), listed with the run's deepest level.
struct RunLoad {
  const uint64_t* keys;
  uint64_t n;
  __device__ unsigned long long operator()(uint64_t i) const {
    const uint64_t k = keys[i];
    return (k != 0 && (i + 1 == n || keys[i + 1] != k)) ? 1ull : 0ull;
  }
};

struct RunSink {
  const uint64_t* keys;
  const uint8_t* levels;
  uint64_t* primal;
  uint8_t* level;
  uint64_t capacity;
  __device__ void operator()(uint64_t i, unsigned long long v, unsigned long long ex) const {
    if (!v || ex >= capacity) return;
    const uint64_t k = keys[i];
    uint8_t lv = levels[i];
    for (uint64_t j = i; j > 0 && keys[j - 1] == k; --j)
      if (levels[j - 1] > lv) lv = levels[j - 1];
    primal[ex] = k;
    level[ex] = lv;
  }
};

// ---- 3. dual cubes ----------------------------------------------------------
struct CubeLoad {
  LhmcNodes N;
  LhmcCubes C;
  int d1;
  unsigned long long* err;
  __device__ unsigned long long operator()(uint64_t p) const {
    const uint64_t key = C.primal[p];
    const int lv = C.primal_level[p];
    uint32_t cube = 0;
    for (int i = 0; i < 8; ++i) {
      uint64_t m = dual_vertex(key, lv, static_cast<unsigned>(c_remap[i]));
      int64_t idx = -1;
      while (m > 1) {  // implicit_octree.rs:87-96
        idx = find_node(N, m);
        if (idx >= 0) break;
        m = m >> 3 > 1 ? m >> 3 : 1;  // Morton::parent
      }
      if (idx < 0) {
        atomicOr(err, kLhmcErrWalk);
        idx = 0;
      }
      C.corner[8 * p + i] = static_cast<uint32_t>(idx);
      if (!(N.values[idx] > 0.0f)) cube |= 1u << i;  // classify_corners
    }
    C.cube_case[p] = static_cast<uint8_t>(cube);
    return c_lt.tri_count[cube];
  }
};

struct NoSink {
  __device__ void operator()(uint64_t, unsigned long long, unsigned long long) const {}
};

// ---- 4. ownership -----------------------------------------------------------
__global__ __launch_bounds__(256) void k_owner(LhmcCubes C, LhmcEdgeTable T,
                                               unsigned long long* err) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < C.n; p += stride) {
    const uint32_t cube = C.cube_case[p];
    const int nslot = 3 * c_lt.tri_count[cube];
    const uint32_t* corner = C.corner + 8 * static_cast<uint64_t>(p);
    for (int j = 0; j < nslot; ++j) {
      const uint64_t h = table_insert(T, edge_key(corner, c_lt.tri[cube][j]));
      if (h == ~0ull) {
        atomicOr(err, kLhmcErrTable);
        continue;
      }
      atomicMin(&T.owner[h], p);
    }
  }
}

struct OwnLoad {
  LhmcCubes C;
  LhmcEdgeTable T;
  __device__ unsigned long long operator()(uint64_t p) const {
    const uint32_t cube = C.cube_case[p];
    const int nslot = 3 * c_lt.tri_count[cube];
    const uint32_t* corner = C.corner + 8 * p;
    unsigned long long seen[12];
    int n_own = 0;
    for (int j = 0; j < nslot; ++j) {
      const unsigned long long k = edge_key(corner, c_lt.tri[cube][j]);
      const uint64_t h = table_find(T, k);
      if (h == ~0ull || T.owner[h] != static_cast<uint32_t>(p)) continue;
      bool dup = false;
      for (int q = 0; q < n_own; ++q) dup |= seen[q] == k;
      if (!dup && n_own < 12) seen[n_own++] = k;
    }
    return static_cast<unsigned long long>(n_own) |
           (static_cast<unsigned long long>(c_lt.tri_count[cube]) << 32);
  }
};

struct OwnSink {
  unsigned long long* base;
  __device__ void operator()(uint64_t p, unsigned long long, unsigned long long ex) const {
    base[p] = ex;
  }
};

// ---- 5. emit ----------------------------------------------------------------
__global__ __launch_bounds__(256) void k_vertices(LhmcNodes N, LhmcCubes C, LhmcEdgeTable T,
                                                  float* __restrict__ vertices, uint64_t vcap) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < C.n; p += stride) {
    const uint32_t cube = C.cube_case[p];
    const int nslot = 3 * c_lt.tri_count[cube];
    if (!nslot) continue;
    const uint32_t* corner = C.corner + 8 * static_cast<uint64_t>(p);
    uint32_t next = static_cast<uint32_t>(C.base[p] & 0xffffffffull);
    unsigned long long seen[12];
    int n_own = 0;
    for (int j = 0; j < nslot; ++j) {
      const int e = c_lt.tri[cube][j];
      const unsigned long long k = edge_key(corner, e);
      const uint64_t h = table_find(T, k);
      if (h == ~0ull || T.owner[h] != p) continue;
      bool dup = false;
      for (int q = 0; q < n_own; ++q) dup |= seen[q] == k;
      if (dup || n_own >= 12) continue;
      seen[n_own++] = k;
      const uint32_t id = next++;
      T.id[h] = id;
      // find_edge_crossings (marching_cubes_impl.rs:39-56) for local edge e
      const uint32_t ia = corner[tables::kEdgeCorners[e][0]];
      const uint32_t ib = corner[tables::kEdgeCorners[e][1]];
      const float a = N.values[ia], b = N.values[ib];
      const F3 pa = center_of(N.keys[ia]), pb = center_of(N.keys[ib]);
      const float delta = b - a;
      const float t = (delta == 0.0f) ? 0.5f : -a / delta;
      const float s = 1.0f - t;
      if (id < vcap) {
        vertices[3 * static_cast<uint64_t>(id) + 0] = pa.x * s + pb.x * t;
        vertices[3 * static_cast<uint64_t>(id) + 1] = pa.y * s + pb.y * t;
        vertices[3 * static_cast<uint64_t>(id) + 2] = pa.z * s + pb.z * t;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_indices(LhmcCubes C, LhmcEdgeTable T,
                                                 uint32_t* __restrict__ indices, uint64_t icap,
                                                 unsigned long long* err) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < C.n; p += stride) {
    const uint32_t cube = C.cube_case[p];
    const int nslot = 3 * c_lt.tri_count[cube];
    const uint32_t* corner = C.corner + 8 * static_cast<uint64_t>(p);
    const uint64_t tb = 3 * (C.base[p] >> 32);
    for (int j = 0; j < nslot; ++j) {
      const uint64_t h = table_find(T, edge_key(corner, c_lt.tri[cube][j]));
      if (h == ~0ull) {
        atomicOr(err, kLhmcErrTable);
        continue;
      }
      if (tb + j < icap) indices[tb + j] = T.id[h];
    }
  }
}

__global__ __launch_bounds__(256) void k_table_clear(LhmcEdgeTable T) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i <= T.mask;
       i += stride) {
    T.keys[i] = kEmpty;
    T.owner[i] = 0xffffffffu;
  }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride)
    p[i] = v;
}

unsigned grid_of(uint64_t n, uint64_t cap = 16384) {
  uint64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return static_cast<unsigned>(b);
}

}  // namespace

hipError_t launch_lhmc_level(const DevOp* ops, int n_ops, int max_depth, uint64_t* keys,
                             float* values, uint8_t* leaf, int64_t off, int64_t n,
                             int64_t next_capacity, const scan::State& st, hipStream_t s) {
  LevelLoad load{ops, n_ops, max_depth, keys, values, leaf, off};
  LevelSink sink{keys, off, n, next_capacity};
  return scan::launch<1>(static_cast<uint64_t>(n), load, sink, st, s);
}

hipError_t launch_lhmc_primal(const LhmcNodes& N, int dense_depth, uint32_t* dense,
                              uint64_t n_dense, const LhmcCubes& C, uint64_t capacity,
                              const scan::State& st, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_of(n_dense)), dim3(256), 0, s, dense, n_dense, 0u);
  const int64_t n_nodes = N.level_off[N.n_levels];
  hipLaunchKernelGGL(k_primal_mark, dim3(grid_of(static_cast<uint64_t>(n_nodes))), dim3(256), 0, s,
                     N, n_nodes, dense_depth, dense);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  PrimalLoad load{dense};
  PrimalSink sink{dense, C.primal, C.primal_level, capacity, dense_depth};
  return scan::launch<scan::kItems>(n_dense, load, sink, st, s);
}

size_t lhmc_sort_temp_bytes(uint64_t n_pairs) {
  size_t bytes = 0;
  if (rocprim::radix_sort_pairs(nullptr, bytes, static_cast<const uint64_t*>(nullptr),
                                static_cast<uint64_t*>(nullptr),
                                static_cast<const uint8_t*>(nullptr),
                                static_cast<uint8_t*>(nullptr), static_cast<size_t>(n_pairs), 0,
                                64) != hipSuccess)
    return 0;
  return bytes;
}

hipError_t launch_lhmc_primal_sorted(const LhmcNodes& N, uint64_t* pair_keys,
                                     uint8_t* pair_levels, uint64_t* sorted_keys,
                                     uint8_t* sorted_levels, void* temp, size_t temp_bytes,
                                     const LhmcCubes& C, uint64_t capacity,
                                     const scan::State& st, hipStream_t s) {
  const int64_t n_nodes = N.level_off[N.n_levels];
  const uint64_t n_pairs = 8 * static_cast<uint64_t>(n_nodes);
  hipLaunchKernelGGL(k_primal_pairs, dim3(grid_of(static_cast<uint64_t>(n_nodes))), dim3(256), 0,
                     s, N, n_nodes, pair_keys, pair_levels);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  size_t bytes = temp_bytes;
  e = rocprim::radix_sort_pairs(temp, bytes, static_cast<const uint64_t*>(pair_keys), sorted_keys,
                                static_cast<const uint8_t*>(pair_levels), sorted_levels,
                                static_cast<size_t>(n_pairs), 0, 64, s);
  if (e != hipSuccess) return e;
  RunLoad load{sorted_keys, n_pairs};
  RunSink sink{sorted_keys, sorted_levels, C.primal, C.primal_level, capacity};
  return scan::launch<scan::kItems>(n_pairs, load, sink, st, s);
}

void lhmc_level_splits(uint64_t split[kLhmcMaxLevels]) {
  // Morton::level, morton.rs:53-59 (the same libm log, f64)
  auto level = [](uint64_t a) {
    return a == 0 ? 0 : static_cast<int>(std::floor(std::log(static_cast<double>(a)) * 0.48089834696));
  };
  split[0] = 0;
  for (int d = 1; d < kLhmcMaxLevels; ++d) {
    const uint64_t base = 1ull << (3 * d);
    // the first offset k at which level(base + k) reaches d (level is
    // monotone in the key, and d at the depth's last key, offset hi)
    uint64_t lo = 0, hi = (d == 21) ? ~0ull - base : 7 * base - 1;
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if (level(base + mid) >= d) hi = mid;
      else lo = mid + 1;
    }
    split[d] = lo;
  }
}

hipError_t launch_lhmc_levels(const uint64_t split[kLhmcMaxLevels], hipStream_t s) {
  return hipMemcpyToSymbolAsync(HIP_SYMBOL(c_level_split), split,
                                kLhmcMaxLevels * sizeof(uint64_t), 0, hipMemcpyHostToDevice, s);
}

hipError_t launch_lhmc_cubes(const LhmcNodes& N, int dense_depth, const LhmcCubes& C,
                             const scan::State& st, hipStream_t s) {
  CubeLoad load{N, C, dense_depth, st.err};
  return scan::launch<1>(C.n, load, NoSink{}, st, s);
}

hipError_t launch_lhmc_own(const LhmcNodes& N, const LhmcCubes& C, const LhmcEdgeTable& T,
                           unsigned long long* err, const scan::State& st, hipStream_t s) {
  (void)N;
  hipLaunchKernelGGL(k_owner, dim3(grid_of(C.n)), dim3(256), 0, s, C, T, err);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  OwnLoad load{C, T};
  OwnSink sink{C.base};
  return scan::launch<1>(C.n, load, sink, st, s);
}

hipError_t launch_lhmc_emit(const LhmcNodes& N, const LhmcCubes& C, const LhmcEdgeTable& T,
                            float* vertices, uint64_t vcap, uint32_t* indices, uint64_t icap,
                            unsigned long long* err, hipStream_t s) {
  hipLaunchKernelGGL(k_vertices, dim3(grid_of(C.n)), dim3(256), 0, s, N, C, T, vertices, vcap);
  hipLaunchKernelGGL(k_indices, dim3(grid_of(C.n)), dim3(256), 0, s, C, T, indices, icap, err);
  return hipGetLastError();
}

hipError_t launch_lhmc_table_clear(const LhmcEdgeTable& T, hipStream_t s) {
  hipLaunchKernelGGL(k_table_clear, dim3(grid_of(T.mask + 1)), dim3(256), 0, s, T);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_of(n)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

}  // namespace iso
