// Marching-cubes kernels for gfx950 (MI355X).
//
// One extraction is three kernels:
//   1. signs  one bit per lattice sample, "inside" = !(v > 0) exactly as
//             classify_corners (marching_cubes_impl.rs:26-37, distance.rs:52-54).
//             From a dense FP32 grid this is a pure HBM stream (16-byte
//             loads, wave64 ballots); from a program it is the SDF
//             evaluation.  The launch also resets the extraction's state.
//   2. count  one thread per 64-cell segment of a cell row: the segment's
//             eight corner bit-planes come from four row words, vertex
//             ownership counts are popcounts of XORed planes, triangles are
//             counted over the mixed cells only.  A single-pass decoupled
//             look-back scan turns the counts into global first ids and
//             writes the list of active (mixed) cells in cell order.
//   3. emit   one thread per active cell: its 3x3x3 sign neighbourhood holds
//             every cell that can own one of its edges; each needed owner's
//             first vertex id is resolved once, the cell's own vertices are
//             written (FP32 crossing points from the grid or the program)
//             and its index triples are written.
//
// Vertex numbering is the reference's first-encounter order
// (MeshTopologyBuilder::add_vertex, mesh.rs:240-251, driven by
// PrimalGrid::traverse's z-major cell order, primal_grid.rs:59-85): a
// lattice edge's vertex belongs to the lowest-numbered cell containing the
// edge, its position is computed along that cell's local edge direction, and
// vertices are numbered by (owner cell, first appearance in the owner's
// triangle list).  The owner of every edge follows from the edge's position
// alone, so no hash table is needed.
#include <hip/hip_runtime.h>

#include "mc_kernels.h"
#include "mc_tables.h"
#include "sdf_eval.h"

namespace iso {

namespace {

// ---------------------------------------------------------------------------
// Tables
// ---------------------------------------------------------------------------
struct EdgeGeom {
  int8_t axis[12];
  int8_t minp[12][3];        // lattice offset of the edge's lower endpoint
  int8_t of_axis_min[3][8];  // [axis][x | y<<1 | z<<2 of lower endpoint] -> edge
};

constexpr EdgeGeom make_geom() {
  EdgeGeom g{};
  for (int a = 0; a < 3; ++a)
    for (int r = 0; r < 8; ++r) g.of_axis_min[a][r] = -1;
  for (int e = 0; e < 12; ++e) {
    const int* u = tables::kCorner[tables::kEdgeCorners[e][0]];
    const int* v = tables::kCorner[tables::kEdgeCorners[e][1]];
    int axis = 0;
    for (int d = 0; d < 3; ++d) {
      if (u[d] != v[d]) axis = d;
      g.minp[e][d] = static_cast<int8_t>(u[d] < v[d] ? u[d] : v[d]);
    }
    g.axis[e] = static_cast<int8_t>(axis);
    g.of_axis_min[axis][g.minp[e][0] | (g.minp[e][1] << 1) | (g.minp[e][2] << 2)] =
        static_cast<int8_t>(e);
  }
  return g;
}

__constant__ tables::McTables c_tables = tables::kTables;

// Per edge, packed for register use: bits 0-2 = lower endpoint (x|y<<1|z<<2),
// bits 3-4 = axis.
constexpr uint64_t pack_edge_geom() {
  EdgeGeom g = make_geom();
  uint64_t r = 0;
  for (int e = 0; e < 12; ++e) {
    uint64_t v = static_cast<uint64_t>(g.minp[e][0] | (g.minp[e][1] << 1) | (g.minp[e][2] << 2) |
                                       (g.axis[e] << 3));
    r |= v << (5 * e);
  }
  return r;
}
constexpr uint64_t kEdgeGeomPacked = pack_edge_geom();

// Edge of a given axis whose lower endpoint sits at corner offset r (3 bits),
// four bits per entry [axis * 8 + r] over two words.
constexpr uint64_t pack_of_axis_min(int half) {
  EdgeGeom g = make_geom();
  uint64_t r = 0;
  for (int i = 0; i < 16; ++i) {
    const int k = 16 * half + i;
    if (k >= 24) break;
    const int e = g.of_axis_min[k / 8][k % 8];
    r |= static_cast<uint64_t>(e < 0 ? 15 : e) << (4 * i);
  }
  return r;
}
constexpr uint64_t kOfAxisMin0 = pack_of_axis_min(0);
constexpr uint64_t kOfAxisMin1 = pack_of_axis_min(1);

__device__ __forceinline__ int edge_of_axis_min(int axis, int rel) {
  const int k = 8 * axis + rel;
  const uint64_t w = k < 16 ? kOfAxisMin0 : kOfAxisMin1;
  return static_cast<int>((w >> (4 * (k & 15))) & 15u);
}

__device__ __forceinline__ void load_tables(tables::McTables* dst) {
  const uint32_t* src = reinterpret_cast<const uint32_t*>(&c_tables);
  uint32_t* d = reinterpret_cast<uint32_t*>(dst);
  constexpr int kWords = sizeof(tables::McTables) / 4;
  static_assert(sizeof(tables::McTables) % 4 == 0, "table size");
  for (int i = threadIdx.x; i < kWords; i += blockDim.x) d[i] = src[i];
}

// ---------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t popc64(uint64_t v) {
  return static_cast<uint32_t>(__popcll(v));
}

__device__ __forceinline__ uint64_t low_mask(int b) { return (1ull << b) - 1ull; }

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = __lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

// Signed::find_crossing_point (distance.rs:64-70): t = d == 0 ? 0.5 : -a/d,
// p = pa*(1-t) + pb*t per component, every operation rounded once.
__device__ __forceinline__ void crossing(float a, float b, float pax, float pay, float paz,
                                         float pbx, float pby, float pbz, float* o) {
  const float delta = b - a;
  const float t = (delta == 0.0f) ? 0.5f : -a / delta;
  const float s = 1.0f - t;
  o[0] = pax * s + pbx * t;
  o[1] = pay * s + pby * t;
  o[2] = paz * s + pbz * t;
}

// Local edges whose vertex the cell (x, y, z) owns, i.e. for which it is the
// first cell in z-major order that contains the edge: e5, e6, e10 always,
// and the low-face edges on the lattice's low faces.
__device__ __forceinline__ uint32_t owned_edges(int64_t x, int64_t y, int64_t z) {
  uint32_t m = (1u << 5) | (1u << 6) | (1u << 10);
  if (y == 0) m |= (1u << 4) | (1u << 9);
  if (z == 0) m |= (1u << 2) | (1u << 1);
  if (x == 0) m |= (1u << 7) | (1u << 11);
  if (y == 0 && z == 0) m |= 1u << 0;
  if (x == 0 && z == 0) m |= 1u << 3;
  if (x == 0 && y == 0) m |= 1u << 8;
  return m;
}

// Row-aligned 64 sign bits starting at flat sample f, and the low bit of the
// next word (sample f + 64) in *next.
__device__ __forceinline__ uint64_t row_word(const uint64_t* __restrict__ S, int64_t f,
                                             uint64_t* next) {
  const int64_t w0 = f >> 6;
  const int s = static_cast<int>(f & 63);
  const uint64_t a = S[w0], b = S[w0 + 1];
  *next = (b >> s) & 1ull;
  return s ? ((a >> s) | (b << (64 - s))) : a;
}

// ---------------------------------------------------------------------------
// Segment words: the eight corner bit-planes of 64 consecutive cells of a
// cell row (bit b of c[i] = "inside" for corner i of cell x0 + b), so every
// per-cell quantity below is computed for 64 cells per 64-bit operation.
// ---------------------------------------------------------------------------
struct SegWords {
  uint64_t c[8];
};

// From the row words of the four point rows around a cell row
// (00=(y,z) 10=(y+1,z) 01=(y,z+1) 11=(y+1,z+1)) and their next bits.
__device__ __forceinline__ SegWords seg_words(uint64_t w00, uint64_t n00, uint64_t w10,
                                              uint64_t n10, uint64_t w01, uint64_t n01,
                                              uint64_t w11, uint64_t n11) {
  SegWords s;
  s.c[0] = w00;
  s.c[1] = (w00 >> 1) | (n00 << 63);
  s.c[2] = (w10 >> 1) | (n10 << 63);
  s.c[3] = w10;
  s.c[4] = w01;
  s.c[5] = (w01 >> 1) | (n01 << 63);
  s.c[6] = (w11 >> 1) | (n11 << 63);
  s.c[7] = w11;
  return s;
}

__device__ __forceinline__ uint32_t case_of(const SegWords& s, int b) {
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) c |= static_cast<uint32_t>((s.c[i] >> b) & 1ull) << i;
  return c;
}

// Cells with a surface crossing (cube index not 0 or 255).
__device__ __forceinline__ uint64_t mixed_cells(const SegWords& s) {
  uint64_t any = 0, all = ~0ull;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    any |= s.c[i];
    all &= s.c[i];
  }
  return any & ~all;
}

// Valid cells (x < nx - 1) of segment w.
__device__ __forceinline__ uint64_t seg_valid(int64_t nx, int64_t w) {
  const int64_t n = nx - 1 - 64 * w;
  return n >= 64 ? ~0ull : (n <= 0 ? 0ull : ((1ull << n) - 1ull));
}

// Vertices owned by the cells selected by `m` (a subset of one segment): a
// crossed edge is the XOR of its endpoints' planes.  `first` marks the
// segment holding cell x = 0.
__device__ __forceinline__ uint32_t owned_count(const SegWords& s, uint64_t m, bool y0, bool z0,
                                                bool first) {
  const uint64_t* c = s.c;
  uint32_t n = popc64(m & (c[6] ^ c[7])) + popc64(m & (c[5] ^ c[6])) + popc64(m & (c[2] ^ c[6]));
  if (y0) n += popc64(m & (c[4] ^ c[5])) + popc64(m & (c[1] ^ c[5]));
  if (z0) n += popc64(m & (c[2] ^ c[3])) + popc64(m & (c[1] ^ c[2]));
  if (y0 && z0) n += popc64(m & (c[0] ^ c[1]));
  if (first && (m & 1ull)) {
    n += static_cast<uint32_t>(((c[4] ^ c[7]) & 1ull) + ((c[3] ^ c[7]) & 1ull));
    if (z0) n += static_cast<uint32_t>((c[0] ^ c[3]) & 1ull);
    if (y0) n += static_cast<uint32_t>((c[0] ^ c[4]) & 1ull);
  }
  return n;
}

// ---------------------------------------------------------------------------
// 1. signs (+ reset of the extraction's scratch state)
// ---------------------------------------------------------------------------
struct ResetArgs {
  uint64_t* signs;  // guard words at [-1], [n_words], [n_words + 1]
  int64_t n_words;
  uint32_t* flags;
  int64_t n_blocks;
  unsigned long long* totals;
};

__device__ __forceinline__ void reset_state(const ResetArgs& r) {
  const int64_t i0 = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = i0; i < r.n_blocks; i += stride) r.flags[i] = 0u;
  if (i0 < kTotSlots) r.totals[i0] = 0ull;
  if (i0 == 0) {
    r.signs[-1] = 0ull;
    r.signs[r.n_words] = 0ull;
    r.signs[r.n_words + 1] = 0ull;
  }
}

constexpr int kSignVec = 4;  // float4 loads per lane per iteration (16 words)

typedef float f32x4 __attribute__((ext_vector_type(4)));

// Moves bit i of a 16-bit value to bit 4i.
__device__ __forceinline__ uint64_t spread4(uint64_t x) {
  x &= 0xFFFFull;
  x = (x | (x << 24)) & 0x000000FF000000FFull;
  x = (x | (x << 12)) & 0x000F000F000F000Full;
  x = (x | (x << 6)) & 0x0303030303030303ull;
  x = (x | (x << 3)) & 0x1111111111111111ull;
  return x;
}

// 16-byte loads: lane l of a wavefront reads samples 4l..4l+3 of a
// 256-sample chunk, so ballot j holds component j of every lane and word k
// of the chunk interleaves the 16-lane slices [16k, 16k+16) of the ballots.
__global__ __launch_bounds__(256) void k_signs_grid4(const f32x4* __restrict__ g4,
                                                     int64_t n_points, int64_t n_words,
                                                     uint64_t* __restrict__ S, ResetArgs R) {
  reset_state(R);
  const int lane = __lane_id();
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int64_t n_waves = static_cast<int64_t>(gridDim.x) * 4;
  const float* g = reinterpret_cast<const float*>(g4);
  constexpr int64_t kSpan = 256 * kSignVec;  // samples per wavefront iteration
  for (int64_t base = wave * kSpan; base < n_points; base += n_waves * kSpan) {
    f32x4 v[kSignVec];
#pragma unroll
    for (int j = 0; j < kSignVec; ++j) {
      const int64_t f = base + 256 * j + 4 * lane;
      if (f + 3 < n_points) {
        v[j] = __builtin_nontemporal_load(g4 + (f >> 2));
      } else {
        v[j].x = f + 0 < n_points ? g[f + 0] : 1.0f;
        v[j].y = f + 1 < n_points ? g[f + 1] : 1.0f;
        v[j].z = f + 2 < n_points ? g[f + 2] : 1.0f;
        v[j].w = 1.0f;
      }
    }
    uint64_t mine = 0;
#pragma unroll
    for (int j = 0; j < kSignVec; ++j) {
      const uint64_t b0 = __ballot(!(v[j].x > 0.0f));
      const uint64_t b1 = __ballot(!(v[j].y > 0.0f));
      const uint64_t b2 = __ballot(!(v[j].z > 0.0f));
      const uint64_t b3 = __ballot(!(v[j].w > 0.0f));
      if ((lane >> 2) == j) {
        const int sh = 16 * (lane & 3);
        mine = spread4(b0 >> sh) | (spread4(b1 >> sh) << 1) | (spread4(b2 >> sh) << 2) |
               (spread4(b3 >> sh) << 3);
      }
    }
    const int64_t w = (base >> 6) + lane;
    if (lane < 4 * kSignVec && w < n_words) S[w] = mine;
  }
}

// Grids that are not 16-byte aligned: one float per lane per ballot.
constexpr int kSignUnroll = 8;

__global__ __launch_bounds__(256) void k_signs_grid(const float* __restrict__ g,
                                                    int64_t n_points, int64_t n_words,
                                                    uint64_t* __restrict__ S, ResetArgs R) {
  reset_state(R);
  const int lane = __lane_id();
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int64_t n_waves = static_cast<int64_t>(gridDim.x) * 4;
  for (int64_t w0 = wave * kSignUnroll; w0 < n_words; w0 += n_waves * kSignUnroll) {
    float v[kSignUnroll];
#pragma unroll
    for (int j = 0; j < kSignUnroll; ++j) {
      const int64_t f = (w0 + j) * 64 + lane;
      v[j] = f < n_points ? g[f] : 1.0f;
    }
    uint64_t mine = 0;
#pragma unroll
    for (int j = 0; j < kSignUnroll; ++j) {
      const uint64_t m = __ballot(!(v[j] > 0.0f));
      if (lane == j) mine = m;
    }
    if (lane < kSignUnroll && w0 + lane < n_words) S[w0 + lane] = mine;
  }
}

__global__ __launch_bounds__(256) void k_signs_program(const DevOp* __restrict__ ops, int n_ops,
                                                       Lattice L, int64_t n_points,
                                                       int64_t n_words,
                                                       uint64_t* __restrict__ S, ResetArgs R) {
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
      const float v = sdf_eval(ops, n_ops, lattice_coord(L, 0, x), lattice_coord(L, 1, y),
                               lattice_coord(L, 2, z));
      inside = !(v > 0.0f);
    }
    const uint64_t m = __ballot(inside);
    if (lane == 0) S[w] = m;
  }
}

// ---------------------------------------------------------------------------
// 2. count + single-pass scan (decoupled look-back)
// ---------------------------------------------------------------------------
// A predecessor publishes its aggregate right after its own counting, so the
// wait is short; the bound only guarantees that a broken invariant ends the
// kernel (with an error flag) instead of spinning forever.
constexpr uint32_t kLookBackSpinLimit = 1u << 22;

__device__ __forceinline__ uint32_t load_acquire(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void store_release(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kCountBlock) void k_count(const uint64_t* __restrict__ S, Lattice L,
                                                       McDims D, McBuffers B) {
  __shared__ uint8_t s_ntri[256];
  __shared__ uint32_t s_wv[4], s_wt[4], s_wa[4];
  __shared__ uint32_t s_block;
  __shared__ ScanTriple s_prefix;
  for (int i = threadIdx.x; i < 256; i += blockDim.x) s_ntri[i] = c_tables.tri_count[i];
  if (threadIdx.x == 0)
    s_block = atomicAdd(reinterpret_cast<unsigned int*>(&B.totals[kTotTicket]), 1u);
  __syncthreads();
  // Workgroups scan in ticket order, so every predecessor is already running.
  const uint32_t blk = s_block;
  const int64_t g = static_cast<int64_t>(blk) * kCountBlock + threadIdx.x;

  uint32_t nv = 0, nt = 0, na = 0;
  uint64_t act = 0;
  if (g < D.n_segs) {
    const int64_t nx = L.nx, ny = L.ny;
    const uint32_t spr = static_cast<uint32_t>(D.segs_per_row);
    const uint32_t cyn = static_cast<uint32_t>(ny - 1);
    const uint32_t gg = static_cast<uint32_t>(g);
    const uint32_t row = gg / spr;
    const uint32_t w = gg - row * spr;
    const uint32_t cz = row / cyn;
    const uint32_t cy = row - cz * cyn;
    const int64_t f00 = (static_cast<int64_t>(cz) * ny + cy) * nx + 64 * static_cast<int64_t>(w);
    uint64_t n00, n10, n01, n11;
    const uint64_t w00 = row_word(S, f00, &n00);
    const uint64_t w10 = row_word(S, f00 + nx, &n10);
    const uint64_t w01 = row_word(S, f00 + nx * ny, &n01);
    const uint64_t w11 = row_word(S, f00 + nx * ny + nx, &n11);
    const SegWords sw = seg_words(w00, n00, w10, n10, w01, n01, w11, n11);
    const uint64_t valid = seg_valid(nx, w);
    nv = owned_count(sw, valid, cy == 0, cz == 0, w == 0);
    act = mixed_cells(sw) & valid;
    na = popc64(act);
    uint64_t p0 = 0, p1 = 0, p2 = 0;  // per-cell triangle counts, bit-sliced
    for (uint64_t m = act; m; m &= m - 1) {
      const int b = __builtin_ctzll(m);
      const uint32_t c = s_ntri[case_of(sw, b)];
      nt += c;
      p0 |= static_cast<uint64_t>(c & 1u) << b;
      p1 |= static_cast<uint64_t>((c >> 1) & 1u) << b;
      p2 |= static_cast<uint64_t>((c >> 2) & 1u) << b;
    }
    B.tplanes[3 * g + 0] = p0;
    B.tplanes[3 * g + 1] = p1;
    B.tplanes[3 * g + 2] = p2;
  }

  // Workgroup scan of (vertices, triangles, active cells).
  const uint32_t iv = wave_incl_scan(nv), it = wave_incl_scan(nt), ia = wave_incl_scan(na);
  const int wv = threadIdx.x >> 6;
  if (__lane_id() == 63) {
    s_wv[wv] = iv;
    s_wt[wv] = it;
    s_wa[wv] = ia;
  }
  __syncthreads();
  uint32_t ov = iv - nv, ot = it - nt, oa = ia - na;
  for (int i = 0; i < wv; ++i) {
    ov += s_wv[i];
    ot += s_wt[i];
    oa += s_wa[i];
  }

  if (threadIdx.x == 0) {
    ScanTriple total;
    total.v = static_cast<unsigned long long>(s_wv[0]) + s_wv[1] + s_wv[2] + s_wv[3];
    total.t = static_cast<unsigned long long>(s_wt[0]) + s_wt[1] + s_wt[2] + s_wt[3];
    total.a = static_cast<unsigned long long>(s_wa[0]) + s_wa[1] + s_wa[2] + s_wa[3];
    ScanTriple prefix{0ull, 0ull, 0ull};
    if (blk == 0) {
      B.incl[0] = total;
      store_release(&B.flags[0], 2u);
    } else {
      B.agg[blk] = total;
      store_release(&B.flags[blk], 1u);
      int64_t j = static_cast<int64_t>(blk) - 1;
      uint32_t spins = 0;
      while (j >= 0) {
        const uint32_t f = load_acquire(&B.flags[j]);
        if (f == 0u) {
          if (++spins > kLookBackSpinLimit) {
            atomicOr(&B.totals[kTotErrors], 2ull);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        const ScanTriple x = (f == 2u) ? B.incl[j] : B.agg[j];
        prefix.v += x.v;
        prefix.t += x.t;
        prefix.a += x.a;
        if (f == 2u) break;
        --j;
      }
      B.incl[blk] = ScanTriple{prefix.v + total.v, prefix.t + total.t, prefix.a + total.a};
      store_release(&B.flags[blk], 2u);
    }
    s_prefix = prefix;
    if (static_cast<int64_t>(blk) == D.n_blocks - 1) {
      B.totals[kTotVertices] = prefix.v + total.v;
      B.totals[kTotTriangles] = prefix.t + total.t;
      B.totals[kTotActive] = prefix.a + total.a;
      if (prefix.v + total.v > 0xffffffffull || 3 * (prefix.t + total.t) > 0xffffffffull)
        atomicOr(&B.totals[kTotErrors], 1ull);
    }
  }
  __syncthreads();

  if (g < D.n_segs) {
    const ScanTriple p = s_prefix;
    B.seg[g] = make_uint2(static_cast<uint32_t>(p.v + ov), static_cast<uint32_t>(p.t + ot));
    uint64_t k = p.a + oa;
    for (uint64_t m = act; m; m &= m - 1, ++k)
      if (k < B.active_capacity)
        B.active[k] = (static_cast<uint32_t>(g) << 6) | static_cast<uint32_t>(__builtin_ctzll(m));
  }
}

// ---------------------------------------------------------------------------
// 3. emit — one thread per active cell
// ---------------------------------------------------------------------------
struct EmitArgs {
  const uint64_t* S;
  const uint2* seg;
  const uint64_t* tplanes;
  const uint32_t* active;
  uint64_t active_capacity;
  const float* grid;
  const DevOp* ops;
  int n_ops;
  const float* positions;  // per lattice point (3 floats), kEmitPositions only
  float* vertices;
  uint32_t* indices;
  uint64_t vcap, icap;
  const unsigned long long* totals;
};

// Where an extraction's sample values and vertex positions come from.
enum EmitSource : int {
  kEmitProgram = 0,    // values from the SDF program, positions from the lattice
  kEmitGrid = 1,       // values from a dense grid, positions from the lattice
  kEmitPositions = 2,  // values from a dense grid, positions per lattice point
};

template <int kSrc>
__device__ __forceinline__ float lattice_value(const EmitArgs& A, const Lattice& L, int64_t x,
                                               int64_t y, int64_t z) {
  if (kSrc != kEmitProgram) {
    return A.grid[(z * L.ny + y) * L.nx + x];
  } else {
    return sdf_eval(A.ops, A.n_ops, lattice_coord(L, 0, x), lattice_coord(L, 1, y),
                    lattice_coord(L, 2, z));
  }
}

// Cube index of the cell whose low corner sits at (dx, dy, dz) of the 3x3x3
// sign neighbourhood `cube` (bit dx + 3 dy + 9 dz).
__device__ __forceinline__ uint32_t case_in_cube(uint32_t cube, int dx, int dy, int dz) {
  const uint32_t s = cube >> (dx + 3 * dy + 9 * dz);
  return (s & 1u) | ((s >> 1) & 1u) << 1 | ((s >> 4) & 1u) << 2 | ((s >> 3) & 1u) << 3 |
         ((s >> 9) & 1u) << 4 | ((s >> 10) & 1u) << 5 | ((s >> 13) & 1u) << 6 |
         ((s >> 12) & 1u) << 7;
}

constexpr int kEmitThreads = 256;

template <int kSrc>
__device__ __forceinline__ void lattice_position(const EmitArgs& A, const Lattice& L, int64_t x,
                                                 int64_t y, int64_t z, float* p) {
  if (kSrc == kEmitPositions) {
    const float* q = A.positions + 3 * ((z * L.ny + y) * L.nx + x);
    p[0] = q[0];
    p[1] = q[1];
    p[2] = q[2];
  } else {
    p[0] = lattice_coord(L, 0, x);
    p[1] = lattice_coord(L, 1, y);
    p[2] = lattice_coord(L, 2, z);
  }
}

template <int kSrc>
__global__ __launch_bounds__(kEmitThreads) void k_emit(EmitArgs A, Lattice L, McDims D) {
  __shared__ tables::McTables s_t;
  // Per thread: first vertex id and (cube index | owned crossed edges << 8)
  // of each of the eight cells that can own one of its edges, indexed
  // rx | ry<<1 | rz<<2 for the cell at (x-rx, y-ry, z-rz).
  __shared__ uint32_t s_obase[8][kEmitThreads];
  __shared__ uint32_t s_oinfo[8][kEmitThreads];
  load_tables(&s_t);
  __syncthreads();

  const int tid = threadIdx.x;
  const int64_t nx = L.nx, ny = L.ny;
  const int64_t nxy = nx * ny;
  const uint32_t spr = static_cast<uint32_t>(D.segs_per_row);
  const uint32_t cyn = static_cast<uint32_t>(ny - 1);
  uint64_t n_act = A.totals[kTotActive];
  if (n_act > A.active_capacity) n_act = A.active_capacity;

  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kEmitThreads + tid; i < n_act;
       i += static_cast<uint64_t>(gridDim.x) * kEmitThreads) {
    const uint32_t rec = A.active[i];
    const uint32_t g = rec >> 6;
    const int b = static_cast<int>(rec & 63u);
    const uint32_t row = g / spr;
    const uint32_t w = g - row * spr;
    const int64_t cz = row / cyn;
    const int64_t cy = row - cz * cyn;
    const int64_t cx = 64 * static_cast<int64_t>(w) + b;
    const bool prev_seg = (b == 0) && (w > 0);  // owners at x-1 live in segment w-1

    // Row words of point rows y-1..y+1, z-1..z+1 at segment w (and w-1 when
    // the cell is its segment's first), indexed [dz][dy].
    uint64_t rw[3][3], rn[3][3], pw[3][3], pn[3][3];
#pragma unroll
    for (int dz = 0; dz < 3; ++dz) {
#pragma unroll
      for (int dy = 0; dy < 3; ++dy) {
        rw[dz][dy] = rn[dz][dy] = pw[dz][dy] = pn[dz][dy] = 0ull;
        const int64_t y = cy - 1 + dy, z = cz - 1 + dz;
        if (y >= 0 && z >= 0) {
          const int64_t f = z * nxy + y * nx + 64 * static_cast<int64_t>(w);
          rw[dz][dy] = row_word(A.S, f, &rn[dz][dy]);
          if (prev_seg) pw[dz][dy] = row_word(A.S, f - 64, &pn[dz][dy]);
        }
      }
    }
    // 3x3x3 sign neighbourhood: every cell that can own one of this cell's
    // edges lies inside it.
    uint32_t cube = 0;
#pragma unroll
    for (int dz = 0; dz < 3; ++dz) {
#pragma unroll
      for (int dy = 0; dy < 3; ++dy) {
        const uint64_t v = rw[dz][dy];
        const uint32_t lo = b ? static_cast<uint32_t>((v >> (b - 1)) & 1ull)
                              : static_cast<uint32_t>(pw[dz][dy] >> 63);
        const uint32_t mid = static_cast<uint32_t>((v >> b) & 1ull);
        const uint32_t hi = b < 63 ? static_cast<uint32_t>((v >> (b + 1)) & 1ull)
                                   : static_cast<uint32_t>(rn[dz][dy]);
        cube |= (lo | (mid << 1) | (hi << 2)) << (3 * dy + 9 * dz);
      }
    }
    const uint32_t c0 = case_in_cube(cube, 1, 1, 1);
    const uint32_t emask = s_t.edge_mask[c0];

    // Which owner cells the cell's crossed edges need.
    uint32_t needed = 0;
    for (uint32_t m = emask; m; m &= m - 1) {
      const int e = __builtin_ctz(m);
      const uint32_t gm = static_cast<uint32_t>(kEdgeGeomPacked >> (5 * e)) & 31u;
      const int axis = static_cast<int>(gm >> 3);
      const int rx = (axis != 0 && (gm & 1u) == 0 && cx >= 1) ? 1 : 0;
      const int ry = (axis != 1 && (gm & 2u) == 0 && cy >= 1) ? 1 : 0;
      const int rz = (axis != 2 && (gm & 4u) == 0 && cz >= 1) ? 1 : 0;
      needed |= 1u << (rx | (ry << 1) | (rz << 2));
    }

    // First vertex id, cube index and owned crossed edges of each needed owner.
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      if (!(needed & (1u << o))) continue;
      const int rx = o & 1, ry = (o >> 1) & 1, rz = (o >> 2) & 1;
      const int64_t ox = cx - rx, oy = cy - ry, oz = cz - rz;
      const uint32_t oc = case_in_cube(cube, 1 - rx, 1 - ry, 1 - rz);
      const uint32_t oown = owned_edges(ox, oy, oz) & s_t.edge_mask[oc];
      // Owner's segment words: point rows oy..oy+1 (dy = 1-ry, 2-ry) and
      // oz..oz+1 (dz = 1-rz, 2-rz), segment w, or w-1 for x-1 at bit 0.
      const bool pseg = rx && b == 0;
      const int y0i = 1 - ry, z0i = 1 - rz;
      SegWords os;
      if (pseg) {
        os = seg_words(pw[z0i][y0i], pn[z0i][y0i], pw[z0i][y0i + 1], pn[z0i][y0i + 1],
                       pw[z0i + 1][y0i], pn[z0i + 1][y0i], pw[z0i + 1][y0i + 1],
                       pn[z0i + 1][y0i + 1]);
      } else {
        os = seg_words(rw[z0i][y0i], rn[z0i][y0i], rw[z0i][y0i + 1], rn[z0i][y0i + 1],
                       rw[z0i + 1][y0i], rn[z0i + 1][y0i], rw[z0i + 1][y0i + 1],
                       rn[z0i + 1][y0i + 1]);
      }
      const int64_t ow = pseg ? static_cast<int64_t>(w) - 1 : static_cast<int64_t>(w);
      const int ob = pseg ? 63 : b - rx;
      const uint64_t og = (static_cast<uint64_t>(oz) * cyn + oy) * spr + ow;
      const uint32_t pre =
          owned_count(os, seg_valid(nx, ow) & low_mask(ob), oy == 0, oz == 0, ow == 0);
      s_obase[o][tid] = A.seg[og].x + pre;
      s_oinfo[o][tid] = oc | (oown << 8);
    }

    // Vertices this cell owns: positions along its own edge directions.
    const uint32_t mine = owned_edges(cx, cy, cz) & emask;
    if (mine) {
      const uint32_t base = s_obase[0][tid];
      for (uint32_t todo = mine; todo; todo &= todo - 1) {
        const int e = __builtin_ctz(todo);
        const uint32_t id = base + static_cast<uint32_t>(__popc(mine & s_t.before[c0][e]));
        const int u = tables::kEdgeCorners[e][0], v = tables::kEdgeCorners[e][1];
        const int64_t ux = cx + tables::kCorner[u][0], uy = cy + tables::kCorner[u][1],
                      uz = cz + tables::kCorner[u][2];
        const int64_t vx = cx + tables::kCorner[v][0], vy = cy + tables::kCorner[v][1],
                      vz = cz + tables::kCorner[v][2];
        const float va = lattice_value<kSrc>(A, L, ux, uy, uz);
        const float vb = lattice_value<kSrc>(A, L, vx, vy, vz);
        float pa[3], pb[3], p[3];
        lattice_position<kSrc>(A, L, ux, uy, uz, pa);
        lattice_position<kSrc>(A, L, vx, vy, vz, pb);
        crossing(va, vb, pa[0], pa[1], pa[2], pb[0], pb[1], pb[2], p);
        if (id < A.vcap) {
          A.vertices[3 * static_cast<uint64_t>(id) + 0] = p[0];
          A.vertices[3 * static_cast<uint64_t>(id) + 1] = p[1];
          A.vertices[3 * static_cast<uint64_t>(id) + 2] = p[2];
        }
      }
    }

    // Triangles: cell order, then table order (mesh.rs:91-100).
    const uint64_t lm = low_mask(b);
    const uint64_t* tp = A.tplanes + 3 * static_cast<uint64_t>(g);
    const uint64_t tbase = A.seg[g].y + popc64(tp[0] & lm) + 2 * popc64(tp[1] & lm) +
                           4 * popc64(tp[2] & lm);
    const uint32_t nidx = 3u * s_t.tri_count[c0];
    for (uint32_t j = 0; j < nidx; ++j) {
      const int e = s_t.tri[c0][j];
      const uint32_t gm = static_cast<uint32_t>(kEdgeGeomPacked >> (5 * e)) & 31u;
      const int axis = static_cast<int>(gm >> 3);
      const int px = gm & 1, py = (gm >> 1) & 1, pz = (gm >> 2) & 1;
      // Owner: one cell back along each non-axis direction where the edge
      // lies on this cell's low face and a lower cell exists.
      const int rx = (axis != 0 && px == 0 && cx >= 1) ? 1 : 0;
      const int ry = (axis != 1 && py == 0 && cy >= 1) ? 1 : 0;
      const int rz = (axis != 2 && pz == 0 && cz >= 1) ? 1 : 0;
      const int o = rx | (ry << 1) | (rz << 2);
      const int eo = edge_of_axis_min(axis, (px | rx) | ((py | ry) << 1) | ((pz | rz) << 2));
      const uint32_t info = s_oinfo[o][tid];
      const uint32_t oc = info & 0xffu, oown = info >> 8;
      const uint32_t id =
          s_obase[o][tid] + static_cast<uint32_t>(__popc(oown & s_t.before[oc][eo]));
      const uint64_t ix = 3 * tbase + j;
      if (ix < A.icap) A.indices[ix] = id;
    }
  }
}

// ---------------------------------------------------------------------------
// Program sampling (iso_program_sample_lattice / iso_program_sample)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sample_lattice(const DevOp* __restrict__ ops, int n_ops,
                                                        Lattice L, int64_t n_points,
                                                        float* __restrict__ out) {
  const uint32_t nx = static_cast<uint32_t>(L.nx);
  const uint32_t nxy = static_cast<uint32_t>(L.nx * L.ny);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; f < n_points;
       f += stride) {
    const uint32_t ff = static_cast<uint32_t>(f);
    const uint32_t z = ff / nxy;
    const uint32_t r = ff - z * nxy;
    const uint32_t y = r / nx;
    const uint32_t x = r - y * nx;
    out[f] = sdf_eval(ops, n_ops, lattice_coord(L, 0, x), lattice_coord(L, 1, y),
                      lattice_coord(L, 2, z));
  }
}

__global__ __launch_bounds__(256) void k_sample_points(const DevOp* __restrict__ ops, int n_ops,
                                                       const float* __restrict__ pts, uint64_t n,
                                                       float* __restrict__ out) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = sdf_eval(ops, n_ops, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
}

int grid_for(int64_t waves_needed) {
  int64_t blocks = (waves_needed + 3) / 4;
  if (blocks < 1) blocks = 1;
  if (blocks > 16384) blocks = 16384;
  return static_cast<int>(blocks);
}

int device_cus() {
  int device = 0, cus = 256;
  if (hipGetDevice(&device) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  return cus > 0 ? cus : 256;
}

ResetArgs reset_args(const McDims& D, const McBuffers& B) {
  ResetArgs r;
  r.signs = B.signs;
  r.n_words = D.n_sign_words;
  r.flags = B.flags;
  r.n_blocks = D.n_blocks;
  r.totals = B.totals;
  return r;
}

}  // namespace

McDims mc_dims(const Lattice& L) {
  McDims D;
  D.n_points = L.nx * L.ny * L.nz;
  D.n_sign_words = (D.n_points + 63) / 64;
  D.n_cell_rows = (L.ny - 1) * (L.nz - 1);
  D.segs_per_row = (L.nx - 1 + 63) / 64;
  D.n_segs = D.n_cell_rows * D.segs_per_row;
  D.n_blocks = (D.n_segs + kCountBlock - 1) / kCountBlock;
  D.n_cells = (L.nx - 1) * (L.ny - 1) * (L.nz - 1);
  return D;
}

hipError_t launch_signs_grid(const float* grid, const Lattice& L, const McDims& D,
                             const McBuffers& B, hipStream_t s) {
  (void)L;
  if (!grid) return hipErrorInvalidValue;
  const ResetArgs R = reset_args(D, B);
  // At least enough threads to clear the look-back flags in one sweep.
  const int64_t min_blocks = (D.n_blocks + 255) / 256;
  if ((reinterpret_cast<uintptr_t>(grid) & 15u) == 0) {
    int64_t blocks = grid_for((D.n_points + 256 * kSignVec - 1) / (256 * kSignVec));
    if (blocks < min_blocks) blocks = min_blocks;
    hipLaunchKernelGGL(k_signs_grid4, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s,
                       reinterpret_cast<const f32x4*>(grid), D.n_points, D.n_sign_words, B.signs,
                       R);
  } else {
    int64_t blocks = grid_for((D.n_sign_words + kSignUnroll - 1) / kSignUnroll);
    if (blocks < min_blocks) blocks = min_blocks;
    hipLaunchKernelGGL(k_signs_grid, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, grid,
                       D.n_points, D.n_sign_words, B.signs, R);
  }
  return hipGetLastError();
}

hipError_t launch_signs_program(const DevOp* ops, int n_ops, const Lattice& L, const McDims& D,
                                const McBuffers& B, hipStream_t s) {
  if (!ops) return hipErrorInvalidValue;
  int64_t blocks = grid_for(D.n_sign_words);
  const int64_t min_blocks = (D.n_blocks + 255) / 256;
  if (blocks < min_blocks) blocks = min_blocks;
  hipLaunchKernelGGL(k_signs_program, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, ops,
                     n_ops, L, D.n_points, D.n_sign_words, B.signs, reset_args(D, B));
  return hipGetLastError();
}

hipError_t launch_count(const Lattice& L, const McDims& D, const McBuffers& B, hipStream_t s) {
  if (D.n_blocks > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_count, dim3(static_cast<unsigned>(D.n_blocks)), dim3(kCountBlock), 0, s,
                     B.signs, L, D, B);
  return hipGetLastError();
}

hipError_t launch_emit(const float* grid, const DevOp* ops, int n_ops, const float* positions,
                       const Lattice& L, const McDims& D, const McBuffers& B, hipStream_t s) {
  if ((positions && !grid) || (!grid && !ops)) return hipErrorInvalidValue;
  EmitArgs A;
  A.S = B.signs;
  A.seg = B.seg;
  A.tplanes = B.tplanes;
  A.active = B.active;
  A.active_capacity = B.active_capacity;
  A.grid = grid;
  A.ops = ops;
  A.n_ops = n_ops;
  A.positions = positions;
  A.vertices = B.vertices;
  A.indices = B.indices;
  A.vcap = B.vertex_capacity;
  A.icap = B.index_capacity;
  A.totals = B.totals;
  // Grid-stride over the active list (its length is only known on the device).
  int64_t blocks = (static_cast<int64_t>(B.active_capacity) + kEmitThreads - 1) / kEmitThreads;
  const int64_t cap = 8LL * device_cus();
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  if (grid && positions) {
    hipLaunchKernelGGL(k_emit<kEmitPositions>, dim3(static_cast<unsigned>(blocks)),
                       dim3(kEmitThreads), 0, s, A, L, D);
  } else if (grid) {
    hipLaunchKernelGGL(k_emit<kEmitGrid>, dim3(static_cast<unsigned>(blocks)), dim3(kEmitThreads),
                       0, s, A, L, D);
  } else {
    hipLaunchKernelGGL(k_emit<kEmitProgram>, dim3(static_cast<unsigned>(blocks)),
                       dim3(kEmitThreads), 0, s, A, L, D);
  }
  return hipGetLastError();
}

hipError_t launch_sample_lattice(const DevOp* ops, int n_ops, const Lattice& L, float* out,
                                 hipStream_t s) {
  const int64_t n = L.nx * L.ny * L.nz;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_sample_lattice, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, ops,
                     n_ops, L, n, out);
  return hipGetLastError();
}

hipError_t launch_sample_points(const DevOp* ops, int n_ops, const float* pts, uint64_t n,
                                float* out, hipStream_t s) {
  const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sample_points, dim3(blocks), dim3(256), 0, s, ops, n_ops, pts, n, out);
  return hipGetLastError();
}

}  // namespace iso
