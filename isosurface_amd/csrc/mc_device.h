// Device helpers shared by the marching-cubes kernels (single lattice and
// batched chunks): edge geometry tables, the bit-sliced segment words of 64
// cells, vertex-ownership counts, sign neighbourhoods, crossing points and
// the look-back primitives.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mc_kernels.h"
#include "mc_tables.h"

namespace iso {
namespace mcd {

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

// Owner of each edge of a cell that has lower neighbours along every axis,
// per case: need[c] = owner slots (rx | ry << 1 | rz << 2 of the cell at
// (x-rx, y-ry, z-rz)) its crossed edges use, ix[c][j] = slot | (the edge's
// local id in the owner) << 3 for entry j of its triangle list.
struct InteriorTables {
  uint8_t ix[256][16];
  uint8_t need[256];
};

// Edge e of an interior cell: its owner's slot | the edge's local id in the
// owner << 3.
constexpr uint8_t interior_code(int e) {
  const EdgeGeom g = make_geom();
  const int axis = g.axis[e];
  const int r[3] = {axis != 0 && g.minp[e][0] == 0 ? 1 : 0,
                    axis != 1 && g.minp[e][1] == 0 ? 1 : 0,
                    axis != 2 && g.minp[e][2] == 0 ? 1 : 0};
  const int slot = r[0] | (r[1] << 1) | (r[2] << 2);
  const int owner_edge = g.of_axis_min[axis][(g.minp[e][0] | r[0]) |
                                             ((g.minp[e][1] | r[1]) << 1) |
                                             ((g.minp[e][2] | r[2]) << 2)];
  return static_cast<uint8_t>(slot | (owner_edge << 3));
}

constexpr InteriorTables make_interior() {
  InteriorTables t{};
  const tables::McTables& T = tables::kTables;
  uint8_t code[12] = {};
  for (int e = 0; e < 12; ++e) code[e] = interior_code(e);
  for (int c = 0; c < 256; ++c) {
    uint32_t need = 0;
    for (int e = 0; e < 12; ++e)
      if ((T.edge_mask[c] >> e) & 1u) need |= 1u << (code[e] & 7);
    t.need[c] = static_cast<uint8_t>(need);
    for (int j = 0; j < 16; ++j) {
      const int e = T.tri[c][j];
      t.ix[c][j] = e < 0 ? 0 : code[e];
    }
  }
  return t;
}

// Corner mask (bit per corner) of each edge's two endpoints, eight bits per
// edge: edges 0-7 in the first word, 8-11 in the second.
constexpr uint64_t pack_edge_corners(int half) {
  uint64_t r = 0;
  for (int i = 0; i < 8; ++i) {
    const int e = 8 * half + i;
    if (e >= 12) break;
    r |= static_cast<uint64_t>((1u << tables::kEdgeCorners[e][0]) |
                               (1u << tables::kEdgeCorners[e][1]))
         << (8 * i);
  }
  return r;
}
constexpr uint64_t kEdgeCornerMask0 = pack_edge_corners(0);
constexpr uint64_t kEdgeCornerMask1 = pack_edge_corners(1);

__device__ __forceinline__ uint32_t edge_corner_mask(int e) {
  const uint64_t w = e < 8 ? kEdgeCornerMask0 : kEdgeCornerMask1;
  return static_cast<uint32_t>((w >> (8 * (e & 7))) & 0xFFu);
}

// The tables emit stages in LDS (the triangle lists themselves stay in
// constant memory: cells away from the low faces read ix instead).
struct EmitTables {
  uint16_t before[256][12];   // edges first used before edge e (by case)
  uint16_t edge_mask[256];    // edges crossed by the case
  uint8_t tri_count[256];     // triangles emitted by the case
  uint8_t need[256];          // InteriorTables::need
  uint8_t ix[256][16];        // InteriorTables::ix
};

constexpr EmitTables make_emit_tables() {
  EmitTables t{};
  const tables::McTables& T = tables::kTables;
  const InteriorTables in = make_interior();
  for (int c = 0; c < 256; ++c) {
    for (int e = 0; e < 12; ++e) t.before[c][e] = T.before[c][e];
    t.edge_mask[c] = T.edge_mask[c];
    t.tri_count[c] = T.tri_count[c];
    t.need[c] = in.need[c];
    for (int j = 0; j < 16; ++j) t.ix[c][j] = in.ix[c][j];
  }
  return t;
}

__device__ __forceinline__ int edge_of_axis_min(int axis, int rel) {
  const int k = 8 * axis + rel;
  const uint64_t w = k < 16 ? kOfAxisMin0 : kOfAxisMin1;
  return static_cast<int>((w >> (4 * (k & 15))) & 15u);
}

// Copies a table into LDS with 16-byte accesses.
template <class T>
__device__ __forceinline__ void load_table16(T* dst, const T* from) {
  static_assert(sizeof(T) % 16 == 0, "table size");
  const uint4* src = reinterpret_cast<const uint4*>(from);
  uint4* d = reinterpret_cast<uint4*>(dst);
  for (int i = threadIdx.x; i < static_cast<int>(sizeof(T) / 16); i += blockDim.x) d[i] = src[i];
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

// Adds the 1-bit-per-cell mask m to the bit-sliced per-cell counter p.
__device__ __forceinline__ void sliced_add(uint64_t* p, int n_planes, uint64_t m) {
#pragma unroll
  for (int k = 0; k < n_planes; ++k) {
    const uint64_t carry = p[k] & m;
    p[k] ^= m;
    m = carry;
  }
}

// Per-cell owned crossed-edge counts of the cells in `valid`, bit-sliced in
// four planes (a cell owns at most 12 edges): the same edges owned_count
// counts, so owned_count(s, m, ...) == sum_k 2^k popc(p[k] & m) for m in valid.
__device__ __forceinline__ void owned_planes(const SegWords& s, uint64_t valid, bool y0, bool z0,
                                             bool first, uint64_t* p) {
  const uint64_t* c = s.c;
  // e6, e5, e10 (every cell): a full adder of three masks
  const uint64_t a = valid & (c[6] ^ c[7]), b = valid & (c[5] ^ c[6]), d = valid & (c[2] ^ c[6]);
  p[0] = a ^ b ^ d;
  p[1] = (a & b) | (d & (a ^ b));
  p[2] = p[3] = 0;
  if (y0) {
    sliced_add(p, 4, valid & (c[4] ^ c[5]));
    sliced_add(p, 4, valid & (c[1] ^ c[5]));
  }
  if (z0) {
    sliced_add(p, 4, valid & (c[2] ^ c[3]));
    sliced_add(p, 4, valid & (c[1] ^ c[2]));
  }
  if (y0 && z0) sliced_add(p, 4, valid & (c[0] ^ c[1]));
  if (first) {
    const uint64_t v0 = valid & 1ull;
    sliced_add(p, 4, v0 & (c[4] ^ c[7]));
    sliced_add(p, 4, v0 & (c[3] ^ c[7]));
    if (z0) sliced_add(p, 4, v0 & (c[0] ^ c[3]));
    if (y0) sliced_add(p, 4, v0 & (c[0] ^ c[4]));
  }
}

// Per-cell triangle counts of the cells in `m` (mixed cells only), bit-sliced
// in three planes, with no table lookup.  For every case other than 0 and
// 255 the triangle table's count equals
//   V - 2 F - 4 H,
// V = corners inside, F = cube faces whose four corners are all inside,
// H = 1 when the two corners outside are antipodal.  The table triangulates
// each surface patch as a fan over its crossed edges, (edges - 2)
// triangles; a cell has 3 V - 2 I crossed edges (I: cube edges with both
// corners inside) and, but for the antipodal case, one patch per connected
// component of its inside corners, V - I + F components on the cube's
// graph (whose cycles there are the full faces, and the hexagon around an
// antipodal pair, for which the table keeps two separate triangles).  The
// identity is checked against the table, case by case, in the tests.
__device__ __forceinline__ void tri_planes(const SegWords& s, uint64_t m, uint64_t* t) {
  const uint64_t* c = s.c;
  auto full_add = [](uint64_t a, uint64_t b, uint64_t d, uint64_t* carry) {
    const uint64_t x = a ^ b;
    *carry = (a & b) | (x & d);
    return x ^ d;
  };
  // V mod 8 (V is 1..7 in a mixed cell): a carry-save sum of the corner planes
  uint64_t k1, k2, k4, k5;
  const uint64_t s1 = full_add(c[0], c[1], c[2], &k1);
  const uint64_t s2 = full_add(c[3], c[4], c[5], &k2);
  const uint64_t v0 = full_add(s1, s2, c[6] ^ c[7], &k4);
  const uint64_t twos = full_add(k1, k2, c[6] & c[7], &k5);
  const uint64_t v1 = twos ^ k4, v2 = k5 ^ (twos & k4);
  // F mod 4 (at most 3 in a mixed cell); faces x=0 {0,3,4,7}, x=1 {1,2,5,6},
  // y=0 {0,1,4,5}, y=1 {2,3,6,7}, z=0 {0,1,2,3}, z=1 {4,5,6,7}
  const uint64_t c04 = c[0] & c[4], c37 = c[3] & c[7], c15 = c[1] & c[5], c26 = c[2] & c[6];
  uint64_t ka, kb;
  const uint64_t sa = full_add(c04 & c37, c15 & c26, c04 & c15, &ka);
  const uint64_t sb = full_add(c26 & c37, c[0] & c[1] & c[2] & c[3], c[4] & c[5] & c[6] & c[7], &kb);
  const uint64_t f0 = sa ^ sb, f1 = ka ^ kb ^ (sa & sb);
  // H: six corners inside, and an antipodal pair (0,6), (1,7), (2,4) or (3,5) outside
  const uint64_t h = v2 & v1 & ~v0 &
                     (~(c[0] | c[6]) | ~(c[1] | c[7]) | ~(c[2] | c[4]) | ~(c[3] | c[5]));
  // bit 0 is V's; bits 1-2 are (V >> 1) - F - 2 H, which lies in 0..2
  const uint64_t d0 = v1 ^ f0, d1 = v2 ^ f1 ^ (~v1 & f0);
  t[0] = v0 & m;
  t[1] = d0 & m;
  t[2] = (d1 ^ h) & m;
}

__device__ __forceinline__ uint32_t planes_count(const uint64_t* p, uint64_t m) {
  return popc64(p[0] & m) + 2 * popc64(p[1] & m) + 4 * popc64(p[2] & m) + 8 * popc64(p[3] & m);
}

// Cube index of the cell whose low corner sits at (dx, dy, dz) of the 3x3x3
// sign neighbourhood `cube` (bit dx + 3 dy + 9 dz).
__device__ __forceinline__ uint32_t case_in_cube(uint32_t cube, int dx, int dy, int dz) {
  // corners 0, 1, 3 sit at bits 0, 1, 3 of s; 2 at 4; 4, 5, 7 at 9, 10, 12;
  // 6 at 13
  const uint32_t s = cube >> (dx + 3 * dy + 9 * dz);
  return (s & 0xBu) | ((s >> 2) & 0x4u) | ((s >> 5) & 0xB0u) | ((s >> 7) & 0x40u);
}

// A predecessor publishes its aggregate right after its own counting, so the
// wait is short; the bound only guarantees that a broken invariant ends the
// kernel (with an error flag) instead of spinning forever.
constexpr uint32_t kLookBackSpinLimit = 1u << 22;

// Look-back words: status << 62 | value (1: aggregate, 2: inclusive),
// relaxed agent-scope atomics; the status travels with the value.
constexpr unsigned long long kSlotAggregate = 1ull << 62;
constexpr unsigned long long kSlotInclusive = 2ull << 62;
constexpr unsigned long long kSlotValue = (1ull << 62) - 1ull;

__device__ __forceinline__ unsigned long long slot_load(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void slot_store(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Per-extraction state cleared by the first kernel of an extraction.
struct ResetArgs {
  uint64_t* signs;  // guard words at [-1], [n_words], [n_words + 1]
  int64_t n_words;
  ScanTriple* agg;
  ScanTriple* incl;
  int64_t n_blocks;
  unsigned long long* totals;
};

__device__ __forceinline__ void reset_state(const ResetArgs& r) {
  const int64_t i0 = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = i0; i < r.n_blocks; i += stride) {
    r.agg[i] = ScanTriple{0ull, 0ull, 0ull};
    r.incl[i] = ScanTriple{0ull, 0ull, 0ull};
  }
  if (i0 < kTotSlots) r.totals[i0] = 0ull;
  if (i0 == 0 && r.signs) {
    r.signs[-1] = 0ull;
    r.signs[r.n_words] = 0ull;
    r.signs[r.n_words + 1] = 0ull;
  }
}

}  // namespace mcd
}  // namespace iso
