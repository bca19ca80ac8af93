// Marching-cubes kernels for gfx950 (MI355X).
//
// Pipeline (one extraction):
//   1. signs   one bit per lattice sample, "inside" = !(v > 0) exactly as
//              classify_corners (marching_cubes_impl.rs:26-37, distance.rs:52-54).
//              From a dense FP32 grid this is a pure HBM stream; from a
//              program it is the SDF evaluation.  Wave64 ballots pack 64
//              consecutive samples (flat x-fastest order) into one word.
//   2. count   one wavefront per 64-cell segment of a cell row: cube indices
//              come from the sign words alone (no FP32 re-read); per segment
//              it counts the vertices the segment's cells own and the
//              triangles they emit.
//   3. scan    exclusive scan of the segment counts in cell order.
//   4. emit    recomputes the cube indices of active segments, writes vertex
//              positions (FP32 crossing points, from the grid or the program)
//              and the u32 index triples.
//
// Vertex numbering is the reference's first-encounter order
// (MeshTopologyBuilder::add_vertex, mesh.rs:240-251, driven by
// PrimalGrid::traverse's z-major cell order, primal_grid.rs:59-85): a lattice
// edge's vertex belongs to the lowest-numbered cell that contains the edge,
// its position is computed with that cell's local edge direction, and
// vertices are numbered by (owner cell, first appearance in the owner's
// triangle list).  The owner of every edge follows from the edge's position
// alone, so each wavefront computes vertex ids without any hash table.
#include <hip/hip_runtime.h>

#include "mc_kernels.h"
#include "mc_tables.h"
#include "sdf_eval.h"

namespace iso {

namespace {

// ---------------------------------------------------------------------------
// Edge geometry derived from the corner / edge tables.
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
__constant__ EdgeGeom c_geom = make_geom();

// Lower endpoint offsets and axis of each edge, packed for register use:
// bits 0-2 = min point (x|y<<1|z<<2), bits 3-4 = axis.
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

__device__ __forceinline__ void load_tables(tables::McTables* dst) {
  const uint32_t* src = reinterpret_cast<const uint32_t*>(&c_tables);
  uint32_t* d = reinterpret_cast<uint32_t*>(dst);
  constexpr int kWords = sizeof(tables::McTables) / 4;
  static_assert(sizeof(tables::McTables) % 4 == 0, "table size");
  for (int i = threadIdx.x; i < kWords; i += blockDim.x) d[i] = src[i];
}

// ---------------------------------------------------------------------------
// Wavefront helpers.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t mbcnt64(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
}

// Exclusive wave prefix and wave total of v < 2^BITS (bit-sliced ballots).
template <int BITS>
__device__ __forceinline__ uint32_t wave_excl_small(uint32_t v, uint32_t* total) {
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int b = 0; b < BITS; ++b) {
    const uint64_t m = __ballot((v >> b) & 1u);
    pre += mbcnt64(m) << b;
    tot += static_cast<uint32_t>(__popcll(m)) << b;
  }
  *total = tot;
  return pre;
}

__device__ __forceinline__ int64_t uniform64(int64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(static_cast<uint64_t>(v) >> 32));
  return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}

// 64 sign bits of a point row starting at flat sample index f:
// cur bit l = sample f+l, hi bit l = sample f+l+1, prev bit l = sample f+l-1.
struct Win {
  uint64_t cur, hi, prev;
};

__device__ __forceinline__ Win load_win(const uint64_t* __restrict__ S, int64_t f) {
  const int64_t w0 = f >> 6;
  const int s = static_cast<int>(f & 63);
  const uint64_t a = S[w0 - 1], b = S[w0], c = S[w0 + 1];
  const uint64_t cur = s ? ((b >> s) | (c << (64 - s))) : b;
  const uint64_t nextbit = (c >> s) & 1ull;
  const uint64_t prevbit = s ? ((b >> (s - 1)) & 1ull) : (a >> 63);
  Win r;
  r.cur = cur;
  r.hi = (cur >> 1) | (nextbit << 63);
  r.prev = (cur << 1) | prevbit;
  return r;
}

__device__ __forceinline__ uint32_t bit(uint64_t w, int l) {
  return static_cast<uint32_t>((w >> l) & 1ull);
}

// Cube index from the x/x+1 sign bits of the four point rows around a cell
// row: 00=(y,z) 10=(y+1,z) 01=(y,z+1) 11=(y+1,z+1).  Corner order is CORNERS
// (marching_cubes_tables.rs:16-25).
__device__ __forceinline__ uint32_t cube_index(uint64_t lo00, uint64_t hi00, uint64_t lo10,
                                               uint64_t hi10, uint64_t lo01, uint64_t hi01,
                                               uint64_t lo11, uint64_t hi11, int l) {
  return bit(lo00, l) | (bit(hi00, l) << 1) | (bit(hi10, l) << 2) | (bit(lo10, l) << 3) |
         (bit(lo01, l) << 4) | (bit(hi01, l) << 5) | (bit(hi11, l) << 6) | (bit(lo11, l) << 7);
}

// Local edges whose vertex the cell (x, y, z) owns, i.e. for which it is the
// first cell in z-major order that contains the edge.
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

__device__ __forceinline__ float lattice_coord(const Lattice& L, int axis, int64_t i) {
  return L.origin[axis] + static_cast<float>(i) * L.spacing[axis];
}

// Signed::find_crossing_point (distance.rs:64-70).
__device__ __forceinline__ void crossing(float a, float b, float pax, float pay, float paz,
                                         float pbx, float pby, float pbz, float* o) {
  const float delta = b - a;
  const float t = (delta == 0.0f) ? 0.5f : -a / delta;
  const float s = 1.0f - t;
  o[0] = pax * s + pbx * t;
  o[1] = pay * s + pby * t;
  o[2] = paz * s + pbz * t;
}

// ---------------------------------------------------------------------------
// 1. signs
// ---------------------------------------------------------------------------
constexpr int kSignUnroll = 8;  // words per wavefront iteration

__global__ __launch_bounds__(256) void k_signs_grid(const float* __restrict__ g,
                                                    int64_t n_points, int64_t n_words,
                                                    uint64_t* __restrict__ S) {
  const int lane = __lane_id();
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int64_t n_waves = static_cast<int64_t>(gridDim.x) * 4;
  for (int64_t w0 = wave * kSignUnroll; w0 < n_words; w0 += n_waves * kSignUnroll) {
    float v[kSignUnroll];
#pragma unroll
    for (int j = 0; j < kSignUnroll; ++j) {
      const int64_t f = (w0 + j) * 64 + lane;
      v[j] = f < n_points ? __builtin_nontemporal_load(g + f) : 1.0f;
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
                                                       uint64_t* __restrict__ S) {
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
// 2. count
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_count(const uint64_t* __restrict__ S, Lattice L,
                                               McDims D, uint2* __restrict__ seg) {
  __shared__ uint16_t s_mask[256];
  __shared__ uint8_t s_ntri[256];
  for (int i = threadIdx.x; i < 256; i += blockDim.x) {
    s_mask[i] = c_tables.edge_mask[i];
    s_ntri[i] = c_tables.tri_count[i];
  }
  __syncthreads();
  const int lane = __lane_id();
  const int64_t nx = L.nx, ny = L.ny;
  const uint32_t spr = static_cast<uint32_t>(D.segs_per_row);
  const uint32_t cyn = static_cast<uint32_t>(ny - 1);
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int64_t n_waves = static_cast<int64_t>(gridDim.x) * 4;
  for (int64_t g = wave; g < D.n_segs; g += n_waves) {
    const uint32_t gg = static_cast<uint32_t>(uniform64(g));
    const uint32_t row = gg / spr;
    const uint32_t w = gg - row * spr;
    const uint32_t cz = row / cyn;
    const uint32_t cy = row - cz * cyn;
    const int64_t x0 = static_cast<int64_t>(w) * 64;
    const int64_t f00 = (static_cast<int64_t>(cz) * ny + cy) * nx + x0;
    const Win r00 = load_win(S, f00);
    const Win r10 = load_win(S, f00 + nx);
    const Win r01 = load_win(S, f00 + nx * ny);
    const Win r11 = load_win(S, f00 + nx * ny + nx);
    const int64_t cx = x0 + lane;
    const bool valid = cx < nx - 1;
    const uint32_t c = cube_index(r00.cur, r00.hi, r10.cur, r10.hi, r01.cur, r01.hi, r11.cur,
                                  r11.hi, lane);
    const uint32_t own = owned_edges(cx, cy, cz) & s_mask[c];
    const uint32_t nv = valid ? static_cast<uint32_t>(__popc(own)) : 0u;
    const uint32_t nt = valid ? s_ntri[c] : 0u;
    uint32_t tv, tt;
    wave_excl_small<4>(nv, &tv);
    wave_excl_small<3>(nt, &tt);
    if (lane == 0) seg[gg] = make_uint2(tv, tt);
  }
}

// ---------------------------------------------------------------------------
// 3. scan (two levels: blocks of kScanBlock segments, then block totals)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = __lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

__device__ __forceinline__ uint64_t wave_incl_scan64(uint64_t v) {
  const int lane = __lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

__global__ __launch_bounds__(256) void k_scan_blocks(uint2* __restrict__ seg, int64_t n_segs,
                                                     uint64_t* __restrict__ blocks,
                                                     int64_t n_blocks) {
  __shared__ uint32_t s_v[4], s_t[4];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanBlock + threadIdx.x * kScanItems;
  uint2 it[kScanItems];
  uint32_t sv = 0, st = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const int64_t k = base + i;
    it[i] = k < n_segs ? seg[k] : make_uint2(0u, 0u);
    sv += it[i].x;
    st += it[i].y;
  }
  const uint32_t iv = wave_incl_scan(sv), ivt = wave_incl_scan(st);
  const int wv = threadIdx.x >> 6, lane = __lane_id();
  if (lane == 63) {
    s_v[wv] = iv;
    s_t[wv] = ivt;
  }
  __syncthreads();
  uint32_t off_v = 0, off_t = 0;
  for (int i = 0; i < wv; ++i) {
    off_v += s_v[i];
    off_t += s_t[i];
  }
  uint32_t pv = off_v + iv - sv, pt = off_t + ivt - st;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const int64_t k = base + i;
    if (k < n_segs) seg[k] = make_uint2(pv, pt);
    pv += it[i].x;
    pt += it[i].y;
  }
  if (threadIdx.x == 255) {
    blocks[2 * blockIdx.x] = pv;
    blocks[2 * blockIdx.x + 1] = pt;
  }
}

// One workgroup: exclusive scan of the block totals, grand totals.
__global__ __launch_bounds__(1024) void k_scan_top(uint64_t* __restrict__ blocks,
                                                   int64_t n_blocks,
                                                   unsigned long long* __restrict__ totals) {
  __shared__ uint64_t s_v[16], s_t[16];
  const int64_t per = (n_blocks + 1023) / 1024;
  const int64_t b0 = threadIdx.x * per;
  uint64_t sv = 0, st = 0;
  for (int64_t i = 0; i < per; ++i) {
    const int64_t k = b0 + i;
    if (k < n_blocks) {
      sv += blocks[2 * k];
      st += blocks[2 * k + 1];
    }
  }
  const uint64_t iv = wave_incl_scan64(sv), it = wave_incl_scan64(st);
  const int wv = threadIdx.x >> 6, lane = __lane_id();
  if (lane == 63) {
    s_v[wv] = iv;
    s_t[wv] = it;
  }
  __syncthreads();
  uint64_t off_v = 0, off_t = 0;
  for (int i = 0; i < wv; ++i) {
    off_v += s_v[i];
    off_t += s_t[i];
  }
  uint64_t pv = off_v + iv - sv, pt = off_t + it - st;
  for (int64_t i = 0; i < per; ++i) {
    const int64_t k = b0 + i;
    if (k < n_blocks) {
      const uint64_t v = blocks[2 * k], t = blocks[2 * k + 1];
      blocks[2 * k] = pv;
      blocks[2 * k + 1] = pt;
      pv += v;
      pt += t;
    }
  }
  if (threadIdx.x == 1023) {
    totals[0] = pv;
    totals[1] = pt;
  }
}

// ---------------------------------------------------------------------------
// 4. emit
// ---------------------------------------------------------------------------
struct EmitArgs {
  const uint64_t* S;
  const uint2* seg;
  const uint64_t* blocks;
  const float* grid;
  const DevOp* ops;
  int n_ops;
  float* vertices;
  uint32_t* indices;
  uint64_t vcap, icap;
  unsigned long long* totals;
};

template <bool kFromGrid>
__device__ __forceinline__ float lattice_value(const EmitArgs& A, const Lattice& L, int64_t x,
                                               int64_t y, int64_t z) {
  if (kFromGrid) {
    return A.grid[(z * L.ny + y) * L.nx + x];
  } else {
    return sdf_eval(A.ops, A.n_ops, lattice_coord(L, 0, x), lattice_coord(L, 1, y),
                    lattice_coord(L, 2, z));
  }
}

// Per-lane view of one owner cell row: cube index, owned-and-crossed edges
// and first vertex id of the lane's cell (x) and its left neighbour (x-1).
struct RowOwn {
  uint32_t case_x, case_xm;
  uint32_t own_x, own_xm;
  uint64_t base_x, base_xm;
};

// Branch-free choice among the eight (z, y, x) owner candidates, indexed
// [z][y][x] with 0 = one cell back, 1 = this cell's coordinate.
template <typename T>
__device__ __forceinline__ T pick(bool x, bool y, bool z, T zyx000, T zyx001, T zyx010, T zyx011,
                                  T zyx100, T zyx101, T zyx110, T zyx111) {
  const T a = y ? (x ? zyx011 : zyx010) : (x ? zyx001 : zyx000);
  const T b = y ? (x ? zyx111 : zyx110) : (x ? zyx101 : zyx100);
  return z ? b : a;
}

template <bool kFromGrid>
__global__ __launch_bounds__(256) void k_emit(EmitArgs A, Lattice L, McDims D) {
  __shared__ tables::McTables s_t;
  load_tables(&s_t);
  __syncthreads();

  const int lane = __lane_id();
  const int64_t nx = L.nx, ny = L.ny;
  const uint32_t spr = static_cast<uint32_t>(D.segs_per_row);
  const uint32_t cyn = static_cast<uint32_t>(ny - 1);
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int64_t n_waves = static_cast<int64_t>(gridDim.x) * 4;

  for (int64_t g = wave; g < D.n_segs; g += n_waves) {
    const uint32_t gg = static_cast<uint32_t>(uniform64(g));
    const uint32_t row = gg / spr;
    const uint32_t w = gg - row * spr;
    const int64_t cz = row / cyn;
    const int64_t cy = row - cz * cyn;
    const int64_t x0 = static_cast<int64_t>(w) * 64;
    const int64_t cx = x0 + lane;
    const bool valid = cx < nx - 1;

    // Point rows y in {cy-1, cy, cy+1} x z in {cz-1, cz, cz+1}; index [dz+1][dy+1].
    Win P[3][3];
#pragma unroll
    for (int dz = -1; dz <= 1; ++dz) {
#pragma unroll
      for (int dy = -1; dy <= 1; ++dy) {
        Win r{0ull, 0ull, 0ull};
        if (cy + dy >= 0 && cz + dz >= 0) r = load_win(A.S, ((cz + dz) * ny + cy + dy) * nx + x0);
        P[dz + 1][dy + 1] = r;
      }
    }

    const uint32_t c0 = cube_index(P[1][1].cur, P[1][1].hi, P[1][2].cur, P[1][2].hi, P[2][1].cur,
                                   P[2][1].hi, P[2][2].cur, P[2][2].hi, lane);
    const uint32_t ntri = valid ? s_t.tri_count[c0] : 0u;
    if (__ballot(ntri > 0) == 0ull) continue;

    // Owner rows (oy, oz) = (cy + ry, cz + rz), ry, rz in {-1, 0}; index [rz+1][ry+1].
    RowOwn R[2][2];
#pragma unroll
    for (int rz = -1; rz <= 0; ++rz) {
#pragma unroll
      for (int ry = -1; ry <= 0; ++ry) {
        RowOwn o{};
        const int64_t oy = cy + ry, oz = cz + rz;
        if (oy >= 0 && oz >= 0) {
          const Win& q00 = P[rz + 1][ry + 1];
          const Win& q10 = P[rz + 1][ry + 2];
          const Win& q01 = P[rz + 2][ry + 1];
          const Win& q11 = P[rz + 2][ry + 2];
          o.case_x = cube_index(q00.cur, q00.hi, q10.cur, q10.hi, q01.cur, q01.hi, q11.cur,
                                q11.hi, lane);
          o.case_xm = cube_index(q00.prev, q00.cur, q10.prev, q10.cur, q01.prev, q01.cur,
                                 q11.prev, q11.cur, lane);
          o.own_x = valid ? (owned_edges(cx, oy, oz) & s_t.edge_mask[o.case_x]) : 0u;
          o.own_xm = (cx >= 1) ? (owned_edges(cx - 1, oy, oz) & s_t.edge_mask[o.case_xm]) : 0u;
          uint32_t tot;
          const uint32_t pre = wave_excl_small<4>(static_cast<uint32_t>(__popc(o.own_x)), &tot);
          const uint64_t sidx = (static_cast<uint64_t>(oz) * cyn + oy) * spr + w;
          const uint64_t seg_base =
              A.seg[sidx].x + A.blocks[2 * (sidx / kScanBlock)];
          o.base_x = seg_base + pre;
          o.base_xm = o.base_x - static_cast<uint32_t>(__popc(o.own_xm));
        }
        R[rz + 1][ry + 1] = o;
      }
    }

    // Triangle numbering: cell order, then table order (mesh.rs:91-100).
    uint32_t tri_tot;
    const uint32_t tri_pre = wave_excl_small<3>(ntri, &tri_tot);
    const uint64_t sidx0 = static_cast<uint64_t>(row) * spr + w;
    const uint64_t tri_base = A.seg[sidx0].y + A.blocks[2 * (sidx0 / kScanBlock) + 1] + tri_pre;

    if (ntri == 0) continue;  // inactive lane (its wave still helped with the prefixes)

    // --- vertices owned by this cell (mesh.rs:240-251: extract_vertex on first use)
    const RowOwn& me = R[1][1];
    {
      uint32_t todo = me.own_x;
      while (todo) {
        const int e = __builtin_ctz(todo);
        todo &= todo - 1;
        const uint64_t id = me.base_x + static_cast<uint32_t>(__popc(me.own_x & s_t.before[c0][e]));
        const int u = tables::kEdgeCorners[e][0], v = tables::kEdgeCorners[e][1];
        const int64_t ux = cx + tables::kCorner[u][0], uy = cy + tables::kCorner[u][1],
                      uz = cz + tables::kCorner[u][2];
        const int64_t vx = cx + tables::kCorner[v][0], vy = cy + tables::kCorner[v][1],
                      vz = cz + tables::kCorner[v][2];
        const float va = lattice_value<kFromGrid>(A, L, ux, uy, uz);
        const float vb = lattice_value<kFromGrid>(A, L, vx, vy, vz);
        float p[3];
        crossing(va, vb, lattice_coord(L, 0, ux), lattice_coord(L, 1, uy), lattice_coord(L, 2, uz),
                 lattice_coord(L, 0, vx), lattice_coord(L, 1, vy), lattice_coord(L, 2, vz), p);
        if (id < A.vcap) {
          A.vertices[3 * id + 0] = p[0];
          A.vertices[3 * id + 1] = p[1];
          A.vertices[3 * id + 2] = p[2];
        }
        if (id > 0xffffffffull) atomicOr(&A.totals[2], 1ull);
      }
    }

    // --- triangles (march_cube, marching_cubes_impl.rs:102-117)
    for (uint32_t j = 0; j < ntri; ++j) {
      const uint64_t tix = tri_base + j;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int e = s_t.tri[c0][3 * j + k];
        const uint32_t gm = static_cast<uint32_t>(kEdgeGeomPacked >> (5 * e)) & 31u;
        const int axis = static_cast<int>(gm >> 3);
        const int px = gm & 1, py = (gm >> 1) & 1, pz = (gm >> 2) & 1;
        // Owner of the lattice edge: step back by one cell along each
        // non-axis direction where the edge lies on this cell's low face
        // and a lower cell exists.
        const int rx = (axis != 0 && px == 0 && cx >= 1) ? -1 : 0;
        const int ry = (axis != 1 && py == 0 && cy >= 1) ? -1 : 0;
        const int rz = (axis != 2 && pz == 0 && cz >= 1) ? -1 : 0;
        const int relx = px - rx, rely = py - ry, relz = pz - rz;
        const int eo = c_geom.of_axis_min[axis][relx | (rely << 1) | (relz << 2)];
        // Select the owner row / x without dynamic register indexing.
        const bool y0 = ry == 0, z0 = rz == 0, x0s = rx == 0;
        const uint32_t oc = pick(x0s, y0, z0, R[0][0].case_xm, R[0][0].case_x, R[0][1].case_xm,
                                 R[0][1].case_x, R[1][0].case_xm, R[1][0].case_x,
                                 R[1][1].case_xm, R[1][1].case_x);
        const uint32_t oown = pick(x0s, y0, z0, R[0][0].own_xm, R[0][0].own_x, R[0][1].own_xm,
                                   R[0][1].own_x, R[1][0].own_xm, R[1][0].own_x,
                                   R[1][1].own_xm, R[1][1].own_x);
        const uint64_t obase = pick(x0s, y0, z0, R[0][0].base_xm, R[0][0].base_x,
                                    R[0][1].base_xm, R[0][1].base_x, R[1][0].base_xm,
                                    R[1][0].base_x, R[1][1].base_xm, R[1][1].base_x);
        const uint64_t id = obase + static_cast<uint32_t>(__popc(oown & s_t.before[oc][eo]));
        if (3 * tix + k < A.icap) A.indices[3 * tix + k] = static_cast<uint32_t>(id);
      }
    }
  }
}

// Dense program samples over a lattice (iso_program_sample_lattice).
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

// Program samples at arbitrary points (iso_program_sample).
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

}  // namespace

McDims mc_dims(const Lattice& L) {
  McDims D;
  D.n_points = L.nx * L.ny * L.nz;
  D.n_sign_words = (D.n_points + 63) / 64;
  D.n_cell_rows = (L.ny - 1) * (L.nz - 1);
  D.segs_per_row = (L.nx - 1 + 63) / 64;
  D.n_segs = D.n_cell_rows * D.segs_per_row;
  D.n_scan_blocks = (D.n_segs + kScanBlock - 1) / kScanBlock;
  return D;
}

hipError_t launch_signs_grid(const float* grid, const Lattice& L, const McDims& D,
                             uint64_t* signs, hipStream_t s) {
  const int blocks = grid_for((D.n_sign_words + kSignUnroll - 1) / kSignUnroll);
  hipLaunchKernelGGL(k_signs_grid, dim3(blocks), dim3(256), 0, s, grid, D.n_points,
                     D.n_sign_words, signs);
  return hipGetLastError();
}

hipError_t launch_signs_program(const DevOp* ops, int n_ops, const Lattice& L, const McDims& D,
                                uint64_t* signs, hipStream_t s) {
  const int blocks = grid_for(D.n_sign_words);
  hipLaunchKernelGGL(k_signs_program, dim3(blocks), dim3(256), 0, s, ops, n_ops, L, D.n_points,
                     D.n_sign_words, signs);
  return hipGetLastError();
}

hipError_t launch_count(const Lattice& L, const McDims& D, const McBuffers& B, hipStream_t s) {
  hipLaunchKernelGGL(k_count, dim3(grid_for(D.n_segs)), dim3(256), 0, s, B.signs, L, D, B.seg);
  return hipGetLastError();
}

hipError_t launch_scan_blocks(const McDims& D, const McBuffers& B, hipStream_t s) {
  hipLaunchKernelGGL(k_scan_blocks, dim3(static_cast<unsigned>(D.n_scan_blocks)), dim3(256), 0, s,
                     B.seg, D.n_segs, B.blocks, D.n_scan_blocks);
  return hipGetLastError();
}

hipError_t launch_scan_top(const McDims& D, const McBuffers& B, hipStream_t s) {
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, B.blocks, D.n_scan_blocks, B.totals);
  return hipGetLastError();
}

hipError_t launch_emit(const float* grid, const DevOp* ops, int n_ops, const Lattice& L,
                       const McDims& D, const McBuffers& B, hipStream_t s) {
  EmitArgs A;
  A.S = B.signs;
  A.seg = B.seg;
  A.blocks = B.blocks;
  A.grid = grid;
  A.ops = ops;
  A.n_ops = n_ops;
  A.vertices = B.vertices;
  A.indices = B.indices;
  A.vcap = B.vertex_capacity;
  A.icap = B.index_capacity;
  A.totals = B.totals;
  const int blocks = grid_for(D.n_segs);
  if (grid) {
    hipLaunchKernelGGL(k_emit<true>, dim3(blocks), dim3(256), 0, s, A, L, D);
  } else {
    hipLaunchKernelGGL(k_emit<false>, dim3(blocks), dim3(256), 0, s, A, L, D);
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
