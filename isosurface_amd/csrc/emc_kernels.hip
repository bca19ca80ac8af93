// ExtendedMarchingCubes kernels (extended_marching_cubes.rs:59-186); the
// stages are described in emc_kernels.h.  Every f32 operation follows the
// reference's order (the library is built with -ffp-contract=off), so the
// vertices, the feature points (TangentPlanes + MinimiseQEF, the f64 SVD of
// qef_device.h over the cell's crossings) and the index lists are
// bit-identical to the reference's.
#include <hip/hip_runtime.h>

#include "emc_kernels.h"
#include "mc_device.h"
#include "mc_tables.h"
#include "qef_device.h"
#include "sdf_eval.h"

namespace iso {
namespace {

using namespace mcd;

__constant__ __align__(16) tables::McTables c_mct = tables::kTables;
__constant__ __align__(16) tables::LoopTables c_loops = tables::kLoops;

struct CellAt {
  int32_t x, y, z;
  int64_t f;  // flat index of corner 0
};

__device__ __forceinline__ CellAt decode(const EmcArgs& A, uint32_t rec) {
  const uint32_t g = rec >> 6;
  const uint32_t spr = A.segs_per_row;
  const uint32_t cyn = static_cast<uint32_t>(A.L.ny - 1);
  const uint32_t row = g / spr;
  const uint32_t w = g - row * spr;
  CellAt c;
  c.z = static_cast<int32_t>(row / cyn);
  c.y = static_cast<int32_t>(row - static_cast<uint32_t>(c.z) * cyn);
  c.x = static_cast<int32_t>(64 * w + (rec & 63u));
  c.f = (static_cast<int64_t>(c.z) * A.L.ny + c.y) * A.L.nx + c.x;
  return c;
}

__device__ __forceinline__ uint32_t encode(const EmcArgs& A, int32_t x, int32_t y, int32_t z) {
  const uint32_t cyn = static_cast<uint32_t>(A.L.ny - 1);
  const uint32_t row = static_cast<uint32_t>(z) * cyn + static_cast<uint32_t>(y);
  const uint32_t g = row * A.segs_per_row + (static_cast<uint32_t>(x) >> 6);
  return (g << 6) | (static_cast<uint32_t>(x) & 63u);
}

// Index of a cell in the active list (sorted by cell), or -1.
__device__ __forceinline__ int64_t find_cell(const EmcArgs& A, uint32_t key) {
  uint64_t lo = 0, hi = A.n_act;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (A.active[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < A.n_act && A.active[lo] == key) ? static_cast<int64_t>(lo) : -1;
}

__device__ __forceinline__ float point_coord(const EmcArgs& A, int axis, int32_t i) {
  return lattice_coord(A.L, axis, i);
}

// The Directed crossing on local edge e (distance.rs:89-102): the component
// of the axis (p_a - p_b).abs().max_component_index() is interpolated.
__device__ __forceinline__ void edge_crossing(const EmcArgs& A, const CellAt& c, int e, float* p) {
  const int u = tables::kEdgeCorners[e][0], v = tables::kEdgeCorners[e][1];
  const int64_t nx = A.L.nx, nxy = A.L.nx * A.L.ny;
  const float pax = point_coord(A, 0, c.x + tables::kCorner[u][0]);
  const float pay = point_coord(A, 1, c.y + tables::kCorner[u][1]);
  const float paz = point_coord(A, 2, c.z + tables::kCorner[u][2]);
  const float pbx = point_coord(A, 0, c.x + tables::kCorner[v][0]);
  const float pby = point_coord(A, 1, c.y + tables::kCorner[v][1]);
  const float pbz = point_coord(A, 2, c.z + tables::kCorner[v][2]);
  const float dx = __builtin_fabsf(pax - pbx), dy = __builtin_fabsf(pay - pby),
              dz = __builtin_fabsf(paz - pbz);
  const int axis = (dx > dy && dx > dz) ? 0 : (dy > dz ? 1 : 2);
  const int64_t fu = c.f + tables::kCorner[u][2] * nxy + tables::kCorner[u][1] * nx +
                     tables::kCorner[u][0];
  const int64_t fv = c.f + tables::kCorner[v][2] * nxy + tables::kCorner[v][1] * nx +
                     tables::kCorner[v][0];
  crossing(A.values[3 * fu + axis], A.values[3 * fv + axis], pax, pay, paz, pbx, pby, pbz, p);
}

// Rank of the cell's vertex for local edge e among the vertices it creates:
// a feature cell creates its feature vertex first, then its edges in the
// order its fans first use them; a planar cell in the order its triangles
// first use them (mesh.rs:240-251).
__device__ __forceinline__ uint32_t rank_in_cell(uint32_t cube, bool feature, uint32_t own,
                                                 int e) {
  return feature ? 1u + static_cast<uint32_t>(__popc(own & c_loops.before[cube][e]))
                 : static_cast<uint32_t>(__popc(own & c_mct.before[cube][e]));
}

// Global vertex id of local edge e of cell c (active index i): resolved in
// the cell that owns the edge (one back along each non-axis direction where
// the edge lies on the low face and a lower cell exists).
__device__ __forceinline__ uint32_t vertex_id(const EmcArgs& A, const CellAt& c, uint64_t i,
                                              uint32_t cube, bool feature, uint32_t own, int e) {
  const uint32_t gm = static_cast<uint32_t>(kEdgeGeomPacked >> (5 * e)) & 31u;
  const int axis = static_cast<int>(gm >> 3);
  const int px = gm & 1, py = (gm >> 1) & 1, pz = (gm >> 2) & 1;
  const int rx = (axis != 0 && px == 0 && c.x >= 1) ? 1 : 0;
  const int ry = (axis != 1 && py == 0 && c.y >= 1) ? 1 : 0;
  const int rz = (axis != 2 && pz == 0 && c.z >= 1) ? 1 : 0;
  if (!(rx | ry | rz)) return A.vfirst[i] + rank_in_cell(cube, feature, own, e);
  const int32_t ox = c.x - rx, oy = c.y - ry, oz = c.z - rz;
  const int eo = edge_of_axis_min(axis, (px + rx) | ((py + ry) << 1) | ((pz + rz) << 2));
  const int64_t j = find_cell(A, encode(A, ox, oy, oz));
  if (j < 0) return 0xffffffffu;  // not reached: the owner holds a crossed edge
  const uint32_t info = A.cells[j].info;
  const uint32_t cube_o = info & 255u;
  const bool feature_o = (info >> 8) & 1u;
  const uint32_t own_o = owned_edges(ox, oy, oz) & c_mct.edge_mask[cube_o];
  return A.vfirst[j] + rank_in_cell(cube_o, feature_o, own_o, eo);
}

// ---------------------------------------------------------------------------
// 1. per-cell tangent planes and feature points, fused with the id scan
// ---------------------------------------------------------------------------
// Cube index of an active cell from its corners' directed values
// (Directed::is_positive, distance.rs:76-79).
__device__ __forceinline__ uint32_t cell_cube(const EmcArgs& A, const CellAt& c) {
  const int64_t nx = A.L.nx, nxy = A.L.nx * A.L.ny;
  uint32_t cube = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int64_t f = c.f + tables::kCorner[k][2] * nxy + tables::kCorner[k][1] * nx +
                      tables::kCorner[k][0];
    const float* v = A.values + 3 * f;
    if (!(v[0] > 0.0f || v[1] > 0.0f || v[2] > 0.0f)) cube |= 1u << k;
  }
  return cube;
}

// The crossing and normal of edge e of active cell i, one thread each
// (marching_cubes_impl.rs:39-90, TangentPlanes::from_edge_crossings,
// feature/mod.rs:77-89), so the program evaluations of a cell run side by
// side instead of in its scan item's chain.
__global__ __launch_bounds__(256) void k_emc_crossings(EmcArgs A) {
  const uint64_t t = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (t >= 12 * A.n_act) return;
  const uint64_t i = t / 12;
  const int e = static_cast<int>(t - 12 * i);
  const CellAt c = decode(A, A.active[i]);
  if (!((c_mct.edge_mask[cell_cube(A, c)] >> e) & 1u)) return;
  float p[3], raw[3];
  edge_crossing(A, c, e, p);
  if (A.epsilon > 0.0f)
    sdf_central_difference(A.ops, A.n_ops, p[0], p[1], p[2], A.epsilon, raw);
  else
    sdf_eval_hermite(A.ops, A.n_ops, p[0], p[1], p[2], raw);
  const qef::Vec3f n = qef::normalised_or_zero(qef::Vec3f{raw[0], raw[1], raw[2]});
  float* q = A.crossings + 6 * t;
  q[0] = p[0];
  q[1] = p[1];
  q[2] = p[2];
  q[3] = n.x;
  q[4] = n.y;
  q[5] = n.z;
}

struct CellLoad {
  EmcArgs A;
  __device__ unsigned long long operator()(uint64_t i) const {
    const CellAt c = decode(A, A.active[i]);
    const uint32_t cube = cell_cube(A, c);
    const uint32_t mask = c_mct.edge_mask[cube];
    // Crossings and their normals in edge order.
    qef::Vec3f pts[12], nrm[12];
    int m = 0;
    for (int e = 0; e < 12; ++e) {
      if (!((mask >> e) & 1u)) continue;
      const float* q = A.crossings + 6 * (12 * i + e);
      pts[m] = qef::Vec3f{q[0], q[1], q[2]};
      nrm[m] = qef::Vec3f{q[3], q[4], q[5]};
      ++m;
    }
    bool feature = false, stalled = false;
    qef::Vec3f fp{0.0f, 0.0f, 0.0f};
    if (m > 0) fp = qef::place_feature_n<qef::kMaxRows>(pts, nrm, m, &stalled, &feature);
    if (stalled) atomicOr(&A.counters[kEmcStalled], 1ull);
    const uint32_t own = owned_edges(c.x, c.y, c.z) & mask;
    const uint32_t nv = static_cast<uint32_t>(__popc(own)) + (feature ? 1u : 0u);
    const uint32_t nt = feature ? c_loops.total[cube] : c_mct.tri_count[cube];
    EmcCell out;
    out.p[0] = fp.x;
    out.p[1] = fp.y;
    out.p[2] = fp.z;
    out.info = cube | (feature ? 1u << 8 : 0u) | (nv << 9) | (nt << 14);
    A.cells[i] = out;
    return (static_cast<unsigned long long>(nt) << 32) | nv;
  }
};

struct CellSink {
  uint32_t* vfirst;
  uint32_t* tfirst;
  __device__ void operator()(uint64_t i, unsigned long long, unsigned long long ex) const {
    vfirst[i] = static_cast<uint32_t>(ex & 0xffffffffull);
    tfirst[i] = static_cast<uint32_t>(ex >> 32);
  }
};

// ---------------------------------------------------------------------------
// 2. vertices and faces (march_cube / the fans of march_cube_extended)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_emc_emit(EmcArgs A, uint64_t n_tris) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  const uint64_t tid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  for (uint64_t t = tid; t < n_tris; t += stride) A.flips[t] = make_uint4(0u, 0u, 0u, 0u);
  for (uint64_t i = tid; i < A.n_act; i += stride) {
    const CellAt c = decode(A, A.active[i]);
    const EmcCell cell = A.cells[i];
    const uint32_t cube = cell.info & 255u;
    const bool feature = (cell.info >> 8) & 1u;
    const uint32_t mask = c_mct.edge_mask[cube];
    const uint32_t own = owned_edges(c.x, c.y, c.z) & mask;
    const uint32_t base = A.vfirst[i];
    const uint64_t tb = A.tfirst[i];
    if (feature && base < A.vcap) {
      A.vertices[3 * static_cast<uint64_t>(base) + 0] = cell.p[0];
      A.vertices[3 * static_cast<uint64_t>(base) + 1] = cell.p[1];
      A.vertices[3 * static_cast<uint64_t>(base) + 2] = cell.p[2];
    }
    for (uint32_t todo = own; todo; todo &= todo - 1) {
      const int e = __builtin_ctz(todo);
      const uint64_t id = base + rank_in_cell(cube, feature, own, e);
      if (id < A.vcap) {
        float p[3];
        edge_crossing(A, c, e, p);
        A.vertices[3 * id + 0] = p[0];
        A.vertices[3 * id + 1] = p[1];
        A.vertices[3 * id + 2] = p[2];
      }
    }
    uint32_t ids[12];
    for (int e = 0; e < 12; ++e)
      ids[e] = ((mask >> e) & 1u) ? vertex_id(A, c, i, cube, feature, own, e) : 0u;
    if (!feature) {
      const uint32_t n = 3u * c_mct.tri_count[cube];
      for (uint32_t k = 0; k < n; ++k) {
        const uint64_t at = 3 * tb + k;
        if (at < A.icap) A.indices[at] = ids[c_mct.tri[cube][k]];
      }
    } else {
      uint64_t face = tb;
      int off = 0;
      for (int l = 0; l < c_loops.count[cube]; ++l) {
        const int size = c_loops.size[cube][l];
        for (int q = 0; q < size; ++q) {
          const int e0 = c_loops.edge[cube][off + q];
          const int e1 = c_loops.edge[cube][off + (q + 1 == size ? 0 : q + 1)];
          if (3 * face + 2 < A.icap) {
            A.indices[3 * face + 0] = ids[e0];
            A.indices[3 * face + 1] = ids[e1];
            A.indices[3 * face + 2] = base;
          }
          ++face;
        }
        off += size;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// 3. flip_feature_edges (extended_marching_cubes.rs:158-185 with
//    MeshTopology::rotate_edge, mesh.rs:137-170)
// ---------------------------------------------------------------------------

// The cube face shared by local edges e0 and e1 (axis k, side s), if any.
__device__ __forceinline__ bool common_face(int e0, int e1, int* k_out, int* side) {
  const uint32_t g0 = static_cast<uint32_t>(kEdgeGeomPacked >> (5 * e0)) & 31u;
  const uint32_t g1 = static_cast<uint32_t>(kEdgeGeomPacked >> (5 * e1)) & 31u;
  const int a0 = static_cast<int>(g0 >> 3), a1 = static_cast<int>(g1 >> 3);
  for (int k = 0; k < 3; ++k) {
    if (k == a0 || k == a1) continue;
    const int s0 = (g0 >> k) & 1, s1 = (g1 >> k) & 1;
    if (s0 == s1) {
      *k_out = k;
      *side = s0;
      return true;
    }
  }
  return false;
}

struct FaceRef {
  uint64_t t;
  uint32_t opp;
  bool feature_opp;
};

// Faces of the range [t0, t0 + n) that contain the mesh edge {a, b}.
__device__ __forceinline__ void collect(const EmcArgs& A, uint64_t t0, uint32_t n, uint32_t a,
                                        uint32_t b, bool cell_feature, uint32_t centre,
                                        FaceRef* list, int* cnt) {
  for (uint32_t k = 0; k < n; ++k) {
    const uint64_t t = t0 + k;
    const uint32_t f0 = A.indices[3 * t], f1 = A.indices[3 * t + 1], f2 = A.indices[3 * t + 2];
    const bool ha = f0 == a || f1 == a || f2 == a;
    const bool hb = f0 == b || f1 == b || f2 == b;
    if (!(ha && hb)) continue;
    // Face::vertex_opposite: the first vertex that is neither a nor b
    const uint32_t opp = (f0 != a && f0 != b) ? f0 : ((f1 != a && f1 != b) ? f1 : f2);
    if (*cnt < 2) list[*cnt] = FaceRef{t, opp, cell_feature && opp == centre};
    ++*cnt;
  }
}

__global__ __launch_bounds__(256) void k_emc_flip_find(EmcArgs A) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < A.n_act;
       i += stride) {
    const uint32_t info = A.cells[i].info;
    if (!((info >> 8) & 1u)) continue;  // only fan faces can be flipped
    const CellAt c = decode(A, A.active[i]);
    const uint32_t cube = info & 255u;
    const uint32_t centre = A.vfirst[i];
    const uint64_t tb = A.tfirst[i];
    const uint32_t nt = c_loops.total[cube];
    uint64_t t_self = tb;
    int off = 0;
    for (int l = 0; l < c_loops.count[cube]; ++l) {
      const int size = c_loops.size[cube][l];
      for (int q = 0; q < size; ++q, ++t_self) {
        const int e0 = c_loops.edge[cube][off + q];
        const int e1 = c_loops.edge[cube][off + (q + 1 == size ? 0 : q + 1)];
        const uint32_t a = A.indices[3 * t_self], b = A.indices[3 * t_self + 1];
        FaceRef list[2];
        int cnt = 0;
        collect(A, tb, nt, a, b, true, centre, list, &cnt);
        // The only other cell holding both vertices: across the cube face
        // the two edges share, if they share one and it is inside.
        int k, side;
        if (common_face(e0, e1, &k, &side)) {
          int32_t q3[3] = {c.x, c.y, c.z};
          q3[k] += side ? 1 : -1;
          const int64_t lim[3] = {A.L.nx - 1, A.L.ny - 1, A.L.nz - 1};
          if (q3[k] >= 0 && q3[k] < lim[k]) {
            const int64_t j = find_cell(A, encode(A, q3[0], q3[1], q3[2]));
            if (j >= 0) {
              const uint32_t qi = A.cells[j].info;
              collect(A, A.tfirst[j], (qi >> 14) & 31u, a, b, (qi >> 8) & 1u, A.vfirst[j], list,
                      &cnt);
            }
          }
        }
        if (cnt != 2 || !list[0].feature_opp || !list[1].feature_opp) continue;
        if (list[1].t < list[0].t) {
          const FaceRef tmp = list[0];
          list[0] = list[1];
          list[1] = tmp;
        }
        if (list[0].t != t_self) continue;  // the other face's thread rotates it
        // Edge::new orders the handles; Face::matches_winding_direction
        const uint32_t s = a < b ? a : b, e = a < b ? b : a;
        const uint64_t ta = list[0].t;
        const uint32_t fa[3] = {A.indices[3 * ta], A.indices[3 * ta + 1], A.indices[3 * ta + 2]};
        bool matches = false;
        for (int x = 0; x < 3; ++x) {
          const int y = x == 2 ? 0 : x + 1;
          if (s == fa[x] && e == fa[y]) {
            matches = true;
            break;
          }
          if (s == fa[y] && e == fa[x]) break;
        }
        const uint32_t u = matches ? e : s, v = matches ? s : e;
        const uint32_t cc = list[0].opp, d = list[1].opp;
        A.flips[list[0].t] = make_uint4(cc, d, u, 1u);
        A.flips[list[1].t] = make_uint4(cc, v, d, 1u);
      }
      off += size;
    }
  }
}

__global__ __launch_bounds__(256) void k_emc_flip_apply(EmcArgs A, uint64_t n_tris) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n_tris;
       t += stride) {
    const uint4 r = A.flips[t];
    if (r.w) {
      A.indices[3 * t + 0] = r.x;
      A.indices[3 * t + 1] = r.y;
      A.indices[3 * t + 2] = r.z;
    }
  }
}

unsigned grid_of(uint64_t n) {
  uint64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > 16384) b = 16384;
  return static_cast<unsigned>(b);
}

}  // namespace

hipError_t launch_emc_crossings(const EmcArgs& A, hipStream_t s) {
  if (A.n_act == 0) return hipSuccess;
  if (!A.crossings) return hipErrorInvalidValue;
  const uint64_t blocks = (12 * A.n_act + 255) / 256;
  if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_emc_crossings, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, A);
  return hipGetLastError();
}

hipError_t launch_emc_cells(const EmcArgs& A, const scan::State& st, hipStream_t s) {
  return scan::launch<1>(A.n_act, CellLoad{A}, CellSink{A.vfirst, A.tfirst}, st, s);
}

hipError_t launch_emc_emit(const EmcArgs& A, uint64_t n_triangles, hipStream_t s) {
  const uint64_t n = A.n_act > n_triangles ? A.n_act : n_triangles;
  hipLaunchKernelGGL(k_emc_emit, dim3(grid_of(n)), dim3(256), 0, s, A, n_triangles);
  return hipGetLastError();
}

hipError_t launch_emc_flips(const EmcArgs& A, uint64_t n_triangles, hipStream_t s) {
  hipLaunchKernelGGL(k_emc_flip_find, dim3(grid_of(A.n_act)), dim3(256), 0, s, A);
  hipLaunchKernelGGL(k_emc_flip_apply, dim3(grid_of(n_triangles)), dim3(256), 0, s, A,
                     n_triangles);
  return hipGetLastError();
}

}  // namespace iso
