// Dual-contouring kernels for gfx950 (MI355X).
//
// DualContouring::extract (dual_contouring.rs:65-105) over
// DualGrid::traverse (traversal/dual_grid.rs:52-95) in three kernels before
// the marching-cubes pass over the dual lattice:
//   1. sample   the program at every primal lattice point;
//   2. dual     one thread per primal cell: dual value
//               lerp(v0, v6, 0.5) = 0.5*v0 + 0.5*v6, and either the centre
//               lerp(c0, c6, 0.5) as dual vertex (cube index 0 or 255) or an
//               entry in the mixed-cell list, in ascending cell order by a
//               single-pass scan;
//   3. feature  one thread per listed mixed cell: corner normals by central
//               differences of the program (source.rs:82-94), or with
//               epsilon 0 the program's own HermiteSource normals
//               (implicit/*.rs sample_normal, csg.rs:47-51), normalised or
//               zero (marching_cubes_impl.rs:58-69), then TangentPlanes and
//               the QEF solve (feature/mod.rs:91-148, feature/qef.rs:36-68)
//               with the reference's f64 SVD.
// Every result is written to its cell's own slot.
#include <hip/hip_runtime.h>

#include "dc_kernels.h"
#include "mc_tables.h"
#include "qef_device.h"
#include "sdf_eval.h"

namespace iso {

namespace {

__global__ __launch_bounds__(256) void k_dc_sample(const DevOp* __restrict__ ops, int n_ops,
                                                   Lattice L, int64_t n_points,
                                                   float* __restrict__ out,
                                                   unsigned long long* __restrict__ counters) {
  if (blockIdx.x == 0 && threadIdx.x < kDcSlots) counters[threadIdx.x] = 0ull;
  const uint32_t nx = static_cast<uint32_t>(L.nx);
  const uint32_t nxy = static_cast<uint32_t>(L.nx * L.ny);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  // Whole wavefronts step together (the culled evaluator works on all 64
  // lanes); lanes past the end evaluate the last point and store nothing.
  const int lane = static_cast<int>(threadIdx.x & 63u);
  for (int64_t f0 = static_cast<int64_t>(blockIdx.x) * blockDim.x + (threadIdx.x & ~63u);
       f0 < n_points; f0 += stride) {
    const int64_t f = f0 + lane;
    const uint32_t ff = static_cast<uint32_t>(f < n_points ? f : n_points - 1);
    const uint32_t z = ff / nxy;
    const uint32_t r = ff - z * nxy;
    const uint32_t y = r / nx;
    const uint32_t x = r - y * nx;
    const float v = sdf_eval_culled(ops, n_ops, lattice_coord(L, 0, x), lattice_coord(L, 1, y),
                                    lattice_coord(L, 2, z));
    if (f < n_points) out[f] = v;
  }
}

// One primal cell: its dual value and, unless mixed, its centre; returns
// whether it is mixed.
struct DualCell {
  Lattice L;
  const float* samples;
  float* dual_values;
  float* dual_points;
  __device__ bool operator()(uint64_t c) const {
    const int64_t nx = L.nx, ny = L.ny;
    const uint32_t cnx = static_cast<uint32_t>(nx - 1);
    const uint32_t cnxy = static_cast<uint32_t>((nx - 1) * (ny - 1));
    const uint32_t cc = static_cast<uint32_t>(c);
    const uint32_t z = cc / cnxy;
    const uint32_t r = cc - z * cnxy;
    const uint32_t y = r / cnx;
    const uint32_t x = r - y * cnx;
    float v[8];
    uint32_t cube = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int64_t px = x + tables::kCorner[i][0], py = y + tables::kCorner[i][1],
                    pz = z + tables::kCorner[i][2];
      v[i] = samples[(pz * ny + py) * nx + px];
      if (!(v[i] > 0.0f)) cube |= 1u << i;  // classify_corners, distance.rs:52-54
    }
    // Signed::lerp (distance.rs:56-58) with factor 0.5: (1 - 0.5) * a + 0.5 * b
    dual_values[c] = 0.5f * v[0] + 0.5f * v[6];
    const bool is_mixed = cube != 0u && cube != 255u;
    if (!is_mixed) {
      // Vec3::lerp (vector.rs:325-333) of corners 0 and 6
      const float c0x = lattice_coord(L, 0, x), c0y = lattice_coord(L, 1, y),
                  c0z = lattice_coord(L, 2, z);
      const float c6x = lattice_coord(L, 0, x + 1), c6y = lattice_coord(L, 1, y + 1),
                  c6z = lattice_coord(L, 2, z + 1);
      dual_points[3 * c + 0] = 0.5f * c0x + 0.5f * c6x;
      dual_points[3 * c + 1] = 0.5f * c0y + 0.5f * c6y;
      dual_points[3 * c + 2] = 0.5f * c0z + 0.5f * c6z;
    }
    return is_mixed;
  }
};

// Thread per primal cell (coalesced): dual values, centres, mixed flags.
__global__ __launch_bounds__(256) void k_dc_dual(DualCell cell, uint64_t n_cells,
                                                 uint8_t* __restrict__ mixed_flag) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t c = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < n_cells;
       c += stride)
    mixed_flag[c] = cell(c) ? 1u : 0u;
}

// Scan load: the mixed flag.
struct MixedLoad {
  const uint8_t* flag;
  __device__ unsigned long long operator()(uint64_t c) const { return flag[c]; }
};

// Scan sink: the mixed cells' list slots and, at the last cell, the count.
struct DualSink {
  uint64_t n;
  uint32_t* mixed;
  unsigned long long* counters;
  __device__ void operator()(uint64_t c, unsigned long long v, unsigned long long ex) const {
    if (v) mixed[ex] = static_cast<uint32_t>(c);
    if (c == n - 1) counters[kDcMixed] = ex + v;
  }
};

// CentralDifference::sample_normal (source.rs:82-94) at p, or with eps 0
// the program's HermiteSource::sample_normal, then
// normalised().unwrap_or_default() (marching_cubes_impl.rs:63-67).
__device__ __forceinline__ qef::Vec3f central_normal(const DevOp* ops, int n_ops, float px,
                                                     float py, float pz, float eps) {
  if (eps == 0.0f) {
    float raw[3];
    sdf_eval_hermite(ops, n_ops, px, py, pz, raw);
    return qef::normalised_or_zero(qef::Vec3f{raw[0], raw[1], raw[2]});
  }
  const float vx = sdf_eval(ops, n_ops, px + eps, py + 0.0f, pz + 0.0f) -
                   sdf_eval(ops, n_ops, px - eps, py - 0.0f, pz - 0.0f);
  const float vy = sdf_eval(ops, n_ops, px + 0.0f, py + eps, pz + 0.0f) -
                   sdf_eval(ops, n_ops, px - 0.0f, py - eps, pz - 0.0f);
  const float vz = sdf_eval(ops, n_ops, px + 0.0f, py + 0.0f, pz + eps) -
                   sdf_eval(ops, n_ops, px - 0.0f, py - 0.0f, pz - eps);
  const float e2 = 2.0f * eps;
  return qef::normalised_or_zero(qef::Vec3f{vx / e2, vy / e2, vz / e2});
}

// CentralDifference::sample_normal (source.rs:82-94) as the normal sinks use
// it (extractor.rs:46-69,97-127): not normalised.
__global__ __launch_bounds__(256) void k_sample_normals(const DevOp* __restrict__ ops, int n_ops,
                                                        const float* __restrict__ pts, uint64_t n,
                                                        float eps, float* __restrict__ out) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
  const float vx = sdf_eval(ops, n_ops, px + eps, py + 0.0f, pz + 0.0f) -
                   sdf_eval(ops, n_ops, px - eps, py - 0.0f, pz - 0.0f);
  const float vy = sdf_eval(ops, n_ops, px + 0.0f, py + eps, pz + 0.0f) -
                   sdf_eval(ops, n_ops, px - 0.0f, py - eps, pz - 0.0f);
  const float vz = sdf_eval(ops, n_ops, px + 0.0f, py + 0.0f, pz + eps) -
                   sdf_eval(ops, n_ops, px - 0.0f, py - 0.0f, pz - eps);
  const float e2 = 2.0f * eps;
  out[3 * i + 0] = vx / e2;
  out[3 * i + 1] = vy / e2;
  out[3 * i + 2] = vz / e2;
}

// Corner k of mixed cell c (z-major cell index over (nx-1) x (ny-1) cells).
__device__ __forceinline__ qef::Vec3f cell_corner(const Lattice& L, uint32_t c, int k) {
  const uint32_t cnx = static_cast<uint32_t>(L.nx - 1);
  const uint32_t cnxy = static_cast<uint32_t>((L.nx - 1) * (L.ny - 1));
  const uint32_t z = c / cnxy;
  const uint32_t r = c - z * cnxy;
  const uint32_t y = r / cnx;
  const uint32_t x = r - y * cnx;
  return {lattice_coord(L, 0, x + tables::kCorner[k][0]),
          lattice_coord(L, 1, y + tables::kCorner[k][1]),
          lattice_coord(L, 2, z + tables::kCorner[k][2])};
}

// Flat lattice index of corner k of cell c.
__device__ __forceinline__ uint32_t corner_point(const Lattice& L, uint32_t c, int k) {
  const uint32_t cnx = static_cast<uint32_t>(L.nx - 1);
  const uint32_t cnxy = static_cast<uint32_t>((L.nx - 1) * (L.ny - 1));
  const uint32_t z = c / cnxy;
  const uint32_t r = c - z * cnxy;
  const uint32_t y = r / cnx;
  const uint32_t x = r - y * cnx;
  return ((z + tables::kCorner[k][2]) * static_cast<uint32_t>(L.ny) + y + tables::kCorner[k][1]) *
             static_cast<uint32_t>(L.nx) +
         x + tables::kCorner[k][0];
}

// Scan load: whether lattice point p is a corner of a mixed cell (one of
// the up to eight cells around it is flagged).
struct PointLoad {
  Lattice L;
  const uint8_t* mixed_flag;
  __device__ unsigned long long operator()(uint64_t p) const {
    const uint32_t nx = static_cast<uint32_t>(L.nx), nxy = static_cast<uint32_t>(L.nx * L.ny);
    const uint32_t pp = static_cast<uint32_t>(p);
    const uint32_t z = pp / nxy;
    const uint32_t r = pp - z * nxy;
    const uint32_t y = r / nx;
    const uint32_t x = r - y * nx;
    const uint32_t cnx = nx - 1, cny = static_cast<uint32_t>(L.ny - 1),
                   cnz = static_cast<uint32_t>(L.nz - 1);
    unsigned long long any = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
      if ((dx && x == 0) || (dy && y == 0) || (dz && z == 0)) continue;
      const uint32_t cx = x - dx, cy = y - dy, cz = z - dz;
      if (cx >= cnx || cy >= cny || cz >= cnz) continue;
      any |= mixed_flag[(cz * cny + cy) * cnx + cx];
    }
    return any ? 1ull : 0ull;
  }
};

// Scan sink: the points' list slots and index map, and at the last point
// the count.
struct PointSink {
  uint64_t n;
  uint32_t* plist;
  uint32_t* pindex;
  unsigned long long* counters;
  __device__ void operator()(uint64_t p, unsigned long long v, unsigned long long ex) const {
    if (v) {
      plist[ex] = static_cast<uint32_t>(p);
      pindex[p] = static_cast<uint32_t>(ex);
    }
    if (p == n - 1) counters[kDcPoints] = ex + v;
  }
};

// The corner normals of the mixed cells, one thread per (cell, corner):
// normals[24 i + 3 k + a] for corner k of mixed cell i (no shared-corner
// list: for normals that cost one evaluation, listing the corners costs
// more than evaluating each cell's own).
__global__ __launch_bounds__(256) void k_dc_cell_normals(const DevOp* __restrict__ ops,
                                                         int n_ops, Lattice L, float eps,
                                                         const uint32_t* __restrict__ mixed,
                                                         uint64_t n_mixed,
                                                         float* __restrict__ normals) {
  const uint64_t t = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (t >= 8 * n_mixed) return;
  const qef::Vec3f p = cell_corner(L, mixed[t >> 3], static_cast<int>(t & 7));
  const qef::Vec3f n = central_normal(ops, n_ops, p.x, p.y, p.z, eps);
  normals[3 * t + 0] = n.x;
  normals[3 * t + 1] = n.y;
  normals[3 * t + 2] = n.z;
}

// The normal at each listed lattice point, one thread per point.
__global__ __launch_bounds__(256) void k_dc_normals(const DevOp* __restrict__ ops, int n_ops,
                                                    Lattice L, float eps,
                                                    const uint32_t* __restrict__ plist,
                                                    uint64_t n_points, float* __restrict__ normals) {
  const uint64_t u = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (u >= n_points) return;
  const uint32_t nx = static_cast<uint32_t>(L.nx), nxy = static_cast<uint32_t>(L.nx * L.ny);
  const uint32_t p = plist[u];
  const uint32_t z = p / nxy;
  const uint32_t r = p - z * nxy;
  const uint32_t y = r / nx;
  const uint32_t x = r - y * nx;
  const qef::Vec3f n = central_normal(ops, n_ops, lattice_coord(L, 0, x), lattice_coord(L, 1, y),
                                      lattice_coord(L, 2, z), eps);
  normals[3 * u + 0] = n.x;
  normals[3 * u + 1] = n.y;
  normals[3 * u + 2] = n.z;
}

constexpr int kFeatureThreads = 128;

__global__ __launch_bounds__(kFeatureThreads) void k_dc_feature(
    Lattice L, const uint32_t* __restrict__ mixed, uint64_t n_mixed,
    const uint32_t* __restrict__ pindex, const float* __restrict__ pnormals,
    float* __restrict__ dual_points, unsigned long long* __restrict__ counters) {
  bool stalled = false;
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kFeatureThreads + threadIdx.x;
  if (i < n_mixed) {
    const uint32_t c = mixed[i];
    qef::Vec3f corners[8], normals[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      corners[k] = cell_corner(L, c, k);
      const uint64_t slot = pindex ? pindex[corner_point(L, c, k)] : 8 * i + k;
      const float* q = pnormals + 3 * slot;
      normals[k] = {q[0], q[1], q[2]};
    }
    const qef::Vec3f p = qef::place_feature(corners, normals, &stalled);
    dual_points[3 * static_cast<uint64_t>(c) + 0] = p.x;
    dual_points[3 * static_cast<uint64_t>(c) + 1] = p.y;
    dual_points[3 * static_cast<uint64_t>(c) + 2] = p.z;
  }
  if (stalled) atomicOr(&counters[kDcStalled], 1ull);
}

}  // namespace

hipError_t launch_dc_sample(const DevOp* ops, int n_ops, const Lattice& L, const DcBuffers& B,
                            hipStream_t s) {
  if (!ops || !B.samples || !B.counters) return hipErrorInvalidValue;
  const int64_t n = L.nx * L.ny * L.nz;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_dc_sample, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, ops, n_ops,
                     L, n, B.samples, B.counters);
  return hipGetLastError();
}

hipError_t launch_dc_dual(const Lattice& L, const DcBuffers& B, scan::State st, hipStream_t s) {
  const uint64_t n_cells = static_cast<uint64_t>((L.nx - 1) * (L.ny - 1) * (L.nz - 1));
  if (n_cells == 0 || n_cells > 0xffffffffull || !st.slot || !st.ctl) return hipErrorInvalidValue;
  if (!B.mixed_flag) return hipErrorInvalidValue;
  st.err = B.counters + kDcScanErr;
  const DualCell cell{L, B.samples, B.dual_values, B.dual_points};
  uint64_t blocks = (n_cells + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_dc_dual, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, cell, n_cells,
                     B.mixed_flag);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const MixedLoad load{B.mixed_flag};
  const DualSink sink{n_cells, B.mixed, B.counters};
  return scan::launch<scan::kItems>(n_cells, load, sink, st, s);
}

hipError_t launch_sample_normals(const DevOp* ops, int n_ops, const float* points, uint64_t n,
                                 float epsilon, float* normals, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (!ops || !points || !normals) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_sample_normals, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                     s, ops, n_ops, points, n, epsilon, normals);
  return hipGetLastError();
}

hipError_t launch_dc_corner_points(const Lattice& L, const DcBuffers& B, uint32_t* plist,
                                   uint32_t* pindex, scan::State st, hipStream_t s) {
  const uint64_t n_points = static_cast<uint64_t>(L.nx * L.ny * L.nz);
  if (n_points == 0 || n_points > 0xffffffffull || !plist || !pindex || !B.mixed_flag ||
      !st.slot || !st.ctl)
    return hipErrorInvalidValue;
  st.err = B.counters + kDcScanErr;
  const PointLoad load{L, B.mixed_flag};
  const PointSink sink{n_points, plist, pindex, B.counters};
  return scan::launch<scan::kItems>(n_points, load, sink, st, s);
}

hipError_t launch_dc_normals(const DevOp* ops, int n_ops, const Lattice& L, float epsilon,
                             const uint32_t* plist, uint64_t n_points, float* normals,
                             hipStream_t s) {
  if (!ops || !plist || !normals) return hipErrorInvalidValue;
  if (n_points == 0) return hipSuccess;
  const uint64_t blocks = (n_points + 255) / 256;
  if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_dc_normals, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, ops,
                     n_ops, L, epsilon, plist, n_points, normals);
  return hipGetLastError();
}

hipError_t launch_dc_cell_normals(const DevOp* ops, int n_ops, const Lattice& L, float epsilon,
                                  const DcBuffers& B, uint64_t n_mixed, float* normals,
                                  hipStream_t s) {
  if (!ops || !normals) return hipErrorInvalidValue;
  if (n_mixed == 0) return hipSuccess;
  const uint64_t blocks = (8 * n_mixed + 255) / 256;
  if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_dc_cell_normals, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, ops,
                     n_ops, L, epsilon, B.mixed, n_mixed, normals);
  return hipGetLastError();
}

hipError_t launch_dc_feature(const Lattice& L, const DcBuffers& B, uint64_t n_mixed,
                             const uint32_t* pindex, const float* normals, hipStream_t s) {
  if (!normals) return hipErrorInvalidValue;
  if (n_mixed == 0) return hipSuccess;
  const uint64_t blocks = (n_mixed + kFeatureThreads - 1) / kFeatureThreads;
  if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_dc_feature, dim3(static_cast<unsigned>(blocks)), dim3(kFeatureThreads), 0,
                     s, L, B.mixed, n_mixed, pindex, normals, B.dual_points, B.counters);
  return hipGetLastError();
}

}  // namespace iso
