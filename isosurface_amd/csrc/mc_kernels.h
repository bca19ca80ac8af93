// Launch interface of the marching-cubes kernels (host side).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sdf_program.h"

namespace iso {

// Sample lattice as the kernels see it.
struct Lattice {
  int64_t nx, ny, nz;  // sample dims
  float origin[3];
  float spacing[3];
};

// Sizes derived from a lattice.
struct McDims {
  int64_t n_points;      // nx*ny*nz
  int64_t n_sign_words;  // ceil(n_points/64)
  int64_t n_cell_rows;   // (ny-1)*(nz-1)
  int64_t segs_per_row;  // ceil((nx-1)/64)
  int64_t n_segs;        // n_cell_rows * segs_per_row
  int64_t n_blocks;      // count workgroups (kCountBlock segments each)
  int64_t n_cells;       // (nx-1)*(ny-1)*(nz-1)
};

constexpr int kCountBlock = 256;  // segments per count workgroup

// Running totals of the decoupled look-back scan.
struct ScanTriple {
  unsigned long long v, t, a;  // vertices, triangles, active cells
};

// Slots of McBuffers::totals.
enum : int {
  kTotVertices = 0,
  kTotTriangles = 1,
  kTotActive = 2,
  kTotErrors = 3,  // bit 0: u32 id overflow, bit 1: look-back gave up
  kTotTicket = 4,  // workgroup ticket of the count scan
  kTotSlots = 8,
};

McDims mc_dims(const Lattice& L);

// Coordinate of lattice index i along an axis: origin + i * spacing, the
// reference's `x as f32 * one_over_size` when origin is 0
// (primal_grid.rs:50-51, 62-63).
__device__ __forceinline__ float lattice_coord(const Lattice& L, int axis, int64_t i) {
  return L.origin[axis] + static_cast<float>(i) * L.spacing[axis];
}

// Buffers of one extraction.
struct McBuffers {
  uint64_t* signs;          // n_sign_words; signs[-1], signs[n], signs[n+1] are guard words
  uint2* seg;               // n_segs: global first (vertex, triangle) id of each segment
  uint64_t* tplanes;        // 3 * n_segs: per-cell triangle counts, bit-sliced
  uint32_t* flags;          // n_blocks: look-back status (0, 1 aggregate, 2 inclusive)
  ScanTriple* agg;          // n_blocks
  ScanTriple* incl;         // n_blocks
  uint32_t* active;         // active-cell list: (segment << 6) | bit, cell order
  uint64_t active_capacity;
  unsigned long long* totals;  // kTotSlots
  float* vertices;
  uint32_t* indices;
  uint64_t vertex_capacity;
  uint64_t index_capacity;
};

// Each returns hipSuccess or the launch error.  The signs launchers also
// reset the per-extraction state (guard words, totals, look-back flags).
hipError_t launch_signs_grid(const float* grid, const Lattice& L, const McDims& D,
                             const McBuffers& B, hipStream_t s);
hipError_t launch_signs_program(const DevOp* ops, int n_ops, const Lattice& L, const McDims& D,
                                const McBuffers& B, hipStream_t s);
hipError_t launch_count(const Lattice& L, const McDims& D, const McBuffers& B, hipStream_t s);
// With `positions` (3 floats per lattice point, requires `grid`) vertices
// interpolate between those positions instead of lattice coordinates.
hipError_t launch_emit(const float* grid, const DevOp* ops, int n_ops, const float* positions,
                       const Lattice& L, const McDims& D, const McBuffers& B, hipStream_t s);
hipError_t launch_sample_lattice(const DevOp* ops, int n_ops, const Lattice& L, float* out,
                                 hipStream_t s);
hipError_t launch_sample_points(const DevOp* ops, int n_ops, const float* pts, uint64_t n,
                                float* out, hipStream_t s);

}  // namespace iso
