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
  int64_t n_points;     // nx*ny*nz
  int64_t n_sign_words; // ceil(n_points/64)
  int64_t n_cell_rows;  // (ny-1)*(nz-1)
  int64_t segs_per_row; // ceil((nx-1)/64)
  int64_t n_segs;       // n_cell_rows * segs_per_row
  int64_t n_scan_blocks;
};

constexpr int kScanItems = 8;
constexpr int kScanBlock = 256 * kScanItems;  // segments per scan block

McDims mc_dims(const Lattice& L);

// Buffers of one extraction.  `signs` has one zero word before element 0 and
// two after the last.
struct McBuffers {
  uint64_t* signs;          // n_sign_words (+ padding)
  uint2* seg;               // n_segs: counts, then in-block exclusive prefix
  uint64_t* blocks;         // 2 * n_scan_blocks: exclusive block prefix (verts, tris)
  unsigned long long* totals;  // [0] verts, [1] tris, [2] error flags
  float* vertices;
  uint32_t* indices;
  uint64_t vertex_capacity;
  uint64_t index_capacity;
};

// Each returns hipSuccess or the launch error.
hipError_t launch_signs_grid(const float* grid, const Lattice& L, const McDims& D,
                             uint64_t* signs, hipStream_t s);
hipError_t launch_signs_program(const DevOp* ops, int n_ops, const Lattice& L, const McDims& D,
                                uint64_t* signs, hipStream_t s);
hipError_t launch_count(const Lattice& L, const McDims& D, const McBuffers& B, hipStream_t s);
hipError_t launch_scan_blocks(const McDims& D, const McBuffers& B, hipStream_t s);
hipError_t launch_scan_top(const McDims& D, const McBuffers& B, hipStream_t s);
hipError_t launch_emit(const float* grid, const DevOp* ops, int n_ops, const Lattice& L,
                       const McDims& D, const McBuffers& B, hipStream_t s);
hipError_t launch_sample_lattice(const DevOp* ops, int n_ops, const Lattice& L, float* out,
                                 hipStream_t s);
hipError_t launch_sample_points(const DevOp* ops, int n_ops, const float* pts, uint64_t n,
                                float* out, hipStream_t s);

}  // namespace iso
