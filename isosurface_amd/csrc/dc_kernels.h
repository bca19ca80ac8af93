// Launch interface of the dual-contouring kernels (host side).
//
// Dual contouring is marching cubes on the dual lattice: one dual point per
// primal cell, carrying the cell's dual value (the lerp of its corners 0 and
// 6) and its dual vertex (the cell centre, or for a mixed cell the feature
// point placed by the QEF).  These kernels build that lattice; the
// marching-cubes kernels then extract it with per-point positions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mc_kernels.h"
#include "scan.h"
#include "sdf_program.h"

namespace iso {

// Slots of DcBuffers::counters.
enum : int {
  kDcMixed = 0,    // number of mixed primal cells listed
  kDcStalled = 1,  // nonzero when a QEF's SVD hit the sweep bound
  kDcScanErr = 2,  // nonzero when a scan's look-back gave up
  kDcPoints = 3,   // number of lattice points that are a corner of a mixed cell
  kDcSlots = 4,
};

struct DcBuffers {
  float* samples;       // primal lattice samples, nx*ny*nz
  float* dual_values;   // one per primal cell, z-major
  float* dual_points;   // 3 per primal cell
  uint32_t* mixed;      // mixed primal cells, capacity = number of cells
  uint8_t* mixed_flag;  // one per primal cell
  unsigned long long* counters;  // kDcSlots
};

// Samples the program over the primal lattice and clears the counters.
hipError_t launch_dc_sample(const DevOp* ops, int n_ops, const Lattice& L, const DcBuffers& B,
                            hipStream_t s);
// Dual values, centre dual vertices and mixed flags (thread per cell), then
// the mixed-cell list (ascending, by a single-pass scan over the flags whose
// state `st` holds scan::blocks_for(cells, 1) slots; its err word is
// redirected to counters[kDcScanErr]).
hipError_t launch_dc_dual(const Lattice& L, const DcBuffers& B, scan::State st, hipStream_t s);
// CentralDifference::sample_normal of the program at n points (3 floats
// each, device), unnormalised, into normals (3 floats each, device).
hipError_t launch_sample_normals(const DevOp* ops, int n_ops, const float* points, uint64_t n,
                                 float epsilon, float* normals, hipStream_t s);
// The lattice points that are a corner of a mixed cell (from the mixed
// flags), ascending: plist (capacity n_points) gets their indices, pindex
// (n_points) each one's position in plist, counters[kDcPoints] their count;
// a single-pass scan whose state `st` holds scan::blocks_for(n_points,
// scan::kItems) slots (its err word is redirected to counters[kDcScanErr]).
hipError_t launch_dc_corner_points(const Lattice& L, const DcBuffers& B, uint32_t* plist,
                                   uint32_t* pindex, scan::State st, hipStream_t s);
// The normal at each of the n_points listed points (central differences of
// step epsilon, or the program's analytic normals when epsilon is 0,
// normalised), 3 floats each, one thread per point: each corner shared by
// several mixed cells is evaluated once.
hipError_t launch_dc_normals(const DevOp* ops, int n_ops, const Lattice& L, float epsilon,
                             const uint32_t* plist, uint64_t n_points, float* normals,
                             hipStream_t s);
// The eight corner normals of each of the first n_mixed listed mixed cells,
// 24 floats per cell, one thread per (cell, corner).
hipError_t launch_dc_cell_normals(const DevOp* ops, int n_ops, const Lattice& L, float epsilon,
                                  const DcBuffers& B, uint64_t n_mixed, float* normals,
                                  hipStream_t s);
// Feature points of the first n_mixed listed mixed cells from their
// corners' normals (per listed point through pindex, or with pindex null
// per cell corner as launch_dc_cell_normals writes them), TangentPlanes and
// the QEF solve, one thread per cell.
hipError_t launch_dc_feature(const Lattice& L, const DcBuffers& B, uint64_t n_mixed,
                             const uint32_t* pindex, const float* normals, hipStream_t s);

}  // namespace iso
