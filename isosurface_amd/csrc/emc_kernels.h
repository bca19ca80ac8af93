// ExtendedMarchingCubes on the GPU (extended_marching_cubes.rs:59-186).
//
// After the directed signs and count of marching cubes (the active list is
// the mixed cells in cell order):
//   1. per active cell, fused with an exclusive scan: the crossings of its
//      edges, their normals, the tangent-plane classification and, for a
//      feature cell, the QEF feature point; the cell's vertex count (owned
//      crossed edges, plus one for a feature vertex) and triangle count
//      (its triangle table, or one fan triangle per loop edge) are scanned
//      into first ids, which reproduce the reference's first-encounter
//      numbering (a feature vertex comes first in its cell, then the cell's
//      edges in the order its fans or triangles first use them);
//   2. emit: vertices and index triples, each edge's vertex resolved in the
//      cell that owns it (found by binary search of the active list);
//   3. the edge flips of flip_feature_edges: a fan edge whose two faces
//      both have feature vertices opposite is rotated; every such edge lies
//      in one cell or across one cube face, so each is found from one of its
//      faces (read-only pass into per-face records) and applied after.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mc_kernels.h"
#include "scan.h"

namespace iso {

// Per active cell: feature point and packed counts.
struct EmcCell {
  float p[3];
  uint32_t info;  // case | feature << 8 | vertices << 9 | triangles << 14
};

// counters[] slots
constexpr int kEmcStalled = 0;  // an SVD hit its sweep bound
constexpr int kEmcSlots = 4;

struct EmcArgs {
  Lattice L;
  uint32_t segs_per_row;
  const float* values;     // directed values, 3 per lattice point
  const uint32_t* active;  // mixed cells, (segment << 6) | bit, cell order
  uint64_t n_act;
  const DevOp* ops;
  int32_t n_ops;
  float epsilon;           // > 0: CentralDifference normals; 0: HermiteSource
  // 72 per active cell: per local edge e, the crossing point (floats 6e..6e+2)
  // and its normalised normal (6e+3..6e+5), written for crossed edges only
  float* crossings;
  EmcCell* cells;          // n_act
  uint32_t* vfirst;        // n_act: first vertex id of each cell
  uint32_t* tfirst;        // n_act: first triangle of each cell
  float* vertices;
  uint32_t* indices;
  uint64_t vcap, icap;     // in vertices / indices
  uint4* flips;            // per triangle: replacement (x, y, z, 1) or 0
  unsigned long long* counters;
};

// Stage 0: the crossing points and normals of every active cell's crossed
// edges, one thread per (cell, edge).
hipError_t launch_emc_crossings(const EmcArgs& A, hipStream_t s);
// Stage 1: cells + scan (state for n_act items at one item per thread).
// The totals land in scan state ctl[1]: triangles << 32 | vertices.
hipError_t launch_emc_cells(const EmcArgs& A, const scan::State& st, hipStream_t s);
// Stage 2 (needs n_triangles for the flip records it clears).
hipError_t launch_emc_emit(const EmcArgs& A, uint64_t n_triangles, hipStream_t s);
// Stage 3.
hipError_t launch_emc_flips(const EmcArgs& A, uint64_t n_triangles, hipStream_t s);

}  // namespace iso
