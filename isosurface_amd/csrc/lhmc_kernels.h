// Launch interface of the linear-hashed octree marching-cubes kernels.
//
// The octree is built breadth first, one launch per depth; its nodes live in
// one array sorted by Morton key (depth d's keys all lie in [8^d, 8^(d+1)),
// and children of sorted parents come out sorted).  Primal vertices are
// listed in ascending key order: for shallow trees by marking a dense
// lattice at depth max_depth + 1 indexed by Morton code and compacting it,
// for deeper ones by sorting the leaves' (vertex, level) pairs.  Dual cubes
// are then processed in that order with an open-addressing table over their
// edges.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "scan.h"
#include "sdf_program.h"

namespace iso {

// Nodes reach depth max_depth + 1, whose keys need 3 (max_depth + 1) + 1
// bits: 20 is the deepest whose keys fit 64 bits (the reference's MAX_LEVEL
// is 21).
constexpr int kLhmcMaxDepth = 20;
constexpr int kLhmcMaxLevels = kLhmcMaxDepth + 2;  // depths 0 .. max(2, max_depth) + 1
// Trees whose dense primal lattice (8^(max(2, max_depth) + 1) words) would
// exceed this many levels take the sorted primal path.
constexpr int kLhmcDenseDepth = 9;

// Status bits written by the LHMC kernels.
enum : unsigned long long {
  kLhmcErrScan = 2ull,        // a look-back gave up (scan::State::err)
  kLhmcErrWalk = 4ull,        // a dual corner found no octree node
  kLhmcErrTable = 8ull,       // edge table full
  kLhmcErrOverflow = 16ull,   // ids beyond u32
};

struct LhmcNodes {
  uint64_t* keys;    // sorted Morton keys
  float* values;     // SDF at each node's (quirky) centre
  uint8_t* leaf;     // 1 if not refined
  int64_t level_off[kLhmcMaxLevels + 1];  // nodes of depth d: [level_off[d], level_off[d+1])
  int n_levels;
};

struct LhmcEdgeTable {
  unsigned long long* keys;  // (node a << 32) | node b, a < b; ~0 = empty
  uint32_t* owner;           // first cube (in order) using the edge
  uint32_t* id;              // vertex id
  uint64_t mask;             // capacity - 1 (power of two)
};

struct LhmcCubes {
  uint32_t n;                  // number of dual cubes (= primal vertices)
  uint64_t* primal;            // Morton key (MAX_LEVEL) of each primal vertex, ascending
  uint8_t* primal_level;       // its deepest touching leaf level
  uint32_t* corner;            // 8 node indices per cube
  uint8_t* cube_case;          // cube index
  unsigned long long* base;    // exclusive (vertices | triangles << 32) per cube
};

// Morton::level() (morton.rs:53-59) takes floor(ln(key) * 0.48089834696)
// in f64; the truncated constant puts the first split[d] keys of depth d,
// [8^d, 8^d + split[d]), one level low.  lhmc_level_splits() finds split[d]
// on the host with the same libm log the reference uses (1 for d <= 10, then
// 2, 11, 93, ... up to 2502362113 at d = 21); launch_lhmc_levels hands the
// table to the kernels.
void lhmc_level_splits(uint64_t split[kLhmcMaxLevels]);
hipError_t launch_lhmc_levels(const uint64_t split[kLhmcMaxLevels], hipStream_t s);

// Evaluates the nodes of one depth and writes the next depth's keys
// (children of the refined nodes, in order).  `next_capacity` bounds the
// writes; the number of children is 8 * st.ctl[1] after the scan.
hipError_t launch_lhmc_level(const DevOp* ops, int n_ops, int max_depth, uint64_t* keys,
                             float* values, uint8_t* leaf, int64_t off, int64_t n,
                             int64_t next_capacity, const scan::State& st, hipStream_t s);
// Marks primal vertices on the dense lattice of depth `dense_depth` =
// max(2, max_depth) + 1, the deepest a node can be (levels + 1, 0 = none),
// and lists them in ascending order; the count is st.ctl[1] afterwards.
hipError_t launch_lhmc_primal(const LhmcNodes& N, int dense_depth, uint32_t* dense,
                              uint64_t n_dense, const LhmcCubes& C, uint64_t capacity,
                              const scan::State& st, hipStream_t s);
// The same list without a dense lattice: every node writes 8 (vertex,
// level) pairs (key 0 for none) into pair_keys / pair_levels (8 * n_nodes
// each), a radix sort orders them into sorted_keys / sorted_levels, and the
// last pair of each run of equal keys (with the run's deepest level) is
// listed.  `temp` holds lhmc_sort_temp_bytes(8 * n_nodes) bytes.
size_t lhmc_sort_temp_bytes(uint64_t n_pairs);
hipError_t launch_lhmc_primal_sorted(const LhmcNodes& N, uint64_t* pair_keys,
                                     uint8_t* pair_levels, uint64_t* sorted_keys,
                                     uint8_t* sorted_levels, void* temp, size_t temp_bytes,
                                     const LhmcCubes& C, uint64_t capacity,
                                     const scan::State& st, hipStream_t s);
// Cube corners and cases; triangle counts summed into st.ctl[1].
hipError_t launch_lhmc_cubes(const LhmcNodes& N, int dense_depth, const LhmcCubes& C,
                             const scan::State& st, hipStream_t s);
// Edge ownership (first cube), then owned-vertex counts scanned with the
// triangle counts into C.base; totals in st.ctl[1].
hipError_t launch_lhmc_own(const LhmcNodes& N, const LhmcCubes& C, const LhmcEdgeTable& T,
                           unsigned long long* err, const scan::State& st, hipStream_t s);
// Vertex positions and index triples.
hipError_t launch_lhmc_emit(const LhmcNodes& N, const LhmcCubes& C, const LhmcEdgeTable& T,
                            float* vertices, uint64_t vcap, uint32_t* indices, uint64_t icap,
                            unsigned long long* err, hipStream_t s);
hipError_t launch_lhmc_table_clear(const LhmcEdgeTable& T, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);

}  // namespace iso
