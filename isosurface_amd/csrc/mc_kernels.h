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

constexpr int kCountBlock = 1024;  // threads per count workgroup (16 waves)
constexpr int kCountRounds = 1;    // segments per count thread
constexpr int kCountSegs = kCountBlock * kCountRounds;  // segments per count workgroup
// Narrow-band bricks: 2^shift cells per edge, shift 2 (4 x 4 x 4, the
// default) or 3 (8 x 8 x 8).
constexpr int kBandBrickShiftDefault = 2;
constexpr int kBandBrickShiftMax = 3;

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

// One lattice as the count and emit kernels see it.  A single extraction
// passes one as a kernel argument (its fields stay scalar); a batch keeps
// one per chunk in device memory and finds a segment's chunk by its global
// index.  The first vertex id of the lattice is the first id of its first
// segment, which the kernels subtract from the ids they write into index
// lists (zero for a single lattice), so a chunk's indices are chunk-local.
struct CellSpace {
  int64_t nx, ny, nz;
  float origin[3];
  float spacing[3];
  const uint64_t* signs;    // the lattice's flat sign bits
  uint32_t segs_per_row;
  uint32_t seg_off;         // global index of the lattice's first segment
  const DevOp* ops;         // program source
  int32_t n_ops;
  int32_t reserved;
  const float* grid;        // dense samples (a batch's stored program samples), or null
  const float* positions;   // per-point vertex positions, or null
};

// Buffers of one extraction.
struct McBuffers {
  uint64_t* signs;          // n_sign_words; signs[-1], signs[n], signs[n+1] are guard words
  uint2* seg;               // n_segs: global first (vertex, triangle) id of each segment
  uint64_t* tplanes;        // 3 * n_segs: per-cell triangle counts, bit-sliced
  uint64_t* vplanes;        // 4 * n_segs: per-cell owned-vertex counts (0..12), bit-sliced
  ScanTriple* agg;          // n_blocks: look-back aggregates, each word status << 62 | value
  ScanTriple* incl;         // n_blocks: look-back inclusive prefixes, same encoding
  uint32_t* active;         // active-cell list: (segment << 6) | bit, cell order
  uint64_t active_capacity;
  unsigned long long* totals;  // kTotSlots
  float* vertices;
  uint32_t* indices;
  uint64_t vertex_capacity;
  uint64_t index_capacity;
  // When set (host-coherent memory, device address): emit copies the
  // extraction's totals there, so the host reads them without a copy.
  unsigned long long* host_totals;
};

// Kernel timing: the next launch by one of the single-kernel launchers below
// (signs, count, emit) records `start` and `stop` at the kernel's own start
// and end (hipExtLaunchKernelGGL) and takes the pair; a launch that does not
// take it leaves it pending (launch_events_pending) for the caller to clear.
void set_launch_events(hipEvent_t start, hipEvent_t stop);
bool launch_events_pending();
void clear_launch_events();
// Takes the pending pair (false if none), for launchers in other files.
bool take_launch_events(hipEvent_t* start, hipEvent_t* stop);

// Each returns hipSuccess or the launch error.  The signs launchers also
// reset the per-extraction state (guard words, totals, look-back words).
hipError_t launch_signs_grid(const float* grid, const Lattice& L, const McDims& D,
                             const McBuffers& B, hipStream_t s);
// With `values` (n_points floats) the program's samples are stored too.
hipError_t launch_signs_program(const DevOp* ops, int n_ops, const Lattice& L, const McDims& D,
                                const McBuffers& B, float* values, hipStream_t s);
hipError_t launch_count(const Lattice& L, const McDims& D, const McBuffers& B, hipStream_t s);
// With `positions` (3 floats per lattice point, requires `grid`) vertices
// interpolate between those positions instead of lattice coordinates.
hipError_t launch_emit(const float* grid, const DevOp* ops, int n_ops, const float* positions,
                       const Lattice& L, const McDims& D, const McBuffers& B, hipStream_t s);
// MarchingCubes::<Directed>: sign bits from directed values (3 floats per
// lattice point, x fastest) — evaluated from `ops` and stored into `values`,
// or, with ops null, read from `values` — and emit interpolating along each
// edge's axis component.
hipError_t launch_signs_directed(const DevOp* ops, int n_ops, const Lattice& L, const McDims& D,
                                 const McBuffers& B, float* values, hipStream_t s);
hipError_t launch_emit_directed(const float* values, const Lattice& L, const McDims& D,
                                const McBuffers& B, hipStream_t s);
// Directed distances of a program at points (3 floats in and out per point).
hipError_t launch_sample_vector_points(const DevOp* ops, int n_ops, const float* pts, uint64_t n,
                                       float* out, hipStream_t s);
// PointCloud points (3 floats each, up to `capacity`) of the mixed cells in
// the active list, after launch_count.
hipError_t launch_point_cloud(const Lattice& L, const McDims& D, const McBuffers& B,
                              float* points, uint64_t capacity, hipStream_t s);
// Emit of a narrow-band extraction: corner values from `values` (stored by
// the band's exact sampling) for cells of bricks flagged exact (bit 0 of
// brick_flags, bricks of 2^brick_shift cells per edge, nbx x nby per z
// layer), the program elsewhere.
hipError_t launch_emit_band(const float* values, const uint8_t* brick_flags, int64_t nbx,
                            int64_t nby, int brick_shift, const DevOp* ops, int n_ops, const Lattice& L,
                            const McDims& D, const McBuffers& B, hipStream_t s);
hipError_t launch_sample_lattice(const DevOp* ops, int n_ops, const Lattice& L, float* out,
                                 hipStream_t s);

// A batch of independent chunk lattices extracted in one pass: the chunks'
// sign words and segments are concatenated, one count scan runs over all
// segments, and chunk c's vertex / triangle ids start at the first ids of
// its first segment.
struct BatchPlan {
  const CellSpace* chunks;     // device, n_chunks (signs = the chunk's first data word)
  const uint32_t* seg_start;   // device, n_chunks + 1
  const uint64_t* word_start;  // device, n_chunks + 1: first sign word of each chunk
  uint32_t n_chunks;
  int64_t n_words;   // sign words of all chunks
  int64_t n_segs;    // segments of all chunks
  // The chunks' `grid` pointers name a buffer of 64 samples per sign word
  // (the chunk's word w holds samples 64 w ..): the signs stage stores the
  // program's samples there and emit reads them.  False: no such buffer,
  // emit evaluates the programs at the edge endpoints.
  bool stored_values;
};

// Count workgroups of a batch at `rounds` segments per thread (1 or 2).
int64_t batch_count_blocks(const BatchPlan& P, int rounds);
// Program sign bits of every chunk (and the extraction's state reset, sized
// for the finest count tile); the grid gives each wavefront about
// words_per_wave runs of consecutive sign words.
hipError_t launch_batch_signs(const BatchPlan& P, const McBuffers& B, int words_per_wave,
                              hipStream_t s);
hipError_t launch_batch_count(const BatchPlan& P, const McBuffers& B, int rounds,
                              hipStream_t s);
hipError_t launch_batch_emit(const BatchPlan& P, const McBuffers& B, int blocks_per_cu,
                             hipStream_t s);
// voff / ioff (device, n_chunks + 1): each chunk's first vertex and index.
hipError_t launch_batch_offsets(const BatchPlan& P, const McBuffers& B, uint64_t* voff,
                                uint64_t* ioff, hipStream_t s);
hipError_t launch_sample_points(const DevOp* ops, int n_ops, const float* pts, uint64_t n,
                                float* out, hipStream_t s);

// Single-pass marching cubes over a dense grid (mc_fused.hip): tiles of R
// consecutive cell rows of one cell layer, TL tiles per layer, taken in
// order by persistent 512-thread workgroups; TL more tickets produce the
// sign rows of sample layer 0.
struct FusedPlan {
  int32_t R;             // cell rows per tile
  int32_t TL;            // tiles per cell layer
  int64_t n_tiles;       // TL * (nz - 1)
  int64_t n_tickets;     // n_tiles + TL
  int32_t wpr;           // sign words per sample row: ceil(nx / 64)
  int32_t spr;           // 64-cell segments per cell row: ceil((nx - 1) / 64)
  uint64_t lds_bytes;    // dynamic LDS per workgroup
  uint64_t sign_words;   // nz * ny * wpr
  uint64_t n_rows;       // cell rows: (nz - 1) * (ny - 1)
};
// Dynamic LDS bound: with the static tables and owner slots (29 KB) it
// keeps a workgroup within 64 KB (the 257^3 lattice's tiles need 31 KB).
constexpr uint64_t kFusedMaxLds = 34 * 1024;

// Slots of FusedBuffers::ctl.
enum : int {
  kFusedTicket = 0,     // two ticket counters: extraction e draws from kFusedTicket + (e & 1)
  kFusedVertices = 2,
  kFusedTriangles = 3,
  kFusedErrors = 4,     // epoch << 32 | bit 0: u32 id overflow, bit 1: a wait gave up
  kFusedCtlSlots = 8,
};

struct FusedBuffers {
  uint64_t* signs;               // sign_words: row-aligned sign bits (x fastest)
  unsigned long long* flags;     // n_tickets: epoch once the ticket's sign rows are published
  unsigned long long* slots;     // 4 * n_tickets: look-back aggregates, then inclusive prefixes
  unsigned long long* rowbase;   // n_rows: tagged first vertex id of each cell row
  unsigned long long* ctl;       // kFusedCtlSlots
  float* vertices;
  uint32_t* indices;
  uint64_t vcap, icap;
  // When set (host-coherent memory, kTotSlots words): the last tile writes
  // the vertex and triangle totals there (kTotVertices, kTotTriangles); 1
  // goes to kTotErrors on an id overflow and to kTotErrors + 1 when a wait
  // between tiles gave up (the caller zeroes both before the launch).
  unsigned long long* host_totals;
};

// False when the lattice does not fit the tiling (no cells, a row wider
// than 65,536 cells, or more LDS than kFusedMaxLds).
bool fused_plan(const Lattice& L, FusedPlan* P);
// One extraction; `epoch` must differ from the previous extraction's on
// the same buffers (scratch starts zeroed, so the first epoch is nonzero).
hipError_t launch_mc_fused(const float* grid, const Lattice& L, const FusedPlan& P,
                           const FusedBuffers& B, uint32_t epoch, hipStream_t s);

}  // namespace iso
