// Marching-cubes kernels for gfx950 (MI355X).
//
// One extraction is three kernels:
//   1. signs  one bit per lattice sample, "inside" = !(v > 0) exactly as
//             classify_corners (marching_cubes_impl.rs:26-37, distance.rs:52-54).
//             From a dense FP32 grid this is a pure HBM stream (each load
//             instruction reads 256 contiguous bytes per wavefront and its
//             ballot is one sign word); from a program it is the SDF
//             evaluation.  The launch also resets the extraction's state.
//   2. count  one thread per 64-cell segment of a cell row: the segment's
//             eight corner bit-planes come from four row words, vertex
//             ownership counts are popcounts of XORed planes, triangles are
//             counted over the mixed cells only.  A single-pass decoupled
//             look-back scan turns the counts into global first ids and
//             writes the list of active (mixed) cells in cell order.
//   3. emit   one thread per active cell: its 3x3x3 sign neighbourhood holds
//             every cell that can own one of its edges; each needed owner's
//             first vertex id is resolved once, the cell's own vertices are
//             written (FP32 crossing points from the grid or the program)
//             and its index triples are written.
//
// Vertex numbering is the reference's first-encounter order
// (MeshTopologyBuilder::add_vertex, mesh.rs:240-251, driven by
// PrimalGrid::traverse's z-major cell order, primal_grid.rs:59-85): a
// lattice edge's vertex belongs to the lowest-numbered cell containing the
// edge, its position is computed along that cell's local edge direction, and
// vertices are numbered by (owner cell, first appearance in the owner's
// triangle list).  The owner of every edge follows from the edge's position
// alone, so no hash table is needed.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "mc_device.h"
#include "mc_kernels.h"
#include "mc_tables.h"
#include "sdf_eval.h"

namespace iso {

namespace {

using namespace mcd;

__constant__ __align__(16) tables::McTables c_tables = tables::kTables;
__constant__ __align__(16) EmitTables c_emit = make_emit_tables();

// ---------------------------------------------------------------------------
// 1. signs (+ reset of the extraction's scratch state)
// ---------------------------------------------------------------------------
constexpr int kSignWords = 16;  // sign words per wavefront iteration (1024 samples)

// Lane l of a wavefront reads samples 64j + l of a 1024-sample chunk (one
// coalesced 256-byte load per j), so the ballot of load j is sign word j
// itself: no bit interleaving, a compare and a lane select per word.  Works
// for any 4-byte aligned grid.
__global__ __launch_bounds__(256) void k_signs_grid(const float* __restrict__ g,
                                                    int64_t n_points, int64_t n_words,
                                                    uint64_t* __restrict__ S, ResetArgs R) {
  reset_state(R);
  const int lane = __lane_id();
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int64_t n_waves = static_cast<int64_t>(gridDim.x) * 4;
  for (int64_t w0 = wave * kSignWords; w0 < n_words; w0 += n_waves * kSignWords) {
    const float* p = g + w0 * 64 + lane;
    float v[kSignWords];
    if ((w0 + kSignWords) * 64 <= n_points) {  // uniform: the whole chunk is in the grid
#pragma unroll
      for (int j = 0; j < kSignWords; ++j) v[j] = __builtin_nontemporal_load(p + 64 * j);
    } else {
#pragma unroll
      for (int j = 0; j < kSignWords; ++j)
        v[j] = (w0 + j) * 64 + lane < n_points ? p[64 * j] : 1.0f;
    }
    uint64_t mine = 0;
#pragma unroll
    for (int j = 0; j < kSignWords; ++j) {
      const uint64_t m = __ballot(!(v[j] > 0.0f));  // classify_corners, distance.rs:52-54
      if (lane == j) mine = m;
    }
    if (lane < kSignWords && w0 + lane < n_words) S[w0 + lane] = mine;
  }
}

// Program signs; with `values` the samples are also stored (x fastest), so
// emit reads the endpoint values instead of evaluating the program again.
__global__ __launch_bounds__(256) void k_signs_program(const DevOp* __restrict__ ops, int n_ops,
                                                       Lattice L, int64_t n_points,
                                                       int64_t n_words,
                                                       uint64_t* __restrict__ S,
                                                       float* __restrict__ values, ResetArgs R) {
  reset_state(R);
  const int lane = __lane_id();
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int64_t n_waves = static_cast<int64_t>(gridDim.x) * 4;
  const uint32_t nx = static_cast<uint32_t>(L.nx);
  const uint32_t nxy = static_cast<uint32_t>(L.nx * L.ny);
  for (int64_t w = wave; w < n_words; w += n_waves) {
    const int64_t f = w * 64 + lane;
    // Every lane evaluates (the last word's spare lanes at the last point):
    // the culled evaluator works on the whole wavefront.
    const uint32_t ff = static_cast<uint32_t>(f < n_points ? f : n_points - 1);
    const uint32_t z = ff / nxy;
    const uint32_t r = ff - z * nxy;
    const uint32_t y = r / nx;
    const uint32_t x = r - y * nx;
    const float v = sdf_eval_culled(ops, n_ops, lattice_coord(L, 0, x), lattice_coord(L, 1, y),
                                    lattice_coord(L, 2, z));
    const bool inside = f < n_points && !(v > 0.0f);
    if (values && f < n_points) values[f] = v;
    const uint64_t m = __ballot(inside);
    if (lane == 0) S[w] = m;
  }
}

// Directed signs (MarchingCubes::<Directed>): a lattice point is inside when
// no component of its directed distance is positive (Directed::is_positive,
// distance.rs:76-79).  From a program (`ops` set: the values are evaluated
// and stored, 3 floats per point, for emit) or from stored values.
// Held to 64 VGPRs for eight wavefronts per SIMD (116 unconstrained VGPRs
// allow four): 147 us instead of 188 us for MarchingCubes::<Directed> at 128^3.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void
k_signs_directed(const DevOp* __restrict__ ops, int n_ops,
                 Lattice L, int64_t n_points,
                 int64_t n_words,
                 uint64_t* __restrict__ S,
                 float* __restrict__ values, ResetArgs R) {
  reset_state(R);
  const int lane = __lane_id();
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int64_t n_waves = static_cast<int64_t>(gridDim.x) * 4;
  const uint32_t nx = static_cast<uint32_t>(L.nx);
  const uint32_t nxy = static_cast<uint32_t>(L.nx * L.ny);
  for (int64_t w = wave; w < n_words; w += n_waves) {
    const int64_t f = w * 64 + lane;
    bool inside = false;
    if (f < n_points) {
      float v[3];
      if (ops) {
        const uint32_t ff = static_cast<uint32_t>(f);
        const uint32_t z = ff / nxy;
        const uint32_t r = ff - z * nxy;
        const uint32_t y = r / nx;
        const uint32_t x = r - y * nx;
        sdf_eval_vector(ops, n_ops, lattice_coord(L, 0, x), lattice_coord(L, 1, y),
                        lattice_coord(L, 2, z), v);
        values[3 * f + 0] = v[0];
        values[3 * f + 1] = v[1];
        values[3 * f + 2] = v[2];
      } else {
        v[0] = values[3 * f + 0];
        v[1] = values[3 * f + 1];
        v[2] = values[3 * f + 2];
      }
      inside = !(v[0] > 0.0f || v[1] > 0.0f || v[2] > 0.0f);
    }
    const uint64_t m = __ballot(inside);
    if (lane == 0) S[w] = m;
  }
}

// Batched chunks: a wavefront takes runs of kBatchRun consecutive sign words
// of the concatenated word numbering; it finds the first word's chunk by
// binary search and steps to the next chunk as the run crosses into it, so
// each word is evaluated with a single chunk's program and the search (a
// chain of dependent loads) is paid once per run.  Thread i < n_chunks also
// clears chunk i's guard words.  A chunk with a `grid` buffer gets its
// samples stored there for emit.
constexpr int kBatchRun = 8;

__global__ __launch_bounds__(256) void k_batch_signs(const CellSpace* __restrict__ chunks,
                                                     const uint64_t* __restrict__ word_start,
                                                     uint32_t n_chunks, ResetArgs R) {
  reset_state(R);
  const int64_t tid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (tid < n_chunks) {
    uint64_t* S = const_cast<uint64_t*>(chunks[tid].signs);
    const int64_t nw = static_cast<int64_t>(word_start[tid + 1] - word_start[tid]);
    S[-1] = 0ull;
    S[nw] = 0ull;
    S[nw + 1] = 0ull;
  }
  const int lane = __lane_id();
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int64_t n_waves = static_cast<int64_t>(gridDim.x) * 4;
  const int64_t n_words = static_cast<int64_t>(word_start[n_chunks]);
  // A word's chunk is the same for the whole wavefront: reading the table
  // entries through readfirstlane keeps the chunk index in a scalar
  // register, so the chunk record and its program are scalar loads and the
  // program's operands scalar operands.
  auto first_word = [&](uint32_t i) -> int64_t {
    const uint64_t v = word_start[i];
    const uint32_t lo32 = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
    const uint32_t hi32 = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
    return static_cast<int64_t>((static_cast<uint64_t>(hi32) << 32) | lo32);
  };
  for (int64_t w0 = wave * kBatchRun; w0 < n_words; w0 += n_waves * kBatchRun) {
    uint32_t lo = 0, hi = n_chunks;  // last chunk whose first word <= w0
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (first_word(mid) <= w0) lo = mid;
      else hi = mid;
    }
    const int64_t w1 = w0 + kBatchRun < n_words ? w0 + kBatchRun : n_words;
    for (int64_t w = w0; w < w1; ++w) {
      while (lo + 1 < n_chunks && first_word(lo + 1) <= w) ++lo;
      lo = __builtin_amdgcn_readfirstlane(lo);
      const CellSpace& C = chunks[lo];
      const int64_t lw = w - first_word(lo);
      const int64_t f = lw * 64 + lane;
      const int64_t n_points = C.nx * C.ny * C.nz;
      bool inside = false;
      if (f < n_points) {
        const uint32_t nx = static_cast<uint32_t>(C.nx);
        const uint32_t nxy = static_cast<uint32_t>(C.nx * C.ny);
        const uint32_t ff = static_cast<uint32_t>(f);
        const uint32_t z = ff / nxy;
        const uint32_t r = ff - z * nxy;
        const uint32_t y = r / nx;
        const uint32_t x = r - y * nx;
        const float v =
            sdf_eval(C.ops, C.n_ops, C.origin[0] + static_cast<float>(x) * C.spacing[0],
                     C.origin[1] + static_cast<float>(y) * C.spacing[1],
                     C.origin[2] + static_cast<float>(z) * C.spacing[2]);
        if (C.grid) const_cast<float*>(C.grid)[f] = v;
        inside = !(v > 0.0f);
      }
      const uint64_t m = __ballot(inside);
      if (lane == 0) const_cast<uint64_t*>(C.signs)[lw] = m;
    }
  }
}

// ---------------------------------------------------------------------------
// 2. count + single-pass scan (decoupled look-back)
// ---------------------------------------------------------------------------
// The lattice of one extraction, passed by value: every field is uniform.
struct SingleSpace {
  CellSpace C;
  __device__ __forceinline__ const CellSpace& at(uint32_t g, uint32_t* local) const {
    *local = g;
    return C;
  }
};

// A batch: the chunk owning global segment g is found by binary search over
// the chunks' first segments.
struct BatchSpace {
  const CellSpace* chunks;
  const uint32_t* seg_start;  // n_chunks + 1 ascending first segments
  uint32_t n_chunks;
  __device__ __forceinline__ CellSpace at(uint32_t g, uint32_t* local) const {
    uint32_t lo = 0, hi = n_chunks;  // last chunk with seg_start <= g
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (seg_start[mid] <= g) lo = mid;
      else hi = mid;
    }
    *local = g - seg_start[lo];
    return chunks[lo];
  }
};

template <class Space, int kCountRounds>
__global__ __launch_bounds__(kCountBlock) void k_count(Space sp, int64_t n_segs, int64_t n_blocks,
                                                       McBuffers B) {
  constexpr int kCountSegs = kCountBlock * kCountRounds;  // segments per workgroup
  constexpr int kWaves = kCountBlock / 64;
  __shared__ uint32_t s_wv[kCountRounds][kWaves], s_wt[kCountRounds][kWaves],
      s_wa[kCountRounds][kWaves];
  __shared__ uint32_t s_block;
  __shared__ ScanTriple s_prefix;
  if (threadIdx.x == 0)
    s_block = atomicAdd(reinterpret_cast<unsigned int*>(&B.totals[kTotTicket]), 1u);
  __syncthreads();
  // Workgroups scan in ticket order, so every predecessor is already running.
  const uint32_t blk = s_block;
  const int wv = threadIdx.x >> 6;

  // Round r handles segments blk * kCountSegs + r * kCountBlock + thread, so
  // the loads of a round are coalesced and the segments stay in order.
  uint32_t ov[kCountRounds], ot[kCountRounds], oa[kCountRounds];
  uint64_t act[kCountRounds];
#pragma unroll
  for (int r = 0; r < kCountRounds; ++r) {
    const int64_t g = static_cast<int64_t>(blk) * kCountSegs + r * kCountBlock + threadIdx.x;
    uint32_t nv = 0, nt = 0, na = 0;
    act[r] = 0;
    if (g < n_segs) {
      uint32_t gg;
      const CellSpace& C = sp.at(static_cast<uint32_t>(g), &gg);
      const uint64_t* __restrict__ S = C.signs;
      const int64_t nx = C.nx, ny = C.ny;
      const uint32_t spr = C.segs_per_row;
      const uint32_t cyn = static_cast<uint32_t>(ny - 1);
      const uint32_t row = gg / spr;
      const uint32_t w = gg - row * spr;
      const uint32_t cz = row / cyn;
      const uint32_t cy = row - cz * cyn;
      const int64_t f00 =
          (static_cast<int64_t>(cz) * ny + cy) * nx + 64 * static_cast<int64_t>(w);
      uint64_t n00, n10, n01, n11;
      const uint64_t w00 = row_word(S, f00, &n00);
      const uint64_t w10 = row_word(S, f00 + nx, &n10);
      const uint64_t w01 = row_word(S, f00 + nx * ny, &n01);
      const uint64_t w11 = row_word(S, f00 + nx * ny + nx, &n11);
      const SegWords sw = seg_words(w00, n00, w10, n10, w01, n01, w11, n11);
      const uint64_t valid = seg_valid(nx, w);
      uint64_t vp[4];
      owned_planes(sw, valid, cy == 0, cz == 0, w == 0, vp);
      nv = planes_count(vp, ~0ull);
      act[r] = mixed_cells(sw) & valid;
      na = popc64(act[r]);
      // Only segments with a mixed cell are ever read back: emit reads the
      // planes of its own segment and of owner cells, which hold a crossed
      // edge and so are mixed.
      if (na) {
#pragma unroll
        for (int k = 0; k < 4; ++k) B.vplanes[4 * g + k] = vp[k];
        uint64_t tp[3];  // per-cell triangle counts, bit-sliced
        tri_planes(sw, act[r], tp);
        nt = popc64(tp[0]) + 2 * popc64(tp[1]) + 4 * popc64(tp[2]);
        B.tplanes[3 * g + 0] = tp[0];
        B.tplanes[3 * g + 1] = tp[1];
        B.tplanes[3 * g + 2] = tp[2];
      }
    }
    const uint32_t iv = wave_incl_scan(nv), it = wave_incl_scan(nt), ia = wave_incl_scan(na);
    if (__lane_id() == 63) {
      s_wv[r][wv] = iv;
      s_wt[r][wv] = it;
      s_wa[r][wv] = ia;
    }
    ov[r] = iv - nv;
    ot[r] = it - nt;
    oa[r] = ia - na;
  }
  __syncthreads();
  // Workgroup-exclusive offsets: earlier rounds, then earlier waves of this round.
  {
    uint32_t bv = 0, bt = 0, ba = 0;
#pragma unroll
    for (int r = 0; r < kCountRounds; ++r) {
      uint32_t xv = bv, xt = bt, xa = ba;
      for (int i = 0; i < kWaves; ++i) {
        if (i < wv) {
          xv += s_wv[r][i];
          xt += s_wt[r][i];
          xa += s_wa[r][i];
        }
        bv += s_wv[r][i];
        bt += s_wt[r][i];
        ba += s_wa[r][i];
      }
      ov[r] += xv;
      ot[r] += xt;
      oa[r] += xa;
    }
  }

  // Look-back by the first wave: its 64 lanes inspect 64 predecessors at a
  // time, so a chain of published aggregates is summed in one step.
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    ScanTriple total{0ull, 0ull, 0ull};
    for (int r = 0; r < kCountRounds; ++r)
      for (int i = 0; i < kWaves; ++i) {
        total.v += s_wv[r][i];
        total.t += s_wt[r][i];
        total.a += s_wa[r][i];
      }
    ScanTriple prefix{0ull, 0ull, 0ull};
    if (blk == 0) {
      if (lane == 0) {
        slot_store(&B.incl[0].v, kSlotInclusive | total.v);
        slot_store(&B.incl[0].t, kSlotInclusive | total.t);
        slot_store(&B.incl[0].a, kSlotInclusive | total.a);
      }
    } else {
      if (lane == 0) {
        slot_store(&B.agg[blk].v, kSlotAggregate | total.v);
        slot_store(&B.agg[blk].t, kSlotAggregate | total.t);
        slot_store(&B.agg[blk].a, kSlotAggregate | total.a);
      }
      // Each step inspects kWindows * 64 predecessors at once.  One window:
      // wider steps (4 windows measured) made the 256^3 count slower, as a
      // wider window more often holds a predecessor that has not published.
      constexpr int kWindows = 1;
      int64_t j = static_cast<int64_t>(blk) - 1;  // nearest predecessor of the step
      uint32_t spins = 0;
      while (true) {
        unsigned long long xv[kWindows], xt[kWindows], xa[kWindows];
        int state[kWindows];
#pragma unroll
        for (int k = 0; k < kWindows; ++k) {
          const int64_t idx = j - lane - 64 * k;
          // A predecessor is inclusive when all three inclusive words carry
          // the inclusive status, aggregate when all three aggregate words
          // do.  Blocks before the first one act as an inclusive zero.
          xv[k] = 0ull;
          xt[k] = 0ull;
          xa[k] = 0ull;
          state[k] = 2;
          if (idx >= 0) {
            xv[k] = slot_load(&B.incl[idx].v);
            xt[k] = slot_load(&B.incl[idx].t);
            xa[k] = slot_load(&B.incl[idx].a);
            if (((xv[k] & xt[k] & xa[k]) >> 62) != 2ull) {
              xv[k] = slot_load(&B.agg[idx].v);
              xt[k] = slot_load(&B.agg[idx].t);
              xa[k] = slot_load(&B.agg[idx].a);
              state[k] = ((xv[k] >> 62) == 1ull && (xt[k] >> 62) == 1ull &&
                          (xa[k] >> 62) == 1ull) ? 1 : 0;
            }
          }
        }
        // The nearest inclusive predecessor (window sk, lane stop); every
        // nearer one must have published its aggregate.
        int sk = kWindows, stop = 64;
        bool waiting = false;
#pragma unroll
        for (int k = 0; k < kWindows; ++k) {
          const uint64_t incl = __ballot(state[k] == 2);
          const uint64_t zero = __ballot(state[k] == 0);
          const int st = incl ? __builtin_ctzll(incl) : 64;
          const uint64_t upto = st < 64 ? (2ull << st) - 1ull : ~0ull;
          if (zero & upto) {
            waiting = true;
            break;
          }
          if (st < 64) {
            sk = k;
            stop = st;
            break;
          }
        }
        if (waiting) {
          if (++spins > kLookBackSpinLimit) {
            if (lane == 0) atomicOr(&B.totals[kTotErrors], 2ull);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        ScanTriple x{0ull, 0ull, 0ull};
#pragma unroll
        for (int k = 0; k < kWindows; ++k) {
          if (k < sk || (k == sk && lane <= stop)) {
            x.v += xv[k] & kSlotValue;
            x.t += xt[k] & kSlotValue;
            x.a += xa[k] & kSlotValue;
          }
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
          x.v += __shfl_xor(x.v, d, 64);
          x.t += __shfl_xor(x.t, d, 64);
          x.a += __shfl_xor(x.a, d, 64);
        }
        prefix.v += x.v;
        prefix.t += x.t;
        prefix.a += x.a;
        if (sk < kWindows) break;
        j -= 64 * kWindows;
      }
      if (lane == 0) {
        slot_store(&B.incl[blk].v, kSlotInclusive | (prefix.v + total.v));
        slot_store(&B.incl[blk].t, kSlotInclusive | (prefix.t + total.t));
        slot_store(&B.incl[blk].a, kSlotInclusive | (prefix.a + total.a));
      }
    }
    if (lane == 0) {
      s_prefix = prefix;
      if (static_cast<int64_t>(blk) == n_blocks - 1) {
        B.totals[kTotVertices] = prefix.v + total.v;
        B.totals[kTotTriangles] = prefix.t + total.t;
        B.totals[kTotActive] = prefix.a + total.a;
        if (prefix.v + total.v > 0xffffffffull || 3 * (prefix.t + total.t) > 0xffffffffull)
          atomicOr(&B.totals[kTotErrors], 1ull);
      }
    }
  }
  __syncthreads();

  const ScanTriple p = s_prefix;
#pragma unroll
  for (int r = 0; r < kCountRounds; ++r) {
    const int64_t g = static_cast<int64_t>(blk) * kCountSegs + r * kCountBlock + threadIdx.x;
    if (g < n_segs) {
      B.seg[g] =
          make_uint2(static_cast<uint32_t>(p.v + ov[r]), static_cast<uint32_t>(p.t + ot[r]));
      uint64_t k = p.a + oa[r];
      for (uint64_t m = act[r]; m; m &= m - 1, ++k)
        if (k < B.active_capacity)
          B.active[k] =
              (static_cast<uint32_t>(g) << 6) | static_cast<uint32_t>(__builtin_ctzll(m));
    }
  }
}

// ---------------------------------------------------------------------------
// 3. emit — one thread per active cell
// ---------------------------------------------------------------------------
struct EmitArgs {
  const uint2* seg;
  const uint64_t* tplanes;
  const uint64_t* vplanes;
  const uint32_t* active;
  uint64_t active_capacity;
  float* vertices;
  uint32_t* indices;
  uint64_t vcap, icap;
  const unsigned long long* totals;
  unsigned long long* host_totals;  // or null (McBuffers::host_totals)
  // kEmitBand: brick flags (bit 0: exact samples), bricks per x and y row
  const uint8_t* brick_flags;
  int64_t nbx, nby;
  int brick_shift;  // bricks of 2^brick_shift cells per edge
};

// Where an extraction's sample values and vertex positions come from.
enum EmitSource : int {
  kEmitProgram = 0,    // values from the SDF program, positions from the lattice
  kEmitGrid = 1,       // values from a dense grid, positions from the lattice
  kEmitPositions = 2,  // values from a dense grid, positions per lattice point
  kEmitBand = 3,       // values stored for the exact bricks of a narrow band, the
                       // SDF program elsewhere; positions from the lattice
  kEmitDirected = 4,   // directed values (3 floats per point, in `grid`),
                       // positions from the lattice
};

// Lattice coordinate i (< 2^31) along an axis: the same float as a 64-bit
// index would give (both conversions round the integer once).
__device__ __forceinline__ float space_coord(const CellSpace& C, int axis, int32_t i) {
  return C.origin[axis] + static_cast<float>(i) * C.spacing[axis];
}

// Sample value at lattice point (x, y, z), flat index f.
template <int kSrc>
__device__ __forceinline__ float space_value(const CellSpace& C, int32_t x, int32_t y, int32_t z,
                                             int64_t f) {
  if (kSrc != kEmitProgram && kSrc != kEmitBand) {
    return C.grid[f];
  } else {
    return sdf_eval(C.ops, C.n_ops, space_coord(C, 0, x), space_coord(C, 1, y),
                    space_coord(C, 2, z));
  }
}

constexpr int kEmitThreads = 256;
constexpr uint32_t kXcds = 8;  // accelerator complex dies of an MI355X

template <int kSrc>
__device__ __forceinline__ void space_position(const CellSpace& C, int32_t x, int32_t y,
                                               int32_t z, int64_t f, float* p) {
  if (kSrc == kEmitPositions) {
    const float* q = C.positions + 3 * f;
    p[0] = q[0];
    p[1] = q[1];
    p[2] = q[2];
  } else {
    p[0] = space_coord(C, 0, x);
    p[1] = space_coord(C, 1, y);
    p[2] = space_coord(C, 2, z);
  }
}

template <int kSrc, class Space>
__global__ __launch_bounds__(kEmitThreads) void k_emit(EmitArgs A, Space sp) {
  // LDS stays near 20 KB per workgroup so eight fit a CU: the tables, and
  // per thread the first vertex id and the cube index of each of the seven
  // cells that can own one of its edges, slot rx | ry<<1 | rz<<2 for the
  // cell at (x-rx, y-ry, z-rz).
  __shared__ __align__(16) EmitTables s_t;
  __shared__ uint32_t s_obase[7][kEmitThreads];
  __shared__ uint8_t s_ocase[7][kEmitThreads];

  const int tid = threadIdx.x;
  // The count kernel's totals are final here (the previous launch): hand
  // them to the host with the mesh.
  if (A.host_totals && blockIdx.x == 0 && tid < kTotSlots) A.host_totals[tid] = A.totals[tid];
  uint64_t n_act = A.totals[kTotActive];
  if (n_act > A.active_capacity) n_act = A.active_capacity;

  // The active list in 256-cell chunks, one per thread of a workgroup; a
  // grid larger than the chunk count leaves its spare workgroups idle before
  // they stage any table.  Workgroups are dealt to the eight XCDs
  // round-robin; the working ones of each XCD take a contiguous share of the
  // chunks (neighbouring cells, which share sign words, corner values and
  // segment entries, stay in one L2).
  const uint64_t n_chunks = (n_act + kEmitThreads - 1) / kEmitThreads;
  if (blockIdx.x >= n_chunks) return;
  const uint32_t nw = n_chunks < gridDim.x ? static_cast<uint32_t>(n_chunks) : gridDim.x;
  const uint32_t xcd = blockIdx.x % kXcds, k_in = blockIdx.x / kXcds;
  const uint32_t per = nw / kXcds, rem = nw % kXcds;
  const uint64_t vblock = static_cast<uint64_t>(xcd) * per + (xcd < rem ? xcd : rem) + k_in;
  const uint64_t chunks_per = (n_chunks + nw - 1) / nw;
  const uint64_t first = vblock * chunks_per * kEmitThreads;
  uint64_t last = first + chunks_per * kEmitThreads;
  if (last > n_act) last = n_act;
  // The first active entry is in flight while the tables are staged.
  const uint64_t i0 = first + tid;
  const uint32_t rec0 = i0 < last ? A.active[i0] : 0u;
  load_table16(&s_t, &c_emit);
  __syncthreads();

  for (uint64_t i = i0; i < last; i += kEmitThreads) {
    const uint32_t rec = i == i0 ? rec0 : A.active[i];
    const uint32_t g = rec >> 6;  // global segment
    const int b = static_cast<int>(rec & 63u);
    uint32_t lg;                  // segment within its lattice
    const CellSpace& C = sp.at(g, &lg);
    const uint64_t* __restrict__ S = C.signs;
    const int64_t nx = C.nx, ny = C.ny;
    const int64_t nxy = nx * ny;
    const uint32_t spr = C.segs_per_row;
    const uint32_t cyn = static_cast<uint32_t>(ny - 1);
    // ids written into index lists are relative to the lattice's first vertex
    const uint32_t vbase = A.seg[C.seg_off].x;
    const uint32_t row = lg / spr;
    const uint32_t w = lg - row * spr;
    const int32_t cz = static_cast<int32_t>(row / cyn);
    const int32_t cy = static_cast<int32_t>(row - static_cast<uint32_t>(cz) * cyn);
    const int32_t cx = static_cast<int32_t>(64 * w) + b;
    // Flat index of the cell's corner 0; the other corners and the sign
    // neighbourhood are fixed offsets of it (nx and nxy are per lattice).
    const int64_t fc = (static_cast<int64_t>(cz) * ny + cy) * nx + cx;

    // Loads that depend only on the cell's position go out together with the
    // sign words below: the first vertex id of every cell that can own one
    // of its edges (segment of (x-rx, y-ry, z-rz): rz layers and ry rows of
    // segments back, and one segment back for x-1 at bit 0; slot 7, a step
    // back along all three axes, holds no edge of the cell), and the cell's
    // triangle planes and first triangle id.
    const uint32_t layer_segs = cyn * spr;
    uint32_t obase_pre[8];
    obase_pre[7] = 0;
#pragma unroll
    for (int o = 0; o < 7; ++o) {
      const int rx = o & 1, ry = (o >> 1) & 1, rz = (o >> 2) & 1;
      obase_pre[o] = 0;
      if ((rx && cx < 1) || (ry && cy < 1) || (rz && cz < 1)) continue;
      const uint32_t og = g - (rz ? layer_segs : 0u) - (ry ? spr : 0u) - ((rx && b == 0) ? 1u : 0u);
      obase_pre[o] = A.seg[og].x;
    }
    const uint64_t* tp = A.tplanes + 3 * static_cast<uint64_t>(g);
    const uint64_t tp0 = tp[0], tp1 = tp[1], tp2 = tp[2];
    const uint32_t tseg = A.seg[g].y;
    // A band cell's corners all lie in its brick's closed sample box, whose
    // samples were evaluated exactly and stored; a culled brick's cells are
    // only active when a predicted sign was wrong, and evaluate the program.
    bool stored = kSrc != kEmitProgram && kSrc != kEmitDirected;
    if (kSrc == kEmitBand)
      stored = (A.brick_flags[((cz >> A.brick_shift) * A.nby + (cy >> A.brick_shift)) * A.nbx +
                              (cx >> A.brick_shift)] &
                1u) != 0;
    const bool interior = cx >= 1 && cy >= 1 && cz >= 1;
    const bool deep = cx >= 2 && cy >= 2 && cz >= 2;  // so are all its owners

    // 3x3x3 sign neighbourhood: every cell that can own one of this cell's
    // edges lies inside it.  Row (dy, dz) contributes the bits of samples
    // x-1, x, x+1, read straight from the flat sign array (the guard words
    // cover x-1 at the array's start; bits of rows or columns outside the
    // lattice are never used).
    uint32_t cube = 0;
#pragma unroll
    for (int dz = 0; dz < 3; ++dz) {
#pragma unroll
      for (int dy = 0; dy < 3; ++dy) {
        if (cy - 1 + dy >= 0 && cz - 1 + dz >= 0) {
          const int64_t q = fc - 1 + (dy - 1) * nx + (dz - 1) * nxy;  // sample x-1 (>= -1)
          // The three bits q..q+2 lie in the 64 bits that start at q's
          // 32-bit word: one 8-byte load (4-byte aligned) per row.
          uint64_t v;
          __builtin_memcpy(&v, reinterpret_cast<const uint32_t*>(S) + (q >> 5), sizeof(v));
          v >>= static_cast<int>(q & 31);
          cube |= static_cast<uint32_t>(v & 7ull) << (3 * dy + 9 * dz);
        }
      }
    }
    const uint32_t c0 = case_in_cube(cube, 1, 1, 1);
    const uint32_t emask = s_t.edge_mask[c0];
    // Vertices this cell owns, and the corner values their crossings read
    // (only the endpoints of those edges are loaded).
    const uint32_t mine = owned_edges(cx, cy, cz) & emask;
    float cv[8];
    if (stored) {
      uint32_t need_c = 0;
      for (uint32_t m = mine; m; m &= m - 1) need_c |= edge_corner_mask(__builtin_ctz(m));
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        cv[k] = 0.0f;
        if ((need_c >> k) & 1u)
          cv[k] = C.grid[fc + tables::kCorner[k][2] * nxy + tables::kCorner[k][1] * nx +
                         tables::kCorner[k][0]];
      }
    }

    // Which owner cells the cell's crossed edges need.
    uint32_t needed = interior ? s_t.need[c0] : 0u;
    if (!interior) {
      for (uint32_t m = emask; m; m &= m - 1) {
        const int e = __builtin_ctz(m);
        const uint32_t gm = static_cast<uint32_t>(kEdgeGeomPacked >> (5 * e)) & 31u;
        const int axis = static_cast<int>(gm >> 3);
        const int rx = (axis != 0 && (gm & 1u) == 0 && cx >= 1) ? 1 : 0;
        const int ry = (axis != 1 && (gm & 2u) == 0 && cy >= 1) ? 1 : 0;
        const int rz = (axis != 2 && (gm & 4u) == 0 && cz >= 1) ? 1 : 0;
        needed |= 1u << (rx | (ry << 1) | (rz << 2));
      }
    }

    // First vertex id, cube index and owned crossed edges of each needed
    // owner: the owner segment's first id plus the owned vertices of its
    // cells before the owner, from the segment's bit-sliced counts.
    // Owners 2k and 2k+1 (x and x-1 of the same row and layer) share a
    // segment unless the cell is first in its own: one planes load per pair.
#pragma unroll
    for (int o = 0; o < 7; ++o) s_obase[o][tid] = obase_pre[o];
#pragma unroll 1
    for (int k = 0; k < 4; ++k) {
      const uint32_t pair = (needed >> (2 * k)) & (k == 3 ? 1u : 3u);
      if (!pair) continue;
      const int ry = k & 1, rz = k >> 1;
      const uint32_t og = g - (rz ? layer_segs : 0u) - (ry ? spr : 0u);
      const ulonglong2* vq =
          reinterpret_cast<const ulonglong2*>(A.vplanes + 4 * static_cast<uint64_t>(og));
      const ulonglong2 v01 = vq[0], v23 = vq[1];
      if (pair & 1u) {  // the cell (x, y-ry, z-rz)
        const uint64_t lm = low_mask(b);
        s_obase[2 * k][tid] += popc64(v01.x & lm) + 2 * popc64(v01.y & lm) +
                               4 * popc64(v23.x & lm) + 8 * popc64(v23.y & lm);
        s_ocase[2 * k][tid] = static_cast<uint8_t>(case_in_cube(cube, 1, 1 - ry, 1 - rz));
      }
      if (pair & 2u) {  // the cell (x-1, y-ry, z-rz)
        uint64_t p0 = v01.x, p1 = v01.y, p2 = v23.x, p3 = v23.y;
        int ob = b - 1;
        if (b == 0) {  // last cell of the previous segment
          const ulonglong2* wq = reinterpret_cast<const ulonglong2*>(
              A.vplanes + 4 * static_cast<uint64_t>(og - 1));
          const ulonglong2 w01 = wq[0], w23 = wq[1];
          p0 = w01.x;
          p1 = w01.y;
          p2 = w23.x;
          p3 = w23.y;
          ob = 63;
        }
        const uint64_t lm = low_mask(ob);
        s_obase[2 * k + 1][tid] +=
            popc64(p0 & lm) + 2 * popc64(p1 & lm) + 4 * popc64(p2 & lm) + 8 * popc64(p3 & lm);
        s_ocase[2 * k + 1][tid] = static_cast<uint8_t>(case_in_cube(cube, 0, 1 - ry, 1 - rz));
      }
    }

    // Vertices this cell owns: positions along its own edge directions.
    // (A cell away from the low faces owns only e5, e6 and e10, whose corner
    // values are picked at compile time; others loop over their edges.)
    const uint32_t vbase_cell = s_obase[0][tid];
    auto vertex = [&](int e, int u, int v, float cu, float cvv) {
      const uint32_t id = vbase_cell + static_cast<uint32_t>(__popc(mine & s_t.before[c0][e]));
      const int32_t ux = cx + tables::kCorner[u][0], uy = cy + tables::kCorner[u][1],
                    uz = cz + tables::kCorner[u][2];
      const int32_t vx = cx + tables::kCorner[v][0], vy = cy + tables::kCorner[v][1],
                    vz = cz + tables::kCorner[v][2];
      const int64_t fu = fc + tables::kCorner[u][2] * nxy + tables::kCorner[u][1] * nx +
                         tables::kCorner[u][0];
      const int64_t fv = fc + tables::kCorner[v][2] * nxy + tables::kCorner[v][1] * nx +
                         tables::kCorner[v][0];
      float pa[3], pb[3], p[3];
      space_position<kSrc>(C, ux, uy, uz, fu, pa);
      space_position<kSrc>(C, vx, vy, vz, fv, pb);
      float va, vb;
      if (kSrc == kEmitDirected) {
        // Directed::find_crossing_point (distance.rs:89-102): the component
        // of the axis (p_a - p_b).abs().max_component_index() (vector.rs:
        // 255-263) — the edge's own axis on any non-degenerate lattice.
        const float dx = __builtin_fabsf(pa[0] - pb[0]), dy = __builtin_fabsf(pa[1] - pb[1]),
                    dz = __builtin_fabsf(pa[2] - pb[2]);
        const int axis = (dx > dy && dx > dz) ? 0 : (dy > dz ? 1 : 2);
        va = C.grid[3 * fu + axis];
        vb = C.grid[3 * fv + axis];
      } else {
        va = stored ? cu : space_value<kSrc>(C, ux, uy, uz, fu);
        vb = stored ? cvv : space_value<kSrc>(C, vx, vy, vz, fv);
      }
      crossing(va, vb, pa[0], pa[1], pa[2], pb[0], pb[1], pb[2], p);
      if (id < A.vcap) {
        A.vertices[3 * static_cast<uint64_t>(id) + 0] = p[0];
        A.vertices[3 * static_cast<uint64_t>(id) + 1] = p[1];
        A.vertices[3 * static_cast<uint64_t>(id) + 2] = p[2];
      }
    };
    if (interior) {
      if (mine & (1u << 5)) vertex(5, 5, 6, cv[5], cv[6]);
      if (mine & (1u << 6)) vertex(6, 6, 7, cv[6], cv[7]);
      if (mine & (1u << 10)) vertex(10, 2, 6, cv[2], cv[6]);
    } else {
      for (uint32_t todo = mine; todo; todo &= todo - 1) {
        const int e = __builtin_ctz(todo);
        const int u = tables::kEdgeCorners[e][0], v = tables::kEdgeCorners[e][1];
        vertex(e, u, v, cv[u], cv[v]);
      }
    }

    // Triangles: cell order, then table order (mesh.rs:91-100).
    const uint64_t lm = low_mask(b);
    const uint64_t tbase =
        tseg + popc64(tp0 & lm) + 2 * popc64(tp1 & lm) + 4 * popc64(tp2 & lm);
    const uint32_t nidx = 3u * s_t.tri_count[c0];
    for (uint32_t j = 0; j < nidx; ++j) {
      int o, eo;
      if (interior) {
        const uint32_t ix = s_t.ix[c0][j];
        o = static_cast<int>(ix & 7u);
        eo = static_cast<int>(ix >> 3);
      } else {
        const int e = c_tables.tri[c0][j];
        const uint32_t gm = static_cast<uint32_t>(kEdgeGeomPacked >> (5 * e)) & 31u;
        const int axis = static_cast<int>(gm >> 3);
        const int px = gm & 1, py = (gm >> 1) & 1, pz = (gm >> 2) & 1;
        // Owner: one cell back along each non-axis direction where the edge
        // lies on this cell's low face and a lower cell exists.
        const int rx = (axis != 0 && px == 0 && cx >= 1) ? 1 : 0;
        const int ry = (axis != 1 && py == 0 && cy >= 1) ? 1 : 0;
        const int rz = (axis != 2 && pz == 0 && cz >= 1) ? 1 : 0;
        o = rx | (ry << 1) | (rz << 2);
        eo = edge_of_axis_min(axis, (px | rx) | ((py | ry) << 1) | ((pz | rz) << 2));
      }
      // Owners with no coordinate 0 own e5, e6 and e10 only.
      const int rx = o & 1, ry = (o >> 1) & 1, rz = (o >> 2) & 1;
      const uint32_t oc = s_ocase[o][tid];
      const uint32_t oedges = deep ? 0x460u : owned_edges(cx - rx, cy - ry, cz - rz);
      const uint32_t oown = oedges & s_t.edge_mask[oc];
      const uint32_t id =
          s_obase[o][tid] + static_cast<uint32_t>(__popc(oown & s_t.before[oc][eo]));
      const uint64_t ix = 3 * tbase + j;
      if (ix < A.icap) A.indices[ix] = id - vbase;
    }
  }
}

// PointCloud::extract (point_cloud.rs:50-63): per mixed cell, in cell order,
// corners[0].lerp(corners[6], 0.5) (vector.rs:325-333, 0.5 * a + 0.5 * b per
// component).  One thread per entry of the active list, which holds exactly
// the mixed cells in cell order.
__global__ __launch_bounds__(256) void k_point_cloud(Lattice L, McDims D,
                                                     const uint32_t* __restrict__ active,
                                                     uint64_t active_capacity,
                                                     const unsigned long long* __restrict__ totals,
                                                     float* __restrict__ points, uint64_t capacity) {
  uint64_t n = totals[kTotActive];
  if (n > active_capacity) n = active_capacity;
  if (n > capacity) n = capacity;
  const uint32_t spr = static_cast<uint32_t>(D.segs_per_row);
  const uint32_t cyn = static_cast<uint32_t>(L.ny - 1);
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t rec = active[i];
    const uint32_t g = rec >> 6;
    const uint32_t row = g / spr;
    const uint32_t w = g - row * spr;
    const int64_t cz = row / cyn, cy = row - (row / cyn) * cyn;
    const int64_t cx = 64 * static_cast<int64_t>(w) + (rec & 63u);
    const float c0[3] = {lattice_coord(L, 0, cx), lattice_coord(L, 1, cy), lattice_coord(L, 2, cz)};
    const float c6[3] = {lattice_coord(L, 0, cx + 1), lattice_coord(L, 1, cy + 1),
                         lattice_coord(L, 2, cz + 1)};
#pragma unroll
    for (int a = 0; a < 3; ++a) points[3 * i + a] = 0.5f * c0[a] + 0.5f * c6[a];
  }
}

// Per chunk first vertex / index of the batch output (n_chunks + 1 entries).
__global__ __launch_bounds__(256) void k_batch_offsets(const uint2* __restrict__ seg,
                                                       const uint32_t* __restrict__ seg_start,
                                                       uint32_t n_chunks,
                                                       const unsigned long long* __restrict__ totals,
                                                       uint64_t* __restrict__ voff,
                                                       uint64_t* __restrict__ ioff) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > n_chunks) return;
  if (c < n_chunks) {
    const uint2 f = seg[seg_start[c]];
    voff[c] = f.x;
    ioff[c] = 3ull * f.y;
  } else {
    voff[c] = totals[kTotVertices];
    ioff[c] = 3ull * totals[kTotTriangles];
  }
}

// ---------------------------------------------------------------------------
// Program sampling (iso_program_sample_lattice / iso_program_sample)
// ---------------------------------------------------------------------------
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

__global__ __launch_bounds__(256) void k_sample_vector_points(const DevOp* __restrict__ ops,
                                                              int n_ops,
                                                              const float* __restrict__ pts,
                                                              uint64_t n,
                                                              float* __restrict__ out) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) sdf_eval_vector(ops, n_ops, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], out + 3 * i);
}

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

int device_cus() {
  int device = 0, cus = 256;
  if (hipGetDevice(&device) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  return cus > 0 ? cus : 256;
}

ResetArgs reset_args(const McDims& D, const McBuffers& B) {
  ResetArgs r;
  r.signs = B.signs;
  r.n_words = D.n_sign_words;
  r.agg = B.agg;
  r.incl = B.incl;
  r.n_blocks = D.n_blocks;
  r.totals = B.totals;
  return r;
}

}  // namespace

McDims mc_dims(const Lattice& L) {
  McDims D;
  D.n_points = L.nx * L.ny * L.nz;
  D.n_sign_words = (D.n_points + 63) / 64;
  D.n_cell_rows = (L.ny - 1) * (L.nz - 1);
  D.segs_per_row = (L.nx - 1 + 63) / 64;
  D.n_segs = D.n_cell_rows * D.segs_per_row;
  D.n_blocks = (D.n_segs + kCountSegs - 1) / kCountSegs;
  D.n_cells = (L.nx - 1) * (L.ny - 1) * (L.nz - 1);
  return D;
}

namespace {
thread_local hipEvent_t t_ev_start = nullptr;
thread_local hipEvent_t t_ev_stop = nullptr;

// A launch that takes the pending timing events, if any.
template <typename K, typename... Args>
void launch_k(K kernel, dim3 grid, dim3 block, hipStream_t s, Args... args) {
  if (t_ev_start) {
    hipEvent_t a = t_ev_start, z = t_ev_stop;
    t_ev_start = nullptr;
    t_ev_stop = nullptr;
    hipExtLaunchKernelGGL(kernel, grid, block, 0, s, a, z, 0, args...);
  } else {
    hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
  }
}
}  // namespace

void set_launch_events(hipEvent_t start, hipEvent_t stop) {
  t_ev_start = start;
  t_ev_stop = stop;
}
bool launch_events_pending() { return t_ev_start != nullptr; }
bool take_launch_events(hipEvent_t* start, hipEvent_t* stop) {
  if (!t_ev_start) return false;
  *start = t_ev_start;
  *stop = t_ev_stop;
  t_ev_start = nullptr;
  t_ev_stop = nullptr;
  return true;
}
void clear_launch_events() {
  t_ev_start = nullptr;
  t_ev_stop = nullptr;
}

hipError_t launch_signs_grid(const float* grid, const Lattice& L, const McDims& D,
                             const McBuffers& B, hipStream_t s) {
  (void)L;
  if (!grid) return hipErrorInvalidValue;
  const ResetArgs R = reset_args(D, B);
  // At least enough threads to clear the look-back words in one sweep.
  const int64_t min_blocks = (D.n_blocks + 255) / 256;
  int64_t blocks = grid_for((D.n_sign_words + kSignWords - 1) / kSignWords);
  if (blocks < min_blocks) blocks = min_blocks;
  launch_k(k_signs_grid, dim3(static_cast<unsigned>(blocks)), dim3(256), s, grid,
           D.n_points, D.n_sign_words, B.signs, R);
  return hipGetLastError();
}

hipError_t launch_signs_program(const DevOp* ops, int n_ops, const Lattice& L, const McDims& D,
                                const McBuffers& B, float* values, hipStream_t s) {
  if (!ops) return hipErrorInvalidValue;
  int64_t blocks = grid_for(D.n_sign_words);
  const int64_t min_blocks = (D.n_blocks + 255) / 256;
  if (blocks < min_blocks) blocks = min_blocks;
  launch_k(k_signs_program, dim3(static_cast<unsigned>(blocks)), dim3(256), s, ops,
           n_ops, L, D.n_points, D.n_sign_words, B.signs, values, reset_args(D, B));
  return hipGetLastError();
}

namespace {

SingleSpace single_space(const Lattice& L, const McDims& D, const uint64_t* signs,
                         const DevOp* ops, int n_ops, const float* grid,
                         const float* positions) {
  SingleSpace sp;
  sp.C.nx = L.nx;
  sp.C.ny = L.ny;
  sp.C.nz = L.nz;
  for (int i = 0; i < 3; ++i) {
    sp.C.origin[i] = L.origin[i];
    sp.C.spacing[i] = L.spacing[i];
  }
  sp.C.signs = signs;
  sp.C.segs_per_row = static_cast<uint32_t>(D.segs_per_row);
  sp.C.seg_off = 0;
  sp.C.ops = ops;
  sp.C.n_ops = n_ops;
  sp.C.reserved = 0;
  sp.C.grid = grid;
  sp.C.positions = positions;
  return sp;
}

EmitArgs emit_args(const McBuffers& B) {
  EmitArgs A;
  A.seg = B.seg;
  A.tplanes = B.tplanes;
  A.vplanes = B.vplanes;
  A.active = B.active;
  A.active_capacity = B.active_capacity;
  A.vertices = B.vertices;
  A.indices = B.indices;
  A.vcap = B.vertex_capacity;
  A.icap = B.index_capacity;
  A.totals = B.totals;
  A.host_totals = B.host_totals;
  A.brick_flags = nullptr;
  A.nbx = A.nby = 0;
  A.brick_shift = 0;
  return A;
}

// Grid-stride over the active list (its length is only known on the device):
// at most `per_cu` resident workgroups per CU (0: as many as fit).
template <class K>
unsigned emit_blocks(const McBuffers& B, int per_cu, K kernel, int64_t n_cells = -1) {
  (void)kernel;
  // One 256-cell chunk of the active list per workgroup (the list's length
  // is only known on the device: workgroups past it exit at once), or at
  // most per_cu resident workgroups per CU, each taking several chunks.
  // The list holds at most one entry per cell of the lattice: a list grown
  // by a larger lattice does not enlarge the grid (tens of thousands of idle
  // workgroups cost tens of microseconds).
  int64_t items = static_cast<int64_t>(B.active_capacity);
  if (n_cells >= 0 && n_cells < items) items = n_cells > 0 ? n_cells : 1;
  int64_t blocks = (items + kEmitThreads - 1) / kEmitThreads;
  if (per_cu > 0) {
    const int64_t cap = static_cast<int64_t>(per_cu) * device_cus();
    if (blocks > cap) blocks = cap;
  }
  if (blocks > 0x7fffffff) blocks = 0x7fffffff;
  if (blocks < 1) blocks = 1;
  return static_cast<unsigned>(blocks);
}

}  // namespace

hipError_t launch_count(const Lattice& L, const McDims& D, const McBuffers& B, hipStream_t s) {
  if (D.n_blocks > 0x7fffffff) return hipErrorInvalidValue;
  const SingleSpace sp = single_space(L, D, B.signs, nullptr, 0, nullptr, nullptr);
  launch_k(k_count<SingleSpace, kCountRounds>, dim3(static_cast<unsigned>(D.n_blocks)),
           dim3(kCountBlock), s, sp, D.n_segs, D.n_blocks, B);
  return hipGetLastError();
}

hipError_t launch_emit(const float* grid, const DevOp* ops, int n_ops, const float* positions,
                       const Lattice& L, const McDims& D, const McBuffers& B, hipStream_t s) {
  if ((positions && !grid) || (!grid && !ops)) return hipErrorInvalidValue;
  const EmitArgs A = emit_args(B);
  const SingleSpace sp = single_space(L, D, B.signs, ops, n_ops, grid, positions);
  if (grid && positions) {
    auto k = k_emit<kEmitPositions, SingleSpace>;
    launch_k(k, dim3(emit_blocks(B, 0, k, D.n_cells)), dim3(kEmitThreads), s, A, sp);
  } else if (grid) {
    auto k = k_emit<kEmitGrid, SingleSpace>;
    launch_k(k, dim3(emit_blocks(B, 0, k, D.n_cells)), dim3(kEmitThreads), s, A, sp);
  } else {
    auto k = k_emit<kEmitProgram, SingleSpace>;
    launch_k(k, dim3(emit_blocks(B, 0, k, D.n_cells)), dim3(kEmitThreads), s, A, sp);
  }
  return hipGetLastError();
}

hipError_t launch_signs_directed(const DevOp* ops, int n_ops, const Lattice& L, const McDims& D,
                                 const McBuffers& B, float* values, hipStream_t s) {
  if (!values) return hipErrorInvalidValue;
  int64_t blocks = grid_for(D.n_sign_words);
  const int64_t min_blocks = (D.n_blocks + 255) / 256;
  if (blocks < min_blocks) blocks = min_blocks;
  launch_k(k_signs_directed, dim3(static_cast<unsigned>(blocks)), dim3(256), s, ops,
           n_ops, L, D.n_points, D.n_sign_words, B.signs, values, reset_args(D, B));
  return hipGetLastError();
}

hipError_t launch_emit_directed(const float* values, const Lattice& L, const McDims& D,
                                const McBuffers& B, hipStream_t s) {
  if (!values) return hipErrorInvalidValue;
  const EmitArgs A = emit_args(B);
  const SingleSpace sp = single_space(L, D, B.signs, nullptr, 0, values, nullptr);
  auto k = k_emit<kEmitDirected, SingleSpace>;
  launch_k(k, dim3(emit_blocks(B, 0, k, D.n_cells)), dim3(kEmitThreads), s, A, sp);
  return hipGetLastError();
}

hipError_t launch_point_cloud(const Lattice& L, const McDims& D, const McBuffers& B,
                              float* points, uint64_t capacity, hipStream_t s) {
  if (!points || capacity == 0) return hipSuccess;
  int64_t blocks = (static_cast<int64_t>(B.active_capacity) + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_point_cloud, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, L, D,
                     B.active, B.active_capacity, B.totals, points, capacity);
  return hipGetLastError();
}

hipError_t launch_emit_band(const float* values, const uint8_t* brick_flags, int64_t nbx,
                            int64_t nby, int brick_shift, const DevOp* ops, int n_ops,
                            const Lattice& L, const McDims& D, const McBuffers& B,
                            hipStream_t s) {
  if (!values || !brick_flags || !ops || nbx < 1 || nby < 1 || brick_shift < 1 ||
      brick_shift > kBandBrickShiftMax)
    return hipErrorInvalidValue;
  EmitArgs A = emit_args(B);
  A.brick_flags = brick_flags;
  A.nbx = nbx;
  A.nby = nby;
  A.brick_shift = brick_shift;
  const SingleSpace sp = single_space(L, D, B.signs, ops, n_ops, values, nullptr);
  auto k = k_emit<kEmitBand, SingleSpace>;
  launch_k(k, dim3(emit_blocks(B, 0, k, D.n_cells)), dim3(kEmitThreads), s, A, sp);
  return hipGetLastError();
}

hipError_t launch_batch_signs(const BatchPlan& P, const McBuffers& B, int words_per_wave,
                              hipStream_t s) {
  // The state reset covers the look-back words of the finest count tile.
  const int64_t n_blocks = batch_count_blocks(P, 1);
  ResetArgs R;
  R.signs = nullptr;
  R.n_words = 0;
  R.agg = B.agg;
  R.incl = B.incl;
  R.n_blocks = n_blocks;
  R.totals = B.totals;
  const int64_t wpw = words_per_wave > 0 ? words_per_wave : 1;
  const int64_t per_block = 4 * kBatchRun * wpw;  // sign words per workgroup
  int64_t blocks = (P.n_words + per_block - 1) / per_block;
  if (blocks < 1) blocks = 1;
  if (blocks > 65536) blocks = 65536;
  const int64_t min_blocks = (n_blocks + 255) / 256;
  const int64_t chunk_blocks = (static_cast<int64_t>(P.n_chunks) + 255) / 256;
  if (blocks < min_blocks) blocks = min_blocks;
  if (blocks < chunk_blocks) blocks = chunk_blocks;
  hipLaunchKernelGGL(k_batch_signs, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, P.chunks,
                     P.word_start, P.n_chunks, R);
  return hipGetLastError();
}

int64_t batch_count_blocks(const BatchPlan& P, int rounds) {
  const int64_t per = static_cast<int64_t>(kCountBlock) * rounds;
  return (P.n_segs + per - 1) / per;
}

hipError_t launch_batch_count(const BatchPlan& P, const McBuffers& B, int rounds,
                              hipStream_t s) {
  const int64_t nb = batch_count_blocks(P, rounds);
  if (nb > 0x7fffffff) return hipErrorInvalidValue;
  BatchSpace sp;
  sp.chunks = P.chunks;
  sp.seg_start = P.seg_start;
  sp.n_chunks = P.n_chunks;
  const dim3 grid(static_cast<unsigned>(nb)), block(kCountBlock);
  switch (rounds) {
    case 1:
      hipLaunchKernelGGL((k_count<BatchSpace, 1>), grid, block, 0, s, sp, P.n_segs, nb, B);
      break;
    case 2:
      hipLaunchKernelGGL((k_count<BatchSpace, 2>), grid, block, 0, s, sp, P.n_segs, nb, B);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_batch_emit(const BatchPlan& P, const McBuffers& B, int blocks_per_cu,
                             hipStream_t s) {
  BatchSpace sp;
  sp.chunks = P.chunks;
  sp.seg_start = P.seg_start;
  sp.n_chunks = P.n_chunks;
  if (P.stored_values) {
    auto k = k_emit<kEmitGrid, BatchSpace>;
    hipLaunchKernelGGL(k, dim3(emit_blocks(B, blocks_per_cu, k)), dim3(kEmitThreads), 0, s,
                       emit_args(B), sp);
  } else {
    auto k = k_emit<kEmitProgram, BatchSpace>;
    hipLaunchKernelGGL(k, dim3(emit_blocks(B, blocks_per_cu, k)), dim3(kEmitThreads), 0, s,
                       emit_args(B), sp);
  }
  return hipGetLastError();
}

hipError_t launch_batch_offsets(const BatchPlan& P, const McBuffers& B, uint64_t* voff,
                                uint64_t* ioff, hipStream_t s) {
  const unsigned blocks = (P.n_chunks + 1 + 255) / 256;
  hipLaunchKernelGGL(k_batch_offsets, dim3(blocks), dim3(256), 0, s, B.seg, P.seg_start,
                     P.n_chunks, B.totals, voff, ioff);
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

hipError_t launch_sample_vector_points(const DevOp* ops, int n_ops, const float* pts, uint64_t n,
                                       float* out, hipStream_t s) {
  const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sample_vector_points, dim3(blocks), dim3(256), 0, s, ops, n_ops, pts, n,
                     out);
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
