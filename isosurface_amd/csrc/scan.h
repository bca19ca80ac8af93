// Single-pass exclusive scan over n items for gfx950 (device templates).
//
// A workgroup takes a ticket, sums its tile of kThreads * I items (I per
// thread, consecutive; I = 1 for loads that do heavy work), publishes the
// aggregate and looks back over its
// predecessors' published values (decoupled look-back).  Items come from a
// load functor and results go to a sink functor, so a scan can be fused with
// the work that produces its counts and with the writes that consume them:
//   Load: __device__ unsigned long long operator()(uint64_t i) const
//   Sink: __device__ void operator()(uint64_t i, unsigned long long value,
//                                    unsigned long long exclusive) const
// Each workgroup publishes one 64-bit word whose top two bits say what the
// low 62 bits hold (1: its aggregate, 2: its inclusive prefix), written and
// read with relaxed agent-scope atomics: the status travels with the value,
// so no release/acquire fence (and no L2 writeback or invalidation across
// the XCDs) is needed.  The wait is bounded: a look-back that never
// completes sets bit 1 of *err and the kernel still finishes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace iso {
namespace scan {

constexpr int kThreads = 256;
constexpr int kItems = 8;  // items per thread for light loads
constexpr uint32_t kSpinLimit = 1u << 22;

// Device state of one scan; launch() clears it before every scan.
struct State {
  unsigned long long* slot;  // n_blocks: status << 62 | value
  unsigned long long* ctl;   // [0] workgroup ticket, [1] total
  unsigned long long* err;   // bit 1: look-back gave up
};

constexpr unsigned long long kAggregate = 1ull << 62;
constexpr unsigned long long kInclusive = 2ull << 62;
constexpr unsigned long long kValueMask = (1ull << 62) - 1ull;

// Workgroups of a scan of n items at I items per thread (at least 1).
inline int64_t blocks_for(uint64_t n, int items) {
  const uint64_t tile = static_cast<uint64_t>(kThreads) * items;
  return n ? static_cast<int64_t>((n + tile - 1) / tile) : 1;
}

__device__ __forceinline__ unsigned long long slot_load(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void slot_store(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int kUnused = 0>
__global__ void k_reset(State st, int64_t n_blocks) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n_blocks;
       i += stride)
    st.slot[i] = 0ull;
  if (blockIdx.x == 0 && threadIdx.x < 2) st.ctl[threadIdx.x] = 0ull;
}

template <int I, class Load, class Sink>
__global__ __launch_bounds__(kThreads) void k_scan(uint64_t n, Load load, Sink sink, State st) {
  constexpr uint64_t kTile = static_cast<uint64_t>(kThreads) * I;
  __shared__ uint32_t s_block;
  __shared__ unsigned long long s_wave[kThreads / 64];
  __shared__ unsigned long long s_prefix;
  if (threadIdx.x == 0) s_block = atomicAdd(reinterpret_cast<unsigned int*>(&st.ctl[0]), 1u);
  __syncthreads();
  // Workgroups scan in ticket order, so every predecessor is already running.
  const uint64_t blk = s_block;
  const uint64_t first = blk * kTile + static_cast<uint64_t>(threadIdx.x) * I;
  unsigned long long v[I];
  unsigned long long sum = 0ull;
#pragma unroll
  for (int k = 0; k < I; ++k) {
    const uint64_t i = first + k;
    v[k] = i < n ? load(i) : 0ull;
    sum += v[k];
  }
  const int lane = __lane_id();
  unsigned long long inc = sum;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned long long o = __shfl_up(inc, d, 64);
    if (lane >= d) inc += o;
  }
  const int wv = threadIdx.x >> 6;
  if (lane == 63) s_wave[wv] = inc;
  __syncthreads();
  unsigned long long off = inc - sum;
  for (int i = 0; i < wv; ++i) off += s_wave[i];

  // Look-back by the first wave: 64 predecessors per step.
  if (threadIdx.x < 64) {
    unsigned long long total = 0ull;
    for (int i = 0; i < kThreads / 64; ++i) total += s_wave[i];
    unsigned long long prefix = 0ull;
    if (blk == 0) {
      if (lane == 0) slot_store(&st.slot[0], kInclusive | total);
    } else {
      if (lane == 0) slot_store(&st.slot[blk], kAggregate | total);
      int64_t j = static_cast<int64_t>(blk) - 1;
      uint32_t spins = 0;
      while (true) {
        const int64_t idx = j - lane;
        // Blocks before the first one act as an inclusive zero.
        const unsigned long long w = idx >= 0 ? slot_load(&st.slot[idx]) : kInclusive;
        const uint64_t incl = __ballot((w >> 62) == 2ull);
        const uint64_t zero = __ballot((w >> 62) == 0ull);
        const int stop = incl ? __builtin_ctzll(incl) : 64;
        const uint64_t upto = stop < 64 ? (2ull << stop) - 1ull : ~0ull;
        if (zero & upto) {
          if (++spins > kSpinLimit) {
            if (lane == 0) atomicOr(st.err, 2ull);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        unsigned long long x = lane <= stop ? (w & kValueMask) : 0ull;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
        prefix += x;
        if (stop < 64) break;
        j -= 64;
      }
      if (lane == 0) slot_store(&st.slot[blk], kInclusive | (prefix + total));
    }
    if (lane == 0) {
      s_prefix = prefix;
      if (blk == (n ? (n - 1) / kTile : 0)) st.ctl[1] = prefix + total;
    }
  }
  __syncthreads();
  unsigned long long run = s_prefix + off;
#pragma unroll
  for (int k = 0; k < I; ++k) {
    const uint64_t i = first + k;
    if (i < n) sink(i, v[k], run);
    run += v[k];
  }
}

// Clears the state for a scan of n items and runs it on stream s; the
// state must hold blocks_for(n, I) slots.
template <int I, class Load, class Sink>
hipError_t launch(uint64_t n, const Load& load, const Sink& sink, const State& st,
                  hipStream_t s) {
  const int64_t nb = blocks_for(n, I);
  if (nb > 0x7fffffff) return hipErrorInvalidValue;
  int64_t rb = (nb + 255) / 256;
  if (rb > 4096) rb = 4096;
  hipLaunchKernelGGL(k_reset<0>, dim3(static_cast<unsigned>(rb)), dim3(256), 0, s, st, nb);
  hipLaunchKernelGGL((k_scan<I, Load, Sink>), dim3(static_cast<unsigned>(nb)), dim3(kThreads), 0, s,
                     n, load, sink, st);
  return hipGetLastError();
}

}  // namespace scan
}  // namespace iso
