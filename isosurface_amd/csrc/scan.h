// Single-pass exclusive scan over n items for gfx950 (device templates).
//
// A workgroup takes a ticket, sums its tile of kTile items (kItems per
// thread, consecutive), publishes the aggregate and looks back over its
// predecessors' published values (decoupled look-back).  Items come from a
// load functor and results go to a sink functor, so a scan can be fused with
// the work that produces its counts and with the writes that consume them:
//   Load: __device__ unsigned long long operator()(uint64_t i) const
//   Sink: __device__ void operator()(uint64_t i, unsigned long long value,
//                                    unsigned long long exclusive) const
// The wait is bounded: a look-back that never completes sets bit 1 of
// *err and the kernel still finishes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace iso {
namespace scan {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int64_t kTile = static_cast<int64_t>(kThreads) * kItems;
constexpr uint32_t kSpinLimit = 1u << 22;

// Device state of one scan; launch() clears it before every scan.
struct State {
  uint32_t* flags;           // n_blocks: 0 none, 1 aggregate, 2 inclusive
  unsigned long long* agg;   // n_blocks
  unsigned long long* incl;  // n_blocks
  unsigned long long* ctl;   // [0] workgroup ticket, [1] total
  unsigned long long* err;   // bit 1: look-back gave up
};

inline int64_t blocks_for(uint64_t n) {
  return n ? static_cast<int64_t>((n + kTile - 1) / kTile) : 1;
}

__device__ __forceinline__ uint32_t flag_acquire(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void flag_release(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

template <int kUnused = 0>
__global__ void k_reset(State st, int64_t n_blocks) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n_blocks;
       i += stride)
    st.flags[i] = 0u;
  if (blockIdx.x == 0 && threadIdx.x < 2) st.ctl[threadIdx.x] = 0ull;
}

template <class Load, class Sink>
__global__ __launch_bounds__(kThreads) void k_scan(uint64_t n, Load load, Sink sink, State st) {
  __shared__ uint32_t s_block;
  __shared__ unsigned long long s_wave[kThreads / 64];
  __shared__ unsigned long long s_prefix;
  if (threadIdx.x == 0) s_block = atomicAdd(reinterpret_cast<unsigned int*>(&st.ctl[0]), 1u);
  __syncthreads();
  // Workgroups scan in ticket order, so every predecessor is already running.
  const uint64_t blk = s_block;
  const uint64_t first = blk * kTile + static_cast<uint64_t>(threadIdx.x) * kItems;
  unsigned long long v[kItems];
  unsigned long long sum = 0ull;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
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

  if (threadIdx.x == 0) {
    unsigned long long total = 0ull;
    for (int i = 0; i < kThreads / 64; ++i) total += s_wave[i];
    unsigned long long prefix = 0ull;
    if (blk == 0) {
      st.incl[0] = total;
      flag_release(&st.flags[0], 2u);
    } else {
      st.agg[blk] = total;
      flag_release(&st.flags[blk], 1u);
      int64_t j = static_cast<int64_t>(blk) - 1;
      uint32_t spins = 0;
      while (j >= 0) {
        const uint32_t f = flag_acquire(&st.flags[j]);
        if (f == 0u) {
          if (++spins > kSpinLimit) {
            atomicOr(st.err, 2ull);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        prefix += (f == 2u) ? st.incl[j] : st.agg[j];
        if (f == 2u) break;
        --j;
      }
      st.incl[blk] = prefix + total;
      flag_release(&st.flags[blk], 2u);
    }
    s_prefix = prefix;
    if (blk == (n ? (n - 1) / kTile : 0)) st.ctl[1] = prefix + total;
  }
  __syncthreads();
  unsigned long long run = s_prefix + off;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const uint64_t i = first + k;
    if (i < n) sink(i, v[k], run);
    run += v[k];
  }
}

// Clears the state for a scan of n items and runs it on stream s.
template <class Load, class Sink>
hipError_t launch(uint64_t n, const Load& load, const Sink& sink, const State& st,
                  hipStream_t s) {
  const int64_t nb = blocks_for(n);
  if (nb > 0x7fffffff) return hipErrorInvalidValue;
  int64_t rb = (nb + 255) / 256;
  if (rb > 4096) rb = 4096;
  hipLaunchKernelGGL(k_reset<0>, dim3(static_cast<unsigned>(rb)), dim3(256), 0, s, st, nb);
  hipLaunchKernelGGL((k_scan<Load, Sink>), dim3(static_cast<unsigned>(nb)), dim3(kThreads), 0, s,
                     n, load, sink, st);
  return hipGetLastError();
}

}  // namespace scan
}  // namespace iso
