// Single-pass marching cubes over a dense FP32 grid (gfx950).
//
// One launch does what the signs / count / emit kernels do in three, and
// reads the grid once.  The cells are cut into tiles of R consecutive cell
// rows of one cell layer (R * ceil((nx-1)/64) <= 1024 segments of 64 cells),
// taken in z-major order by ticket, so every tile's predecessors are running
// or done.  A tile:
//
//   1. produces the sign words of its own sample rows (rows y0+1..y1 of the
//      layer above its cells; tile 0 of a layer also row 0, and the tiles
//      of cell layer 0 also the same rows of sample layer 0): coalesced
//      256-byte loads, one ballot per 64 samples, row-aligned words kept in
//      LDS and published for the tiles that need them;
//   2. acquires the other sign rows its cells and their owner cells touch
//      (rows y0-1..y1 of three sample layers) from the tiles that produced
//      them;
//   3. counts, per 64-cell segment, the vertices its cells own and the
//      triangles they emit (bit-sliced planes of the segment's corner
//      bits, as the count kernel does) and scans them in the workgroup;
//   4. turns the tile totals into global first ids with a decoupled
//      look-back over the tiles before it, and publishes the first vertex
//      id of each of its cell rows;
//   5. emits: its mixed cells in cell order write the vertices they own and
//      their index triples.  An edge's vertex id is its owner cell's first
//      id (own rows: the tile's scan; the row above in the previous tile
//      and the previous cell layer: that row's published first id plus the
//      counts of the row's cells before the owner, recomputed from the
//      sign rows in LDS) plus its rank in the owner's triangle list.
//
// Numbering is the reference's first-encounter order exactly as in the
// three-kernel path (MeshTopologyBuilder::add_vertex, mesh.rs:240-251,
// driven by PrimalGrid::traverse's z-major cell order, primal_grid.rs:59-85;
// classification !(v > 0), marching_cubes_impl.rs:26-37, distance.rs:52-54;
// crossing points distance.rs:64-70).
//
// Cross-tile data (sign words, first ids, look-back words) is written and
// read with relaxed agent-scope atomics, which the gfx950 memory system
// keeps coherent across the eight XCDs' L2 caches (write-through stores and
// coherent loads).  A tile publishes its sign rows, waits for the stores to
// complete (s_waitcnt vmcnt(0)) and then sets its flag; readers load the
// flag, then the words.  Flags, first ids and look-back words carry the
// extraction's epoch, so no state is cleared between extractions; the
// ticket counter carries it too.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "mc_device.h"
#include "mc_kernels.h"
#include "mc_tables.h"

namespace iso {

namespace {

using namespace mcd;

__constant__ __align__(16) tables::McTables f_tables = tables::kTables;
__constant__ __align__(16) EmitTables f_emit = make_emit_tables();

constexpr int kThreads = 512;
constexpr int kWaves = kThreads / 64;
constexpr int kLoadBatch = 16;  // 256-byte loads in flight per wavefront
constexpr int kAcquireBatch = 4;  // sign words a thread loads at once from other tiles

// Look-back words: tag (22 bits of the epoch) | status (2) | value (40).
constexpr int kValueBits = 40;
constexpr unsigned long long kValueMask = (1ull << kValueBits) - 1ull;
constexpr unsigned long long kStatusAggregate = 1ull;
constexpr unsigned long long kStatusInclusive = 2ull;

__device__ __forceinline__ unsigned long long atom_load(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void atom_store(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits until this wavefront's stores have completed.
__device__ __forceinline__ void wait_stores() { __builtin_amdgcn_s_waitcnt(0x0F70); }

__device__ __forceinline__ unsigned long long slot_word(uint32_t tag, unsigned long long status,
                                                        unsigned long long v) {
  return (static_cast<unsigned long long>(tag & 0x3FFFFFu) << 42) | (status << kValueBits) |
         (v & kValueMask);
}

// 0: not published in this extraction, else the status.
__device__ __forceinline__ unsigned long long slot_status(unsigned long long w, uint32_t tag) {
  return (w >> 42) == (tag & 0x3FFFFFu) ? (w >> kValueBits) & 3ull : 0ull;
}

// Row first ids: tag (24 bits of the epoch) | value (40).
__device__ __forceinline__ unsigned long long row_word_tagged(uint32_t tag, unsigned long long v) {
  return (static_cast<unsigned long long>(tag & 0xFFFFFFu) << 40) | (v & kValueMask);
}

struct Lds {
  // Offsets (in bytes) into the dynamic LDS of the arrays below; see
  // fused_lds_layout.
  uint32_t sg, vpre, mpre, mixed, pvpre, pbase, uvpre, flat, total;
};

__host__ __device__ inline uint32_t align16(uint32_t v) { return (v + 15u) & ~15u; }

__host__ __device__ inline Lds fused_lds_layout(int R, int wpr, int spr) {
  Lds o;
  uint32_t at = 0;
  o.sg = at;  // 3 layers x (R + 2) rows x wpr words
  at = align16(at + 3u * static_cast<uint32_t>(R + 2) * wpr * 8u);
  o.mixed = at;  // R * spr mixed-cell masks
  at = align16(at + static_cast<uint32_t>(R) * spr * 8u);
  o.vpre = at;  // R * spr exclusive vertex counts (tile-local)
  at = align16(at + static_cast<uint32_t>(R) * spr * 4u);
  o.mpre = at;  // R * spr + 1 exclusive mixed-cell counts
  at = align16(at + (static_cast<uint32_t>(R) * spr + 1u) * 4u);
  o.pvpre = at;  // (R + 1) * spr in-row vertex counts of the previous layer's rows
  at = align16(at + static_cast<uint32_t>(R + 1) * spr * 4u);
  o.pbase = at;  // R + 1 first vertex ids of those rows
  at = align16(at + static_cast<uint32_t>(R + 1) * 4u);
  o.uvpre = at;  // spr in-row vertex counts of the row above the tile
  at = align16(at + static_cast<uint32_t>(spr) * 4u);
  o.flat = at;  // flat sign words of the produced rows: (R + 1) rows + 2 words
  at = align16(at + (static_cast<uint32_t>(R + 1) * wpr + 2u) * 8u);
  o.total = at;
  return o;
}

// Exclusive scan of one value per thread over the workgroup; returns the
// thread's prefix and sets *total.  `ws` holds kWaves entries.
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t* ws, uint32_t* total) {
  const int lane = __lane_id(), wv = threadIdx.x >> 6;
  const uint32_t inc = wave_incl_scan(v);
  if (lane == 63) ws[wv] = inc;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int i = 0; i < kWaves; ++i) {
    const uint32_t s = ws[i];
    if (i < wv) before += s;
    all += s;
  }
  __syncthreads();
  *total = all;
  return before + inc - v;
}

// The same over 64-bit values; `ws` holds kWaves entries.
__device__ __forceinline__ unsigned long long block_scan64(unsigned long long v,
                                                          unsigned long long* ws,
                                                          unsigned long long* total) {
  const int lane = __lane_id(), wv = threadIdx.x >> 6;
  unsigned long long inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned long long o = __shfl_up(inc, d, 64);
    if (lane >= d) inc += o;
  }
  if (lane == 63) ws[wv] = inc;
  __syncthreads();
  unsigned long long before = 0, all = 0;
#pragma unroll
  for (int i = 0; i < kWaves; ++i) {
    const unsigned long long x = ws[i];
    if (i < wv) before += x;
    all += x;
  }
  __syncthreads();
  *total = all;
  return before + inc - v;
}

struct Tile {
  int32_t lz, y0, y1;   // cell layer, cell rows [y0, y1)
  int32_t R, wpr, spr;
  int64_t nx, ny, nz;
  const uint64_t* sg;   // LDS sign rows
  // sign row (slot 0: layer lz-1, 1: lz, 2: lz+1) of sample row y
  __device__ __forceinline__ const uint64_t* row(int slot, int y) const {
    return sg + (static_cast<int64_t>(slot) * (R + 2) + (y - (y0 - 1))) * wpr;
  }
  // Segment words of 64-cell segment s of the cell row (y, layer of slot lo)
  __device__ __forceinline__ SegWords seg(int lo, int y, int s) const {
    const uint64_t* a = row(lo, y);
    const uint64_t* b = row(lo, y + 1);
    const uint64_t* c = row(lo + 1, y);
    const uint64_t* d = row(lo + 1, y + 1);
    const bool nx1 = s + 1 < wpr;
    return seg_words(a[s], nx1 ? a[s + 1] & 1ull : 0ull, b[s], nx1 ? b[s + 1] & 1ull : 0ull, c[s],
                     nx1 ? c[s + 1] & 1ull : 0ull, d[s], nx1 ? d[s + 1] & 1ull : 0ull);
  }
  // Owned vertices of the cells before x in the cell row (y, cell layer
  // lo == 0 ? lz-1 : lz) within x's segment.
  __device__ __forceinline__ uint32_t owned_before(int lo, int y, int x) const {
    const int s = x >> 6;
    const SegWords sw = seg(lo, y, s);
    uint64_t vp[4];
    const int cz = lo == 0 ? lz - 1 : lz;
    owned_planes(sw, seg_valid(nx, s), y == 0, cz == 0, s == 0, vp);
    return planes_count(vp, low_mask(x & 63));
  }
  // Three sign bits x-1, x, x+1 of sample row y in slot `slot` (bits
  // outside the lattice read as zero and are never used).
  __device__ __forceinline__ uint32_t bits3(int slot, int y, int x) const {
    const uint64_t* r = row(slot, y);
    const int q = x - 1;
    if (q < 0) return static_cast<uint32_t>((r[0] << 1) & 7ull);
    const int w = q >> 6, b = q & 63;
    uint64_t v = r[w] >> b;
    if (b > 61 && w + 1 < wpr) v |= r[w + 1] << (64 - b);
    return static_cast<uint32_t>(v & 7ull);
  }
};

__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4, 4))) void
k_mc_fused(const float* __restrict__ g, Lattice L,
                                                       FusedPlan P, FusedBuffers B,
                                                       uint32_t epoch) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ __align__(16) EmitTables s_t;
  __shared__ uint32_t s_ws[3][kWaves];
  __shared__ unsigned long long s_ws64[kWaves];
  __shared__ uint32_t s_tile, s_ubase;
  // per thread of the emit: first vertex id and case of each owner cell slot
  __shared__ uint32_t s_obase[7][kThreads];
  __shared__ uint8_t s_ocase[7][kThreads];
  __shared__ unsigned long long s_vb, s_tb;

  const int tid = threadIdx.x;
  const int lane = __lane_id();
  const int wave = tid >> 6;
  const int R = P.R, wpr = P.wpr, spr = P.spr, TL = P.TL;
  const int64_t nx = L.nx, ny = L.ny;
  const Lds lo = fused_lds_layout(R, wpr, spr);
  uint64_t* sg = reinterpret_cast<uint64_t*>(smem + lo.sg);
  uint64_t* s_mixed = reinterpret_cast<uint64_t*>(smem + lo.mixed);
  uint32_t* s_vpre = reinterpret_cast<uint32_t*>(smem + lo.vpre);
  uint32_t* s_mpre = reinterpret_cast<uint32_t*>(smem + lo.mpre);
  uint32_t* s_pvpre = reinterpret_cast<uint32_t*>(smem + lo.pvpre);
  uint32_t* s_pbase = reinterpret_cast<uint32_t*>(smem + lo.pbase);
  uint32_t* s_uvpre = reinterpret_cast<uint32_t*>(smem + lo.uvpre);
  uint64_t* s_flat = reinterpret_cast<uint64_t*>(smem + lo.flat);
  const uint32_t tag = epoch;
  unsigned long long* agg = B.slots;
  unsigned long long* inc = B.slots + 2 * P.n_tickets;
  const int64_t rows_per_layer = ny - 1;
  const int64_t nxy = nx * ny;

  // Error bits of this extraction: the word carries the epoch, so bits of
  // an earlier extraction read as none.  Host-coherent copies: slot
  // kTotErrors for an id overflow, kTotErrors + 1 for a wait that gave up.
  const unsigned long long ep_hi = static_cast<unsigned long long>(epoch) << 32;
  auto raise_error = [&](unsigned long long bit) {
    unsigned long long* w = &B.ctl[kFusedErrors];
    unsigned long long old = atom_load(w);
    while (true) {
      const unsigned long long nw = (((old >> 32) == epoch) ? old : ep_hi) | bit;
      if (nw == old) break;
      const unsigned long long prev = atomicCAS(w, old, nw);
      if (prev == old) break;
      old = prev;
    }
    if (B.host_totals) B.host_totals[bit == 1ull ? kTotErrors : kTotErrors + 1] = 1ull;
  };
  // Aborts every wait of the extraction once one has timed out.
  auto give_up = [&](uint32_t& spins) -> bool {
    if (++spins > kLookBackSpinLimit) {
      raise_error(2ull);
      return true;
    }
    if ((spins & 255u) == 0u) {
      const unsigned long long w = atom_load(&B.ctl[kFusedErrors]);
      if ((w >> 32) == epoch && (w & 2ull)) return true;
    }
    __builtin_amdgcn_s_sleep(1);
    return false;
  };

  load_table16(&s_t, &f_emit);
  // Tickets come from counter (epoch & 1); this launch clears the other
  // one, which the next extraction uses (the previous one cleared this).
  unsigned long long* ticket_ctr = &B.ctl[kFusedTicket + (epoch & 1u)];
  if (blockIdx.x == 0 && tid == 0) atom_store(&B.ctl[kFusedTicket + ((epoch + 1u) & 1u)], 0ull);

  // Persistent: tickets in order until the extraction's are all drawn.
  // Tickets [0, TL) produce sample layer 0; ticket (lz + 1) TL + j is the
  // tile of cell layer lz, rows [j R, j R + R).
  while (true) {
    __syncthreads();  // the previous tile's LDS reads are done
    if (tid == 0) s_tile = static_cast<uint32_t>(atomicAdd(ticket_ctr, 1ull));
    __syncthreads();
    const uint32_t t = s_tile;
    if (static_cast<int64_t>(t) >= P.n_tickets) break;
    const int lz = static_cast<int>(t / static_cast<uint32_t>(TL)) - 1;
    const int j = static_cast<int>(t % static_cast<uint32_t>(TL));
    const int y0 = j * R;
    const int y1 = y0 + R < static_cast<int>(ny) - 1 ? y0 + R : static_cast<int>(ny) - 1;
    const int ylo = j == 0 ? 0 : y0 + 1;  // produced sample rows [ylo, y1]
    const Tile T{lz, y0, y1, R, wpr, spr, nx, ny, L.nz, sg};
    // A producer of sample layer 0 has no cells: its look-back words are
    // an inclusive zero from the start.
    if (lz < 0 && tid == 0) {
      atom_store(&inc[2 * t], slot_word(tag, kStatusInclusive, 0));
      atom_store(&inc[2 * t + 1], slot_word(tag, kStatusInclusive, 0));
    }

    // ---- 1. produce: sign rows [ylo, y1] of sample layer lz + 1 ----------
    // The rows are contiguous in the grid: stream them flat, one aligned
    // 256-byte load per 64 samples (ballot = flat sign word), into LDS,
    // then cut the flat bits into row-aligned words.
    {
      const int64_t z = lz + 1;
      const int nrow = y1 - ylo + 1;
      const int64_t f0 = (z * ny + ylo) * nx;               // first produced sample
      const int64_t f1 = f0 + static_cast<int64_t>(nrow) * nx;
      const int64_t k0 = f0 >> 6;
      const int n_flat = static_cast<int>(((f1 - 1) >> 6) - k0 + 1);
      for (int b0 = wave * kLoadBatch; b0 < n_flat; b0 += kWaves * kLoadBatch) {
        // Every load is issued unconditionally (addresses clamped into the
        // produced range), so the wavefront has all of them in flight at
        // once; samples outside the range are replaced afterwards.
        // Offsets relative to the batch's first sample (uniform base, 32-bit
        // lane offsets): the produced range is [rel0, rel1).
        const int64_t base = 64 * (k0 + b0);
        const float* gb = g + base;
        const int rel0 = static_cast<int>(f0 - base);
        const int rel1 = static_cast<int>((f1 < base + 64 * kLoadBatch ? f1 : base + 64 * kLoadBatch) - base);
        float v[kLoadBatch];
#pragma unroll
        for (int i = 0; i < kLoadBatch; ++i) {
          const int off = 64 * i + lane;
          const int oc = off < rel0 ? rel0 : (off >= rel1 ? rel1 - 1 : off);
          v[i] = __builtin_nontemporal_load(gb + oc);
        }
#pragma unroll
        for (int i = 0; i < kLoadBatch; ++i) {
          const int off = 64 * i + lane;
          if (off < rel0 || off >= rel1) v[i] = 1.0f;
        }
        uint64_t mine = 0;
#pragma unroll
        for (int i = 0; i < kLoadBatch; ++i) {
          const uint64_t m = __ballot(!(v[i] > 0.0f));  // classify_corners, distance.rs:52-54
          if (lane == i) mine = m;
        }
        if (lane < kLoadBatch && b0 + lane < n_flat) s_flat[b0 + lane] = mine;
      }
      __syncthreads();
      // Row-aligned words (bits past the row's end cleared), kept in LDS and
      // published: the rows are contiguous in the row-aligned sign array.
      const int sh0 = static_cast<int>(f0 & 63);
      unsigned long long* dst =
          reinterpret_cast<unsigned long long*>(B.signs) + (z * ny + ylo) * wpr;
      uint64_t* rows = sg + (2 * static_cast<int64_t>(R + 2) + (ylo - (y0 - 1))) * wpr;
      for (int i = tid; i < nrow * wpr; i += kThreads) {
        const int ri = i / wpr, w = i - ri * wpr;
        const int64_t pos = sh0 + static_cast<int64_t>(ri) * nx + 64 * w;
        const int64_t wi = pos >> 6;
        const int b = static_cast<int>(pos & 63);
        uint64_t word = s_flat[wi] >> b;
        if (b && wi + 1 < n_flat) word |= s_flat[wi + 1] << (64 - b);
        const int64_t valid = nx - 64 * static_cast<int64_t>(w);
        if (valid < 64) word &= (1ull << valid) - 1ull;
        rows[i] = word;
        atom_store(dst + i, word);
      }
      wait_stores();
      __syncthreads();
      if (tid == 0) atom_store(&B.flags[t], epoch);
    }
    if (lz < 0) continue;

    // ---- 2. acquire the sign rows other tiles produce ---------------------
    // Slot s holds sample layer lz - 1 + s, rows y0-1 .. y1.  Missing rows:
    // slot 2 below ylo, slots 0 and 1 all.
    {
      const int nr = R + 2;
      auto missing = [&](int slot, int y) -> bool {
        if (y < 0 || lz - 1 + slot < 0) return false;
        return slot < 2 || y < ylo;
      };
      for (int i = tid; i < 3 * nr; i += kThreads) {
        const int slot = i / nr, y = y0 - 1 + (i - slot * nr);
        if (missing(slot, y)) {
          // producer ticket of sample row y of layer z: z TL + (y - 1) / R
          const int64_t p = static_cast<int64_t>(lz - 1 + slot) * TL + (y == 0 ? 0 : (y - 1) / R);
          uint32_t spins = 0;
          while (atom_load(&B.flags[p]) != epoch)
            if (give_up(spins)) break;
        } else if (y < 0 || lz - 1 + slot < 0) {
          for (int w = 0; w < wpr; ++w)
            sg[(static_cast<int64_t>(slot) * nr + (y - (y0 - 1))) * wpr + w] = 0;
        }
      }
      __syncthreads();
      // Up to kAcquireBatch words per thread per round, all loads issued
      // before any is used (a word that is not missing loads word 0 and is
      // dropped).
      const unsigned long long* gs = reinterpret_cast<const unsigned long long*>(B.signs);
      const int n_words = 3 * nr * wpr;
      for (int i0 = 0; i0 < n_words; i0 += kThreads * kAcquireBatch) {
        unsigned long long wv[kAcquireBatch];
        bool want[kAcquireBatch];
#pragma unroll
        for (int k = 0; k < kAcquireBatch; ++k) {
          const int i = i0 + k * kThreads + tid;
          const int rr = i / wpr, w = i - rr * wpr;
          const int slot = rr / nr, y = y0 - 1 + (rr - slot * nr);
          want[k] = i < n_words && missing(slot, y);
          const int64_t z = lz - 1 + slot;
          wv[k] = atom_load(gs + (want[k] ? (z * ny + y) * wpr + w : 0));
        }
#pragma unroll
        for (int k = 0; k < kAcquireBatch; ++k)
          if (want[k]) sg[i0 + k * kThreads + tid] = wv[k];
      }
      __syncthreads();
    }

    // ---- 3. count own segments --------------------------------------------
    const int n_pairs = (y1 - y0) * spr;
    uint32_t tot_v = 0, tot_t = 0, tot_a = 0;
    {
      uint32_t nv = 0, nt = 0, na = 0;
      uint64_t mix = 0;
      const int p = tid;
      if (p < n_pairs) {
        const int r = p / spr, s = p - (p / spr) * spr;
        const int y = y0 + r;
        const SegWords sw = T.seg(1, y, s);
        const uint64_t valid = seg_valid(nx, s);
        uint64_t vp[4];
        owned_planes(sw, valid, y == 0, lz == 0, s == 0, vp);
        nv = planes_count(vp, ~0ull);
        mix = mixed_cells(sw) & valid;
        na = popc64(mix);
        if (mix) {
          uint64_t tp[3];
          tri_planes(sw, mix, tp);
          nt = popc64(tp[0]) + 2 * popc64(tp[1]) + 4 * popc64(tp[2]);
        }
      }
      // One scan of the three counts packed in 64 bits (per tile: mixed
      // cells < 2^16, triangles < 2^18, vertices < 2^19).
      unsigned long long tot;
      const unsigned long long ex = block_scan64(
          (static_cast<unsigned long long>(nv) << 34) | (static_cast<unsigned long long>(nt) << 16) |
              na,
          s_ws64, &tot);
      const uint32_t ev = static_cast<uint32_t>(ex >> 34);
      const uint32_t ea = static_cast<uint32_t>(ex & 0xFFFFu);
      tot_v = static_cast<uint32_t>(tot >> 34);
      tot_t = static_cast<uint32_t>((tot >> 16) & 0x3FFFFu);
      tot_a = static_cast<uint32_t>(tot & 0xFFFFu);
      if (p < n_pairs) {
        s_vpre[p] = ev;
        s_mpre[p] = ea;
        s_mixed[p] = mix;
      }
      if (tid == 0) {
        s_mpre[n_pairs] = tot_a;
        // this tile's aggregate, as early as its successors can use it
        atom_store(&agg[2 * t], slot_word(tag, kStatusAggregate, tot_v));
        atom_store(&agg[2 * t + 1], slot_word(tag, kStatusAggregate, tot_t));
      }
    }

    // ---- 4a. the emit's cell decode, corner loads and local scans (chunk c0)
    // The first chunk is prepared before the look-back, so that the corner
    // loads' latency overlaps it.
    uint32_t c_e = 0;
    bool live = false;
    int x = 0, y = 0;
    uint32_t cube = 0, cs = 0, mine = 0, lv = 0, lt = 0, cv_total = 0, ct_total = 0;
    float cv[8];
    auto prep = [&](uint32_t c0) {
      const uint32_t e = c0 + static_cast<uint32_t>(tid);
      c_e = e;
      live = e < tot_a;
      x = 0;
      y = 0;
      cube = 0;
      cs = 0;
      mine = 0;
      uint32_t nvo = 0, ntri = 0;
#pragma unroll
      for (int k = 0; k < 8; ++k) cv[k] = 0.0f;
      if (live) {
        // the pair holding mixed cell e, then its bit
        int lo_p = 0, hi_p = n_pairs;  // last p with mpre[p] <= e
        while (hi_p - lo_p > 1) {
          const int mid = (lo_p + hi_p) >> 1;
          if (s_mpre[mid] <= e) lo_p = mid;
          else hi_p = mid;
        }
        const int pp = lo_p;
        uint64_t m = s_mixed[pp];
        for (uint32_t k = e - s_mpre[pp]; k; --k) m &= m - 1;
        const int r = pp / spr;
        x = 64 * (pp - r * spr) + __builtin_ctzll(m);
        y = y0 + r;
        // 3x3x3 sign neighbourhood (bit dx + 3 dy + 9 dz)
#pragma unroll
        for (int dz = 0; dz < 3; ++dz)
#pragma unroll
          for (int dy = 0; dy < 3; ++dy) cube |= T.bits3(dz, y - 1 + dy, x) << (3 * dy + 9 * dz);
        cs = case_in_cube(cube, 1, 1, 1);
        mine = owned_edges(x, y, lz) & s_t.edge_mask[cs];
        nvo = __popc(mine);
        ntri = s_t.tri_count[cs];
        // All eight corners, unconditionally: no branch between the loads,
        // so they are all in flight together (every corner is in the grid).
        const int64_t fc = (static_cast<int64_t>(lz) * ny + y) * nx + x;
#pragma unroll
        for (int k = 0; k < 8; ++k)
          cv[k] = g[fc + tables::kCorner[k][2] * nxy + tables::kCorner[k][1] * nx +
                    tables::kCorner[k][0]];
      }
      lv = block_scan(nvo, s_ws[0], &cv_total);
      lt = block_scan(ntri, s_ws[1], &ct_total);
    };
    __syncthreads();  // the count's per-pair arrays are complete
    prep(0);

    // ---- 4. look-back (wave 0) | previous-layer row counts (other waves) --
    if (wave == 0) {
      unsigned long long pv = 0, pt = 0;
      {
        int64_t jj = static_cast<int64_t>(t) - 1;
        uint32_t spins = 0;
        while (true) {
          const int64_t idx = jj - lane;
          unsigned long long xv = 0, xt = 0;
          int state = 2;  // tiles before the first act as an inclusive zero
          if (idx >= 0) {
            xv = atom_load(&inc[2 * idx]);
            xt = atom_load(&inc[2 * idx + 1]);
            if (slot_status(xv, tag) == kStatusInclusive && slot_status(xt, tag) == kStatusInclusive) {
              state = 2;
            } else {
              xv = atom_load(&agg[2 * idx]);
              xt = atom_load(&agg[2 * idx + 1]);
              state = (slot_status(xv, tag) == kStatusAggregate &&
                       slot_status(xt, tag) == kStatusAggregate) ? 1 : 0;
            }
          }
          const uint64_t incl = __ballot(state == 2);
          const uint64_t zero = __ballot(state == 0);
          const int st = incl ? __builtin_ctzll(incl) : 64;
          const uint64_t upto = st < 64 ? (2ull << st) - 1ull : ~0ull;
          if (zero & upto) {
            if (give_up(spins)) break;
            continue;
          }
          unsigned long long sv = 0, stt = 0;
          if (lane <= st) {
            sv = xv & kValueMask;
            stt = xt & kValueMask;
          }
  #pragma unroll
          for (int d = 32; d >= 1; d >>= 1) {
            sv += __shfl_xor(sv, d, 64);
            stt += __shfl_xor(stt, d, 64);
          }
          pv += sv;
          pt += stt;
          if (st < 64) break;
          jj -= 64;
        }
        if (lane == 0) {
          atom_store(&inc[2 * t], slot_word(tag, kStatusInclusive, pv + tot_v));
          atom_store(&inc[2 * t + 1], slot_word(tag, kStatusInclusive, pt + tot_t));
        }
      }
      if (lane == 0) {
        s_vb = pv;
        s_tb = pt;
        if (static_cast<int64_t>(t) == P.n_tickets - 1) {
          const unsigned long long V = pv + tot_v, Tt = pt + tot_t;
          if (V > 0xFFFFFFFFull || 3ull * Tt > 0xFFFFFFFFull) raise_error(1ull);
          B.ctl[kFusedVertices] = V;
          B.ctl[kFusedTriangles] = Tt;
          if (B.host_totals) {
            B.host_totals[kTotVertices] = V;
            B.host_totals[kTotTriangles] = Tt;
          }
        }
      }
    } else {
      // In-row vertex counts of the previous cell layer's rows y0-1 .. y1-1
      // and of the row above the tile: one thread per row, segments in order.
      const int ya = y0 > 0 ? y0 - 1 : 0;
      const int nprev = lz > 0 ? y1 - ya : 0;
      const int n_items = nprev + (y0 > 0 ? 1 : 0);
      for (int i = tid - 64; i < n_items; i += kThreads - 64) {
        const bool up = i == nprev;  // the row above the tile, in this layer
        const int y = up ? y0 - 1 : ya + i;
        const int slot = up ? 1 : 0;
        const int cz = up ? lz : lz - 1;
        uint32_t* out = up ? s_uvpre : s_pvpre + i * spr;
        uint32_t run = 0;
        for (int s = 0; s < spr; ++s) {
          out[s] = run;
          const SegWords sw = T.seg(slot, y, s);
          uint64_t vp[4];
          owned_planes(sw, seg_valid(nx, s), y == 0, cz == 0, s == 0, vp);
          run += planes_count(vp, ~0ull);
        }
      }
    }
    __syncthreads();
    const unsigned long long VB = s_vb, TB = s_tb;

    // ---- publish own row first ids, then wait for the owner rows' ----------
    for (int r = tid; r < y1 - y0; r += kThreads)
      atom_store(&B.rowbase[static_cast<int64_t>(lz) * rows_per_layer + y0 + r],
                 row_word_tagged(tag, VB + s_vpre[r * spr]));
    {
      const int ya = y0 > 0 ? y0 - 1 : 0;
      const int nprev = lz > 0 ? y1 - ya : 0;
      const uint32_t want = tag & 0xFFFFFFu;
      const int n_items = nprev + (y0 > 0 ? 1 : 0);
      for (int i = tid; i < n_items; i += kThreads) {
        const bool up = i == nprev;
        const int64_t row = up ? static_cast<int64_t>(lz) * rows_per_layer + (y0 - 1)
                               : static_cast<int64_t>(lz - 1) * rows_per_layer + ya + i;
        unsigned long long w = atom_load(&B.rowbase[row]);
        uint32_t spins = 0;
        while ((w >> 40) != want) {
          if (give_up(spins)) break;
          w = atom_load(&B.rowbase[row]);
        }
        if (up) s_ubase = static_cast<uint32_t>(w & kValueMask);
        else s_pbase[i] = static_cast<uint32_t>(w & kValueMask);
      }
    }
    __syncthreads();

    // ---- 5. emit the mixed cells in cell order ------------------------------
    const int ya = y0 > 0 ? y0 - 1 : 0;
    unsigned long long run_v = VB, run_t = TB;
    for (uint32_t c0 = 0; c0 < tot_a; c0 += kThreads) {
      if (c0) prep(c0);
      if (live) {
        const uint32_t vbase_cell = static_cast<uint32_t>(run_v + lv);
        const bool interior = x >= 1 && y >= 1 && lz >= 1;
        const bool deep = x >= 2 && y >= 2 && lz >= 2;
        // Vertices this cell owns: positions along its own edge directions
        // (a cell away from the low faces owns only e5, e6 and e10).
        auto vertex = [&](int ed, int u, int v, float va, float vb) {
          const uint32_t id = vbase_cell + static_cast<uint32_t>(__popc(mine & s_t.before[cs][ed]));
          float pos[3];
          crossing(va, vb, lattice_coord(L, 0, x + tables::kCorner[u][0]),
                   lattice_coord(L, 1, y + tables::kCorner[u][1]),
                   lattice_coord(L, 2, lz + tables::kCorner[u][2]),
                   lattice_coord(L, 0, x + tables::kCorner[v][0]),
                   lattice_coord(L, 1, y + tables::kCorner[v][1]),
                   lattice_coord(L, 2, lz + tables::kCorner[v][2]), pos);
          if (id < B.vcap) {
            B.vertices[3 * static_cast<uint64_t>(id) + 0] = pos[0];
            B.vertices[3 * static_cast<uint64_t>(id) + 1] = pos[1];
            B.vertices[3 * static_cast<uint64_t>(id) + 2] = pos[2];
          }
        };
        if (interior) {
          if (mine & (1u << 5)) vertex(5, 5, 6, cv[5], cv[6]);
          if (mine & (1u << 6)) vertex(6, 6, 7, cv[6], cv[7]);
          if (mine & (1u << 10)) vertex(10, 2, 6, cv[2], cv[6]);
        } else {
          for (uint32_t todo = mine; todo; todo &= todo - 1) {
            const int ed = __builtin_ctz(todo);
            const int u = tables::kEdgeCorners[ed][0], v = tables::kEdgeCorners[ed][1];
            vertex(ed, u, v, cv[u], cv[v]);
          }
        }
        // First vertex id and case of every owner cell the cell's edges
        // need, slot rx | ry << 1 | rz << 2 for the cell (x-rx, y-ry, z-rz).
        uint32_t needed = 0;
        if (interior) {
          needed = s_t.need[cs];
        } else {
          for (uint32_t mm = s_t.edge_mask[cs]; mm; mm &= mm - 1) {
            const int ed = __builtin_ctz(mm);
            const uint32_t gm = static_cast<uint32_t>(kEdgeGeomPacked >> (5 * ed)) & 31u;
            const int axis = static_cast<int>(gm >> 3);
            const int rx = (axis != 0 && (gm & 1u) == 0 && x >= 1) ? 1 : 0;
            const int ry = (axis != 1 && (gm & 2u) == 0 && y >= 1) ? 1 : 0;
            const int rz = (axis != 2 && (gm & 4u) == 0 && lz >= 1) ? 1 : 0;
            needed |= 1u << (rx | (ry << 1) | (rz << 2));
          }
        }
        s_obase[0][tid] = vbase_cell;
        s_ocase[0][tid] = static_cast<uint8_t>(cs);
#pragma unroll 1
        for (int o = 1; o < 7; ++o) {
          if (!((needed >> o) & 1u)) continue;
          const int rx = o & 1, ry = (o >> 1) & 1, rz = (o >> 2) & 1;
          const int ox = x - rx, oy = y - ry;
          uint32_t base;
          if (rz) {
            base = s_pbase[oy - ya] + s_pvpre[(oy - ya) * spr + (ox >> 6)] +
                   T.owned_before(0, oy, ox);
          } else if (oy >= y0) {
            base = static_cast<uint32_t>(VB) + s_vpre[(oy - y0) * spr + (ox >> 6)] +
                   T.owned_before(1, oy, ox);
          } else {
            base = s_ubase + s_uvpre[ox >> 6] + T.owned_before(1, oy, ox);
          }
          s_obase[o][tid] = base;
          s_ocase[o][tid] = static_cast<uint8_t>(case_in_cube(cube, 1 - rx, 1 - ry, 1 - rz));
        }
        // Triangles: cell order, then table order (mesh.rs:91-100).
        const uint64_t tbase = run_t + lt;
        const uint32_t nidx = 3u * s_t.tri_count[cs];
        for (uint32_t jj = 0; jj < nidx; ++jj) {
          int o, eo;
          if (interior) {
            const uint32_t ix = s_t.ix[cs][jj];
            o = static_cast<int>(ix & 7u);
            eo = static_cast<int>(ix >> 3);
          } else {
            const int ed = f_tables.tri[cs][jj];
            const uint32_t gm = static_cast<uint32_t>(kEdgeGeomPacked >> (5 * ed)) & 31u;
            const int axis = static_cast<int>(gm >> 3);
            const int px = gm & 1, py = (gm >> 1) & 1, pz = (gm >> 2) & 1;
            const int rx = (axis != 0 && px == 0 && x >= 1) ? 1 : 0;
            const int ry = (axis != 1 && py == 0 && y >= 1) ? 1 : 0;
            const int rz = (axis != 2 && pz == 0 && lz >= 1) ? 1 : 0;
            o = rx | (ry << 1) | (rz << 2);
            eo = edge_of_axis_min(axis, (px | rx) | ((py | ry) << 1) | ((pz | rz) << 2));
          }
          const int rx = o & 1, ry = (o >> 1) & 1, rz = (o >> 2) & 1;
          const uint32_t oc = s_ocase[o][tid];
          const uint32_t oedges = deep ? 0x460u : owned_edges(x - rx, y - ry, lz - rz);
          const uint32_t oown = oedges & s_t.edge_mask[oc];
          const uint32_t id = s_obase[o][tid] + static_cast<uint32_t>(__popc(oown & s_t.before[oc][eo]));
          const uint64_t ix = 3 * tbase + jj;
          if (ix < B.icap) B.indices[ix] = id;
        }
      }
      run_v += cv_total;
      run_t += ct_total;
    }
  }
}

}  // namespace

bool fused_plan(const Lattice& L, FusedPlan* P) {
  if (L.nx < 2 || L.ny < 2 || L.nz < 2) return false;
  if (L.nx > (int64_t(1) << 30) || L.ny > (int64_t(1) << 30) || L.nz > (int64_t(1) << 30))
    return false;
  const int64_t spr = (L.nx - 1 + 63) / 64;
  const int64_t wpr = (L.nx + 63) / 64;
  if (spr > kThreads) return false;
  int64_t R = kThreads / spr;
  if (R > L.ny - 1) R = L.ny - 1;
  const Lds lo = fused_lds_layout(static_cast<int>(R), static_cast<int>(wpr),
                                  static_cast<int>(spr));
  if (lo.total > kFusedMaxLds) return false;
  P->R = static_cast<int32_t>(R);
  P->TL = static_cast<int32_t>((L.ny - 1 + R - 1) / R);
  P->n_tiles = static_cast<int64_t>(P->TL) * (L.nz - 1);
  P->n_tickets = P->n_tiles + P->TL;
  if (P->n_tickets > 0x7FFFFFFF) return false;
  P->wpr = static_cast<int32_t>(wpr);
  P->spr = static_cast<int32_t>(spr);
  P->lds_bytes = lo.total;
  P->sign_words = static_cast<uint64_t>(L.nz) * L.ny * wpr;
  P->n_rows = static_cast<uint64_t>(L.nz - 1) * (L.ny - 1);
  return true;
}

hipError_t launch_mc_fused(const float* grid, const Lattice& L, const FusedPlan& P,
                           const FusedBuffers& B, uint32_t epoch, hipStream_t s) {
  if (!grid || !B.signs || !B.flags || !B.slots || !B.rowbase || !B.ctl)
    return hipErrorInvalidValue;
  // Persistent workgroups: as many as are resident at once (tickets are
  // drawn in order, so a tile's predecessors are always held by running
  // workgroups), at most one per ticket.
  int device = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&device);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mc_fused, kThreads,
                                                     static_cast<size_t>(P.lds_bytes));
  if (e != hipSuccess) return e;
  if (per_cu < 1) return hipErrorInvalidConfiguration;
  int64_t n_groups = static_cast<int64_t>(cus) * per_cu;
  if (n_groups > P.n_tickets) n_groups = P.n_tickets;
  const dim3 blocks(static_cast<unsigned>(n_groups)), threads(kThreads);
  const unsigned lds = static_cast<unsigned>(P.lds_bytes);
  hipEvent_t a = nullptr, z = nullptr;
  if (take_launch_events(&a, &z))
    hipExtLaunchKernelGGL(k_mc_fused, blocks, threads, lds, s, a, z, 0, grid, L, P, B, epoch);
  else
    hipLaunchKernelGGL(k_mc_fused, blocks, threads, lds, s, grid, L, P, B, epoch);
  return hipGetLastError();
}

}  // namespace iso
