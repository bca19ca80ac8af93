// C ABI of isosurface_amd: contexts, sources, extraction, timing.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/isosurface.h"
#include "band_kernels.h"
#include "dc_kernels.h"
#include "emc_kernels.h"
#include "lhmc_kernels.h"
#include "mc_kernels.h"
#include "mlp_kernels.h"
#include "program.h"

namespace {

// Device allocation that grows on demand and is reused across calls.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

struct KernelStat {
  uint64_t launches = 0;
  double total_ms = 0.0;
};

struct PendingTiming {
  const char* name;
  hipEvent_t start, stop;
};

}  // namespace

// What a captured single-lattice extraction was captured for: every kernel
// argument of its launches derives from these.
struct MeshGraphKey {
  iso::Lattice L;
  const float* grid;
  const void* ops;
  const float* values;
  int64_t n_prog;
  iso::McBuffers B;
};

struct iso_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;

  // current program
  DevBuf prog;
  int n_prog = 0;
  bool prog_directed = false;  // the program has a directed form (no smooth union)
  bool prog_hermite = false;   // analytic normals exist (shapes, unions, translations)

  // current grid
  DevBuf grid_owned;
  const float* grid = nullptr;
  int64_t gnx = 0, gny = 0, gnz = 0;

  // scratch
  DevBuf signs, seg, tplanes, vplanes, agg, incl, active, totals, vout, iout, pts, vals;
  DevBuf pvals;  // program samples of the current extraction (emit reads them)
  DevBuf dense_check;  // dense sign words a band extraction is checked against
  DevBuf nrm;    // normals of iso_sample_normals (host-memory calls)
  // directed values: the uploaded grid, and a program's values of the
  // current extraction (3 floats per lattice point)
  DevBuf dgrid, dvals;
  int64_t dgnx = 0, dgny = 0, dgnz = 0;
  // extended marching cubes: per active cell state, first ids, flip records
  DevBuf emc_cells, emc_first, emc_flips, emc_ctr, emc_cross;
  // dual contouring: primal samples, dual values and points, mixed list, counters
  DevBuf dc_normals, dc_plist, dc_pindex, dc_samples, dc_values, dc_points, dc_mixed, dc_flags, dc_counters;
  unsigned long long* h_dc = nullptr;  // pinned
  // octree marching cubes: nodes, dense primal lattice (or the sorted path's
  // pairs and sort scratch), cubes, edge table, scan state
  DevBuf oc_keys, oc_values, oc_leaf, oc_dense, oc_primal, oc_plevel, oc_corner, oc_case,
      oc_base, oc_tkeys, oc_towner, oc_tid, sc_slot, sc_ctl, oc_pkeys, oc_plevels, oc_skeys,
      oc_slevels, oc_sort;
  unsigned long long* h_stage = nullptr;  // pinned, 8 words

  // batched chunks: lowered programs, chunk table, sign words, offsets
  DevBuf b_ops, b_chunks, b_segstart, b_wordstart, b_signs, b_voff, b_ioff, b_values;
  iso::BatchPlan plan{};
  uint64_t b_cells = 0;
  bool b_ready = false;
  // launch shapes (chosen by the autotuner when asked)
  int b_rounds = 1, b_words_per_wave = 1, b_emit_per_cu = 0;
  bool b_tuned = false;
  // learned narrow band: network, training batch, brick buffers, counters
  DevBuf nn_theta, nn_m, nn_v, nn_grad, nn_frags, nn_loss, nn_x, nn_h, nn_dz, nn_g, nn_pts, nn_t,
      nn_w, nn_out, br_centres, br_hd, br_pred, br_flags, br_list, br_count, br_coarse, br_proof;
  bool nn_ready = false;
  uint32_t nn_step = 0;  // Adam steps taken
  int64_t nn_batch = 0;  // batch the activation buffers are sized for
  iso_grid_desc band_desc{};  // lattice of the last band extraction
  bool band_valid = false;

  // single-pass grid extraction: row-aligned signs, tile flags, look-back
  // words, row first ids and control words (all epoch-tagged), and the epoch
  DevBuf f_signs, f_flags, f_slots, f_rowbase, f_ctl;
  uint32_t f_epoch = 0;

  // captured extraction and the buffers it was captured with
  hipGraphExec_t b_graph = nullptr;
  iso::McBuffers g_buffers{};
  // captured single-lattice extraction (ISO_MESH_GRAPH)
  hipGraphExec_t m_graph = nullptr;
  MeshGraphKey m_key{};
  uint64_t active_capacity = 0;
  unsigned long long* h_totals = nullptr;  // pinned
  // host-coherent totals that emit writes (host and device addresses)
  unsigned long long* h_coh = nullptr;
  unsigned long long* d_coh = nullptr;

  // mesh retained after a host-memory extraction whose buffers were too small
  uint64_t kept_vertices = 0, kept_indices = 0;
  bool kept = false;

  // timing
  bool timing = false;
  std::vector<hipEvent_t> event_pool;
  std::vector<PendingTiming> pending;
  std::map<std::string, KernelStat> stats;
};

namespace {

int fail(iso_ctx* c, int status, const std::string& msg) {
  if (c) c->err = msg;
  return status;
}

#define ISO_HIP(ctx, expr)                                                              \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail((ctx), e_ == hipErrorOutOfMemory ? ISO_ERR_OUT_OF_MEMORY : ISO_ERR_DEVICE, \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                    \
  } while (0)

hipError_t ensure(DevBuf* b, size_t bytes) {
  if (b->bytes >= bytes && b->ptr) return hipSuccess;
  if (b->ptr) {
    hipError_t e = hipFree(b->ptr);
    if (e != hipSuccess) return e;
    b->ptr = nullptr;
    b->bytes = 0;
  }
  size_t want = bytes < 256 ? 256 : bytes;
  hipError_t e = hipMalloc(&b->ptr, want);
  if (e != hipSuccess) {
    b->ptr = nullptr;
    return e;
  }
  b->bytes = want;
  return hipSuccess;
}

void release(DevBuf* b) {
  if (b->ptr) (void)hipFree(b->ptr);
  b->ptr = nullptr;
  b->bytes = 0;
}

hipEvent_t take_event(iso_ctx* c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Times one launch when timing is on.  The MarchingCubes signs, count and
// emit launchers take the event pair into the dispatch itself, so the pair
// spans exactly the kernel's execution; any other launch is bracketed by
// stream records, which also count the gaps the stream leaves around it.
template <typename F>
hipError_t timed(iso_ctx* c, const char* name, F&& launch) {
  if (!c->timing) return launch();
  hipEvent_t a = take_event(c), b = take_event(c);
  if (!a || !b) return launch();
  (void)hipEventRecord(a, c->stream);
  iso::set_launch_events(a, b);
  hipError_t e = launch();
  if (iso::launch_events_pending()) {
    iso::clear_launch_events();
    (void)hipEventRecord(b, c->stream);
  }
  c->pending.push_back({name, a, b});
  return e;
}

// After the stream has been synchronised: fold pending event pairs into stats.
void collect_timing(iso_ctx* c) {
  for (const PendingTiming& p : c->pending) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
      KernelStat& s = c->stats[p.name];
      s.launches++;
      s.total_ms += ms;
    }
    c->event_pool.push_back(p.start);
    c->event_pool.push_back(p.stop);
  }
  c->pending.clear();
}

int check_lattice(iso_ctx* c, const iso_grid_desc* d, iso::Lattice* L) {
  if (!d) return fail(c, ISO_ERR_INVALID_ARGUMENT, "null grid descriptor");
  if (d->nx < 2 || d->ny < 2 || d->nz < 2)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "lattice needs at least 2 samples per axis");
  const double n = static_cast<double>(d->nx) * static_cast<double>(d->ny) * static_cast<double>(d->nz);
  if (n >= 4294967296.0)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "lattice has 2^32 or more samples");
  if (d->source != ISO_SOURCE_PROGRAM && d->source != ISO_SOURCE_GRID)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "unknown source kind");
  if (d->flags != 0) return fail(c, ISO_ERR_INVALID_ARGUMENT, "flags must be 0");
  L->nx = d->nx;
  L->ny = d->ny;
  L->nz = d->nz;
  for (int i = 0; i < 3; ++i) {
    L->origin[i] = d->origin[i];
    L->spacing[i] = d->spacing[i];
  }
  return ISO_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Scratch of the extraction that allocates nothing (iso_extract_mc_noalloc).
// The single pass's scratch words live in one file-scope device array in
// fixed regions, each only ever holding its own kind of word, so a word of
// one kind is never read as another; every word but the sign rows carries
// the extraction's epoch, and the epochs of one device only grow, so the
// array is never cleared between calls (only when the epoch's 22 tag bits
// wrap).  Each GPU holds its own copy of the array; its device address is
// looked up once per device, at context creation, and used from then on.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kNaSignBytes = size_t(28) << 20;  // row-aligned sign words
constexpr uint64_t kNaMaxTickets = uint64_t(1) << 16;
constexpr uint64_t kNaMaxRows = uint64_t(1) << 18;
constexpr size_t kNaFlagsOff = kNaSignBytes;
constexpr size_t kNaSlotsOff = kNaFlagsOff + kNaMaxTickets * 8;
constexpr size_t kNaRowbaseOff = kNaSlotsOff + 4 * kNaMaxTickets * 8;
constexpr size_t kNaCtlOff = kNaRowbaseOff + kNaMaxRows * 8;
constexpr size_t kNaBytes = kNaCtlOff + 256;
constexpr int kNaMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {
struct NoAllocDevice {
  unsigned char* ptr = nullptr;  // this device's copy of `scratch`
  uint32_t epoch = 0;
  bool looked_up = false;
  hipError_t status = hipSuccess;
};
std::mutex g_na_mutex;
NoAllocDevice g_na[kNaMaxDevices];

// With `device` current: the address of its copy of the array, once.
void noalloc_setup(int device) {
  if (device < 0 || device >= kNaMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_na_mutex);
  NoAllocDevice& d = g_na[device];
  if (d.looked_up) return;
  void* p = nullptr;
  d.status = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  d.ptr = d.status == hipSuccess ? static_cast<unsigned char*>(p) : nullptr;
  d.looked_up = true;
}
}  // namespace

extern "C" {

int iso_abi_version(void) { return ISO_ABI_VERSION; }

const char* iso_status_string(int status) {
  switch (status) {
    case ISO_OK: return "ok";
    case ISO_ERR_INVALID_ARGUMENT: return "invalid argument";
    case ISO_ERR_CAPACITY: return "output capacity too small";
    case ISO_ERR_DEVICE: return "device error";
    case ISO_ERR_OUT_OF_MEMORY: return "out of device memory";
    case ISO_ERR_PROGRAM: return "malformed program";
    case ISO_ERR_NO_DEVICE: return "no such device";
    case ISO_ERR_NO_SOURCE: return "no source bound";
    case ISO_ERR_OVERFLOW: return "mesh exceeds u32 index range";
  }
  return "unknown status";
}

const char* iso_ctx_last_error(const iso_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int iso_ctx_create(int device, void* stream, iso_ctx** out) {
  if (!out) return ISO_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return ISO_ERR_NO_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return ISO_ERR_DEVICE;
  iso_ctx* c = new iso_ctx();
  c->device = device;
  noalloc_setup(device);
  if (stream) {
    c->stream = static_cast<hipStream_t>(stream);
  } else {
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
      delete c;
      return ISO_ERR_DEVICE;
    }
    c->own_stream = true;
  }
  if (hipHostMalloc(reinterpret_cast<void**>(&c->h_totals), iso::kTotSlots * sizeof(unsigned long long),
                    hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&c->h_dc), iso::kDcSlots * sizeof(unsigned long long),
                    hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&c->h_stage), 8 * sizeof(unsigned long long),
                    hipHostMallocDefault) != hipSuccess) {
    if (c->h_totals) (void)hipHostFree(c->h_totals);
    if (c->h_dc) (void)hipHostFree(c->h_dc);
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return ISO_ERR_OUT_OF_MEMORY;
  }
  // Host-coherent totals for device-output extractions (optional: without
  // them the totals are copied back after emit).
  if (hipHostMalloc(reinterpret_cast<void**>(&c->h_coh), iso::kTotSlots * sizeof(unsigned long long),
                    hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
      hipHostGetDevicePointer(reinterpret_cast<void**>(&c->d_coh), c->h_coh, 0) != hipSuccess) {
    if (c->h_coh) (void)hipHostFree(c->h_coh);
    c->h_coh = nullptr;
    c->d_coh = nullptr;
    (void)hipGetLastError();
  }
  *out = c;
  return ISO_OK;
}

int iso_ctx_destroy(iso_ctx* c) {
  if (!c) return ISO_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  collect_timing(c);
  for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
  DevBuf* bufs[] = {&c->prog,       &c->grid_owned, &c->signs,     &c->seg,      &c->tplanes, &c->vplanes,
                    &c->agg,        &c->incl,       &c->active,    &c->totals,
                    &c->vout,       &c->iout,       &c->pts,       &c->vals,     &c->dc_normals, &c->dc_plist, &c->dc_pindex, &c->dc_samples,
                    &c->dc_values,  &c->dc_points,  &c->dc_mixed,  &c->dc_flags,  &c->dc_counters,
                    &c->oc_keys,    &c->oc_values,  &c->oc_leaf,   &c->oc_dense, &c->oc_primal,
                    &c->oc_plevel,  &c->oc_corner,  &c->oc_case,   &c->oc_base,  &c->oc_tkeys,
                    &c->oc_towner,  &c->oc_tid,     &c->sc_slot,   &c->sc_ctl,
                    &c->oc_pkeys,   &c->oc_plevels, &c->oc_skeys,  &c->oc_slevels, &c->oc_sort,
                    &c->b_ops,      &c->b_chunks,   &c->b_segstart, &c->b_wordstart,
                    &c->b_signs,    &c->b_voff,     &c->b_ioff,    &c->b_values,
                    &c->nn_theta,   &c->nn_m,
                    &c->nn_v,       &c->nn_grad,    &c->nn_frags,  &c->nn_loss,  &c->nn_x,
                    &c->nn_h,       &c->nn_dz,      &c->nn_g,      &c->nn_pts,   &c->nn_t,
                    &c->nn_w,       &c->nn_out,     &c->br_centres, &c->br_hd,   &c->br_pred,
                    &c->br_flags,   &c->br_list,    &c->br_count,  &c->pvals,  &c->dense_check,  &c->nrm,
                    &c->dgrid,      &c->dvals,      &c->emc_cells, &c->emc_first, &c->emc_cross,
                    &c->emc_flips,  &c->emc_ctr,    &c->br_coarse, &c->br_proof,
                    &c->f_signs,    &c->f_flags,    &c->f_slots,   &c->f_rowbase, &c->f_ctl};
  if (c->b_graph) (void)hipGraphExecDestroy(c->b_graph);
  if (c->m_graph) (void)hipGraphExecDestroy(c->m_graph);
  for (DevBuf* b : bufs) release(b);
  if (c->h_totals) (void)hipHostFree(c->h_totals);
  if (c->h_coh) (void)hipHostFree(c->h_coh);
  if (c->h_dc) (void)hipHostFree(c->h_dc);
  if (c->h_stage) (void)hipHostFree(c->h_stage);
  if (c->own_stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return ISO_OK;
}

int iso_ctx_synchronize(iso_ctx* c) {
  if (!c) return ISO_ERR_INVALID_ARGUMENT;
  ISO_HIP(c, hipSetDevice(c->device));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  return ISO_OK;
}

int iso_program_validate(const iso_node* nodes, uint32_t n_nodes, char* message,
                         size_t message_len) {
  std::vector<iso::DevOp> ops;
  std::string err;
  const int st = iso::lower_program(nodes, n_nodes, &ops, &err);
  if (message && message_len) std::snprintf(message, message_len, "%s", err.c_str());
  return st;
}

int iso_program_upload(iso_ctx* c, const iso_node* nodes, uint32_t n_nodes) {
  if (!c) return ISO_ERR_INVALID_ARGUMENT;
  std::vector<iso::DevOp> ops;
  std::string err;
  int st = iso::lower_program(nodes, n_nodes, &ops, &err);
  if (st != ISO_OK) return fail(c, st, err);
  ISO_HIP(c, hipSetDevice(c->device));
  ISO_HIP(c, ensure(&c->prog, ops.size() * sizeof(iso::DevOp)));
  ISO_HIP(c, hipMemcpyAsync(c->prog.ptr, ops.data(), ops.size() * sizeof(iso::DevOp),
                            hipMemcpyHostToDevice, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  c->n_prog = static_cast<int>(ops.size());
  c->prog_directed = true;
  c->prog_hermite = true;
  for (uint32_t i = 0; i < n_nodes; ++i) {
    if (nodes[i].op == ISO_OP_SMOOTH_UNION) c->prog_directed = false;
    if (nodes[i].op == ISO_OP_SMOOTH_UNION || nodes[i].op == ISO_OP_INTERSECTION ||
        nodes[i].op == ISO_OP_DIFFERENCE)
      c->prog_hermite = false;
  }
  c->band_valid = false;  // a band extraction is checked against its own program
  return ISO_OK;
}

int iso_program_sample(iso_ctx* c, const float* points, uint64_t n, float* values) {
  if (!c || (n && (!points || !values))) return fail(c, ISO_ERR_INVALID_ARGUMENT, "null buffer");
  if (c->n_prog == 0) return fail(c, ISO_ERR_NO_SOURCE, "no program uploaded");
  if (n == 0) return ISO_OK;
  ISO_HIP(c, hipSetDevice(c->device));
  ISO_HIP(c, ensure(&c->pts, n * 3 * sizeof(float)));
  ISO_HIP(c, ensure(&c->vals, n * sizeof(float)));
  ISO_HIP(c, hipMemcpyAsync(c->pts.ptr, points, n * 3 * sizeof(float), hipMemcpyHostToDevice,
                            c->stream));
  ISO_HIP(c, timed(c, "sample_points", [&] {
            return iso::launch_sample_points(static_cast<const iso::DevOp*>(c->prog.ptr),
                                             c->n_prog, static_cast<const float*>(c->pts.ptr), n,
                                             static_cast<float*>(c->vals.ptr), c->stream);
          }));
  ISO_HIP(c, hipMemcpyAsync(values, c->vals.ptr, n * sizeof(float), hipMemcpyDeviceToHost,
                            c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  return ISO_OK;
}

int iso_sample_normals(iso_ctx* c, const float* points, uint64_t n, float epsilon, float* normals,
                       int32_t memory) {
  if (!c || (n && (!points || !normals))) return fail(c, ISO_ERR_INVALID_ARGUMENT, "null buffer");
  if (c->n_prog == 0) return fail(c, ISO_ERR_NO_SOURCE, "no program uploaded");
  if (!(epsilon > 0.0f) || !(epsilon < 3.0e38f))
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "epsilon must be positive and finite");
  if (memory != ISO_MEMORY_HOST && memory != ISO_MEMORY_DEVICE)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "unknown memory kind");
  if (n == 0) return ISO_OK;
  if (n > (uint64_t(1) << 38)) return fail(c, ISO_ERR_INVALID_ARGUMENT, "too many points");
  ISO_HIP(c, hipSetDevice(c->device));
  const float* dp = points;
  float* dn = normals;
  if (memory == ISO_MEMORY_HOST) {
    ISO_HIP(c, ensure(&c->pts, n * 3 * sizeof(float)));
    ISO_HIP(c, ensure(&c->nrm, n * 3 * sizeof(float)));
    ISO_HIP(c, hipMemcpyAsync(c->pts.ptr, points, n * 3 * sizeof(float), hipMemcpyHostToDevice,
                              c->stream));
    dp = static_cast<const float*>(c->pts.ptr);
    dn = static_cast<float*>(c->nrm.ptr);
  }
  ISO_HIP(c, timed(c, "sample_normals", [&] {
            return iso::launch_sample_normals(static_cast<const iso::DevOp*>(c->prog.ptr),
                                              c->n_prog, dp, n, epsilon, dn, c->stream);
          }));
  if (memory == ISO_MEMORY_HOST)
    ISO_HIP(c, hipMemcpyAsync(normals, dn, n * 3 * sizeof(float), hipMemcpyDeviceToHost,
                              c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  return ISO_OK;
}

int iso_program_sample_lattice(iso_ctx* c, const iso_grid_desc* desc, float* out,
                               int32_t memory) {
  if (!c || !out) return fail(c, ISO_ERR_INVALID_ARGUMENT, "null argument");
  iso::Lattice L;
  int st = check_lattice(c, desc, &L);
  if (st != ISO_OK) return st;
  if (c->n_prog == 0) return fail(c, ISO_ERR_NO_SOURCE, "no program uploaded");
  if (memory != ISO_MEMORY_HOST && memory != ISO_MEMORY_DEVICE)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "unknown output memory kind");
  ISO_HIP(c, hipSetDevice(c->device));
  const size_t bytes = static_cast<size_t>(L.nx * L.ny * L.nz) * sizeof(float);
  float* dst = out;
  if (memory == ISO_MEMORY_HOST) {
    ISO_HIP(c, ensure(&c->vals, bytes));
    dst = static_cast<float*>(c->vals.ptr);
  }
  ISO_HIP(c, timed(c, "sample_lattice", [&] {
            return iso::launch_sample_lattice(static_cast<const iso::DevOp*>(c->prog.ptr),
                                              c->n_prog, L, dst, c->stream);
          }));
  if (memory == ISO_MEMORY_HOST)
    ISO_HIP(c, hipMemcpyAsync(out, dst, bytes, hipMemcpyDeviceToHost, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  return ISO_OK;
}

int iso_grid_upload(iso_ctx* c, const float* values, int64_t nx, int64_t ny, int64_t nz) {
  if (!c || !values || nx < 1 || ny < 1 || nz < 1)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "bad grid upload arguments");
  const size_t bytes = static_cast<size_t>(nx) * ny * nz * sizeof(float);
  ISO_HIP(c, hipSetDevice(c->device));
  ISO_HIP(c, ensure(&c->grid_owned, bytes));
  ISO_HIP(c, hipMemcpyAsync(c->grid_owned.ptr, values, bytes, hipMemcpyHostToDevice, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  c->grid = static_cast<const float*>(c->grid_owned.ptr);
  c->gnx = nx;
  c->gny = ny;
  c->gnz = nz;
  return ISO_OK;
}

int iso_grid_upload_directed(iso_ctx* c, const float* xyz, int64_t nx, int64_t ny, int64_t nz) {
  if (!c || !xyz || nx < 1 || ny < 1 || nz < 1)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "bad grid upload arguments");
  const size_t bytes = static_cast<size_t>(nx) * ny * nz * 3 * sizeof(float);
  ISO_HIP(c, hipSetDevice(c->device));
  ISO_HIP(c, ensure(&c->dgrid, bytes));
  ISO_HIP(c, hipMemcpyAsync(c->dgrid.ptr, xyz, bytes, hipMemcpyHostToDevice, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  c->dgnx = nx;
  c->dgny = ny;
  c->dgnz = nz;
  return ISO_OK;
}

int iso_grid_bind_device(iso_ctx* c, const float* dev, int64_t nx, int64_t ny, int64_t nz) {
  if (!c || !dev || nx < 1 || ny < 1 || nz < 1)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "bad grid bind arguments");
  c->grid = dev;
  c->gnx = nx;
  c->gny = ny;
  c->gnz = nz;
  return ISO_OK;
}

int iso_grid_desc_reference(int64_t size, int32_t source, iso_grid_desc* out) {
  if (!out || size < 2) return ISO_ERR_INVALID_ARGUMENT;
  std::memset(out, 0, sizeof(*out));
  out->nx = size;
  out->ny = size;
  out->nz = size + 1;
  // one_over_size = 1.0 / (size_minus_one as f32)  (primal_grid.rs:44-45)
  const float ovs = 1.0f / static_cast<float>(size - 1);
  for (int i = 0; i < 3; ++i) {
    out->origin[i] = 0.0f;
    out->spacing[i] = ovs;
  }
  out->source = source;
  return ISO_OK;
}

namespace {

// Sizes the context's scratch for a lattice; the active-cell list starts at
// 1/16 of the cells (at least 2^20) and grows when a mesh needs more.
int prepare_mc(iso_ctx* c, const iso::McDims& D, uint64_t min_active, iso::McBuffers* B) {
  ISO_HIP(c, ensure(&c->signs, (D.n_sign_words + 3) * sizeof(uint64_t)));
  ISO_HIP(c, ensure(&c->seg, D.n_segs * sizeof(uint2)));
  ISO_HIP(c, ensure(&c->tplanes, 3 * D.n_segs * sizeof(uint64_t)));
  ISO_HIP(c, ensure(&c->vplanes, 4 * D.n_segs * sizeof(uint64_t)));
  ISO_HIP(c, ensure(&c->agg, D.n_blocks * sizeof(iso::ScanTriple)));
  ISO_HIP(c, ensure(&c->incl, D.n_blocks * sizeof(iso::ScanTriple)));
  ISO_HIP(c, ensure(&c->totals, iso::kTotSlots * sizeof(unsigned long long)));
  uint64_t cap = static_cast<uint64_t>(D.n_cells / 16);
  if (cap < (1u << 20)) cap = 1u << 20;
  if (cap < min_active) cap = min_active;
  if (cap > static_cast<uint64_t>(D.n_cells)) cap = static_cast<uint64_t>(D.n_cells);
  if (cap < 1) cap = 1;
  if (c->active_capacity < cap || !c->active.ptr) {
    ISO_HIP(c, ensure(&c->active, cap * sizeof(uint32_t)));
    c->active_capacity = c->active.bytes / sizeof(uint32_t);
  }
  B->signs = static_cast<uint64_t*>(c->signs.ptr) + 1;
  B->seg = static_cast<uint2*>(c->seg.ptr);
  B->tplanes = static_cast<uint64_t*>(c->tplanes.ptr);
  B->vplanes = static_cast<uint64_t*>(c->vplanes.ptr);
  B->agg = static_cast<iso::ScanTriple*>(c->agg.ptr);
  B->incl = static_cast<iso::ScanTriple*>(c->incl.ptr);
  B->active = static_cast<uint32_t*>(c->active.ptr);
  B->active_capacity = c->active_capacity;
  B->totals = static_cast<unsigned long long*>(c->totals.ptr);
  return ISO_OK;
}

// signs (with state reset) + count/scan; samples from `grid` (device) or,
// when it is null, from the context's program.
// A replacement for the signs stage (the learned band's); `values`, when
// not null, receives the exact samples it evaluates.
using SignsStage = std::function<int(const iso::McBuffers&, float* values)>;

int classify_and_count(iso_ctx* c, const float* grid, const iso::Lattice& L, const iso::McDims& D,
                       const iso::McBuffers& B, const SignsStage* signs = nullptr,
                       float* values = nullptr) {
  const iso::DevOp* ops = static_cast<const iso::DevOp*>(c->prog.ptr);
  if (signs) {
    const int st = (*signs)(B, values);
    if (st != ISO_OK) return st;
  } else if (grid) {
    ISO_HIP(c, timed(c, "mc_signs_grid",
                     [&] { return iso::launch_signs_grid(grid, L, D, B, c->stream); }));
  } else {
    ISO_HIP(c, timed(c, "mc_signs_program", [&] {
              return iso::launch_signs_program(ops, c->n_prog, L, D, B, values, c->stream);
            }));
  }
  ISO_HIP(c, timed(c, "mc_count", [&] { return iso::launch_count(L, D, B, c->stream); }));
  return ISO_OK;
}

// After an emit launched with B.host_totals set: the totals it handed over.
int read_emitted_totals(iso_ctx* c) {
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  const volatile unsigned long long* h = c->h_coh;
  for (int i = 0; i < iso::kTotSlots; ++i) c->h_totals[i] = h[i];
  if (c->h_totals[iso::kTotErrors] & 2ull)
    return fail(c, ISO_ERR_DEVICE, "count scan look-back did not complete");
  return ISO_OK;
}

int read_totals(iso_ctx* c, const iso::McBuffers& B) {
  ISO_HIP(c, hipMemcpyAsync(c->h_totals, B.totals, iso::kTotSlots * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  if (c->h_totals[iso::kTotErrors] & 2ull)
    return fail(c, ISO_ERR_DEVICE, "count scan look-back did not complete");
  return ISO_OK;
}

// Marching cubes over lattice L whose samples come from `grid` (device) or,
// when grid is null, from the context's program; with `positions` the
// vertices interpolate between per-point positions.  Writes the mesh to out.
int prepare_scan(iso_ctx* c, uint64_t n, iso::scan::State* st);

// The brick flags of a narrow-band extraction, for its emit stage.
struct BandEmit {
  const uint8_t* flags;
  int64_t nbx, nby;
  int shift;  // bricks of 2^shift cells per edge
};

// Grows b to `bytes`; a new allocation is zero-filled (the single-pass
// kernel's epoch-tagged words must start below every epoch).
int ensure_zeroed(iso_ctx* c, DevBuf* b, size_t bytes) {
  const void* before = b->ptr;
  ISO_HIP(c, ensure(b, bytes));
  if (b->ptr != before) ISO_HIP(c, hipMemsetAsync(b->ptr, 0, b->bytes, c->stream));
  return ISO_OK;
}

int fused_launch_and_read(iso_ctx* c, const iso::Lattice& L, const float* grid,
                          const iso::FusedPlan& P, iso::FusedBuffers B, uint32_t epoch,
                          bool timing, uint64_t* nv, uint64_t* nt);

// One single-pass extraction (mc_fused.hip) of a grid source into `vout` /
// `iout` (device) with the given capacities; sets *nv, *nt.
int run_fused(iso_ctx* c, const iso::Lattice& L, const float* grid, const iso::FusedPlan& P,
              float* vout, uint32_t* iout, uint64_t vcap, uint64_t icap, uint64_t* nv,
              uint64_t* nt) {
  int st;
  if ((st = ensure_zeroed(c, &c->f_signs, P.sign_words * sizeof(uint64_t))) != ISO_OK) return st;
  if ((st = ensure_zeroed(c, &c->f_flags, P.n_tickets * sizeof(uint64_t))) != ISO_OK) return st;
  if ((st = ensure_zeroed(c, &c->f_slots, 4 * P.n_tickets * sizeof(uint64_t))) != ISO_OK)
    return st;
  if ((st = ensure_zeroed(c, &c->f_rowbase, P.n_rows * sizeof(uint64_t))) != ISO_OK) return st;
  if ((st = ensure_zeroed(c, &c->f_ctl, iso::kFusedCtlSlots * sizeof(uint64_t))) != ISO_OK)
    return st;
  // Tags keep 22 bits of the epoch: clear the words before it wraps.
  c->f_epoch += 1;
  if ((c->f_epoch & 0x3FFFFFu) == 0u) {
    for (DevBuf* b : {&c->f_flags, &c->f_slots, &c->f_rowbase, &c->f_ctl})
      ISO_HIP(c, hipMemsetAsync(b->ptr, 0, b->bytes, c->stream));
    c->f_epoch = 1;
  }
  iso::FusedBuffers B{};
  B.signs = static_cast<uint64_t*>(c->f_signs.ptr);
  B.flags = static_cast<unsigned long long*>(c->f_flags.ptr);
  B.slots = static_cast<unsigned long long*>(c->f_slots.ptr);
  B.rowbase = static_cast<unsigned long long*>(c->f_rowbase.ptr);
  B.ctl = static_cast<unsigned long long*>(c->f_ctl.ptr);
  B.vertices = vout;
  B.indices = iout;
  B.vcap = vout ? vcap : 0;
  B.icap = iout ? icap : 0;
  return fused_launch_and_read(c, L, grid, P, B, c->f_epoch, true, nv, nt);
}

// Launches the single pass over buffers B with `epoch` and reads back its
// totals; `timing`: the launch may take kernel-timing events.
int fused_launch_and_read(iso_ctx* c, const iso::Lattice& L, const float* grid,
                          const iso::FusedPlan& P, iso::FusedBuffers B, uint32_t epoch,
                          bool timing, uint64_t* nv, uint64_t* nt) {
  B.host_totals = c->d_coh;
  // The kernel's last tile writes the totals, and any tile its error flags,
  // into host-coherent memory (without it: copied from the control words).
  if (c->h_coh) {
    volatile unsigned long long* h = c->h_coh;
    h[iso::kTotVertices] = ~0ull;
    h[iso::kTotTriangles] = ~0ull;
    h[iso::kTotErrors] = 0ull;
    h[iso::kTotErrors + 1] = 0ull;
  }
  if (timing) {
    ISO_HIP(c, timed(c, "mc_fused", [&] {
              return iso::launch_mc_fused(grid, L, P, B, epoch, c->stream);
            }));
  } else {
    ISO_HIP(c, iso::launch_mc_fused(grid, L, P, B, epoch, c->stream));
  }
  bool overflow = false, gave_up = false;
  if (c->h_coh) {
    ISO_HIP(c, hipStreamSynchronize(c->stream));
    const volatile unsigned long long* h = c->h_coh;
    *nv = h[iso::kTotVertices];
    *nt = h[iso::kTotTriangles];
    overflow = h[iso::kTotErrors] != 0ull;
    gave_up = h[iso::kTotErrors + 1] != 0ull;
  } else {
    ISO_HIP(c, hipMemcpyAsync(c->h_totals, B.ctl + iso::kFusedVertices,
                              3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    ISO_HIP(c, hipStreamSynchronize(c->stream));
    *nv = c->h_totals[0];
    *nt = c->h_totals[1];
    const unsigned long long e = c->h_totals[2];
    const bool current = (e >> 32) == epoch;
    overflow = current && (e & 1ull);
    gave_up = current && (e & 2ull);
  }
  if (timing) collect_timing(c);
  if (gave_up) return fail(c, ISO_ERR_DEVICE, "single-pass extraction: a wait between tiles gave up");
  if (*nv == ~0ull || *nt == ~0ull)
    return fail(c, ISO_ERR_DEVICE, "single-pass extraction: totals not written");
  if (overflow) return fail(c, ISO_ERR_OVERFLOW, "mesh exceeds u32 indices");
  return ISO_OK;
}

// Marching cubes of a grid source in one launch (see mc_fused.hip).
int extract_fused(iso_ctx* c, const iso::Lattice& L, const float* grid, const iso::FusedPlan& P,
                  iso_mesh_out* out) {
  int st;
  uint64_t nv = 0, nt = 0;
  if (out->memory == ISO_MEMORY_DEVICE) {
    if ((st = run_fused(c, L, grid, P, out->vertices, out->indices,
                        out->vertices ? out->vertex_capacity : 0,
                        out->indices ? out->index_capacity : 0, &nv, &nt)) != ISO_OK)
      return st;
    out->vertex_count = nv;
    out->index_count = 3 * nt;
    if (nv > (out->vertices ? out->vertex_capacity : 0) ||
        3 * nt > (out->indices ? out->index_capacity : 0))
      return fail(c, ISO_ERR_CAPACITY, "output buffers too small");
    return ISO_OK;
  }
  // Host memory: extract into the context's device buffers (sized from the
  // last extraction), again with room if they were too small, then copy.
  for (int attempt = 0; attempt < 2; ++attempt) {
    const uint64_t vcap = c->vout.bytes / (3 * sizeof(float));
    const uint64_t icap = c->iout.bytes / sizeof(uint32_t);
    if ((st = run_fused(c, L, grid, P, static_cast<float*>(c->vout.ptr),
                        static_cast<uint32_t*>(c->iout.ptr), c->vout.ptr ? vcap : 0,
                        c->iout.ptr ? icap : 0, &nv, &nt)) != ISO_OK)
      return st;
    if (nv <= vcap && 3 * nt <= icap && (nv == 0 || c->vout.ptr) && (nt == 0 || c->iout.ptr))
      break;
    ISO_HIP(c, ensure(&c->vout, (nv ? nv : 1) * 3 * sizeof(float)));
    ISO_HIP(c, ensure(&c->iout, (nt ? 3 * nt : 1) * sizeof(uint32_t)));
  }
  const uint64_t ni = 3 * nt;
  out->vertex_count = nv;
  out->index_count = ni;
  if ((nv && (!out->vertices || out->vertex_capacity < nv)) ||
      (ni && (!out->indices || out->index_capacity < ni))) {
    c->kept = true;
    c->kept_vertices = nv;
    c->kept_indices = ni;
    return fail(c, ISO_ERR_CAPACITY, "output buffers too small; mesh kept for iso_mesh_fetch");
  }
  if (nv)
    ISO_HIP(c, hipMemcpyAsync(out->vertices, c->vout.ptr, nv * 3 * sizeof(float),
                              hipMemcpyDeviceToHost, c->stream));
  if (ni)
    ISO_HIP(c, hipMemcpyAsync(out->indices, c->iout.ptr, ni * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  return ISO_OK;
}

int extract_lattice(iso_ctx* c, const iso::Lattice& L, const float* grid, const float* positions,
                    iso_mesh_out* out, const SignsStage* signs = nullptr,
                    const BandEmit* band = nullptr, const float* directed = nullptr) {
  int st = ISO_OK;
  c->band_valid = false;  // the context's sign words are about to change
  // The single pass over a grid source, when asked for (not with a captured
  // graph, whose launches the single pass, keyed on a per-extraction epoch,
  // cannot replay).
  if (grid && !positions && !signs && !band && !directed &&
      (out->flags & ISO_MESH_SINGLE_PASS) && !(out->flags & ISO_MESH_GRAPH)) {
    iso::FusedPlan P;
    if (iso::fused_plan(L, &P)) return extract_fused(c, L, grid, P, out);
  }
  const iso::McDims D = iso::mc_dims(L);
  if (D.n_segs >= (int64_t(1) << 26))
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "lattice has 2^26 or more 64-cell segments");
  iso::McBuffers B{};
  if ((st = prepare_mc(c, D, 0, &B)) != ISO_OK) return st;
  const iso::DevOp* ops = static_cast<const iso::DevOp*>(c->prog.ptr);
  uint64_t nv = 0, ni = 0;
  // A program source stores its samples during classification and emit
  // reads them (the same bits a re-evaluation would give); without the
  // memory for them, emit evaluates the program at the edge endpoints.  A
  // narrow band stores only its exact bricks' samples.
  float* values = nullptr;
  if (!grid && !directed &&
      ensure(&c->pvals, static_cast<size_t>(D.n_points) * sizeof(float)) == hipSuccess)
    values = static_cast<float*>(c->pvals.ptr);
  auto emit = [&]() -> hipError_t {
    if (directed) return iso::launch_emit_directed(directed, L, D, B, c->stream);
    if (band && values)
      return iso::launch_emit_band(values, band->flags, band->nbx, band->nby, band->shift, ops,
                                   c->n_prog, L, D, B, c->stream);
    const float* src = grid ? grid : (band ? nullptr : values);
    return iso::launch_emit(src, ops, c->n_prog, positions, L, D, B, c->stream);
  };

  if (out->memory == ISO_MEMORY_DEVICE) {
    B.vertices = out->vertices;
    B.indices = out->indices;
    B.vertex_capacity = out->vertices ? out->vertex_capacity : 0;
    B.index_capacity = out->indices ? out->index_capacity : 0;
    // Emit hands the totals to host-coherent memory: no copy after it.
    B.host_totals = c->d_coh;
    auto totals_after_emit = [&]() {
      return B.host_totals ? read_emitted_totals(c) : read_totals(c, B);
    };
    // Graph replay: the launches below depend only on the key's fields
    // (sample and program memory is read when the graph runs), so a capture
    // replays for any later call with an equal key.
    // (The legacy null stream cannot be captured: such contexts extract eagerly.)
    const bool graph = (out->flags & ISO_MESH_GRAPH) && !signs && !band && !positions &&
                       !c->timing && c->stream != nullptr;
    auto key_of = [&](const iso::McBuffers& b) {
      MeshGraphKey k;
      std::memset(&k, 0, sizeof(k));
      k.L = L;
      k.grid = grid;
      k.ops = grid ? nullptr : ops;
      k.values = values;
      k.n_prog = grid ? 0 : c->n_prog;
      k.B = b;
      return k;
    };
    bool replayed = false;
    if (graph && c->m_graph) {
      const MeshGraphKey k = key_of(B);
      if (std::memcmp(&k, &c->m_key, sizeof(k)) == 0) {
        ISO_HIP(c, hipGraphLaunch(c->m_graph, c->stream));
        if ((st = totals_after_emit()) != ISO_OK) return st;
        const uint64_t n_act = c->h_totals[iso::kTotActive];
        replayed = n_act <= B.active_capacity;
        // A longer active-cell list changes the key: extract and capture anew.
        if (!replayed && (st = prepare_mc(c, D, n_act, &B)) != ISO_OK) return st;
      }
    }
    for (int attempt = 0; attempt < 2 && !replayed; ++attempt) {
      if ((st = classify_and_count(c, grid, L, D, B, signs, values)) != ISO_OK) return st;
      ISO_HIP(c, timed(c, "mc_emit", emit));
      if ((st = totals_after_emit()) != ISO_OK) return st;
      const uint64_t n_act = c->h_totals[iso::kTotActive];
      if (n_act <= B.active_capacity) break;
      // The active-cell list was too short: grow it and extract again.
      if ((st = prepare_mc(c, D, n_act, &B)) != ISO_OK) return st;
    }
    nv = c->h_totals[iso::kTotVertices];
    ni = 3 * c->h_totals[iso::kTotTriangles];
    out->vertex_count = nv;
    out->index_count = ni;
    if (c->h_totals[iso::kTotErrors] & 1ull) return fail(c, ISO_ERR_OVERFLOW, "mesh exceeds u32 indices");
    if (nv > B.vertex_capacity || ni > B.index_capacity)
      return fail(c, ISO_ERR_CAPACITY, "output buffers too small");
    if (graph && !replayed) {
      if (c->m_graph) {
        (void)hipGraphExecDestroy(c->m_graph);
        c->m_graph = nullptr;
      }
      hipGraph_t g = nullptr;
      ISO_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      hipError_t le = hipSuccess;
      st = classify_and_count(c, grid, L, D, B, nullptr, values);
      if (st == ISO_OK) le = emit();
      const hipError_t ce = hipStreamEndCapture(c->stream, &g);
      if (st != ISO_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
      }
      if (le != hipSuccess && g) (void)hipGraphDestroy(g);
      ISO_HIP(c, le);
      ISO_HIP(c, ce);
      const hipError_t ie = hipGraphInstantiate(&c->m_graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) c->m_graph = nullptr;
      ISO_HIP(c, ie);
      c->m_key = key_of(B);
    }
    return ISO_OK;
  }

  // Host memory: size the context's device buffers from the scan totals.
  if ((st = classify_and_count(c, grid, L, D, B, signs, values)) != ISO_OK) return st;
  if ((st = read_totals(c, B)) != ISO_OK) return st;
  const uint64_t n_act = c->h_totals[iso::kTotActive];
  if (n_act > B.active_capacity) {
    if ((st = prepare_mc(c, D, n_act, &B)) != ISO_OK) return st;
    if ((st = classify_and_count(c, grid, L, D, B, signs, values)) != ISO_OK) return st;
    if ((st = read_totals(c, B)) != ISO_OK) return st;
  }
  nv = c->h_totals[iso::kTotVertices];
  ni = 3 * c->h_totals[iso::kTotTriangles];
  out->vertex_count = nv;
  out->index_count = ni;
  if (c->h_totals[iso::kTotErrors] & 1ull) return fail(c, ISO_ERR_OVERFLOW, "mesh exceeds u32 indices");
  ISO_HIP(c, ensure(&c->vout, (nv ? nv : 1) * 3 * sizeof(float)));
  ISO_HIP(c, ensure(&c->iout, (ni ? ni : 1) * sizeof(uint32_t)));
  B.vertices = static_cast<float*>(c->vout.ptr);
  B.indices = static_cast<uint32_t*>(c->iout.ptr);
  B.vertex_capacity = nv;
  B.index_capacity = ni;
  ISO_HIP(c, timed(c, "mc_emit", emit));
  if ((nv && (!out->vertices || out->vertex_capacity < nv)) ||
      (ni && (!out->indices || out->index_capacity < ni))) {
    ISO_HIP(c, hipStreamSynchronize(c->stream));
    collect_timing(c);
    c->kept = true;
    c->kept_vertices = nv;
    c->kept_indices = ni;
    return fail(c, ISO_ERR_CAPACITY, "output buffers too small; mesh kept for iso_mesh_fetch");
  }
  if (nv)
    ISO_HIP(c, hipMemcpyAsync(out->vertices, B.vertices, nv * 3 * sizeof(float),
                              hipMemcpyDeviceToHost, c->stream));
  if (ni)
    ISO_HIP(c, hipMemcpyAsync(out->indices, B.indices, ni * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  return ISO_OK;
}

}  // namespace

int iso_extract_mc(iso_ctx* c, const iso_grid_desc* desc, iso_mesh_out* out) {
  if (!c || !out) return fail(c, ISO_ERR_INVALID_ARGUMENT, "null argument");
  iso::Lattice L;
  int st = check_lattice(c, desc, &L);
  if (st != ISO_OK) return st;
  const bool from_grid = desc->source == ISO_SOURCE_GRID;
  if (from_grid) {
    if (!c->grid) return fail(c, ISO_ERR_NO_SOURCE, "no grid bound");
    if (c->gnx != L.nx || c->gny != L.ny || c->gnz != L.nz)
      return fail(c, ISO_ERR_INVALID_ARGUMENT, "lattice dims differ from the bound grid");
  } else if (c->n_prog == 0) {
    return fail(c, ISO_ERR_NO_SOURCE, "no program uploaded");
  }
  if (out->memory != ISO_MEMORY_HOST && out->memory != ISO_MEMORY_DEVICE)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "unknown output memory kind");
  c->kept = false;
  ISO_HIP(c, hipSetDevice(c->device));
  return extract_lattice(c, L, from_grid ? c->grid : nullptr, nullptr, out);
}

int iso_extract_mc_noalloc(iso_ctx* c, const iso_grid_desc* desc, iso_mesh_out* out) {
  if (!c || !out) return fail(c, ISO_ERR_INVALID_ARGUMENT, "null argument");
  iso::Lattice L;
  int st = check_lattice(c, desc, &L);
  if (st != ISO_OK) return st;
  if (desc->source != ISO_SOURCE_GRID)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "the no-allocation extraction takes a grid source");
  if (!c->grid) return fail(c, ISO_ERR_NO_SOURCE, "no grid bound");
  if (c->gnx != L.nx || c->gny != L.ny || c->gnz != L.nz)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "lattice dims differ from the bound grid");
  if (out->memory != ISO_MEMORY_DEVICE)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "the no-allocation extraction writes device memory");
  iso::FusedPlan P;
  if (!iso::fused_plan(L, &P))
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "lattice rows are too wide for the single pass");
  if (P.sign_words * sizeof(uint64_t) > kNaSignBytes ||
      static_cast<uint64_t>(P.n_tickets) > kNaMaxTickets || P.n_rows > kNaMaxRows)
    return fail(c, ISO_ERR_INVALID_ARGUMENT,
                "lattice exceeds the no-allocation scratch (28 MiB of sign words, about 513^3 "
                "samples; 65,536 tiles; 262,144 cell rows)");
  if (c->device < 0 || c->device >= kNaMaxDevices)
    return fail(c, ISO_ERR_DEVICE, "device index beyond the no-allocation table");
  ISO_HIP(c, hipSetDevice(c->device));
  unsigned char* base = nullptr;
  uint32_t epoch = 0;
  bool wrapped = false;
  {
    std::lock_guard<std::mutex> lock(g_na_mutex);
    NoAllocDevice& d = g_na[c->device];
    if (!d.looked_up || !d.ptr)
      return fail(c, ISO_ERR_DEVICE,
                  std::string("no-allocation scratch address: ") + hipGetErrorString(d.status));
    base = d.ptr;
    d.epoch += 1;
    if ((d.epoch & 0x3FFFFFu) == 0u) {  // tags keep 22 bits of the epoch
      d.epoch = 1;
      wrapped = true;
    }
    epoch = d.epoch;
  }
  if (wrapped)  // clear every tagged word (a fill kernel: no memset node, no allocation)
    ISO_HIP(c, iso::launch_fill_u32(reinterpret_cast<uint32_t*>(base + kNaFlagsOff),
                                    (kNaBytes - kNaFlagsOff) / 4, 0u, c->stream));
  iso::FusedBuffers B{};
  B.signs = reinterpret_cast<uint64_t*>(base);
  B.flags = reinterpret_cast<unsigned long long*>(base + kNaFlagsOff);
  B.slots = reinterpret_cast<unsigned long long*>(base + kNaSlotsOff);
  B.rowbase = reinterpret_cast<unsigned long long*>(base + kNaRowbaseOff);
  B.ctl = reinterpret_cast<unsigned long long*>(base + kNaCtlOff);
  B.vertices = out->vertices;
  B.indices = out->indices;
  B.vcap = out->vertices ? out->vertex_capacity : 0;
  B.icap = out->indices ? out->index_capacity : 0;
  c->kept = false;
  c->band_valid = false;
  uint64_t nv = 0, nt = 0;
  if ((st = fused_launch_and_read(c, L, c->grid, P, B, epoch, false, &nv, &nt)) != ISO_OK)
    return st;
  out->vertex_count = nv;
  out->index_count = 3 * nt;
  if (nv > B.vcap || 3 * nt > B.icap) return fail(c, ISO_ERR_CAPACITY, "output buffers too small");
  return ISO_OK;
}

int iso_extract_mc_directed(iso_ctx* c, const iso_grid_desc* desc, iso_mesh_out* out) {
  if (!c || !out) return fail(c, ISO_ERR_INVALID_ARGUMENT, "null argument");
  iso::Lattice L;
  int st = check_lattice(c, desc, &L);
  if (st != ISO_OK) return st;
  const bool from_grid = desc->source == ISO_SOURCE_GRID;
  if (from_grid) {
    if (!c->dgrid.ptr || c->dgnx == 0) return fail(c, ISO_ERR_NO_SOURCE, "no directed grid uploaded");
    if (c->dgnx != L.nx || c->dgny != L.ny || c->dgnz != L.nz)
      return fail(c, ISO_ERR_INVALID_ARGUMENT, "lattice dims differ from the directed grid");
  } else if (c->n_prog == 0) {
    return fail(c, ISO_ERR_NO_SOURCE, "no program uploaded");
  } else if (!c->prog_directed) {
    return fail(c, ISO_ERR_PROGRAM, "the program has no directed form (smooth union)");
  }
  if (out->memory != ISO_MEMORY_HOST && out->memory != ISO_MEMORY_DEVICE)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "unknown output memory kind");
  c->kept = false;
  ISO_HIP(c, hipSetDevice(c->device));
  const iso::McDims D = iso::mc_dims(L);
  float* vals = static_cast<float*>(c->dgrid.ptr);
  if (!from_grid) {
    ISO_HIP(c, ensure(&c->dvals, static_cast<size_t>(D.n_points) * 3 * sizeof(float)));
    vals = static_cast<float*>(c->dvals.ptr);
  }
  const iso::DevOp* ops = from_grid ? nullptr : static_cast<const iso::DevOp*>(c->prog.ptr);
  const SignsStage signs = [&](const iso::McBuffers& B, float*) -> int {
    ISO_HIP(c, timed(c, "mc_signs_directed", [&] {
              return iso::launch_signs_directed(ops, c->n_prog, L, D, B, vals, c->stream);
            }));
    return ISO_OK;
  };
  return extract_lattice(c, L, nullptr, nullptr, out, &signs, nullptr, vals);
}

int iso_extract_emc(iso_ctx* c, const iso_grid_desc* desc, float epsilon, iso_mesh_out* out) {
  if (!c || !out) return fail(c, ISO_ERR_INVALID_ARGUMENT, "null argument");
  iso::Lattice L;
  int st = check_lattice(c, desc, &L);
  if (st != ISO_OK) return st;
  if (desc->source != ISO_SOURCE_PROGRAM)
    return fail(c, ISO_ERR_INVALID_ARGUMENT,
                "extended marching cubes samples normals at the crossings: program sources only");
  if (!(epsilon >= 0.0f) || !std::isfinite(epsilon))
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "epsilon must be finite and >= 0");
  if (c->n_prog == 0) return fail(c, ISO_ERR_NO_SOURCE, "no program uploaded");
  if (!c->prog_directed)
    return fail(c, ISO_ERR_PROGRAM, "the program has no directed form (smooth union)");
  if (epsilon == 0.0f && !c->prog_hermite)
    return fail(c, ISO_ERR_PROGRAM,
                "analytic normals need shapes, unions and translations only; pass epsilon > 0 "
                "for central differences");
  if (out->memory != ISO_MEMORY_HOST && out->memory != ISO_MEMORY_DEVICE)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "unknown output memory kind");
  c->kept = false;
  c->band_valid = false;
  ISO_HIP(c, hipSetDevice(c->device));
  const iso::McDims D = iso::mc_dims(L);
  if (D.n_segs >= (int64_t(1) << 26))
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "lattice has 2^26 or more 64-cell segments");
  iso::McBuffers B{};
  if ((st = prepare_mc(c, D, 0, &B)) != ISO_OK) return st;
  ISO_HIP(c, ensure(&c->dvals, static_cast<size_t>(D.n_points) * 3 * sizeof(float)));
  float* vals = static_cast<float*>(c->dvals.ptr);
  const iso::DevOp* ops = static_cast<const iso::DevOp*>(c->prog.ptr);
  // Directed signs and the active (mixed) cells.
  uint64_t n_act = 0;
  for (int attempt = 0; attempt < 2; ++attempt) {
    ISO_HIP(c, timed(c, "mc_signs_directed", [&] {
              return iso::launch_signs_directed(ops, c->n_prog, L, D, B, vals, c->stream);
            }));
    ISO_HIP(c, timed(c, "mc_count", [&] { return iso::launch_count(L, D, B, c->stream); }));
    if ((st = read_totals(c, B)) != ISO_OK) return st;
    n_act = c->h_totals[iso::kTotActive];
    if (n_act <= B.active_capacity) break;
    if ((st = prepare_mc(c, D, n_act, &B)) != ISO_OK) return st;
  }
  // Vertex and triangle totals share one 62-bit scan word (32 + 30 bits).
  if (n_act * 13 >= (1ull << 32) || n_act * 12 >= (1ull << 30))
    return fail(c, ISO_ERR_OVERFLOW, "too many mixed cells for extended marching cubes");
  const uint64_t na = n_act ? n_act : 1;
  ISO_HIP(c, ensure(&c->emc_cells, na * sizeof(iso::EmcCell)));
  ISO_HIP(c, ensure(&c->emc_first, 2 * na * sizeof(uint32_t)));
  ISO_HIP(c, ensure(&c->emc_ctr, iso::kEmcSlots * sizeof(unsigned long long)));
  ISO_HIP(c, hipMemsetAsync(c->emc_ctr.ptr, 0, iso::kEmcSlots * sizeof(unsigned long long),
                            c->stream));
  iso::scan::State S;
  if ((st = prepare_scan(c, n_act, &S)) != ISO_OK) return st;
  ISO_HIP(c, iso::launch_fill_u32(reinterpret_cast<uint32_t*>(S.err), 2, 0u, c->stream));
  iso::EmcArgs A{};
  A.L = L;
  A.segs_per_row = static_cast<uint32_t>(D.segs_per_row);
  A.values = vals;
  A.active = B.active;
  A.n_act = n_act;
  A.ops = ops;
  A.n_ops = c->n_prog;
  A.epsilon = epsilon;
  A.cells = static_cast<iso::EmcCell*>(c->emc_cells.ptr);
  A.vfirst = static_cast<uint32_t*>(c->emc_first.ptr);
  A.tfirst = A.vfirst + na;
  A.counters = static_cast<unsigned long long*>(c->emc_ctr.ptr);
  ISO_HIP(c, ensure(&c->emc_cross, na * 72 * sizeof(float)));
  A.crossings = static_cast<float*>(c->emc_cross.ptr);
  ISO_HIP(c, timed(c, "emc_crossings", [&] { return iso::launch_emc_crossings(A, c->stream); }));
  ISO_HIP(c, timed(c, "emc_cells", [&] { return iso::launch_emc_cells(A, S, c->stream); }));
  ISO_HIP(c, hipMemcpyAsync(c->h_stage, S.ctl, 4 * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  if (c->h_stage[2]) return fail(c, ISO_ERR_DEVICE, "the cell scan did not complete");
  const unsigned long long total = n_act ? c->h_stage[1] : 0ull;
  const uint64_t nv = total & 0xffffffffull, nt = total >> 32, ni = 3 * nt;
  out->vertex_count = nv;
  out->index_count = ni;
  const bool device = out->memory == ISO_MEMORY_DEVICE;
  if (device) {
    if ((nv && (!out->vertices || out->vertex_capacity < nv)) ||
        (ni && (!out->indices || out->index_capacity < ni)))
      return fail(c, ISO_ERR_CAPACITY, "output buffers too small");
    A.vertices = out->vertices;
    A.indices = out->indices;
  } else {
    ISO_HIP(c, ensure(&c->vout, (nv ? nv : 1) * 3 * sizeof(float)));
    ISO_HIP(c, ensure(&c->iout, (ni ? ni : 1) * sizeof(uint32_t)));
    A.vertices = static_cast<float*>(c->vout.ptr);
    A.indices = static_cast<uint32_t*>(c->iout.ptr);
  }
  A.vcap = nv;
  A.icap = ni;
  ISO_HIP(c, ensure(&c->emc_flips, (nt ? nt : 1) * sizeof(uint4)));
  A.flips = static_cast<uint4*>(c->emc_flips.ptr);
  if (n_act) {
    ISO_HIP(c, timed(c, "emc_emit", [&] { return iso::launch_emc_emit(A, nt, c->stream); }));
    ISO_HIP(c, timed(c, "emc_flips", [&] { return iso::launch_emc_flips(A, nt, c->stream); }));
  }
  ISO_HIP(c, hipMemcpyAsync(c->h_stage, c->emc_ctr.ptr, sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  if (c->h_stage[0]) return fail(c, ISO_ERR_DEVICE, "a feature point's SVD did not converge");
  if (device) return ISO_OK;
  if ((nv && (!out->vertices || out->vertex_capacity < nv)) ||
      (ni && (!out->indices || out->index_capacity < ni))) {
    c->kept = true;
    c->kept_vertices = nv;
    c->kept_indices = ni;
    return fail(c, ISO_ERR_CAPACITY, "output buffers too small; mesh kept for iso_mesh_fetch");
  }
  if (nv)
    ISO_HIP(c, hipMemcpyAsync(out->vertices, A.vertices, nv * 3 * sizeof(float),
                              hipMemcpyDeviceToHost, c->stream));
  if (ni)
    ISO_HIP(c, hipMemcpyAsync(out->indices, A.indices, ni * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  return ISO_OK;
}

int iso_program_sample_vector(iso_ctx* c, const float* points, uint64_t n, float* values) {
  if (!c || (n && (!points || !values))) return fail(c, ISO_ERR_INVALID_ARGUMENT, "null buffer");
  if (c->n_prog == 0) return fail(c, ISO_ERR_NO_SOURCE, "no program uploaded");
  if (!c->prog_directed)
    return fail(c, ISO_ERR_PROGRAM, "the program has no directed form (smooth union)");
  if (n == 0) return ISO_OK;
  ISO_HIP(c, hipSetDevice(c->device));
  ISO_HIP(c, ensure(&c->pts, n * 3 * sizeof(float)));
  ISO_HIP(c, ensure(&c->nrm, n * 3 * sizeof(float)));
  ISO_HIP(c, hipMemcpyAsync(c->pts.ptr, points, n * 3 * sizeof(float), hipMemcpyHostToDevice,
                            c->stream));
  ISO_HIP(c, timed(c, "sample_vector_points", [&] {
            return iso::launch_sample_vector_points(static_cast<const iso::DevOp*>(c->prog.ptr),
                                                    c->n_prog, static_cast<const float*>(c->pts.ptr),
                                                    n, static_cast<float*>(c->nrm.ptr), c->stream);
          }));
  ISO_HIP(c, hipMemcpyAsync(values, c->nrm.ptr, n * 3 * sizeof(float), hipMemcpyDeviceToHost,
                            c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  return ISO_OK;
}

int iso_extract_point_cloud(iso_ctx* c, const iso_grid_desc* desc, iso_mesh_out* out) {
  if (!c || !out) return fail(c, ISO_ERR_INVALID_ARGUMENT, "null argument");
  iso::Lattice L;
  int st = check_lattice(c, desc, &L);
  if (st != ISO_OK) return st;
  const bool from_grid = desc->source == ISO_SOURCE_GRID;
  if (from_grid) {
    if (!c->grid) return fail(c, ISO_ERR_NO_SOURCE, "no grid bound");
    if (c->gnx != L.nx || c->gny != L.ny || c->gnz != L.nz)
      return fail(c, ISO_ERR_INVALID_ARGUMENT, "lattice dims differ from the bound grid");
  } else if (c->n_prog == 0) {
    return fail(c, ISO_ERR_NO_SOURCE, "no program uploaded");
  }
  if (out->memory != ISO_MEMORY_HOST && out->memory != ISO_MEMORY_DEVICE)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "unknown output memory kind");
  c->kept = false;
  c->band_valid = false;
  ISO_HIP(c, hipSetDevice(c->device));
  const iso::McDims D = iso::mc_dims(L);
  if (D.n_segs >= (int64_t(1) << 26))
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "lattice has 2^26 or more 64-cell segments");
  iso::McBuffers B{};
  if ((st = prepare_mc(c, D, 0, &B)) != ISO_OK) return st;
  // Classification and the count's active list (the mixed cells in cell
  // order) are all a point cloud needs.
  const float* grid = from_grid ? c->grid : nullptr;
  for (int attempt = 0; attempt < 2; ++attempt) {
    if ((st = classify_and_count(c, grid, L, D, B)) != ISO_OK) return st;
    if ((st = read_totals(c, B)) != ISO_OK) return st;
    if (c->h_totals[iso::kTotActive] <= B.active_capacity) break;
    if ((st = prepare_mc(c, D, c->h_totals[iso::kTotActive], &B)) != ISO_OK) return st;
  }
  const uint64_t n = c->h_totals[iso::kTotActive];
  out->vertex_count = n;
  out->index_count = 0;
  if (out->memory == ISO_MEMORY_DEVICE) {
    if (n > (out->vertices ? out->vertex_capacity : 0))
      return fail(c, ISO_ERR_CAPACITY, "output buffers too small");
    ISO_HIP(c, timed(c, "point_cloud", [&] {
              return iso::launch_point_cloud(L, D, B, out->vertices, n, c->stream);
            }));
    ISO_HIP(c, hipStreamSynchronize(c->stream));
    collect_timing(c);
    return ISO_OK;
  }
  ISO_HIP(c, ensure(&c->vout, (n ? n : 1) * 3 * sizeof(float)));
  ISO_HIP(c, timed(c, "point_cloud", [&] {
            return iso::launch_point_cloud(L, D, B, static_cast<float*>(c->vout.ptr), n,
                                           c->stream);
          }));
  if (n && (!out->vertices || out->vertex_capacity < n)) {
    ISO_HIP(c, hipStreamSynchronize(c->stream));
    collect_timing(c);
    c->kept = true;
    c->kept_vertices = n;
    c->kept_indices = 0;
    return fail(c, ISO_ERR_CAPACITY, "output buffers too small; points kept for iso_mesh_fetch");
  }
  if (n)
    ISO_HIP(c, hipMemcpyAsync(out->vertices, c->vout.ptr, n * 3 * sizeof(float),
                              hipMemcpyDeviceToHost, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  return ISO_OK;
}

int iso_extract_dc(iso_ctx* c, const iso_grid_desc* desc, float epsilon, iso_mesh_out* out) {
  if (!c || !out) return fail(c, ISO_ERR_INVALID_ARGUMENT, "null argument");
  iso::Lattice L;
  int st = check_lattice(c, desc, &L);
  if (st != ISO_OK) return st;
  if (desc->source != ISO_SOURCE_PROGRAM)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "dual contouring needs a program source (normals)");
  if (c->n_prog == 0) return fail(c, ISO_ERR_NO_SOURCE, "no program uploaded");
  if (!(epsilon >= 0.0f) || !(epsilon < 3.0e38f))
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "epsilon must be finite and >= 0");
  if (epsilon == 0.0f && !c->prog_hermite)
    return fail(c, ISO_ERR_PROGRAM,
                "analytic normals need shapes, unions and translations only; pass epsilon > 0 "
                "for central differences");
  if (out->memory != ISO_MEMORY_HOST && out->memory != ISO_MEMORY_DEVICE)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "unknown output memory kind");
  c->kept = false;
  ISO_HIP(c, hipSetDevice(c->device));

  // The dual lattice has one point per primal cell; with fewer than two
  // dual points along an axis there is no dual cube and the mesh is empty.
  iso::Lattice Ld = L;
  Ld.nx = L.nx - 1;
  Ld.ny = L.ny - 1;
  Ld.nz = L.nz - 1;
  if (Ld.nx < 2 || Ld.ny < 2 || Ld.nz < 2) {
    out->vertex_count = 0;
    out->index_count = 0;
    return ISO_OK;
  }
  const uint64_t n_points = static_cast<uint64_t>(L.nx * L.ny * L.nz);
  const uint64_t n_cells = static_cast<uint64_t>(Ld.nx * Ld.ny * Ld.nz);
  ISO_HIP(c, ensure(&c->dc_samples, n_points * sizeof(float)));
  ISO_HIP(c, ensure(&c->dc_values, n_cells * sizeof(float)));
  ISO_HIP(c, ensure(&c->dc_points, 3 * n_cells * sizeof(float)));
  ISO_HIP(c, ensure(&c->dc_mixed, n_cells * sizeof(uint32_t)));
  ISO_HIP(c, ensure(&c->dc_flags, n_cells));
  ISO_HIP(c, ensure(&c->dc_counters, iso::kDcSlots * sizeof(unsigned long long)));
  iso::DcBuffers B;
  B.samples = static_cast<float*>(c->dc_samples.ptr);
  B.dual_values = static_cast<float*>(c->dc_values.ptr);
  B.dual_points = static_cast<float*>(c->dc_points.ptr);
  B.mixed = static_cast<uint32_t*>(c->dc_mixed.ptr);
  B.mixed_flag = static_cast<uint8_t*>(c->dc_flags.ptr);
  B.counters = static_cast<unsigned long long*>(c->dc_counters.ptr);
  const iso::DevOp* ops = static_cast<const iso::DevOp*>(c->prog.ptr);
  ISO_HIP(c, timed(c, "dc_sample",
                   [&] { return iso::launch_dc_sample(ops, c->n_prog, L, B, c->stream); }));
  iso::scan::State S{};
  if ((st = prepare_scan(c, n_cells, &S)) != ISO_OK) return st;
  ISO_HIP(c, timed(c, "dc_dual", [&] { return iso::launch_dc_dual(L, B, S, c->stream); }));
  // Central differences evaluate the program six times per normal: the
  // mixed cells' corners are listed once each and their normals shared.
  // Analytic normals cost one evaluation: each cell corner takes its own.
  const bool shared = epsilon > 0.0f;
  uint32_t* plist = nullptr;
  uint32_t* pindex = nullptr;
  if (shared) {
    ISO_HIP(c, ensure(&c->dc_plist, n_points * sizeof(uint32_t)));
    ISO_HIP(c, ensure(&c->dc_pindex, n_points * sizeof(uint32_t)));
    plist = static_cast<uint32_t*>(c->dc_plist.ptr);
    pindex = static_cast<uint32_t*>(c->dc_pindex.ptr);
    if ((st = prepare_scan(c, n_points, &S)) != ISO_OK) return st;
    ISO_HIP(c, timed(c, "dc_points", [&] {
              return iso::launch_dc_corner_points(L, B, plist, pindex, S, c->stream);
            }));
  }
  // The counts size the normals and feature passes.
  ISO_HIP(c, hipMemcpyAsync(c->h_dc, B.counters, iso::kDcSlots * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  if (c->h_dc[iso::kDcScanErr])
    return fail(c, ISO_ERR_DEVICE, "a dual-contouring scan did not complete");
  const uint64_t n_mixed = c->h_dc[iso::kDcMixed];
  const uint64_t n_normals = shared ? c->h_dc[iso::kDcPoints] : 8 * n_mixed;
  ISO_HIP(c, ensure(&c->dc_normals, (n_normals ? n_normals : 1) * 3 * sizeof(float)));
  float* pnormals = static_cast<float*>(c->dc_normals.ptr);
  ISO_HIP(c, timed(c, "dc_normals", [&] {
            return shared ? iso::launch_dc_normals(ops, c->n_prog, L, epsilon, plist, n_normals,
                                                   pnormals, c->stream)
                          : iso::launch_dc_cell_normals(ops, c->n_prog, L, epsilon, B, n_mixed,
                                                        pnormals, c->stream);
          }));
  ISO_HIP(c, timed(c, "dc_feature", [&] {
            return iso::launch_dc_feature(L, B, n_mixed, pindex, pnormals, c->stream);
          }));
  st = extract_lattice(c, Ld, B.dual_values, B.dual_points, out);
  // The stall flag is final once the stream has passed the feature kernel,
  // which every extract_lattice path has synchronised past.
  ISO_HIP(c, hipMemcpyAsync(c->h_dc, B.counters, iso::kDcSlots * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  if (c->h_dc[iso::kDcStalled])
    return fail(c, ISO_ERR_DEVICE, "a feature point's SVD did not converge");
  return st;
}

}  // extern "C"

namespace {

// Grows a device buffer to `bytes`, keeping its first `keep` bytes.
int grow_keep(iso_ctx* c, DevBuf* b, size_t bytes, size_t keep) {
  if (b->bytes >= bytes && b->ptr) return ISO_OK;
  size_t want = b->bytes ? b->bytes : 256;
  while (want < bytes) want *= 2;
  void* p = nullptr;
  ISO_HIP(c, hipMalloc(&p, want));
  if (keep && b->ptr) {
    hipError_t e = hipMemcpyAsync(p, b->ptr, keep, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      (void)hipFree(p);
      return fail(c, ISO_ERR_DEVICE, std::string("grow copy: ") + hipGetErrorString(e));
    }
  }
  release(b);
  b->ptr = p;
  b->bytes = want;
  return ISO_OK;
}

// Scan state for scans of up to n items at one item per thread; err points
// at the status word.
int prepare_scan(iso_ctx* c, uint64_t n, iso::scan::State* st) {
  const size_t nb = static_cast<size_t>(iso::scan::blocks_for(n, 1));
  ISO_HIP(c, ensure(&c->sc_slot, nb * sizeof(unsigned long long)));
  ISO_HIP(c, ensure(&c->sc_ctl, 4 * sizeof(unsigned long long)));
  st->slot = static_cast<unsigned long long*>(c->sc_slot.ptr);
  st->ctl = static_cast<unsigned long long*>(c->sc_ctl.ptr);
  st->err = st->ctl + 2;
  return ISO_OK;
}

// Reads ctl[0..4) of the scan state (total in [1], status bits in [2]);
// `stage` names the failing pass in the error message.
int read_ctl(iso_ctx* c, const iso::scan::State& st, const char* stage = "octree extraction") {
  ISO_HIP(c, hipMemcpyAsync(c->h_stage, st.ctl, 4 * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  if (c->h_stage[2]) {
    const unsigned long long e = c->h_stage[2];
    return fail(c, ISO_ERR_DEVICE,
                std::string(stage) + " failed:" + ((e & iso::kLhmcErrScan) ? " scan" : "") +
                    ((e & iso::kLhmcErrWalk) ? " dual-corner walk" : "") +
                    ((e & iso::kLhmcErrTable) ? " edge table" : ""));
  }
  return ISO_OK;
}

}  // namespace

extern "C" {

int iso_extract_lhmc(iso_ctx* c, int32_t max_depth, iso_mesh_out* out) {
  if (!c || !out) return fail(c, ISO_ERR_INVALID_ARGUMENT, "null argument");
  if (c->n_prog == 0) return fail(c, ISO_ERR_NO_SOURCE, "no program uploaded");
  if (max_depth < 0 || max_depth > iso::kLhmcMaxDepth)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "max_depth must be in [0, 20]");
  if (out->memory != ISO_MEMORY_HOST && out->memory != ISO_MEMORY_DEVICE)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "unknown output memory kind");
  c->kept = false;
  ISO_HIP(c, hipSetDevice(c->device));
  int st = ISO_OK;
  const iso::DevOp* ops = static_cast<const iso::DevOp*>(c->prog.ptr);
  const int d1 = (max_depth > 2 ? max_depth : 2) + 1;  // deepest possible node
  const bool dense_primal = d1 <= iso::kLhmcDenseDepth;
  const uint64_t n_dense = dense_primal ? 1ull << (3 * d1) : 0;
  // Morton::level's split points, found once with the host's libm log
  static uint64_t splits[iso::kLhmcMaxLevels];
  static std::once_flag splits_once;
  std::call_once(splits_once, [] { iso::lhmc_level_splits(splits); });
  ISO_HIP(c, iso::launch_lhmc_levels(splits, c->stream));
  iso::scan::State S{};
  if ((st = prepare_scan(c, 1u << 16, &S)) != ISO_OK) return st;
  // status word: cleared once, accumulated by every kernel below
  ISO_HIP(c, iso::launch_fill_u32(reinterpret_cast<uint32_t*>(S.err), 2, 0u, c->stream));

  // 1. octree, breadth first
  iso::LhmcNodes N{};
  N.n_levels = d1 + 1;
  int64_t off[iso::kLhmcMaxLevels + 2] = {0};
  off[0] = 0;
  off[1] = 1;
  {
    const size_t cap = 4096;
    if ((st = grow_keep(c, &c->oc_keys, cap * 8, 0)) != ISO_OK) return st;
    c->h_stage[4] = 1ull;  // the root key
    ISO_HIP(c, hipMemcpyAsync(c->oc_keys.ptr, &c->h_stage[4], sizeof(uint64_t),
                              hipMemcpyHostToDevice, c->stream));
    ISO_HIP(c, hipStreamSynchronize(c->stream));
  }
  for (int d = 0; d <= d1; ++d) {
    const int64_t n = off[d + 1] - off[d];
    const int64_t next_cap = d < d1 ? 8 * n : 0;
    const size_t need = static_cast<size_t>(off[d + 1] + next_cap);
    if ((st = grow_keep(c, &c->oc_keys, need * 8, static_cast<size_t>(off[d + 1]) * 8)) != ISO_OK ||
        (st = grow_keep(c, &c->oc_values, need * 4, static_cast<size_t>(off[d]) * 4)) != ISO_OK ||
        (st = grow_keep(c, &c->oc_leaf, need, static_cast<size_t>(off[d]))) != ISO_OK)
      return st;
    if ((st = prepare_scan(c, static_cast<uint64_t>(n), &S)) != ISO_OK) return st;
    ISO_HIP(c, timed(c, "lhmc_level", [&] {
              return iso::launch_lhmc_level(
                  ops, c->n_prog, max_depth, static_cast<uint64_t*>(c->oc_keys.ptr),
                  static_cast<float*>(c->oc_values.ptr), static_cast<uint8_t*>(c->oc_leaf.ptr),
                  off[d], n, next_cap, S, c->stream);
            }));
    if ((st = read_ctl(c, S)) != ISO_OK) return st;
    const int64_t children = 8 * static_cast<int64_t>(c->h_stage[1]);
    if (children > next_cap)
      return fail(c, ISO_ERR_DEVICE, "octree refined past its deepest possible depth");
    off[d + 2] = off[d + 1] + children;
  }
  N.keys = static_cast<uint64_t*>(c->oc_keys.ptr);
  N.values = static_cast<float*>(c->oc_values.ptr);
  N.leaf = static_cast<uint8_t*>(c->oc_leaf.ptr);
  for (int d = 0; d <= N.n_levels; ++d) N.level_off[d] = off[d];
  const int64_t n_nodes = off[N.n_levels];
  if (n_nodes >= 0xffffffffll)
    return fail(c, ISO_ERR_OVERFLOW, "octree has 2^32 or more nodes");

  // 2. primal vertices in ascending Morton order
  iso::LhmcCubes C{};
  const uint64_t n_pairs = 8 * static_cast<uint64_t>(n_nodes);
  uint64_t pv_cap = n_pairs;
  if (dense_primal && pv_cap > n_dense) pv_cap = n_dense;
  ISO_HIP(c, ensure(&c->oc_primal, pv_cap * sizeof(uint64_t)));
  ISO_HIP(c, ensure(&c->oc_plevel, pv_cap));
  C.primal = static_cast<uint64_t*>(c->oc_primal.ptr);
  C.primal_level = static_cast<uint8_t*>(c->oc_plevel.ptr);
  if ((st = prepare_scan(c, dense_primal ? n_dense : n_pairs, &S)) != ISO_OK) return st;
  if (dense_primal) {
    ISO_HIP(c, ensure(&c->oc_dense, n_dense * sizeof(uint32_t)));
    ISO_HIP(c, timed(c, "lhmc_primal", [&] {
              return iso::launch_lhmc_primal(N, d1, static_cast<uint32_t*>(c->oc_dense.ptr),
                                             n_dense, C, pv_cap, S, c->stream);
            }));
  } else {
    const size_t temp = iso::lhmc_sort_temp_bytes(n_pairs);
    if (temp == 0) return fail(c, ISO_ERR_DEVICE, "radix sort size query failed");
    ISO_HIP(c, ensure(&c->oc_pkeys, n_pairs * sizeof(uint64_t)));
    ISO_HIP(c, ensure(&c->oc_skeys, n_pairs * sizeof(uint64_t)));
    ISO_HIP(c, ensure(&c->oc_plevels, n_pairs));
    ISO_HIP(c, ensure(&c->oc_slevels, n_pairs));
    ISO_HIP(c, ensure(&c->oc_sort, temp));
    ISO_HIP(c, timed(c, "lhmc_primal", [&] {
              return iso::launch_lhmc_primal_sorted(
                  N, static_cast<uint64_t*>(c->oc_pkeys.ptr),
                  static_cast<uint8_t*>(c->oc_plevels.ptr),
                  static_cast<uint64_t*>(c->oc_skeys.ptr),
                  static_cast<uint8_t*>(c->oc_slevels.ptr), c->oc_sort.ptr, temp, C, pv_cap, S,
                  c->stream);
            }));
  }
  if ((st = read_ctl(c, S)) != ISO_OK) return st;
  const uint64_t n_pv = c->h_stage[1];
  if (n_pv > pv_cap) return fail(c, ISO_ERR_DEVICE, "primal vertex list overflow");
  if (n_pv > 0x0fffffffull) return fail(c, ISO_ERR_OVERFLOW, "too many dual cubes");
  C.n = static_cast<uint32_t>(n_pv);

  // 3. dual cubes
  ISO_HIP(c, ensure(&c->oc_corner, (n_pv ? n_pv : 1) * 8 * sizeof(uint32_t)));
  ISO_HIP(c, ensure(&c->oc_case, n_pv ? n_pv : 1));
  ISO_HIP(c, ensure(&c->oc_base, (n_pv ? n_pv : 1) * sizeof(unsigned long long)));
  C.corner = static_cast<uint32_t*>(c->oc_corner.ptr);
  C.cube_case = static_cast<uint8_t*>(c->oc_case.ptr);
  C.base = static_cast<unsigned long long*>(c->oc_base.ptr);
  ISO_HIP(c, timed(c, "lhmc_cubes", [&] { return iso::launch_lhmc_cubes(N, d1, C, S, c->stream); }));
  if ((st = read_ctl(c, S)) != ISO_OK) return st;
  const uint64_t n_tri = c->h_stage[1];

  // 4. edge ownership and first ids
  iso::LhmcEdgeTable T{};
  uint64_t tcap = 1024;
  while (tcap < 6 * n_tri) tcap *= 2;
  ISO_HIP(c, ensure(&c->oc_tkeys, tcap * sizeof(unsigned long long)));
  ISO_HIP(c, ensure(&c->oc_towner, tcap * sizeof(uint32_t)));
  ISO_HIP(c, ensure(&c->oc_tid, tcap * sizeof(uint32_t)));
  T.keys = static_cast<unsigned long long*>(c->oc_tkeys.ptr);
  T.owner = static_cast<uint32_t*>(c->oc_towner.ptr);
  T.id = static_cast<uint32_t*>(c->oc_tid.ptr);
  T.mask = tcap - 1;
  ISO_HIP(c, timed(c, "lhmc_own", [&] {
            hipError_t e = iso::launch_lhmc_table_clear(T, c->stream);
            if (e != hipSuccess) return e;
            return iso::launch_lhmc_own(N, C, T, S.err, S, c->stream);
          }));
  if ((st = read_ctl(c, S)) != ISO_OK) return st;
  const uint64_t nv = c->h_stage[1] & 0xffffffffull;
  const uint64_t ni = 3 * (c->h_stage[1] >> 32);
  if ((c->h_stage[1] >> 32) != n_tri) return fail(c, ISO_ERR_DEVICE, "triangle count mismatch");
  if (ni > 0xffffffffull) return fail(c, ISO_ERR_OVERFLOW, "mesh exceeds u32 indices");
  out->vertex_count = nv;
  out->index_count = ni;

  // 5. emit
  float* vdst = nullptr;
  uint32_t* idst = nullptr;
  uint64_t vcap = 0, icap = 0;
  if (out->memory == ISO_MEMORY_DEVICE) {
    vdst = out->vertices;
    idst = out->indices;
    vcap = out->vertices ? out->vertex_capacity : 0;
    icap = out->indices ? out->index_capacity : 0;
  } else {
    ISO_HIP(c, ensure(&c->vout, (nv ? nv : 1) * 3 * sizeof(float)));
    ISO_HIP(c, ensure(&c->iout, (ni ? ni : 1) * sizeof(uint32_t)));
    vdst = static_cast<float*>(c->vout.ptr);
    idst = static_cast<uint32_t*>(c->iout.ptr);
    vcap = nv;
    icap = ni;
  }
  ISO_HIP(c, timed(c, "lhmc_emit", [&] {
            return iso::launch_lhmc_emit(N, C, T, vdst, vcap, idst, icap, S.err, c->stream);
          }));
  if ((st = read_ctl(c, S)) != ISO_OK) return st;
  if (out->memory == ISO_MEMORY_DEVICE) {
    if (nv > vcap || ni > icap) return fail(c, ISO_ERR_CAPACITY, "output buffers too small");
    return ISO_OK;
  }
  if ((nv && (!out->vertices || out->vertex_capacity < nv)) ||
      (ni && (!out->indices || out->index_capacity < ni))) {
    c->kept = true;
    c->kept_vertices = nv;
    c->kept_indices = ni;
    return fail(c, ISO_ERR_CAPACITY, "output buffers too small; mesh kept for iso_mesh_fetch");
  }
  if (nv)
    ISO_HIP(c, hipMemcpyAsync(out->vertices, vdst, nv * 3 * sizeof(float), hipMemcpyDeviceToHost,
                              c->stream));
  if (ni)
    ISO_HIP(c, hipMemcpyAsync(out->indices, idst, ni * sizeof(uint32_t), hipMemcpyDeviceToHost,
                              c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  return ISO_OK;
}

int iso_batch_upload(iso_ctx* c, const iso_node* nodes, uint64_t n_nodes,
                     const iso_chunk_desc* chunks, uint32_t n_chunks) {
  if (!c || !nodes || !chunks || n_chunks == 0)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "empty or null batch");
  ISO_HIP(c, hipSetDevice(c->device));
  c->b_ready = false;
  c->b_tuned = false;
  if (c->b_graph) {
    (void)hipGraphExecDestroy(c->b_graph);
    c->b_graph = nullptr;
  }
  std::vector<iso::DevOp> ops;
  std::vector<uint32_t> op_off(n_chunks), seg_start(n_chunks + 1);
  std::vector<uint64_t> word_start(n_chunks + 1), sign_off(n_chunks);
  uint64_t segs = 0, words = 0, sign_words = 0, cells = 0;
  for (uint32_t i = 0; i < n_chunks; ++i) {
    const iso_chunk_desc& ch = chunks[i];
    iso::Lattice L;
    int st = check_lattice(c, &ch.grid, &L);
    if (st != ISO_OK) return st;
    if (ch.grid.source != ISO_SOURCE_PROGRAM)
      return fail(c, ISO_ERR_INVALID_ARGUMENT, "batch chunks must name ISO_SOURCE_PROGRAM");
    if (static_cast<uint64_t>(ch.node_offset) + ch.node_count > n_nodes || ch.node_count == 0)
      return fail(c, ISO_ERR_INVALID_ARGUMENT,
                  "chunk " + std::to_string(i) + ": program outside the node array");
    std::vector<iso::DevOp> part;
    std::string err;
    st = iso::lower_program(nodes + ch.node_offset, ch.node_count, &part, &err);
    if (st != ISO_OK) return fail(c, st, "chunk " + std::to_string(i) + ": " + err);
    op_off[i] = static_cast<uint32_t>(ops.size());
    ops.insert(ops.end(), part.begin(), part.end());
    const iso::McDims D = iso::mc_dims(L);
    seg_start[i] = static_cast<uint32_t>(segs);
    word_start[i] = words;
    sign_off[i] = sign_words + 1;  // one guard word before, two after
    segs += static_cast<uint64_t>(D.n_segs);
    words += static_cast<uint64_t>(D.n_sign_words);
    sign_words += static_cast<uint64_t>(D.n_sign_words) + 3;
    cells += static_cast<uint64_t>(D.n_cells);
    if (segs >= (1ull << 26))
      return fail(c, ISO_ERR_INVALID_ARGUMENT, "batch has 2^26 or more 64-cell segments");
  }
  seg_start[n_chunks] = static_cast<uint32_t>(segs);
  word_start[n_chunks] = words;
  ISO_HIP(c, ensure(&c->b_ops, ops.size() * sizeof(iso::DevOp)));
  ISO_HIP(c, ensure(&c->b_signs, sign_words * sizeof(uint64_t)));
  ISO_HIP(c, ensure(&c->b_chunks, n_chunks * sizeof(iso::CellSpace)));
  ISO_HIP(c, ensure(&c->b_segstart, (n_chunks + 1) * sizeof(uint32_t)));
  ISO_HIP(c, ensure(&c->b_wordstart, (n_chunks + 1) * sizeof(uint64_t)));
  ISO_HIP(c, ensure(&c->b_voff, (n_chunks + 1) * sizeof(uint64_t)));
  ISO_HIP(c, ensure(&c->b_ioff, (n_chunks + 1) * sizeof(uint64_t)));
  // The programs' samples, 64 per sign word, for emit to read (as the
  // single-lattice program path stores them); without the memory, emit
  // evaluates the programs at the edge endpoints.
  const bool stored = ensure(&c->b_values, words * 64 * sizeof(float)) == hipSuccess;
  if (!stored) (void)hipGetLastError();
  std::vector<iso::CellSpace> table(n_chunks);
  const iso::DevOp* dops = static_cast<const iso::DevOp*>(c->b_ops.ptr);
  const uint64_t* dsigns = static_cast<const uint64_t*>(c->b_signs.ptr);
  const float* dvalues = static_cast<const float*>(c->b_values.ptr);
  for (uint32_t i = 0; i < n_chunks; ++i) {
    const iso_grid_desc& g = chunks[i].grid;
    iso::CellSpace& C = table[i];
    std::memset(&C, 0, sizeof(C));
    C.nx = g.nx;
    C.ny = g.ny;
    C.nz = g.nz;
    for (int k = 0; k < 3; ++k) {
      C.origin[k] = g.origin[k];
      C.spacing[k] = g.spacing[k];
    }
    C.signs = dsigns + sign_off[i];
    C.segs_per_row = static_cast<uint32_t>((g.nx - 1 + 63) / 64);
    C.seg_off = seg_start[i];
    C.ops = dops + op_off[i];
    C.n_ops = static_cast<int32_t>((i + 1 < n_chunks ? op_off[i + 1] : ops.size()) - op_off[i]);
    C.grid = stored ? dvalues + 64 * word_start[i] : nullptr;
  }
  ISO_HIP(c, hipMemcpyAsync(c->b_ops.ptr, ops.data(), ops.size() * sizeof(iso::DevOp),
                            hipMemcpyHostToDevice, c->stream));
  ISO_HIP(c, hipMemcpyAsync(c->b_chunks.ptr, table.data(), n_chunks * sizeof(iso::CellSpace),
                            hipMemcpyHostToDevice, c->stream));
  ISO_HIP(c, hipMemcpyAsync(c->b_segstart.ptr, seg_start.data(), (n_chunks + 1) * sizeof(uint32_t),
                            hipMemcpyHostToDevice, c->stream));
  ISO_HIP(c, hipMemcpyAsync(c->b_wordstart.ptr, word_start.data(),
                            (n_chunks + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  c->plan.chunks = static_cast<const iso::CellSpace*>(c->b_chunks.ptr);
  c->plan.seg_start = static_cast<const uint32_t*>(c->b_segstart.ptr);
  c->plan.word_start = static_cast<const uint64_t*>(c->b_wordstart.ptr);
  c->plan.n_chunks = n_chunks;
  c->plan.n_words = static_cast<int64_t>(words);
  c->plan.n_segs = static_cast<int64_t>(segs);
  c->plan.stored_values = stored;
  c->b_cells = cells;
  c->b_ready = true;
  return ISO_OK;
}

}  // extern "C"

namespace {

// The batch's launches: signs (with state reset), count, emit (unless only
// counting), offsets.
hipError_t batch_launches(iso_ctx* c, const iso::McBuffers& B, bool emit = true) {
  const iso::BatchPlan& P = c->plan;
  hipError_t e = timed(c, "batch_signs", [&] {
    return iso::launch_batch_signs(P, B, c->b_words_per_wave, c->stream);
  });
  if (e == hipSuccess)
    e = timed(c, "batch_count", [&] { return iso::launch_batch_count(P, B, c->b_rounds, c->stream); });
  if (e == hipSuccess && emit)
    e = timed(c, "batch_emit", [&] {
      return iso::launch_batch_emit(P, B, c->b_emit_per_cu, c->stream);
    });
  if (e == hipSuccess)
    e = iso::launch_batch_offsets(P, B, static_cast<uint64_t*>(c->b_voff.ptr),
                                  static_cast<uint64_t*>(c->b_ioff.ptr), c->stream);
  return e;
}

// Times every launch shape on the batch (outputs already sized) and keeps
// the fastest: count tile (segments per thread), sign words per wavefront,
// emit workgroups per CU.
int autotune_batch(iso_ctx* c, const iso::McBuffers& B) {
  const bool timing = c->timing;
  c->timing = false;
  hipEvent_t a = take_event(c), b = take_event(c);
  if (!a || !b) return fail(c, ISO_ERR_DEVICE, "no events for the autotuner");
  const int rounds[] = {1, 2};
  const int wpw[] = {1, 2, 4};
  const int per_cu[] = {0, 2, 4, 8};
  float best = 1e30f;
  int br = c->b_rounds, bw = c->b_words_per_wave, be = c->b_emit_per_cu;
  auto measure = [&](float* ms) -> int {
    float t = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      ISO_HIP(c, hipEventRecord(a, c->stream));
      ISO_HIP(c, batch_launches(c, B));
      ISO_HIP(c, hipEventRecord(b, c->stream));
      ISO_HIP(c, hipEventSynchronize(b));
      float x = 0.0f;
      ISO_HIP(c, hipEventElapsedTime(&x, a, b));
      if (x < t) t = x;
    }
    *ms = t;
    return ISO_OK;
  };
  // Coordinate search: each shape in turn with the others at their best.
  int st = ISO_OK;
  for (int r : rounds) {
    c->b_rounds = r;
    float ms;
    if ((st = measure(&ms)) != ISO_OK) break;
    if (ms < best) best = ms, br = r;
  }
  c->b_rounds = br;
  for (int w : wpw) {
    if (st != ISO_OK) break;
    c->b_words_per_wave = w;
    float ms;
    if ((st = measure(&ms)) != ISO_OK) break;
    if (ms < best) best = ms, bw = w;
  }
  c->b_words_per_wave = bw;
  for (int e : per_cu) {
    if (st != ISO_OK) break;
    c->b_emit_per_cu = e;
    float ms;
    if ((st = measure(&ms)) != ISO_OK) break;
    if (ms < best) best = ms, be = e;
  }
  c->b_emit_per_cu = be;
  c->event_pool.push_back(a);
  c->event_pool.push_back(b);
  c->timing = timing;
  c->b_tuned = st == ISO_OK;
  return st;
}

}  // namespace

extern "C" {

int iso_extract_batch(iso_ctx* c, iso_batch_out* out) {
  if (!c || !out) return fail(c, ISO_ERR_INVALID_ARGUMENT, "null argument");
  if (!c->b_ready) return fail(c, ISO_ERR_NO_SOURCE, "no batch uploaded");
  if (out->memory != ISO_MEMORY_HOST && out->memory != ISO_MEMORY_DEVICE)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "unknown output memory kind");
  if (!out->vertex_offsets || !out->index_offsets)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "offset arrays are required");
  c->kept = false;
  c->band_valid = false;
  ISO_HIP(c, hipSetDevice(c->device));
  const iso::BatchPlan& P = c->plan;
  iso::McDims D{};
  D.n_segs = P.n_segs;
  D.n_blocks = iso::batch_count_blocks(P, 1);  // look-back words for the finest tile
  D.n_cells = static_cast<int64_t>(c->b_cells);
  int st = ISO_OK;
  iso::McBuffers B{};
  if ((st = prepare_mc(c, D, 0, &B)) != ISO_OK) return st;
  const bool device = out->memory == ISO_MEMORY_DEVICE;
  const bool graph = device && (out->flags & ISO_BATCH_GRAPH);

  auto finish = [&](uint64_t nv, uint64_t ni) -> int {
    ISO_HIP(c, hipMemcpyAsync(out->vertex_offsets, c->b_voff.ptr, (P.n_chunks + 1) * sizeof(uint64_t),
                              hipMemcpyDeviceToHost, c->stream));
    ISO_HIP(c, hipMemcpyAsync(out->index_offsets, c->b_ioff.ptr, (P.n_chunks + 1) * sizeof(uint64_t),
                              hipMemcpyDeviceToHost, c->stream));
    ISO_HIP(c, hipStreamSynchronize(c->stream));
    out->vertex_count = nv;
    out->index_count = ni;
    return ISO_OK;
  };

  if (device) {
    B.vertices = out->vertices;
    B.indices = out->indices;
    B.vertex_capacity = out->vertices ? out->vertex_capacity : 0;
    B.index_capacity = out->indices ? out->index_capacity : 0;
    // A graph replays only with exactly the buffers it was captured with.
    const iso::McBuffers& G = c->g_buffers;
    const bool replay = graph && c->b_graph && G.seg == B.seg && G.tplanes == B.tplanes &&
                        G.vplanes == B.vplanes &&
                        G.agg == B.agg && G.incl == B.incl && G.active == B.active &&
                        G.active_capacity == B.active_capacity && G.totals == B.totals &&
                        G.vertices == B.vertices && G.indices == B.indices &&
                        G.vertex_capacity == B.vertex_capacity &&
                        G.index_capacity == B.index_capacity;
    if (replay) {
      ISO_HIP(c, hipGraphLaunch(c->b_graph, c->stream));
      if ((st = read_totals(c, B)) != ISO_OK) return st;
    } else {
      for (int attempt = 0; attempt < 2; ++attempt) {
        ISO_HIP(c, batch_launches(c, B));
        if ((st = read_totals(c, B)) != ISO_OK) return st;
        const uint64_t n_act = c->h_totals[iso::kTotActive];
        if (n_act <= B.active_capacity) break;
        if ((st = prepare_mc(c, D, n_act, &B)) != ISO_OK) return st;
      }
    }
    const uint64_t nv = c->h_totals[iso::kTotVertices];
    const uint64_t ni = 3 * c->h_totals[iso::kTotTriangles];
    if ((st = finish(nv, ni)) != ISO_OK) return st;
    if (c->h_totals[iso::kTotErrors] & 1ull)
      return fail(c, ISO_ERR_OVERFLOW, "mesh exceeds u32 indices");
    if (nv > B.vertex_capacity || ni > B.index_capacity)
      return fail(c, ISO_ERR_CAPACITY, "output buffers too small");
    if ((out->flags & ISO_BATCH_AUTOTUNE) && !c->b_tuned) {
      if ((st = autotune_batch(c, B)) != ISO_OK) return st;
      if (c->b_graph) {
        (void)hipGraphExecDestroy(c->b_graph);
        c->b_graph = nullptr;
      }
    }
    if (graph && !replay) {
      // Capture the launches for this output binding; later calls replay it.
      if (c->b_graph) {
        (void)hipGraphExecDestroy(c->b_graph);
        c->b_graph = nullptr;
      }
      const bool timing = c->timing;
      c->timing = false;
      hipGraph_t g = nullptr;
      ISO_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      const hipError_t le = batch_launches(c, B);
      const hipError_t ce = hipStreamEndCapture(c->stream, &g);
      c->timing = timing;
      ISO_HIP(c, le);
      ISO_HIP(c, ce);
      const hipError_t ie = hipGraphInstantiate(&c->b_graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      ISO_HIP(c, ie);
      c->g_buffers = B;
    }
    return ISO_OK;
  }

  // Host memory: count, size the context's buffers, emit, copy.
  B.vertices = nullptr;
  B.indices = nullptr;
  for (int attempt = 0; attempt < 2; ++attempt) {
    ISO_HIP(c, batch_launches(c, B, false));
    if ((st = read_totals(c, B)) != ISO_OK) return st;
    const uint64_t n_act = c->h_totals[iso::kTotActive];
    if (n_act <= B.active_capacity) break;
    if ((st = prepare_mc(c, D, n_act, &B)) != ISO_OK) return st;
  }
  const uint64_t nv = c->h_totals[iso::kTotVertices];
  const uint64_t ni = 3 * c->h_totals[iso::kTotTriangles];
  if (c->h_totals[iso::kTotErrors] & 1ull) return fail(c, ISO_ERR_OVERFLOW, "mesh exceeds u32 indices");
  ISO_HIP(c, ensure(&c->vout, (nv ? nv : 1) * 3 * sizeof(float)));
  ISO_HIP(c, ensure(&c->iout, (ni ? ni : 1) * sizeof(uint32_t)));
  B.vertices = static_cast<float*>(c->vout.ptr);
  B.indices = static_cast<uint32_t*>(c->iout.ptr);
  B.vertex_capacity = nv;
  B.index_capacity = ni;
  ISO_HIP(c, timed(c, "batch_emit", [&] {
            return iso::launch_batch_emit(P, B, c->b_emit_per_cu, c->stream);
          }));
  if ((st = finish(nv, ni)) != ISO_OK) return st;
  if ((nv && (!out->vertices || out->vertex_capacity < nv)) ||
      (ni && (!out->indices || out->index_capacity < ni))) {
    c->kept = true;
    c->kept_vertices = nv;
    c->kept_indices = ni;
    return fail(c, ISO_ERR_CAPACITY, "output buffers too small; mesh kept for iso_mesh_fetch");
  }
  if (nv)
    ISO_HIP(c, hipMemcpyAsync(out->vertices, B.vertices, nv * 3 * sizeof(float),
                              hipMemcpyDeviceToHost, c->stream));
  if (ni)
    ISO_HIP(c, hipMemcpyAsync(out->indices, B.indices, ni * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  return ISO_OK;
}

}  // extern "C"

namespace {

// Device model of the context (allocated on first use).
int nn_model(iso_ctx* c, iso::mlp::Model* M) {
  const size_t P = iso::mlp::kParams * sizeof(float);
  ISO_HIP(c, ensure(&c->nn_theta, P));
  ISO_HIP(c, ensure(&c->nn_m, P));
  ISO_HIP(c, ensure(&c->nn_v, P));
  ISO_HIP(c, ensure(&c->nn_grad, P));
  ISO_HIP(c, ensure(&c->nn_frags, iso::mlp::kFrags * sizeof(uint16_t)));
  ISO_HIP(c, ensure(&c->nn_loss, 4 * sizeof(float)));
  M->theta = static_cast<float*>(c->nn_theta.ptr);
  M->m = static_cast<float*>(c->nn_m.ptr);
  M->v = static_cast<float*>(c->nn_v.ptr);
  M->grad = static_cast<float*>(c->nn_grad.ptr);
  M->frags = static_cast<uint16_t*>(c->nn_frags.ptr);
  M->loss = static_cast<float*>(c->nn_loss.ptr);
  return ISO_OK;
}

// Uploads theta, clears the Adam state and gradient, packs the fragments.
int nn_load(iso_ctx* c, const float* theta) {
  iso::mlp::Model M;
  int st = nn_model(c, &M);
  if (st != ISO_OK) return st;
  const size_t P = iso::mlp::kParams * sizeof(float);
  ISO_HIP(c, hipMemcpyAsync(M.theta, theta, P, hipMemcpyHostToDevice, c->stream));
  std::vector<float> zeros(iso::mlp::kParams, 0.0f);
  ISO_HIP(c, hipMemcpyAsync(M.m, zeros.data(), P, hipMemcpyHostToDevice, c->stream));
  ISO_HIP(c, hipMemcpyAsync(M.v, zeros.data(), P, hipMemcpyHostToDevice, c->stream));
  ISO_HIP(c, iso::mlp::launch_zero_grad(M, c->stream));
  ISO_HIP(c, iso::mlp::launch_pack(M, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  c->nn_ready = true;
  c->nn_step = 0;
  return ISO_OK;
}

int nn_batch(iso_ctx* c, int64_t n, iso::mlp::Batch* B) {
  ISO_HIP(c, ensure(&c->nn_x, static_cast<size_t>(iso::mlp::kFeatPad) * n * sizeof(uint16_t)));
  ISO_HIP(c, ensure(&c->nn_h, 3 * static_cast<size_t>(iso::mlp::kHidden) * n * sizeof(uint16_t)));
  ISO_HIP(c, ensure(&c->nn_dz, 3 * static_cast<size_t>(iso::mlp::kHidden) * n * sizeof(uint16_t)));
  ISO_HIP(c, ensure(&c->nn_g, n * sizeof(float)));
  uint16_t* h = static_cast<uint16_t*>(c->nn_h.ptr);
  uint16_t* dz = static_cast<uint16_t*>(c->nn_dz.ptr);
  const size_t hn = static_cast<size_t>(iso::mlp::kHidden) * n;
  B->x = static_cast<uint16_t*>(c->nn_x.ptr);
  B->h1 = h;
  B->h2 = h + hn;
  B->h3 = h + 2 * hn;
  B->dz1 = dz;
  B->dz2 = dz + hn;
  B->dz3 = dz + 2 * hn;
  B->g = static_cast<float*>(c->nn_g.ptr);
  B->n = n;
  c->nn_batch = n;
  return ISO_OK;
}

// SplitMix64 (Steele, Lea & Flood 2014): the initialiser's generator.
uint64_t splitmix(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Copies n elements of `bytes` each from `src` in `memory` into a device
// buffer (or returns src itself when it already is device memory).
int stage_in(iso_ctx* c, DevBuf* buf, const void* src, size_t bytes, int32_t memory,
             const void** dev) {
  if (memory == ISO_MEMORY_DEVICE) {
    *dev = src;
    return ISO_OK;
  }
  ISO_HIP(c, ensure(buf, bytes));
  ISO_HIP(c, hipMemcpyAsync(buf->ptr, src, bytes, hipMemcpyHostToDevice, c->stream));
  *dev = buf->ptr;
  return ISO_OK;
}

}  // namespace

extern "C" {

int iso_band_brick_size(uint32_t flags) {
  return 1 << ((flags & ISO_BAND_BRICK8) ? iso::kBandBrickShiftMax : iso::kBandBrickShiftDefault);
}

int iso_band_init(iso_ctx* c, uint32_t seed) {
  if (!c) return ISO_ERR_INVALID_ARGUMENT;
  ISO_HIP(c, hipSetDevice(c->device));
  namespace m = iso::mlp;
  std::vector<float> theta(m::kParams, 0.0f);
  uint64_t s = 0x5EED0000ull ^ seed;
  auto uni = [&](double lim) {
    const double u = static_cast<double>(splitmix(&s) >> 11) * (1.0 / 9007199254740992.0);
    return static_cast<float>((2.0 * u - 1.0) * lim);
  };
  for (int i = 0; i < m::kHidden; ++i)
    for (int k = 0; k < m::kFeat; ++k)
      theta[m::kOffW1 + i * m::kFeatPad + k] = uni(std::sqrt(6.0 / m::kFeat));
  for (int i = 0; i < m::kHidden * m::kHidden; ++i) {
    theta[m::kOffW2 + i] = uni(std::sqrt(6.0 / m::kHidden));
    theta[m::kOffW3 + i] = uni(std::sqrt(6.0 / m::kHidden));
  }
  for (int i = 0; i < m::kHidden; ++i) theta[m::kOffW4 + i] = uni(std::sqrt(6.0 / m::kHidden));
  return nn_load(c, theta.data());
}

int iso_band_set_params(iso_ctx* c, const float* theta, uint64_t n) {
  if (!c || !theta || n != static_cast<uint64_t>(iso::mlp::kParams))
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "expected ISO_BAND_PARAMS floats");
  ISO_HIP(c, hipSetDevice(c->device));
  return nn_load(c, theta);
}

int iso_band_get_params(iso_ctx* c, float* theta, uint64_t n) {
  if (!c || !theta || n != static_cast<uint64_t>(iso::mlp::kParams))
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "expected ISO_BAND_PARAMS floats");
  if (!c->nn_ready) return fail(c, ISO_ERR_NO_SOURCE, "no network initialised");
  ISO_HIP(c, hipSetDevice(c->device));
  ISO_HIP(c, hipMemcpyAsync(theta, c->nn_theta.ptr, n * sizeof(float), hipMemcpyDeviceToHost,
                            c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  return ISO_OK;
}

int iso_band_train(iso_ctx* c, const float* points, const float* sdf, const float* weight,
                   uint64_t n, int32_t memory, const iso_band_train_params* prm,
                   float* last_loss) {
  if (!c || !points || !sdf || !prm || n == 0)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "null argument or no samples");
  if (!c->nn_ready) return fail(c, ISO_ERR_NO_SOURCE, "no network initialised");
  if (prm->batch == 0 || (prm->batch & 15u) || prm->steps == 0)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "batch must be a positive multiple of 16");
  if (!(prm->lr > 0.0f) || !(prm->clamp > 0.0f) || !(prm->overshoot >= 1.0f) ||
      !(prm->overshoot < 1.0e6f))
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "lr and clamp must be positive, overshoot in [1, 1e6)");
  if (memory != ISO_MEMORY_HOST && memory != ISO_MEMORY_DEVICE)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "unknown memory kind");
  ISO_HIP(c, hipSetDevice(c->device));
  int st = ISO_OK;
  iso::mlp::Model M;
  iso::mlp::Batch B;
  if ((st = nn_model(c, &M)) != ISO_OK || (st = nn_batch(c, prm->batch, &B)) != ISO_OK) return st;
  const void *dp = nullptr, *dt = nullptr, *dw = nullptr;
  if ((st = stage_in(c, &c->nn_pts, points, 3 * n * sizeof(float), memory, &dp)) != ISO_OK ||
      (st = stage_in(c, &c->nn_t, sdf, n * sizeof(float), memory, &dt)) != ISO_OK)
    return st;
  if (weight && (st = stage_in(c, &c->nn_w, weight, n * sizeof(float), memory, &dw)) != ISO_OK)
    return st;
  iso::mlp::Samples S;
  S.points = static_cast<const float*>(dp);
  S.sdf = static_cast<const float*>(dt);
  S.weight = static_cast<const float*>(dw);
  S.count = n;
  S.clamp = prm->clamp;
  S.overshoot = prm->overshoot;
  S.seed = prm->seed;
  S.gather = 1;
  ISO_HIP(c, iso::mlp::launch_zero_grad(M, c->stream));
  for (uint32_t i = 0; i < prm->steps; ++i) {
    const double frac = prm->steps > 1 ? static_cast<double>(i) / (prm->steps - 1) : 0.0;
    const float lr = static_cast<float>(prm->lr_min + 0.5 * (prm->lr - prm->lr_min) *
                                                          (1.0 + std::cos(3.14159265358979 * frac)));
    ISO_HIP(c, timed(c, "band_train_batch", [&] {
              return iso::mlp::launch_train_batch(M, B, S, c->nn_step, c->stream);
            }));
    if (i + 1 == prm->steps && last_loss) {
      ISO_HIP(c, hipMemcpyAsync(c->h_stage, M.loss, sizeof(float), hipMemcpyDeviceToHost,
                                c->stream));
    }
    ++c->nn_step;
    ISO_HIP(c, timed(c, "band_adam", [&] {
              return iso::mlp::launch_adam(M, c->nn_step, lr, 0.9f, 0.999f, 1e-8f, c->stream);
            }));
  }
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  if (last_loss) {
    float v;
    std::memcpy(&v, c->h_stage, sizeof(float));
    *last_loss = v / (2.0f * static_cast<float>(prm->batch));
  }
  return ISO_OK;
}

int iso_band_gradient(iso_ctx* c, const float* points, const float* sdf, const float* weight,
                      uint64_t n, float clamp, float overshoot, float* grad, float* loss) {
  if (!c || !points || !sdf || !grad || n == 0 || (n & 15u))
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "n must be a positive multiple of 16");
  if (!(clamp > 0.0f) || !(overshoot >= 1.0f) || !(overshoot < 1.0e6f))
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "clamp must be positive, overshoot in [1, 1e6)");
  if (!c->nn_ready) return fail(c, ISO_ERR_NO_SOURCE, "no network initialised");
  ISO_HIP(c, hipSetDevice(c->device));
  int st = ISO_OK;
  iso::mlp::Model M;
  iso::mlp::Batch B;
  if ((st = nn_model(c, &M)) != ISO_OK || (st = nn_batch(c, static_cast<int64_t>(n), &B)) != ISO_OK)
    return st;
  const void *dp = nullptr, *dt = nullptr, *dw = nullptr;
  if ((st = stage_in(c, &c->nn_pts, points, 3 * n * sizeof(float), ISO_MEMORY_HOST, &dp)) != ISO_OK ||
      (st = stage_in(c, &c->nn_t, sdf, n * sizeof(float), ISO_MEMORY_HOST, &dt)) != ISO_OK)
    return st;
  if (weight && (st = stage_in(c, &c->nn_w, weight, n * sizeof(float), ISO_MEMORY_HOST, &dw)) != ISO_OK)
    return st;
  iso::mlp::Samples S;
  S.points = static_cast<const float*>(dp);
  S.sdf = static_cast<const float*>(dt);
  S.weight = static_cast<const float*>(dw);
  S.count = n;
  S.clamp = clamp;
  S.overshoot = overshoot;
  S.seed = 0;
  S.gather = 0;
  ISO_HIP(c, iso::mlp::launch_zero_grad(M, c->stream));
  ISO_HIP(c, iso::mlp::launch_train_batch(M, B, S, 0, c->stream));
  ISO_HIP(c, hipMemcpyAsync(grad, M.grad, iso::mlp::kParams * sizeof(float), hipMemcpyDeviceToHost,
                            c->stream));
  ISO_HIP(c, hipMemcpyAsync(c->h_stage, M.loss, sizeof(float), hipMemcpyDeviceToHost, c->stream));
  ISO_HIP(c, iso::mlp::launch_zero_grad(M, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  if (loss) {
    float v;
    std::memcpy(&v, c->h_stage, sizeof(float));
    *loss = v / (2.0f * static_cast<float>(n));
  }
  return ISO_OK;
}

int iso_band_predict(iso_ctx* c, const float* points, uint64_t n, float* out, int32_t memory) {
  if (!c || (n && (!points || !out))) return fail(c, ISO_ERR_INVALID_ARGUMENT, "null argument");
  if (!c->nn_ready) return fail(c, ISO_ERR_NO_SOURCE, "no network initialised");
  if (memory != ISO_MEMORY_HOST && memory != ISO_MEMORY_DEVICE)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "unknown memory kind");
  if (n == 0) return ISO_OK;
  ISO_HIP(c, hipSetDevice(c->device));
  int st = ISO_OK;
  iso::mlp::Model M;
  if ((st = nn_model(c, &M)) != ISO_OK) return st;
  const void* dp = nullptr;
  if ((st = stage_in(c, &c->nn_pts, points, 3 * n * sizeof(float), memory, &dp)) != ISO_OK) return st;
  float* dst = out;
  if (memory == ISO_MEMORY_HOST) {
    ISO_HIP(c, ensure(&c->nn_out, n * sizeof(float)));
    dst = static_cast<float*>(c->nn_out.ptr);
  }
  ISO_HIP(c, timed(c, "band_predict", [&] {
            return iso::mlp::launch_forward(M, static_cast<const float*>(dp),
                                            static_cast<int64_t>(n), dst, c->stream);
          }));
  if (memory == ISO_MEMORY_HOST)
    ISO_HIP(c, hipMemcpyAsync(out, dst, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  return ISO_OK;
}

int iso_extract_mc_band(iso_ctx* c, const iso_grid_desc* desc, float margin, uint32_t flags,
                        iso_mesh_out* out, iso_band_stats* stats) {
  if (!c || !out) return fail(c, ISO_ERR_INVALID_ARGUMENT, "null argument");
  iso::Lattice L;
  int st = check_lattice(c, desc, &L);
  if (st != ISO_OK) return st;
  if (desc->source != ISO_SOURCE_PROGRAM)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "band extraction needs a program source");
  if (c->n_prog == 0) return fail(c, ISO_ERR_NO_SOURCE, "no program uploaded");
  if (flags & ~(ISO_BAND_CERTIFY | ISO_BAND_NO_NETWORK | ISO_BAND_BRICK8))
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "unknown band flags");
  const bool use_network = !(flags & ISO_BAND_NO_NETWORK);
  const bool certify = (flags & (ISO_BAND_CERTIFY | ISO_BAND_NO_NETWORK)) != 0;
  if (use_network && !c->nn_ready) return fail(c, ISO_ERR_NO_SOURCE, "no network initialised");
  if (!(margin >= 0.0f) || !(margin < 3.0e38f))
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "margin must be finite and non-negative");
  if (out->memory != ISO_MEMORY_HOST && out->memory != ISO_MEMORY_DEVICE)
    return fail(c, ISO_ERR_INVALID_ARGUMENT, "unknown output memory kind");
  c->kept = false;
  ISO_HIP(c, hipSetDevice(c->device));
  iso::mlp::Model M{};
  if (use_network && (st = nn_model(c, &M)) != ISO_OK) return st;
  const iso::BrickGrid G = iso::brick_grid(
      L, (flags & ISO_BAND_BRICK8) ? iso::kBandBrickShiftMax : iso::kBandBrickShiftDefault);
  const uint64_t nb = static_cast<uint64_t>(G.n);
  ISO_HIP(c, ensure(&c->br_centres, 3 * nb * sizeof(float)));
  ISO_HIP(c, ensure(&c->br_hd, nb * sizeof(float)));
  ISO_HIP(c, ensure(&c->br_pred, nb * sizeof(float)));
  ISO_HIP(c, ensure(&c->br_flags, nb));
  ISO_HIP(c, ensure(&c->br_list, nb * sizeof(uint32_t)));
  ISO_HIP(c, ensure(&c->br_count, 4 * sizeof(unsigned long long)));
  ISO_HIP(c, ensure(&c->br_coarse, static_cast<size_t>(G.nc)));
  ISO_HIP(c, ensure(&c->br_proof, nb));
  float* centres = static_cast<float*>(c->br_centres.ptr);
  float* hd = static_cast<float*>(c->br_hd.ptr);
  float* pred = static_cast<float*>(c->br_pred.ptr);
  uint8_t* bflags = static_cast<uint8_t*>(c->br_flags.ptr);
  uint32_t* list = static_cast<uint32_t*>(c->br_list.ptr);
  uint8_t* coarse = static_cast<uint8_t*>(c->br_coarse.ptr);
  uint8_t* proof = static_cast<uint8_t*>(c->br_proof.ptr);
  // counters: exact samples, (the dense check's), centre samples, overturned culls
  unsigned long long* counters = static_cast<unsigned long long*>(c->br_count.ptr);
  iso::scan::State S{};
  if ((st = prepare_scan(c, nb, &S)) != ISO_OK) return st;
  ISO_HIP(c, iso::launch_fill_u32(reinterpret_cast<uint32_t*>(S.err), 2, 0u, c->stream));
  ISO_HIP(c, iso::launch_fill_u32(reinterpret_cast<uint32_t*>(counters), 8, 0u, c->stream));
  const iso::DevOp* ops = static_cast<const iso::DevOp*>(c->prog.ptr);
  ISO_HIP(c, timed(c, "band_bricks", [&] {
            hipError_t e = iso::launch_brick_centres(L, G, centres, hd, c->stream);
            if (e == hipSuccess && use_network)
              e = iso::mlp::launch_forward(M, centres, static_cast<int64_t>(nb), pred, c->stream);
            if (e == hipSuccess && certify)
              e = iso::launch_coarse_certify(ops, c->n_prog, L, G, coarse, counters + 2,
                                             c->stream);
            if (e == hipSuccess && certify)
              e = iso::launch_brick_certify(ops, c->n_prog, L, G, coarse, centres, hd, proof,
                                            counters + 2, c->stream);
            if (e == hipSuccess)
              e = iso::launch_brick_flags(G, use_network ? pred : nullptr, hd,
                                          certify ? proof : nullptr, margin, bflags, list,
                                          counters + 3, S, c->stream);
            return e;
          }));
  if ((st = read_ctl(c, S, "band brick selection")) != ISO_OK) return st;
  const uint64_t n_band = c->h_stage[1];
  const iso::McDims D = iso::mc_dims(L);
  SignsStage signs = [&](const iso::McBuffers& B, float* values) -> int {
    ISO_HIP(c, timed(c, "band_background", [&] {
              return iso::launch_band_background(L, D, G, bflags, B, c->stream);
            }));
    // The stage reruns when the active-cell list grows: count the exact
    // samples of this run only.
    ISO_HIP(c, iso::launch_fill_u32(reinterpret_cast<uint32_t*>(counters), 2, 0u, c->stream));
    ISO_HIP(c, timed(c, "band_exact", [&] {
              return iso::launch_band_exact(ops, c->n_prog, L, G, bflags, list, n_band, B, values,
                                            counters, c->stream);
            }));
    return ISO_OK;
  };
  c->band_valid = false;
  const BandEmit band{bflags, G.nbx, G.nby, G.shift};
  st = extract_lattice(c, L, nullptr, nullptr, out, &signs, &band);
  if (st != ISO_OK && st != ISO_ERR_CAPACITY) return st;
  c->band_desc = *desc;
  c->band_valid = true;
  if (stats) {
    ISO_HIP(c, hipMemcpyAsync(c->h_stage + 4, counters, 4 * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, c->stream));
    ISO_HIP(c, hipStreamSynchronize(c->stream));
    stats->bricks = nb;
    stats->band_bricks = n_band;
    stats->exact_samples = c->h_stage[4];
    stats->centre_samples = c->h_stage[6];
    stats->overturned_culls = c->h_stage[7];
  }
  return st;
}

int iso_band_missed_cells(iso_ctx* c, const iso_grid_desc* desc, uint64_t* missed) {
  if (!c || !desc || !missed) return fail(c, ISO_ERR_INVALID_ARGUMENT, "null argument");
  if (!c->band_valid || std::memcmp(&c->band_desc, desc, sizeof(*desc)) != 0)
    return fail(c, ISO_ERR_NO_SOURCE, "no band extraction over this lattice");
  iso::Lattice L;
  int st = check_lattice(c, desc, &L);
  if (st != ISO_OK) return st;
  ISO_HIP(c, hipSetDevice(c->device));
  const iso::McDims D = iso::mc_dims(L);
  // The band extraction's sign words are still in the context (any other
  // extraction clears band_valid); the dense ones go to a buffer of their own.
  iso::McBuffers B{};
  if ((st = prepare_mc(c, D, 0, &B)) != ISO_OK) return st;
  ISO_HIP(c, ensure(&c->dense_check, (D.n_sign_words + 3) * sizeof(uint64_t)));
  iso::McBuffers J = B;
  J.signs = static_cast<uint64_t*>(c->dense_check.ptr) + 1;
  const iso::DevOp* ops = static_cast<const iso::DevOp*>(c->prog.ptr);
  ISO_HIP(c, iso::launch_signs_program(ops, c->n_prog, L, D, J, nullptr, c->stream));
  unsigned long long* counters = static_cast<unsigned long long*>(c->br_count.ptr);
  ISO_HIP(c, iso::launch_fill_u32(reinterpret_cast<uint32_t*>(counters + 1), 2, 0u, c->stream));
  ISO_HIP(c, iso::launch_band_diff(L, D, B.signs, J.signs, counters + 1, c->stream));
  ISO_HIP(c, hipMemcpyAsync(c->h_stage + 4, counters + 1, sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  *missed = c->h_stage[4];
  return ISO_OK;
}

int iso_mesh_fetch(iso_ctx* c, iso_mesh_out* out) {
  if (!c || !out) return fail(c, ISO_ERR_INVALID_ARGUMENT, "null argument");
  if (!c->kept) return fail(c, ISO_ERR_NO_SOURCE, "no kept mesh");
  const uint64_t nv = c->kept_vertices, ni = c->kept_indices;
  out->vertex_count = nv;
  out->index_count = ni;
  if ((nv && (!out->vertices || out->vertex_capacity < nv)) ||
      (ni && (!out->indices || out->index_capacity < ni)))
    return fail(c, ISO_ERR_CAPACITY, "output buffers too small");
  const hipMemcpyKind kind =
      out->memory == ISO_MEMORY_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  ISO_HIP(c, hipSetDevice(c->device));
  if (nv)
    ISO_HIP(c, hipMemcpyAsync(out->vertices, c->vout.ptr, nv * 3 * sizeof(float), kind, c->stream));
  if (ni)
    ISO_HIP(c, hipMemcpyAsync(out->indices, c->iout.ptr, ni * sizeof(uint32_t), kind, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  c->kept = false;
  return ISO_OK;
}

int iso_device_alloc(iso_ctx* c, uint64_t bytes, void** out) {
  if (!c || !out) return ISO_ERR_INVALID_ARGUMENT;
  ISO_HIP(c, hipSetDevice(c->device));
  ISO_HIP(c, hipMalloc(out, bytes ? bytes : 1));
  return ISO_OK;
}

int iso_device_free(iso_ctx* c, void* p) {
  if (!c) return ISO_ERR_INVALID_ARGUMENT;
  ISO_HIP(c, hipSetDevice(c->device));
  if (p) ISO_HIP(c, hipFree(p));
  return ISO_OK;
}

int iso_memcpy_to_device(iso_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!c || (bytes && (!dst || !src))) return ISO_ERR_INVALID_ARGUMENT;
  ISO_HIP(c, hipSetDevice(c->device));
  ISO_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  return ISO_OK;
}

int iso_memcpy_to_host(iso_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!c || (bytes && (!dst || !src))) return ISO_ERR_INVALID_ARGUMENT;
  ISO_HIP(c, hipSetDevice(c->device));
  ISO_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  return ISO_OK;
}

int iso_memcpy_device(iso_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!c || (bytes && (!dst || !src))) return ISO_ERR_INVALID_ARGUMENT;
  ISO_HIP(c, hipSetDevice(c->device));
  ISO_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
  ISO_HIP(c, hipStreamSynchronize(c->stream));
  return ISO_OK;
}

int iso_ctx_set_timing(iso_ctx* c, int enable) {
  if (!c) return ISO_ERR_INVALID_ARGUMENT;
  c->timing = enable != 0;
  return ISO_OK;
}

int iso_ctx_kernel_stats(iso_ctx* c, iso_kernel_stat* out, int max, int* n_out) {
  if (!c || !n_out || (max > 0 && !out)) return ISO_ERR_INVALID_ARGUMENT;
  *n_out = static_cast<int>(c->stats.size());
  int i = 0;
  for (const auto& kv : c->stats) {
    if (i >= max) break;
    std::memset(&out[i], 0, sizeof(out[i]));
    std::snprintf(out[i].name, sizeof(out[i].name), "%s", kv.first.c_str());
    out[i].launches = kv.second.launches;
    out[i].total_ms = kv.second.total_ms;
    ++i;
  }
  return ISO_OK;
}

int iso_ctx_reset_kernel_stats(iso_ctx* c) {
  if (!c) return ISO_ERR_INVALID_ARGUMENT;
  c->stats.clear();
  return ISO_OK;
}

}  // extern "C"
