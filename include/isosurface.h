/*
 * isosurface_amd — C ABI of the MI355X (gfx950) isosurface extraction engine.
 *
 * The reference crate (swiftcoder/isosurface, Rust) has no FFI of its own: its
 * callers use generic traits.  This header is the boundary a Rust caller would
 * bind with `extern "C"` to replace the crate's CPU hot path:
 *
 *   MarchingCubes::<Signed>::new(size)            marching_cubes.rs:46-50
 *   MarchingCubes::extract(&source, &mut extractor) marching_cubes.rs:59-82
 *   ScalarSource::sample_scalar                   source.rs:21-28
 *   Extractor::extract_vertex / extract_index     extractor.rs:17-20
 *
 * Conventions
 *   - Every entry point returns an iso_status; nothing throws across the ABI.
 *   - One context per host thread / stream.  A context owns its device
 *     scratch buffers and reuses them across calls, as the reference's
 *     algorithm objects reuse their layer buffers (primal_grid.rs:18-33).
 *   - Sources: the crate's built-in implicit shapes and CSG operators
 *     (src/implicit/) are described by an iso_node program evaluated on the
 *     GPU; any other source is sampled by the caller on the host into a dense
 *     FP32 grid and uploaded (iso_grid_upload / iso_grid_bind_device).
 *   - Output matches the crate's IndexedVertices sink (extractor.rs:72-94):
 *     xyz f32 per vertex, u32 indices, three per triangle, in the order the
 *     reference emits them (first-encounter vertex order, mesh.rs:240-251;
 *     triangles in cell order, then table order).
 */
#ifndef ISOSURFACE_AMD_H
#define ISOSURFACE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ISO_ABI_VERSION 1

typedef enum iso_status {
  ISO_OK = 0,
  ISO_ERR_INVALID_ARGUMENT = 1,
  /* Output buffers too small.  vertex_count / index_count hold the sizes
   * the mesh needs; nothing beyond the capacities was written. */
  ISO_ERR_CAPACITY = 2,
  ISO_ERR_DEVICE = 3,        /* HIP runtime or kernel launch failure */
  ISO_ERR_OUT_OF_MEMORY = 4,
  ISO_ERR_PROGRAM = 5,       /* malformed iso_node program */
  ISO_ERR_NO_DEVICE = 6,
  ISO_ERR_NO_SOURCE = 7,     /* no program / grid bound for the requested source */
  ISO_ERR_OVERFLOW = 8       /* mesh exceeds the u32 index range */
} iso_status;

typedef struct iso_ctx iso_ctx;

/* ---- context ------------------------------------------------------------ */

/* Creates a context on HIP device `device`.  `stream` is a hipStream_t to
 * launch on, or NULL for a stream owned by the context. */
int iso_ctx_create(int device, void* stream, iso_ctx** out);
int iso_ctx_destroy(iso_ctx* ctx);
int iso_ctx_synchronize(iso_ctx* ctx);
/* Human-readable detail of the last failure on this context. */
const char* iso_ctx_last_error(const iso_ctx* ctx);
const char* iso_status_string(int status);
int iso_abi_version(void);

/* ---- sources: expression programs -------------------------------------- */

/* Program opcodes.  A program is a postfix (post-order) list of nodes that
 * leaves exactly one value, mirroring the crate's source types:
 *   SPHERE       p[0]=radius                      implicit/sphere.rs:35-39
 *   TORUS        p[0]=radius, p[1]=tube_radius    implicit/torus.rs:40-46
 *   BOX          p[0..2]=half_extent              implicit/rectangular_prism.rs:36-41
 *   CYLINDER     p[0]=radius, p[1]=half_length    implicit/cylinder.rs:41-48
 *   UNION        min(a, b)                        implicit/csg.rs:35-39
 *   INTERSECTION max(a, b)                        implicit/csg.rs:68-72
 *   DIFFERENCE   max(b, -a)  (b minus a)          implicit/csg.rs:96-100
 *   TRANSLATE    p[0..2]=offset: the nodes up to the matching END_TRANSLATE
 *                are sampled at (p - offset), per component, as the examples'
 *                DemoSource does (examples/common/sources.rs:38-43)
 *   SMOOTH_UNION p[0]=k: polynomial smooth minimum of a and b (no reference
 *                counterpart; used by the learned narrow-band workload)
 * For binary operators, `a` is the earlier operand and `b` the later one,
 * so Difference::new(a, b) is "a-nodes b-nodes DIFFERENCE". */
enum {
  ISO_OP_SPHERE = 1,
  ISO_OP_TORUS = 2,
  ISO_OP_BOX = 3,
  ISO_OP_CYLINDER = 4,
  ISO_OP_UNION = 16,
  ISO_OP_INTERSECTION = 17,
  ISO_OP_DIFFERENCE = 18,
  ISO_OP_SMOOTH_UNION = 19,
  ISO_OP_TRANSLATE = 32,
  ISO_OP_END_TRANSLATE = 33
};

typedef struct iso_node {
  uint32_t op;
  uint32_t reserved; /* must be 0 */
  float p[6];
} iso_node;

/* Checks that `nodes` is a well-formed program the device can evaluate
 * (stack depths within limits) without touching a device.  On failure the
 * reason is written to `message` (NUL-terminated, at most `message_len`). */
int iso_program_validate(const iso_node* nodes, uint32_t n_nodes, char* message,
                         size_t message_len);

/* Validates `nodes` and makes it the context's current program. */
int iso_program_upload(iso_ctx* ctx, const iso_node* nodes, uint32_t n_nodes);

/* Samples the current program at n points (xyz interleaved, host memory):
 * ScalarSource::sample_scalar evaluated on the GPU. */
int iso_program_sample(iso_ctx* ctx, const float* points, uint64_t n, float* values);
/* CentralDifference::new_with_epsilon(program, epsilon).sample_normal at n
 * points (3 floats each) in `memory` (source.rs:82-94): per axis
 * (f(p + e) - f(p - e)) / (2 epsilon), not normalised — what the normal
 * sinks IndexedInterleavedNormals / OnlyInterleavedNormals interleave with
 * each vertex (extractor.rs:46-69,97-127). */
int iso_sample_normals(iso_ctx* ctx, const float* points, uint64_t n, float epsilon,
                       float* normals, int32_t memory);
/* VectorSource::sample_vector of the current program at n points (3 floats
 * in and 3 out per point, host memory; source.rs:32-38 with the shapes'
 * implementations, implicit/sphere.rs:41-55, torus.rs:48-91,
 * cylinder.rs:50-71, rectangular_prism.rs:43-73, csg.rs:41-106): the signed
 * distance along each axis, f32::MAX where that axis line misses the shape.
 * ISO_ERR_PROGRAM if the program has a smooth union (no directed form). */
int iso_program_sample_vector(iso_ctx* ctx, const float* points, uint64_t n, float* values);

/* ---- sources: dense sample grids ---------------------------------------- */

/* Copies a dense FP32 grid (x fastest, then y, then z) from host memory into
 * the context's device buffer and makes it the current grid. */
int iso_grid_upload(iso_ctx* ctx, const float* values, int64_t nx, int64_t ny, int64_t nz);
/* Uses caller-owned device memory as the current grid (no copy). */
int iso_grid_bind_device(iso_ctx* ctx, const float* device_values, int64_t nx, int64_t ny,
                         int64_t nz);
/* Copies a dense grid of directed distances (3 floats per point — the
 * PrimalGrid<Directed> values a VectorSource gives, x fastest) from host
 * memory; iso_extract_mc_directed of a grid source reads it. */
int iso_grid_upload_directed(iso_ctx* ctx, const float* xyz, int64_t nx, int64_t ny, int64_t nz);

/* ---- extraction ----------------------------------------------------------- */

enum { ISO_SOURCE_PROGRAM = 0, ISO_SOURCE_GRID = 1 };
enum { ISO_MEMORY_HOST = 0, ISO_MEMORY_DEVICE = 1 };

/* Sample lattice.  Sample (i, j, k) sits at
 *   (origin[0] + i*spacing[0], origin[1] + j*spacing[1], origin[2] + k*spacing[2])
 * in f32, and cells are the (nx-1)(ny-1)(nz-1) cubes between samples,
 * visited z-major (z, then y, then x) as PrimalGrid::traverse does
 * (traversal/primal_grid.rs:59-85). */
typedef struct iso_grid_desc {
  int64_t nx, ny, nz;
  float origin[3];
  float spacing[3];
  int32_t source; /* ISO_SOURCE_* */
  int32_t flags;  /* reserved, 0 */
} iso_grid_desc;

/* Fills the lattice that MarchingCubes::new(size) traverses: size x size x
 * (size+1) samples at i * (1/(size-1)) — including the reference's extra
 * z layer (primal_grid.rs:44-69). */
int iso_grid_desc_reference(int64_t size, int32_t source, iso_grid_desc* out);

/* Samples the current program over a lattice into a dense FP32 grid (x
 * fastest), in host or device memory (`memory` = ISO_MEMORY_*).  The values
 * are bit-identical to the ones extraction classifies. */
int iso_program_sample_lattice(iso_ctx* ctx, const iso_grid_desc* desc, float* out,
                               int32_t memory);

typedef struct iso_mesh_out {
  float* vertices;          /* 3 floats per vertex */
  uint32_t* indices;        /* 3 per triangle */
  uint64_t vertex_capacity; /* in vertices */
  uint64_t index_capacity;  /* in indices */
  uint64_t vertex_count;    /* out */
  uint64_t index_count;     /* out */
  int32_t memory;           /* ISO_MEMORY_* of vertices / indices */
  int32_t flags;            /* ISO_MESH_* (0: none) */
} iso_mesh_out;

/* iso_mesh_out flags.  ISO_MESH_GRAPH (iso_extract_mc into device memory):
 * the extraction's launches are captured into a HIP graph and replayed by
 * later calls that repeat the lattice, the source (the same grid binding, or
 * the same uploaded program) and the output buffers — one launch instead of
 * three for callers that extract repeatedly into the same buffers.  The
 * graph reads the grid's and the program's memory when it runs, so new
 * samples in the same grid buffer are extracted correctly.  Ignored for
 * host memory, while kernel timing is on, and for a context on the null
 * stream (which cannot be captured).  Other entry points ignore the field. */
#define ISO_MESH_GRAPH 1
/* ISO_MESH_SINGLE_PASS (iso_extract_mc of a grid source): extract in one
 * launch that reads every sample once (per-tile classify, count, look-back
 * scan and emit) instead of the separate classify, count and emit launches;
 * the same output.  Ignored with ISO_MESH_GRAPH and for program sources. */
#define ISO_MESH_SINGLE_PASS 2

/* Marching cubes over the lattice: replaces MarchingCubes::extract with an
 * IndexedVertices extractor (marching_cubes.rs:59-82, extractor.rs:72-94).
 * Vertex order, positions and index lists are the reference's. */
int iso_extract_mc(iso_ctx* ctx, const iso_grid_desc* desc, iso_mesh_out* out);

/* iso_extract_mc of a bound grid (ISO_SOURCE_GRID) into device memory,
 * by the single pass, that allocates nothing: no hipMalloc, hipFree or other
 * allocating call.  No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  Its scratch is one static device array
 * per GPU (32.5 MiB, held until the process ends), so the lattice is
 * limited to 28 MiB of sign words (about 513^3 samples), 65,536 tiles and
 * 262,144 cell rows; larger ones give ISO_ERR_INVALID_ARGUMENT.  Calls on
 * one GPU share the array: consecutive calls are independent, calls that
 * overlap in time on one GPU (other threads, other streams) race.  Same
 * output, counts and ISO_ERR_CAPACITY rule as iso_extract_mc with device
 * outputs; out->flags is ignored. */
int iso_extract_mc_noalloc(iso_ctx* ctx, const iso_grid_desc* desc, iso_mesh_out* out);

/* MarchingCubes::<Directed>::extract with IndexedVertices
 * (marching_cubes.rs:59-82 over distance.rs:72-103): a lattice point is
 * inside when no component of its directed distance is positive, and each
 * crossing interpolates the component along the edge's axis
 * ((p_a - p_b).abs().max_component_index()).  Values from the current
 * program's sample_vector (ISO_SOURCE_PROGRAM; ISO_ERR_PROGRAM for a smooth
 * union) or from iso_grid_upload_directed (ISO_SOURCE_GRID).  Same outputs,
 * capacity and fetch rules as iso_extract_mc (ISO_MESH_GRAPH is ignored). */
int iso_extract_mc_directed(iso_ctx* ctx, const iso_grid_desc* desc, iso_mesh_out* out);

/* ExtendedMarchingCubes::new(size).extract with IndexedVertices
 * (extended_marching_cubes.rs:59-186) over a program lattice: Directed
 * values from the program's sample_vector; the normals at the edge
 * crossings from its HermiteSource (epsilon == 0: shapes, unions and
 * translations, implicit/<shape>.rs sample_normal, csg.rs:47-51) or from
 * CentralDifference::new_with_epsilon(program, epsilon) (epsilon > 0,
 * source.rs:52-94).  A cell whose tangent planes show a feature gets a QEF
 * feature vertex with triangle fans over its edge loops, and
 * flip_feature_edges rotates the fan edges between feature vertices, all as
 * the reference does (vertex order, f32 bits and index lists).  Program
 * sources only (ISO_ERR_INVALID_ARGUMENT for a grid); same outputs, capacity
 * and fetch rules as iso_extract_mc. */
int iso_extract_emc(iso_ctx* ctx, const iso_grid_desc* desc, float epsilon, iso_mesh_out* out);

/* PointCloud::<Signed>::new(size).extract(source, extractor)
 * (point_cloud.rs:50-63): one point per mixed cell (cube index neither 0
 * nor 255) in the lattice's cell order, corners[0].lerp(corners[6], 0.5);
 * vertices only (index_count = 0).  Same sources, capacity and fetch rules
 * as iso_extract_mc. */
int iso_extract_point_cloud(iso_ctx* ctx, const iso_grid_desc* desc, iso_mesh_out* out);

/* Dual contouring with MinimiseQEF feature placement and CentralDifference
 * normals of the current program (step `epsilon`, the reference's default is
 * 1e-6): replaces DualContouring::new(size, MinimiseQEF {}).extract over a
 * CentralDifference::new_with_epsilon source with an IndexedVertices
 * extractor (dual_contouring.rs:65-105, source.rs:52-94).  `desc` is the
 * primal lattice and must name ISO_SOURCE_PROGRAM; the mesh is the
 * reference's: vertex order, positions (feature points from the reference's
 * f64 SVD) and index lists.  ISO_ERR_DEVICE if a cell's SVD fails to
 * converge within a fixed bound of QR sweeps (the reference loops on).
 * epsilon 0: the program's own HermiteSource normals instead (the reference's
 * DualContouring over a source that implements HermiteSource itself,
 * dual_contouring.rs:65-69, source.rs:43-48): shapes, unions and
 * translations only, else ISO_ERR_PROGRAM. */
int iso_extract_dc(iso_ctx* ctx, const iso_grid_desc* desc, float epsilon, iso_mesh_out* out);

/* Marching cubes over a linear-hashed octree of the current program:
 * replaces LinearHashedMarchingCubes::new(max_depth).extract with an
 * IndexedVertices extractor (linear_hashed_marching_cubes.rs:51-84).  The
 * domain is the unit cube, as the reference's Morton centres.  The mesh is
 * the reference's for the primal-vertex visiting order "ascending Morton
 * key" (the reference visits them in std HashMap order, which differs from
 * run to run); Morton::level()'s truncated-constant quirk is reproduced.
 * max_depth in [0, 20] (nodes reach depth max_depth + 1, whose Morton keys
 * use 64 bits at 20). */
int iso_extract_lhmc(iso_ctx* ctx, int32_t max_depth, iso_mesh_out* out);

/* ---- batched chunks ---------------------------------------------------------
 * Many independent chunks (each its own lattice and program, e.g. the chunks
 * of a paged or LOD terrain) extracted in one pass: one sign pass, one count
 * scan and one emit pass over all chunks.  No reference counterpart as a
 * single call: each chunk's mesh is exactly what iso_extract_mc gives for
 * that chunk's lattice and program (MarchingCubes::extract with an
 * IndexedVertices extractor per chunk); the batch only shares the launches. */
typedef struct iso_chunk_desc {
  iso_grid_desc grid;   /* the chunk's lattice; grid.source must be ISO_SOURCE_PROGRAM */
  uint32_t node_offset; /* the chunk's program: nodes[node_offset, node_offset + node_count) */
  uint32_t node_count;
} iso_chunk_desc;

/* Validates and lowers every chunk's program and lays the batch out on the
 * device; replaces any previous batch of the context. */
int iso_batch_upload(iso_ctx* ctx, const iso_node* nodes, uint64_t n_nodes,
                     const iso_chunk_desc* chunks, uint32_t n_chunks);

#define ISO_BATCH_GRAPH 1u    /* capture the extraction in a HIP graph, replay it on later calls */
#define ISO_BATCH_AUTOTUNE 2u /* on the first extraction of a batch, time the launch shapes
                                 (count tile, sign words per wavefront, emit residency) and
                                 keep the fastest */

typedef struct iso_batch_out {
  float* vertices;           /* all chunks' vertices, chunk after chunk */
  uint32_t* indices;         /* all chunks' index lists, chunk after chunk, chunk-local ids */
  uint64_t vertex_capacity;  /* in vertices */
  uint64_t index_capacity;   /* in indices */
  uint64_t* vertex_offsets;  /* host, n_chunks + 1 (out): chunk c owns [off[c], off[c+1]) */
  uint64_t* index_offsets;   /* host, n_chunks + 1 (out) */
  uint64_t vertex_count;     /* out */
  uint64_t index_count;      /* out */
  int32_t memory;            /* ISO_MEMORY_* of vertices / indices */
  uint32_t flags;            /* ISO_BATCH_* */
} iso_batch_out;

/* Marching cubes over every chunk of the uploaded batch.  ISO_ERR_CAPACITY
 * when the outputs are too small (counts and offsets are still written;
 * host-memory results are kept for iso_mesh_fetch). */
int iso_extract_batch(iso_ctx* ctx, iso_batch_out* out);

/* ---- learned narrow band ----------------------------------------------------
 * No reference counterpart: the library's own acceleration of dense program
 * extraction.  A Fourier-feature MLP (99 features: p, sin and cos of
 * pi 2^(k/2) p for k < 16; three hidden layers of 64 ReLU units; bf16 MFMA
 * with f32 accumulation) is trained on (point, SDF) samples to predict the
 * truncated SDF.  iso_extract_mc_band predicts it at the centre of every
 * brick (4x4x4 cells, or 8x8x8 with ISO_BAND_BRICK8), evaluates the program exactly only in bricks whose
 * prediction lies within their half-diagonal plus `margin` of zero, gives
 * every other sample its brick's predicted sign, and extracts as
 * iso_extract_mc does.  The result equals iso_extract_mc's whenever no
 * surface cell lies in a culled brick (iso_band_missed_cells counts them),
 * which the ISO_BAND_CERTIFY and ISO_BAND_NO_NETWORK proofs guarantee. */
#define ISO_BAND_PARAMS 15617 /* length of the parameter vector */

typedef struct iso_band_train_params {
  uint32_t steps;   /* Adam steps */
  uint32_t batch;   /* samples per step, a multiple of 16 */
  float lr;         /* learning rate, cosine-decayed to lr_min over the steps */
  float lr_min;
  float clamp;      /* targets are clamp(sdf, -clamp, clamp) */
  uint32_t seed;    /* batch sampling */
  float overshoot;  /* loss weight (>= 1) of the part of an error that takes the
                       prediction beyond its target, away from zero, or across
                       zero: 1 is plain least squares, larger values train a
                       prediction that errs toward the band */
} iso_band_train_params;

typedef struct iso_band_stats {
  uint64_t bricks;            /* bricks of the lattice */
  uint64_t band_bricks;       /* bricks evaluated exactly */
  uint64_t exact_samples;     /* program evaluations of the band (without emit's) */
  uint64_t centre_samples;    /* program evaluations of the certificate
                                 (super-brick and brick centres) */
  uint64_t overturned_culls;  /* ISO_BAND_CERTIFY: bricks the network culled
                                 that no proof allowed (kept in the band) */
} iso_band_stats;

/* iso_extract_mc_band flags.  Without flags the network's culls are final
 * (a calibrated margin, not a proof: cells can be missed, which
 * iso_band_missed_cells reports).  ISO_BAND_CERTIFY: a brick leaves the band
 * only by a proof, whatever the network predicted: the program is evaluated
 * at the centre of each super-brick (4^3 bricks), then at the centre of each
 * brick no super-brick proves, and |f| above the box's half-diagonal proves
 * the box's sign (every program is 1-Lipschitz); network culls the proofs
 * overturn are counted.  No surface cell is missed.  ISO_BAND_NO_NETWORK:
 * the same proofs without a network (the method without it, for
 * comparison).  ISO_BAND_BRICK8: bricks of 8x8x8 cells instead of 4x4x4
 * (an eighth of the predictions and proofs, a coarser band). */
#define ISO_BAND_CERTIFY 1u
#define ISO_BAND_NO_NETWORK 2u
#define ISO_BAND_BRICK8 4u

/* Cells per edge of a narrow-band brick (the unit the band culls) under
 * iso_extract_mc_band's `flags`: 4, or 8 with ISO_BAND_BRICK8. */
int iso_band_brick_size(uint32_t flags);

/* Deterministic He-uniform initialisation of the context's network. */
int iso_band_init(iso_ctx* ctx, uint32_t seed);
/* Parameter vector (ISO_BAND_PARAMS floats, host memory) in or out:
 * W1[64][112] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] w4[64] b4. */
int iso_band_set_params(iso_ctx* ctx, const float* theta, uint64_t n);
int iso_band_get_params(iso_ctx* ctx, float* theta, uint64_t n);
/* Trains with Adam on n samples (points: 3 floats each; weight may be null)
 * in `memory`.  Per sample, with r = f - t and e the distance from f to the
 * interval between 0 and t (signed like r), the loss is
 * w (r^2 + (overshoot - 1) e^2) / 2, averaged over the batch; *last_loss
 * receives it for the last batch. */
int iso_band_train(iso_ctx* ctx, const float* points, const float* sdf, const float* weight,
                   uint64_t n, int32_t memory, const iso_band_train_params* params,
                   float* last_loss);
/* Loss and gradient of exactly these n samples (host memory, n a multiple
 * of 16), without updating the network. */
int iso_band_gradient(iso_ctx* ctx, const float* points, const float* sdf, const float* weight,
                      uint64_t n, float clamp, float overshoot, float* grad, float* loss);
/* Network output at n points (3 floats each) in `memory`. */
int iso_band_predict(iso_ctx* ctx, const float* points, uint64_t n, float* out, int32_t memory);
/* Band extraction of the current program over `desc` (ISO_SOURCE_PROGRAM);
 * flags: ISO_BAND_* (0: the network and the margin alone decide). */
int iso_extract_mc_band(iso_ctx* ctx, const iso_grid_desc* desc, float margin, uint32_t flags,
                        iso_mesh_out* out, iso_band_stats* stats);
/* Cells whose marching-cubes case differs between the last
 * iso_extract_mc_band over desc and the dense extraction of desc (missed,
 * spurious or altered surface cells): zero means the band's mesh is the
 * dense mesh.  Valid until the next extraction in the context. */
int iso_band_missed_cells(iso_ctx* ctx, const iso_grid_desc* desc, uint64_t* missed);

/* After an ISO_MEMORY_HOST extraction returned ISO_ERR_CAPACITY, the mesh is
 * kept in the context; this copies it into larger buffers without
 * re-extracting (the reference's Vec-growing sink needs no second pass). */
int iso_mesh_fetch(iso_ctx* ctx, iso_mesh_out* out);

/* ---- device memory helpers (for hosts without their own allocator) ------- */

int iso_device_alloc(iso_ctx* ctx, uint64_t bytes, void** out);
int iso_device_free(iso_ctx* ctx, void* ptr);
int iso_memcpy_to_device(iso_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int iso_memcpy_to_host(iso_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int iso_memcpy_device(iso_ctx* ctx, void* dst, const void* src, uint64_t bytes);

/* ---- per-kernel timing ------------------------------------------------------ */

/* When enabled, every kernel launch is bracketed by HIP events on the
 * context's stream; durations accumulate per kernel name. */
int iso_ctx_set_timing(iso_ctx* ctx, int enable);
typedef struct iso_kernel_stat {
  char name[48];
  uint64_t launches;
  double total_ms;
} iso_kernel_stat;
/* Writes up to `max` entries; *n_out receives the number available. */
int iso_ctx_kernel_stats(iso_ctx* ctx, iso_kernel_stat* out, int max, int* n_out);
int iso_ctx_reset_kernel_stats(iso_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* ISOSURFACE_AMD_H */
