/*
 * CPU oracle for isosurface_amd — TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++ restatement of the reference crate's CPU extraction path
 * (swiftcoder/isosurface), used by the test suite, the smoke check and the
 * benchmark's CPU-baseline leg as the checker.  The product never links,
 * loads or calls it.  Built with -O2 -ffp-contract=off so every f32 operation
 * is rounded exactly once, as rustc does.
 *
 * Each function cites the reference file:line it restates.
 */
#ifndef ISO_ORACLE_H
#define ISO_ORACLE_H

#include <stdint.h>

#include "../include/isosurface.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_mesh {
  float* vertices;     /* 3 per vertex, emission order */
  uint32_t* indices;   /* 3 per triangle */
  int64_t* edge_keys;  /* 4 per vertex: min lattice point x,y,z and axis 0/1/2 */
  uint64_t n_vertices;
  uint64_t n_indices;
  /* per-cell classification (cube index), z-major, if requested */
  uint8_t* cases;
  uint64_t n_cases;
} or_mesh;

/* Scalar SDF of a program at one point (ScalarSource::sample_scalar). */
float or_sample(const iso_node* prog, uint32_t n, float x, float y, float z);
/* The same SDF evaluated in f64 (the FP64 twin). */
double or_sample_f64(const iso_node* prog, uint32_t n, double x, double y, double z);
/* Validates a program; returns 0 when well formed. */
int or_program_check(const iso_node* prog, uint32_t n);

/* Dense samples of a program over a lattice (x fastest). */
int or_sample_lattice(const iso_node* prog, uint32_t n, const iso_grid_desc* desc, float* out);
/* Layers [z0, z1) of the same lattice into out ((z1 - z0) * ny * nx floats):
 * the same values as or_sample_lattice, so callers may sample disjoint
 * layer ranges on several threads. */
int or_sample_lattice_layers(const iso_node* prog, uint32_t n, const iso_grid_desc* desc,
                             int64_t z0, int64_t z1, float* out);

/* MarchingCubes::extract with IndexedVertices over a lattice whose samples
 * come from `prog` (if grid == NULL) or from the dense `grid`.
 * want_cases != 0 also records every cell's cube index. */
int or_extract_mc(const iso_node* prog, uint32_t n, const float* grid,
                  const iso_grid_desc* desc, int want_cases, or_mesh* out);

/* Directed distance of a program at one point (VectorSource::sample_vector);
 * ISO_ERR_PROGRAM for a program with no directed form (smooth union). */
int or_sample_vector(const iso_node* prog, uint32_t n, float x, float y, float z, float* out);
/* MarchingCubes::<Directed>::extract with IndexedVertices: values from
 * `prog`'s sample_vector (grid3 == NULL) or from grid3 (3 floats per
 * lattice point, x fastest). */
int or_extract_mc_directed(const iso_node* prog, uint32_t n, const float* grid3,
                           const iso_grid_desc* desc, int want_cases, or_mesh* out);

/* ExtendedMarchingCubes::new(size).extract with IndexedVertices over a
 * program lattice: Directed values from the program's sample_vector, normals
 * from its HermiteSource (epsilon == 0; shapes, Union and translations only)
 * or CentralDifference::new_with_epsilon(program, epsilon) (epsilon > 0).
 * Feature vertices have edge key (-1, -1, -1, 3).  want_features != 0 puts
 * one byte per cell (1: the cell placed a feature vertex) in `cases`. */
int or_extract_emc(const iso_node* prog, uint32_t n, float epsilon, const iso_grid_desc* desc,
                   int want_features, or_mesh* out);

/* HermiteSource::sample_normal of a program (shapes, Union, translations),
 * not normalised; ISO_ERR_PROGRAM for other operators. */
int or_sample_hermite_normal(const iso_node* prog, uint32_t n, float x, float y, float z,
                             float* out);

/* PointCloud::extract over a lattice (samples from `prog` or `grid`): the
 * vertices are the centres of the mixed cells in cell order. */
int or_extract_point_cloud(const iso_node* prog, uint32_t n, const float* grid,
                           const iso_grid_desc* desc, or_mesh* out);
/* CentralDifference::sample_normal (not normalised) of a program at points. */
int or_sample_normals(const iso_node* prog, uint32_t n, const float* points, uint64_t n_points,
                      float epsilon, float* out);

/* f64 crossing point of lattice edges (FP64 twin): for each key
 * (x, y, z, axis) the zero of the linear interpolant between the f64 SDF
 * values at the edge's endpoints. */
int or_edge_points_f64(const iso_node* prog, uint32_t n, const iso_grid_desc* desc,
                       const int64_t* keys, uint64_t n_keys, double* out_xyz);

/* DualContouring::extract with MinimiseQEF and CentralDifference normals
 * (epsilon as in CentralDifference::new_with_epsilon) over a program lattice;
 * epsilon 0: the program's own HermiteSource normals (shapes, unions and
 * translations only, else ISO_ERR_PROGRAM). */
int or_extract_dc(const iso_node* prog, uint32_t n, const iso_grid_desc* desc, float epsilon,
                  or_mesh* out);
/* The reference SVD solve of an n x 3 system (row-major a, n <= 12). */
int or_svd_solve(const double* a, const double* b, int rows, int drop_min, double* x_out,
                 double* d_out);
/* The dual vertex of a mixed cell from its corners and corner normals. */
void or_place_feature(const float* corners, const float* normals, float* out);

/* LinearHashedMarchingCubes::new(max_depth).extract with IndexedVertices.
 * Primal vertices are visited in ascending Morton order (order_seed == 0) or
 * in a seeded shuffle of it; edge_keys holds (a, b, 0, 0) MortonKeys. */
int or_extract_lhmc(const iso_node* prog, uint32_t n, int max_depth, uint64_t order_seed,
                    or_mesh* out);
/* The octree alone: node keys in BFS order, values, leaf flags. */
int or_lhmc_octree(const iso_node* prog, uint32_t n, int max_depth, uint64_t capacity,
                   uint64_t* keys, float* values, uint8_t* is_leaf, uint64_t* n_nodes);
/* Morton helpers (morton.rs), for direct checks. */
int or_morton_level(uint64_t k);
void or_morton_center(uint64_t k, float* out);
float or_morton_size(uint64_t k);
uint64_t or_morton_add(uint64_t a, uint64_t b);
uint64_t or_morton_sub(uint64_t a, uint64_t b);
uint64_t or_morton_primal_vertex(uint64_t k, int level, unsigned which);
uint64_t or_morton_dual_vertex(uint64_t k, int level, unsigned which);

void or_mesh_free(or_mesh* m);

/* Decoded tables, for checksum tests. */
void or_tri_table(int8_t out[256 * 16]);
void or_edge_mask_table(uint16_t out[256]);
/* The edge-loop table in the reference's EDGE_LOOPS row layout
 * (count, sizes, edges, -1 padding; int8[256][17]). */
void or_loop_table(int8_t out[256 * 17]);

#ifdef __cplusplus
}
#endif

#endif
