// CPU oracle, ExtendedMarchingCubes — TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Restates ExtendedMarchingCubes::extract (extended_marching_cubes.rs:59-186)
// with the reference's own data structures: the Directed primal traversal,
// crossings and their normals (marching_cubes_impl.rs:39-90), TangentPlanes
// from the crossings (feature/mod.rs:77-148), the QEF feature point
// (feature/qef.rs:36-68), fans around feature vertices over EDGE_LOOPS, and
// the post-pass MeshTopology edge flips (mesh.rs:71-89, 101-170) over the
// hash-map edge table.  Normals come from the program's HermiteSource
// (implicit/*.rs sample_normal, csg.rs:47-51) or from CentralDifference
// (source.rs:82-94) when epsilon > 0.
#include <cmath>
#include <cstring>
#include <unordered_set>
#include <vector>

#include "oracle.h"
#include "oracle_internal.h"

using orc::F3;

namespace {

// HermiteSource::sample_normal exists for the shapes, Union and the
// DemoSource shift only (Intersection and Difference have none).
bool has_hermite(const orc::Program& t) {
  for (const orc::TreeNode& n : t.nodes)
    if (n.op == ISO_OP_INTERSECTION || n.op == ISO_OP_DIFFERENCE || n.op == ISO_OP_SMOOTH_UNION)
      return false;
  return true;
}

F3 hermite_normal(const orc::Program& t, int id, F3 p) {
  const orc::TreeNode& n = t.nodes[id];
  switch (n.op) {
    case ISO_OP_SPHERE:  // sphere.rs:58-62
      return p;
    case ISO_OP_TORUS: {  // torus.rs:98-107
      F3 ring = orc::unit_or_zero(F3{p.x, p.y, 0.0f});
      const float l = std::sqrt((p.x * p.x + p.y * p.y) + 0.0f * 0.0f);
      if (std::fabs(l) < 1.1920929e-07f) ring = F3{1.0f, 0.0f, 0.0f};  // unwrap_or((1, 0, 0))
      ring = F3{ring.x * n.p[0], ring.y * n.p[0], ring.z * n.p[0]};
      return F3{p.x - ring.x, p.y - ring.y, p.z - ring.z};
    }
    case ISO_OP_CYLINDER: {  // cylinder.rs:74-84
      const float z = std::fabs(p.z) / n.p[1];
      const float r = std::sqrt(p.x * p.x + p.y * p.y) / n.p[0];
      if (z > r) return F3{0.0f, 0.0f, p.z};
      return F3{p.x, p.y, 0.0f};
    }
    case ISO_OP_BOX: {  // rectangular_prism.rs:76-79: clamp_to_cardinal_axis (vector.rs:278-287)
      const F3 a{std::fabs(p.x), std::fabs(p.y), std::fabs(p.z)};
      if (a.x > a.y && a.x > a.z) return F3{p.x, 0.0f, 0.0f};
      if (a.y > a.z) return F3{0.0f, p.y, 0.0f};
      return F3{0.0f, 0.0f, p.z};
    }
    case ISO_OP_UNION: {  // csg.rs:47-51: a.min(b) componentwise
      const F3 a = hermite_normal(t, n.a, p), b = hermite_normal(t, n.b, p);
      return F3{orc::fmn(a.x, b.x), orc::fmn(a.y, b.y), orc::fmn(a.z, b.z)};
    }
    case ISO_OP_TRANSLATE:  // DemoSource::sample_normal (examples/common/sources.rs:52-57)
      return hermite_normal(t, n.a, F3{p.x - n.p[0], p.y - n.p[1], p.z - n.p[2]});
  }
  return F3{0.0f, 0.0f, 0.0f};
}

// CentralDifference::sample_normal (source.rs:82-94), not normalised.
F3 central_normal(const orc::Program& t, F3 q, float e) {
  const float vx = orc::sample(t, F3{q.x + e, q.y + 0.0f, q.z + 0.0f}) -
                   orc::sample(t, F3{q.x - e, q.y - 0.0f, q.z - 0.0f});
  const float vy = orc::sample(t, F3{q.x + 0.0f, q.y + e, q.z + 0.0f}) -
                   orc::sample(t, F3{q.x - 0.0f, q.y - e, q.z - 0.0f});
  const float vz = orc::sample(t, F3{q.x + 0.0f, q.y + 0.0f, q.z + e}) -
                   orc::sample(t, F3{q.x - 0.0f, q.y - 0.0f, q.z - e});
  const float e2 = 2.0f * e;
  return F3{vx / e2, vy / e2, vz / e2};
}

using orc::EdgeHandle;

// Face::vertex_opposite (mesh.rs:193-203).
uint32_t vertex_opposite(const uint32_t* f, EdgeHandle e) {
  for (int i = 0; i < 3; ++i)
    if (f[i] != e.a && f[i] != e.b) return f[i];
  std::abort();  // unreachable!() in the reference
}

// Face::matches_winding_direction (mesh.rs:205-217).
bool matches_winding(const uint32_t* f, EdgeHandle e) {
  for (int i = 0; i < 3; ++i) {
    const int j = (i + 1) % 3;
    if (e.a == f[i] && e.b == f[j]) return true;
    if (e.a == f[j] && e.b == f[i]) return false;
  }
  std::abort();
}

// MeshTopology::rotate_edge (mesh.rs:137-170).
void rotate_edge(orc::MeshBuilder& mb, EdgeHandle edge) {
  auto it = mb.edge_to_face.find(edge);
  if (it == mb.edge_to_face.end() || it->second.size() != 2) return;
  const uint32_t ha = it->second[0], hb = it->second[1];
  uint32_t fa[3], fb[3];
  std::memcpy(fa, &mb.faces[3 * ha], sizeof(fa));
  std::memcpy(fb, &mb.faces[3 * hb], sizeof(fb));
  const uint32_t c = vertex_opposite(fa, edge), d = vertex_opposite(fb, edge);
  uint32_t u, v;
  if (matches_winding(fa, edge)) {
    u = edge.b;
    v = edge.a;
  } else {
    u = edge.a;
    v = edge.b;
  }
  mb.faces[3 * ha + 0] = c;
  mb.faces[3 * ha + 1] = d;
  mb.faces[3 * ha + 2] = u;
  mb.faces[3 * hb + 0] = c;
  mb.faces[3 * hb + 1] = v;
  mb.faces[3 * hb + 2] = d;
  const EdgeHandle e = c > d ? EdgeHandle{d, c} : EdgeHandle{c, d};
  mb.edges.insert(e);
  mb.edge_to_face[e] = {ha, hb};
  mb.edges.erase(edge);
  mb.edge_to_face.erase(edge);
}

// ExtendedMarchingCubes::flip_feature_edges (extended_marching_cubes.rs:158-185).
void flip_feature_edges(orc::MeshBuilder& mb, const std::unordered_set<uint32_t>& features) {
  std::unordered_set<EdgeHandle, orc::EdgeHash> flips;
  for (const EdgeHandle& edge : mb.edges) {
    if (features.count(edge.a) && features.count(edge.b)) continue;
    auto it = mb.edge_to_face.find(edge);
    if (it == mb.edge_to_face.end() || it->second.size() != 2) continue;
    const uint32_t* fa = &mb.faces[3 * it->second[0]];
    const uint32_t* fb = &mb.faces[3 * it->second[1]];
    if (features.count(vertex_opposite(fa, edge)) && features.count(vertex_opposite(fb, edge)))
      flips.insert(edge);
  }
  for (const EdgeHandle& edge : flips) rotate_edge(mb, edge);
}

}  // namespace

extern "C" int or_extract_emc(const iso_node* prog, uint32_t n, float epsilon,
                              const iso_grid_desc* desc, int want_features, or_mesh* out) {
  using iso::tables::kEdgeCorners;
  using iso::tables::kLoops;
  using iso::tables::kTables;
  std::memset(out, 0, sizeof(*out));
  const iso_grid_desc& d = *desc;
  if (d.nx < 2 || d.ny < 2 || d.nz < 2) return ISO_ERR_INVALID_ARGUMENT;
  orc::Program t;
  if (!orc::load_program(prog, n, &t) || !orc::has_vector_form(t)) return ISO_ERR_PROGRAM;
  if (!(epsilon > 0.0f) && !has_hermite(t)) return ISO_ERR_PROGRAM;

  orc::MeshBuilder mb;
  std::unordered_set<uint32_t> features;
  std::vector<uint8_t> cell_feature;
  orc::traverse_primal<orc::DirectedD>(
      d, [&](int64_t, int64_t, int64_t, F3 p) { return orc::sample_vector(t, p); },
      [&](const orc::Pt* keys, const F3* corners, const F3* values) {
        // classify_corners + find_edge_crossings (marching_cubes_impl.rs:26-56)
        int cube = 0;
        for (int i = 0; i < 8; ++i)
          if (!orc::DirectedD::is_positive(values[i])) cube |= 1 << i;
        const uint16_t mask = kTables.edge_mask[cube];
        F3 verts[12], normals[12], pts[12], nrm[12];
        int m = 0;
        for (int e = 0; e < 12; ++e) {
          if (!(mask & (1u << e))) continue;
          const int u = kEdgeCorners[e][0], v = kEdgeCorners[e][1];
          verts[e] = orc::DirectedD::crossing(values[u], values[v], corners[u], corners[v]);
          // sample_normals_at_edge_crossings (marching_cubes_impl.rs:71-90)
          const F3 raw = epsilon > 0.0f ? central_normal(t, verts[e], epsilon)
                                        : hermite_normal(t, t.root, verts[e]);
          normals[e] = orc::unit_or_zero(raw);
          pts[m] = verts[e];
          nrm[m] = normals[e];
          ++m;
        }
        if (want_features) cell_feature.push_back(0);
        if (m == 0) return;  // no crossings: march_cube emits nothing
        // march_cube_extended (extended_marching_cubes.rs:96-156)
        bool feature = false;
        const F3 fp = orc::place_feature_planes(pts, nrm, m, &feature);
        if (!feature) {
          for (int i = 0; i < 5; ++i) {
            const int8_t* tri = &kTables.tri[cube][3 * i];
            if (tri[0] < 0) break;
            const uint32_t a = mb.add_vertex(orc::grid_key(keys, tri[0]), verts[tri[0]]);
            const uint32_t b = mb.add_vertex(orc::grid_key(keys, tri[1]), verts[tri[1]]);
            const uint32_t c = mb.add_vertex(orc::grid_key(keys, tri[2]), verts[tri[2]]);
            mb.add_face(a, b, c);
          }
          return;
        }
        if (want_features) cell_feature.back() = 1;
        const uint32_t centre = mb.add_vertex_unkeyed(fp);
        features.insert(centre);
        int offset = 0;
        for (int l = 0; l < kLoops.count[cube]; ++l) {
          const int size = kLoops.size[cube][l];
          for (int i = 0; i < size; ++i) {
            const int j = (i + 1) % size;
            const int e0 = kLoops.edge[cube][offset + i];
            const uint32_t a = mb.add_vertex(orc::grid_key(keys, e0), verts[e0]);
            const int e1 = kLoops.edge[cube][offset + j];
            const uint32_t b = mb.add_vertex(orc::grid_key(keys, e1), verts[e1]);
            mb.add_face(a, b, centre);
          }
          offset += size;
        }
      });
  flip_feature_edges(mb, features);
  orc::finish(mb, out);
  if (want_features) {
    out->cases = orc::dup(cell_feature);
    out->n_cases = cell_feature.size();
  }
  return ISO_OK;
}

namespace orc {

bool program_has_hermite(const Program& t) { return has_hermite(t); }

F3 program_hermite_normal(const Program& t, F3 p) { return hermite_normal(t, t.root, p); }

}  // namespace orc

// HermiteSource::sample_normal of a program at one point (not normalised).
extern "C" int or_sample_hermite_normal(const iso_node* prog, uint32_t n, float x, float y,
                                        float z, float* out) {
  orc::Program t;
  if (!orc::load_program(prog, n, &t) || !has_hermite(t)) return ISO_ERR_PROGRAM;
  const F3 v = hermite_normal(t, t.root, F3{x, y, z});
  out[0] = v.x;
  out[1] = v.y;
  out[2] = v.z;
  return ISO_OK;
}
