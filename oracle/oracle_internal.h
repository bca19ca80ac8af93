// Shared pieces of the CPU oracle — TEST INFRASTRUCTURE ONLY (see oracle.h).
#pragma once

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../isosurface_amd/csrc/mc_tables.h"
#include "oracle.h"

namespace orc {

// ---------------------------------------------------------------------------
// Vec3 (reference src/math/vector.rs)
// ---------------------------------------------------------------------------
template <typename T>
struct V3 {
  T x, y, z;
};
using F3 = V3<float>;

// len_sq is fold!(+, x*x, y*y, z*z) = (x*x + y*y) + z*z  (vector.rs:45-52,125-127)
template <typename T>
T len_sq(V3<T> v) {
  return (v.x * v.x + v.y * v.y) + v.z * v.z;
}
template <typename T>
T len(V3<T> v) {
  return std::sqrt(len_sq(v));  // vector.rs:130-132
}
// f32::min / f32::max: NaN-ignoring IEEE minNum/maxNum, as fmin/fmax.
template <typename T>
T fmn(T a, T b) {
  return std::fmin(a, b);
}
template <typename T>
T fmx(T a, T b) {
  return std::fmax(a, b);
}

// ---------------------------------------------------------------------------
// Sources: the program is a post-order list; rebuild the tree of the crate's
// source types and evaluate it recursively, as the monomorphised
// sample_scalar calls do.
// ---------------------------------------------------------------------------
struct TreeNode {
  uint32_t op;
  float p[6];
  int a = -1;  // first operand / translated child
  int b = -1;  // second operand
};

struct Program {
  std::vector<TreeNode> nodes;
  int root = -1;
};

inline bool load_program(const iso_node* prog, uint32_t n, Program* t) {
  std::vector<int> st;
  std::vector<std::pair<size_t, int>> scopes;  // (value stack size, translate node)
  t->nodes.clear();
  if (!prog || n == 0) return false;
  for (uint32_t i = 0; i < n; ++i) {
    const iso_node& nd = prog[i];
    if (nd.reserved != 0) return false;
    TreeNode tn;
    tn.op = nd.op;
    std::memcpy(tn.p, nd.p, sizeof(tn.p));
    switch (nd.op) {
      case ISO_OP_SPHERE:
      case ISO_OP_TORUS:
      case ISO_OP_BOX:
      case ISO_OP_CYLINDER:
        t->nodes.push_back(tn);
        st.push_back(static_cast<int>(t->nodes.size()) - 1);
        break;
      case ISO_OP_UNION:
      case ISO_OP_INTERSECTION:
      case ISO_OP_DIFFERENCE:
      case ISO_OP_SMOOTH_UNION: {
        size_t floor_sz = scopes.empty() ? 0 : scopes.back().first;
        if (st.size() < floor_sz + 2) return false;
        tn.b = st.back();
        st.pop_back();
        tn.a = st.back();
        st.pop_back();
        t->nodes.push_back(tn);
        st.push_back(static_cast<int>(t->nodes.size()) - 1);
        break;
      }
      case ISO_OP_TRANSLATE:
        t->nodes.push_back(tn);
        scopes.push_back({st.size(), static_cast<int>(t->nodes.size()) - 1});
        break;
      case ISO_OP_END_TRANSLATE: {
        if (scopes.empty()) return false;
        auto sc = scopes.back();
        scopes.pop_back();
        if (st.size() != sc.first + 1) return false;
        t->nodes[sc.second].a = st.back();
        st.pop_back();
        st.push_back(sc.second);
        break;
      }
      default:
        return false;
    }
  }
  if (!scopes.empty() || st.size() != 1) return false;
  t->root = st.back();
  return true;
}

template <typename T>
T eval(const Program& t, int id, V3<T> p) {
  const TreeNode& n = t.nodes[id];
  const T p0 = static_cast<T>(n.p[0]), p1 = static_cast<T>(n.p[1]), p2 = static_cast<T>(n.p[2]);
  switch (n.op) {
    case ISO_OP_SPHERE:
      // sphere.rs:35-39: p.len() - radius
      return len(p) - p0;
    case ISO_OP_TORUS: {
      // torus.rs:40-46
      T q_x = std::fabs(std::sqrt(p.x * p.x + p.y * p.y)) - p0;
      T l = std::sqrt(q_x * q_x + p.z * p.z);
      return l - p1;
    }
    case ISO_OP_BOX: {
      // rectangular_prism.rs:36-41: q = |p| - h;
      // q.max(0).len() + q.max_component().min(0)
      V3<T> q{std::fabs(p.x) - p0, std::fabs(p.y) - p1, std::fabs(p.z) - p2};
      V3<T> qm{fmx(q.x, T(0)), fmx(q.y, T(0)), fmx(q.z, T(0))};
      T maxc = fmx(q.x, fmx(q.y, q.z));  // vector.rs max_component
      return len(qm) + fmn(maxc, T(0));
    }
    case ISO_OP_CYLINDER: {
      // cylinder.rs:41-48
      T q_x = std::fabs(std::sqrt(p.x * p.x + p.y * p.y)) - p0;  // p.xy().len().abs() - r
      T q_z = std::fabs(p.z) - p1;
      V3<T> d{fmx(q_x, T(0)), fmx(q_z, T(0)), T(0)};
      return fmn(fmx(q_x, q_z), T(0)) + len(d);
    }
    case ISO_OP_UNION:
      // csg.rs:35-39: a.min(b)
      return fmn(eval(t, n.a, p), eval(t, n.b, p));
    case ISO_OP_INTERSECTION:
      // csg.rs:68-72: a.max(b)
      return fmx(eval(t, n.a, p), eval(t, n.b, p));
    case ISO_OP_DIFFERENCE: {
      // csg.rs:96-100: b.max(-a)
      T b = eval(t, n.b, p);
      T a = eval(t, n.a, p);
      return fmx(b, -a);
    }
    case ISO_OP_SMOOTH_UNION: {
      // Polynomial smooth minimum (no reference counterpart):
      //   h = max(k - |a - b|, 0) / k;  min(a, b) - h*h*k*0.25
      T a = eval(t, n.a, p);
      T b = eval(t, n.b, p);
      T h = fmx(p0 - std::fabs(a - b), T(0)) / p0;
      return fmn(a, b) - ((h * h) * p0) * T(0.25);
    }
    case ISO_OP_TRANSLATE: {
      // DemoSource: q = p - Vec3::from_scalar(0.5) (examples/common/sources.rs:38-43)
      V3<T> q{p.x - p0, p.y - p1, p.z - p2};
      return eval(t, n.a, q);
    }
  }
  return T(0);
}

inline float sample(const Program& p, F3 q) { return eval<float>(p, p.root, q); }

// ---------------------------------------------------------------------------
// Directed distances (VectorSource::sample_vector, source.rs:32-38): per
// axis, the signed distance to the surface along that axis, or f32::MAX.
// ---------------------------------------------------------------------------
constexpr float kF32Max = 3.40282347e+38f;

inline bool has_vector_form(const Program& t) {
  for (const TreeNode& n : t.nodes)
    if (n.op == ISO_OP_SMOOTH_UNION) return false;  // no VectorSource counterpart
  return true;
}

inline F3 eval_vector(const Program& t, int id, F3 p) {
  const TreeNode& n = t.nodes[id];
  switch (n.op) {
    case ISO_OP_SPHERE: {
      // sphere.rs:41-55 (Vec2::len_sq is x*x + y*y)
      const F3 a{std::fabs(p.x), std::fabs(p.y), std::fabs(p.z)};
      const float r2 = n.p[0] * n.p[0];
      const float l_yz = r2 - (a.y * a.y + a.z * a.z);
      const float l_xz = r2 - (a.x * a.x + a.z * a.z);
      const float l_xy = r2 - (a.x * a.x + a.y * a.y);
      return {l_yz < 0.0f ? kF32Max : a.x - std::sqrt(l_yz),
              l_xz < 0.0f ? kF32Max : a.y - std::sqrt(l_xz),
              l_xy < 0.0f ? kF32Max : a.z - std::sqrt(l_xy)};
    }
    case ISO_OP_TORUS: {
      // torus.rs:48-91
      const F3 a{std::fabs(p.x), std::fabs(p.y), std::fabs(p.z)};
      const float radius = n.p[0], tube = n.p[1];
      const float l = std::sqrt(a.x * a.x + a.y * a.y);
      const float l_xy = l - radius;
      const float tube_at_z = std::sqrt(tube * tube - a.z * a.z);
      const float rx = radius + tube_at_z, ry = radius - tube_at_z;
      F3 out;
      if (a.z > tube || a.y > radius + tube_at_z)
        out.x = kF32Max;
      else if (a.x == 0.0f)
        out.x = std::fabs(a.y - radius) - tube;
      else
        out.x = fmx(a.x - std::sqrt(rx * rx - a.y * a.y), std::sqrt(ry * ry - a.y * a.y) - a.x);
      if (a.z > tube || a.x > radius + tube_at_z)
        out.y = kF32Max;
      else if (a.y == 0.0f)
        out.y = std::fabs(a.x - radius) - tube;
      else
        out.y = fmx(a.y - std::sqrt(rx * rx - a.x * a.x), std::sqrt(ry * ry - a.x * a.x) - a.y);
      out.z = std::fabs(l_xy) > tube ? kF32Max : a.z - std::sqrt(tube * tube - l_xy * l_xy);
      return out;
    }
    case ISO_OP_BOX: {
      // rectangular_prism.rs:43-73
      const F3 a{std::fabs(p.x), std::fabs(p.y), std::fabs(p.z)};
      const F3 h{n.p[0], n.p[1], n.p[2]};
      const F3 mask{(((a.y - h.y) > 0.0f || (a.z - h.z) > 0.0f) ? 1.0f : -1.0f) * kF32Max,
                    (((a.x - h.x) > 0.0f || (a.z - h.z) > 0.0f) ? 1.0f : -1.0f) * kF32Max,
                    (((a.x - h.x) > 0.0f || (a.y - h.y) > 0.0f) ? 1.0f : -1.0f) * kF32Max};
      const bool inside = a.x < h.x && a.y < h.y && a.z < h.z;
      const F3 c = inside ? F3{fmx(a.x, h.x), fmx(a.y, h.y), fmx(a.z, h.z)}
                          : F3{fmn(a.x, h.x), fmn(a.y, h.y), fmn(a.z, h.z)};
      return {fmx(a.x - c.x, mask.x), fmx(a.y - c.y, mask.y), fmx(a.z - c.z, mask.z)};
    }
    case ISO_OP_CYLINDER: {
      // cylinder.rs:50-71
      const F3 a{std::fabs(p.x), std::fabs(p.y), std::fabs(p.z)};
      const float r = n.p[0], hl = n.p[1];
      return {(a.z > hl || a.y > r) ? kF32Max : a.x - std::sqrt(r * r - a.y * a.y),
              (a.z > hl || a.x > r) ? kF32Max : a.y - std::sqrt(r * r - a.x * a.x),
              (a.x * a.x + a.y * a.y) > r * r ? kF32Max : a.z - hl};
    }
    case ISO_OP_UNION: {
      // csg.rs:41-45: a.min(b) componentwise (vector.rs:305-311)
      const F3 a = eval_vector(t, n.a, p), b = eval_vector(t, n.b, p);
      return {fmn(a.x, b.x), fmn(a.y, b.y), fmn(a.z, b.z)};
    }
    case ISO_OP_INTERSECTION: {
      // csg.rs:74-78: a.max(b) componentwise (vector.rs:315-321)
      const F3 a = eval_vector(t, n.a, p), b = eval_vector(t, n.b, p);
      return {fmx(a.x, b.x), fmx(a.y, b.y), fmx(a.z, b.z)};
    }
    case ISO_OP_DIFFERENCE: {
      // csg.rs:102-106: b.max(-a)
      const F3 b = eval_vector(t, n.b, p), a = eval_vector(t, n.a, p);
      return {fmx(b.x, -a.x), fmx(b.y, -a.y), fmx(b.z, -a.z)};
    }
    case ISO_OP_TRANSLATE: {
      // DemoSource::sample_vector: q = p - 0.5 (examples/common/sources.rs:45-50)
      const F3 q{p.x - n.p[0], p.y - n.p[1], p.z - n.p[2]};
      return eval_vector(t, n.a, q);
    }
  }
  return {0.0f, 0.0f, 0.0f};
}

inline F3 sample_vector(const Program& p, F3 q) { return eval_vector(p, p.root, q); }

// Lattice coordinates (primal_grid.rs:45-67: x as f32 * one_over_size).
inline float coord(const iso_grid_desc& d, int axis, int64_t i) {
  return d.origin[axis] + static_cast<float>(i) * d.spacing[axis];
}

// classify_corners (marching_cubes_impl.rs:26-37), Signed::is_positive = v > 0.
inline int classify(const float v[8]) {
  int cube = 0;
  for (int i = 0; i < 8; ++i)
    if (!(v[i] > 0.0f)) cube |= 1 << i;
  return cube;
}

// Signed::find_crossing_point (distance.rs:64-70).
inline F3 find_crossing_point(float a, float b, F3 pa, F3 pb) {
  float delta = b - a;
  float t = (delta == 0.0f) ? 0.5f : -a / delta;
  float s = 1.0f - t;
  return {pa.x * s + pb.x * t, pa.y * s + pb.y * t, pa.z * s + pb.z * t};
}

// The two Distance types (distance.rs:20-103) as the cell callback uses them.
struct SignedD {
  using Value = float;
  // Signed::is_positive (distance.rs:52-54)
  static bool is_positive(float v) { return v > 0.0f; }
  static F3 crossing(float a, float b, F3 pa, F3 pb) { return find_crossing_point(a, b, pa, pb); }
};

struct DirectedD {
  using Value = F3;
  // Directed::is_positive (distance.rs:76-79): any component positive
  static bool is_positive(F3 v) { return v.x > 0.0f || v.y > 0.0f || v.z > 0.0f; }
  // Directed::find_crossing_point (distance.rs:89-102): the axis is
  // (p_a - p_b).abs().max_component_index() (vector.rs:255-263); the
  // interpolation is Signed's on that component.
  static F3 crossing(F3 a, F3 b, F3 pa, F3 pb) {
    const F3 d{std::fabs(pa.x - pb.x), std::fabs(pa.y - pb.y), std::fabs(pa.z - pb.z)};
    const int axis = (d.x > d.y && d.x > d.z) ? 0 : (d.y > d.z ? 1 : 2);
    const float av = axis == 0 ? a.x : (axis == 1 ? a.y : a.z);
    const float bv = axis == 0 ? b.x : (axis == 1 ? b.y : b.z);
    const float delta = bv - av;
    const float t = (delta == 0.0f) ? 0.5f : -av / delta;
    const float s = 1.0f - t;
    return {pa.x * s + pb.x * t, pa.y * s + pb.y * t, pa.z * s + pb.z * t};
  }
};

// ---------------------------------------------------------------------------
// IndexCache + MeshTopology (index_cache.rs, mesh.rs) + IndexedVertices
// ---------------------------------------------------------------------------
struct Pt {
  int64_t x, y, z;
  bool operator<(const Pt& o) const {
    if (x != o.x) return x < o.x;
    if (y != o.y) return y < o.y;
    return z < o.z;
  }
  bool operator==(const Pt& o) const { return x == o.x && y == o.y && z == o.z; }
};

struct GridKey {  // index_cache.rs:18, 48-61
  Pt a, b;
  bool operator==(const GridKey& o) const { return a == o.a && b == o.b; }
};

struct GridKeyHash {
  size_t operator()(const GridKey& k) const {
    uint64_t h = 1469598103934665603ull;
    const int64_t v[6] = {k.a.x, k.a.y, k.a.z, k.b.x, k.b.y, k.b.z};
    for (int i = 0; i < 6; ++i) {
      h ^= static_cast<uint64_t>(v[i]) + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    }
    return static_cast<size_t>(h);
  }
};

struct EdgeHandle {  // mesh.rs Edge::new orders the two handles
  uint32_t a, b;
  bool operator==(const EdgeHandle& o) const { return a == o.a && b == o.b; }
};
struct EdgeHash {
  size_t operator()(const EdgeHandle& e) const {
    uint64_t k = (static_cast<uint64_t>(e.a) << 32) | e.b;
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    return static_cast<size_t>(k);
  }
};

struct MeshBuilder {
  // MeshTopologyBuilder (mesh.rs:226-260) + MeshTopology (mesh.rs:44-100)
  std::unordered_map<GridKey, uint32_t, GridKeyHash> cache;
  uint32_t next_vertex = 0;
  std::vector<uint32_t> faces;
  std::unordered_set<EdgeHandle, EdgeHash> edges;
  std::unordered_map<EdgeHandle, std::vector<uint32_t>, EdgeHash> edge_to_face;
  // IndexedVertices sink (extractor.rs:72-94)
  std::vector<float> vertices;
  std::vector<int64_t> keys;
  std::vector<uint8_t> cases;
  bool want_cases = false;

  uint32_t add_vertex(const GridKey& key, F3 v) {
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    uint32_t idx = next_vertex++;
    cache.emplace(key, idx);
    vertices.push_back(v.x);
    vertices.push_back(v.y);
    vertices.push_back(v.z);
    int axis = key.a.x != key.b.x ? 0 : (key.a.y != key.b.y ? 1 : 2);
    keys.push_back(key.a.x);
    keys.push_back(key.a.y);
    keys.push_back(key.a.z);
    keys.push_back(axis);
    return idx;
  }

  // add_vertex(None, v) (mesh.rs:240-251): always a new vertex, no cache
  // entry; its key is recorded as (-1, -1, -1, 3).
  uint32_t add_vertex_unkeyed(F3 v) {
    uint32_t idx = next_vertex++;
    vertices.push_back(v.x);
    vertices.push_back(v.y);
    vertices.push_back(v.z);
    keys.push_back(-1);
    keys.push_back(-1);
    keys.push_back(-1);
    keys.push_back(3);
    return idx;
  }

  void add_edge(uint32_t u, uint32_t v, uint32_t face) {
    EdgeHandle e = u > v ? EdgeHandle{v, u} : EdgeHandle{u, v};
    edges.insert(e);
    edge_to_face[e].push_back(face);
  }

  void add_face(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t face = static_cast<uint32_t>(faces.size() / 3);
    faces.push_back(a);
    faces.push_back(b);
    faces.push_back(c);
    add_edge(a, b, face);
    add_edge(b, c, face);
    add_edge(c, a, face);
  }
};

inline GridKey grid_key(const Pt keys[8], int edge) {
  const Pt& a = keys[iso::tables::kEdgeCorners[edge][0]];
  const Pt& b = keys[iso::tables::kEdgeCorners[edge][1]];
  if (b < a) return {b, a};
  return {a, b};
}

// The marching_cubes.rs:66-79 cell callback: classify_corners
// (marching_cubes_impl.rs:26-37), find_edge_crossings (:39-56), march_cube
// with GridKey dedup; D is the Distance type (SignedD or DirectedD).
template <class D = SignedD>
inline void march(MeshBuilder& mb, const Pt keys[8], const F3 corners[8],
                  const typename D::Value values[8]) {
  using iso::tables::kEdgeCorners;
  using iso::tables::kTables;
  int cube = 0;
  for (int i = 0; i < 8; ++i)
    if (!D::is_positive(values[i])) cube |= 1 << i;
  if (mb.want_cases) mb.cases.push_back(static_cast<uint8_t>(cube));
  F3 verts[12];
  const uint16_t mask = kTables.edge_mask[cube];
  for (int e = 0; e < 12; ++e) {
    if (mask & (1u << e)) {
      const int u = kEdgeCorners[e][0], v = kEdgeCorners[e][1];
      verts[e] = D::crossing(values[u], values[v], corners[u], corners[v]);
    }
  }
  for (int i = 0; i < 5; ++i) {  // march_cube (marching_cubes_impl.rs:102-117)
    const int8_t* tri = &kTables.tri[cube][3 * i];
    if (tri[0] < 0) break;
    const uint32_t a = mb.add_vertex(grid_key(keys, tri[0]), verts[tri[0]]);
    const uint32_t b = mb.add_vertex(grid_key(keys, tri[1]), verts[tri[1]]);
    const uint32_t c = mb.add_vertex(grid_key(keys, tri[2]), verts[tri[2]]);
    mb.add_face(a, b, c);
  }
}

// PrimalGrid::traverse (traversal/primal_grid.rs:39-89): two cached layers
// of (position, value), cells in z-major order; value(x, y, z, p) is the
// Sample<D> call, cell(keys, corners, values) the per-cell callback.
template <class D, class ValueAt, class Cell>
void traverse_primal(const iso_grid_desc& d, ValueAt value, Cell cell) {
  using iso::tables::kCorner;
  const int64_t nx = d.nx, ny = d.ny, nz = d.nz;
  struct Sample {
    F3 p;
    typename D::Value v;
  };
  std::vector<Sample> layers[2] = {std::vector<Sample>(nx * ny), std::vector<Sample>(nx * ny)};
  auto sample = [&](int64_t x, int64_t y, int64_t z) -> Sample {
    const F3 p{coord(d, 0, x), coord(d, 1, y), coord(d, 2, z)};
    return {p, value(x, y, z, p)};
  };
  // Cache layer zero (primal_grid.rs:47-53)
  for (int64_t y = 0; y < ny; ++y)
    for (int64_t x = 0; x < nx; ++x) layers[0][y * nx + x] = sample(x, y, 0);
  Pt keys[8];
  F3 corners[8];
  typename D::Value values[8];
  for (int64_t z = 0; z + 1 < nz; ++z) {
    // Cache layer z+1 (primal_grid.rs:60-70)
    for (int64_t y = 0; y < ny; ++y)
      for (int64_t x = 0; x < nx; ++x) layers[1][y * nx + x] = sample(x, y, z + 1);
    // Cells of this layer (primal_grid.rs:72-85)
    for (int64_t y = 0; y + 1 < ny; ++y) {
      for (int64_t x = 0; x + 1 < nx; ++x) {
        for (int i = 0; i < 8; ++i) {
          keys[i] = {x + kCorner[i][0], y + kCorner[i][1], z + kCorner[i][2]};
          const Sample& s = layers[kCorner[i][2]][(y + kCorner[i][1]) * nx + x + kCorner[i][0]];
          corners[i] = s.p;
          values[i] = s.v;
        }
        cell(keys, corners, values);
      }
    }
    std::swap(layers[0], layers[1]);
  }
}

// Feature placement (oracle_dc.cpp): TangentPlanes::new over n points and
// normals + MinimiseQEF::place_feature_with_tangents; *feature = the planes
// are not Planar.  unit_or_zero = Vec3::normalised().unwrap_or_default().
F3 place_feature_planes(const F3* points, const F3* normals, int n, bool* feature);
F3 unit_or_zero(F3 v);

template <typename T>
T* dup(const std::vector<T>& v) {
  T* p = static_cast<T*>(std::malloc(v.size() * sizeof(T) + 1));
  if (!v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
  return p;
}

// MeshTopology::extract_indices (mesh.rs:91-100) into the C result.
inline void finish(const MeshBuilder& mb, or_mesh* out) {
  out->vertices = dup(mb.vertices);
  out->indices = dup(mb.faces);
  out->edge_keys = dup(mb.keys);
  out->n_vertices = mb.vertices.size() / 3;
  out->n_indices = mb.faces.size();
  if (mb.want_cases) {
    out->cases = dup(mb.cases);
    out->n_cases = mb.cases.size();
  }
}

// HermiteSource::sample_normal of the whole program at p (not normalised),
// and whether the program has one (oracle_emc.cpp).
bool program_has_hermite(const Program& t);
F3 program_hermite_normal(const Program& t, F3 p);

}  // namespace orc
