// C++ host mirror of the swiftcoder/isosurface API over the isosurface_amd
// C ABI.  Header-only; link against libisosurface_amd.so.
//
// The reference's front door (marching_cubes.rs:46-82)
//
//   let mut mc = MarchingCubes::<Signed>::new(65);
//   mc.extract(&Sampler::new(&source), &mut IndexedVertices::new(&mut v, &mut i));
//
// reads here as
//
//   isosurface::Context ctx;                      // one per thread / stream
//   isosurface::MarchingCubes mc(65, ctx);
//   isosurface::IndexedVertices out(v, i);
//   mc.extract(source, out);
//
// Sources built from the crate's implicit types (Sphere, Torus,
// RectangularPrism, Cylinder, Union, Intersection, Difference; implicit/*.rs)
// lower to a device program and are evaluated on the GPU.  Any other type
// with `Signed sample_scalar(Vec3) const` (source.rs:21-28) is sampled on the
// host over the reference lattice and uploaded.  Extractors receive
// extract_vertex for every vertex in first-encounter order, then
// extract_index for every index (extractor.rs:17-20, mesh.rs:91-100).
// The other algorithms read the same way: MarchingCubes<Directed> (sources
// with `Directed sample_vector(Vec3) const`), PointCloud, DualContouring
// (over CentralDifference<S>), LinearHashedMarchingCubes and
// ExtendedMarchingCubes; the normal sinks are IndexedInterleavedNormals<S>.
// Failures throw isosurface::Error carrying the ABI status.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/isosurface.h"

namespace isosurface {

struct Vec3 {
  float x, y, z;
};

// Signed scalar distance (distance.rs:41).
struct Signed {
  float value;
};

// Distance along each cardinal axis (distance.rs:43-45): what a VectorSource
// returns from sample_vector.
struct Directed {
  Vec3 value;
};

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

using Program = std::vector<iso_node>;

namespace detail {
inline iso_node node(uint32_t op, float a = 0, float b = 0, float c = 0) {
  iso_node n{};
  n.op = op;
  n.p[0] = a;
  n.p[1] = b;
  n.p[2] = c;
  return n;
}
}  // namespace detail

// ---- implicit shapes and CSG (lowered to programs) -------------------------

struct Sphere {
  float radius;
  void lower(Program& p) const { p.push_back(detail::node(ISO_OP_SPHERE, radius)); }
};

struct Torus {
  float radius, tube_radius;
  void lower(Program& p) const { p.push_back(detail::node(ISO_OP_TORUS, radius, tube_radius)); }
};

struct RectangularPrism {
  Vec3 half_extent;
  void lower(Program& p) const {
    p.push_back(detail::node(ISO_OP_BOX, half_extent.x, half_extent.y, half_extent.z));
  }
};

struct Cylinder {
  float radius, half_length;
  void lower(Program& p) const {
    p.push_back(detail::node(ISO_OP_CYLINDER, radius, half_length));
  }
};

template <class A, class B>
struct Union {
  A a;
  B b;
  void lower(Program& p) const {
    a.lower(p);
    b.lower(p);
    p.push_back(detail::node(ISO_OP_UNION));
  }
};

template <class A, class B>
struct Intersection {
  A a;
  B b;
  void lower(Program& p) const {
    a.lower(p);
    b.lower(p);
    p.push_back(detail::node(ISO_OP_INTERSECTION));
  }
};

// Solid where `b` is solid except where `a` is (csg.rs:96-100).
template <class A, class B>
struct Difference {
  A a;
  B b;
  void lower(Program& p) const {
    a.lower(p);
    b.lower(p);
    p.push_back(detail::node(ISO_OP_DIFFERENCE));
  }
};

// `source` sampled at p - offset.
template <class S>
struct Translate {
  Vec3 offset;
  S source;
  void lower(Program& p) const {
    p.push_back(detail::node(ISO_OP_TRANSLATE, offset.x, offset.y, offset.z));
    source.lower(p);
    p.push_back(detail::node(ISO_OP_END_TRANSLATE));
  }
};

// The examples' recentring wrapper (examples/common/sources.rs:38-43).
template <class S>
Translate<S> DemoSource(S source) {
  return Translate<S>{Vec3{0.5f, 0.5f, 0.5f}, std::move(source)};
}

// A run-time list of sources joined by Union (a left fold).
struct UnionAll {
  std::vector<Program> parts;
  template <class S>
  void push(const S& s) {
    Program p;
    s.lower(p);
    parts.push_back(std::move(p));
  }
  void lower(Program& p) const {
    for (size_t i = 0; i < parts.size(); ++i) {
      p.insert(p.end(), parts[i].begin(), parts[i].end());
      if (i > 0) p.push_back(detail::node(ISO_OP_UNION));
    }
  }
};

template <class T, class = void>
struct is_lowerable : std::false_type {};
template <class T>
struct is_lowerable<T, std::void_t<decltype(std::declval<const T&>().lower(
                           std::declval<Program&>()))>> : std::true_type {};

// ---- extractors (extractor.rs) ------------------------------------------------

class Extractor {
 public:
  virtual ~Extractor() = default;
  virtual void extract_vertex(Vec3 v) = 0;
  virtual void extract_index(size_t index) = 0;
  // Bulk form (n vertices, xyz interleaved): the same calls in order, so a
  // sink that needs device work per vertex can batch it.
  virtual void extract_vertices(const float* xyz, uint64_t n) {
    for (uint64_t v = 0; v < n; ++v)
      extract_vertex(Vec3{xyz[3 * v], xyz[3 * v + 1], xyz[3 * v + 2]});
  }
};

class IndexedVertices : public Extractor {
 public:
  IndexedVertices(std::vector<float>& vertices, std::vector<uint32_t>& indices)
      : vertices_(vertices), indices_(indices) {}
  void extract_vertex(Vec3 v) override {
    vertices_.push_back(v.x);
    vertices_.push_back(v.y);
    vertices_.push_back(v.z);
  }
  void extract_index(size_t index) override { indices_.push_back(static_cast<uint32_t>(index)); }

 private:
  std::vector<float>& vertices_;
  std::vector<uint32_t>& indices_;
};

class OnlyVertices : public Extractor {
 public:
  explicit OnlyVertices(std::vector<float>& vertices) : vertices_(vertices) {}
  void extract_vertex(Vec3 v) override {
    vertices_.push_back(v.x);
    vertices_.push_back(v.y);
    vertices_.push_back(v.z);
  }
  void extract_index(size_t) override {}

 private:
  std::vector<float>& vertices_;
};

// ---- device context ----------------------------------------------------------

class Context {
 public:
  explicit Context(int device = 0, void* stream = nullptr) {
    check(iso_ctx_create(device, stream, &ctx_), "iso_ctx_create");
  }
  ~Context() {
    if (ctx_) iso_ctx_destroy(ctx_);
  }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;

  iso_ctx* get() const { return ctx_; }

  void check(int status, const char* what) const {
    if (status != ISO_OK) {
      std::string msg = std::string(what) + ": " + iso_status_string(status);
      if (ctx_) msg += std::string(" (") + iso_ctx_last_error(ctx_) + ")";
      throw Error(status, msg);
    }
  }

 private:
  iso_ctx* ctx_ = nullptr;
};

// ---- MarchingCubes ------------------------------------------------------------

template <class D = Signed>
class MarchingCubes {
  static_assert(std::is_same_v<D, Signed> || std::is_same_v<D, Directed>,
                "MarchingCubes<D>: D is Signed or Directed");
  static constexpr bool kDirected = std::is_same_v<D, Directed>;

 public:
  // MarchingCubes::<D>::new(size) (marching_cubes.rs:38-50).
  MarchingCubes(size_t size, Context& ctx) : size_(size), ctx_(ctx) {
    if (size < 2) throw Error(ISO_ERR_INVALID_ARGUMENT, "size must be at least 2");
  }

  // MarchingCubes::extract (marching_cubes.rs:59-82).
  template <class S>
  void extract(const S& source, Extractor& extractor) {
    iso_grid_desc desc;
    if constexpr (is_lowerable<S>::value) {
      Program prog;
      source.lower(prog);
      ctx_.check(iso_program_upload(ctx_.get(), prog.data(), static_cast<uint32_t>(prog.size())),
                 "iso_program_upload");
      ctx_.check(iso_grid_desc_reference(static_cast<int64_t>(size_), ISO_SOURCE_PROGRAM, &desc),
                 "iso_grid_desc_reference");
    } else {
      ctx_.check(iso_grid_desc_reference(static_cast<int64_t>(size_), ISO_SOURCE_GRID, &desc),
                 "iso_grid_desc_reference");
      const size_t per = kDirected ? 3 : 1;
      std::vector<float> grid(static_cast<size_t>(desc.nx * desc.ny * desc.nz) * per);
      size_t i = 0;
      for (int64_t z = 0; z < desc.nz; ++z)
        for (int64_t y = 0; y < desc.ny; ++y)
          for (int64_t x = 0; x < desc.nx; ++x) {
            const Vec3 p{desc.origin[0] + static_cast<float>(x) * desc.spacing[0],
                         desc.origin[1] + static_cast<float>(y) * desc.spacing[1],
                         desc.origin[2] + static_cast<float>(z) * desc.spacing[2]};
            if constexpr (kDirected) {  // Sample<Directed> is sample_vector (sampler.rs:43-47)
              const Vec3 v = source.sample_vector(p).value;
              grid[i++] = v.x;
              grid[i++] = v.y;
              grid[i++] = v.z;
            } else {
              grid[i++] = source.sample_scalar(p).value;
            }
          }
      if constexpr (kDirected)
        ctx_.check(iso_grid_upload_directed(ctx_.get(), grid.data(), desc.nx, desc.ny, desc.nz),
                   "iso_grid_upload_directed");
      else
        ctx_.check(iso_grid_upload(ctx_.get(), grid.data(), desc.nx, desc.ny, desc.nz),
                   "iso_grid_upload");
    }
    run(desc, extractor);
  }

 private:
  void run(const iso_grid_desc& desc, Extractor& extractor) {
    iso_mesh_out out{};
    out.memory = ISO_MEMORY_HOST;
    out.vertices = vertices_.data();
    out.indices = indices_.data();
    out.vertex_capacity = vertices_.size() / 3;
    out.index_capacity = indices_.size();
    int st = kDirected ? iso_extract_mc_directed(ctx_.get(), &desc, &out)
                       : iso_extract_mc(ctx_.get(), &desc, &out);
    if (st == ISO_ERR_CAPACITY) {
      vertices_.resize(3 * out.vertex_count);
      indices_.resize(out.index_count);
      out.vertices = vertices_.data();
      out.indices = indices_.data();
      out.vertex_capacity = out.vertex_count;
      out.index_capacity = out.index_count;
      st = iso_mesh_fetch(ctx_.get(), &out);
    }
    ctx_.check(st, "iso_extract_mc");
    extractor.extract_vertices(vertices_.data(), out.vertex_count);
    for (uint64_t i = 0; i < out.index_count; ++i) extractor.extract_index(indices_[i]);
  }

  size_t size_;
  Context& ctx_;
  std::vector<float> vertices_;      // reused across calls, like the
  std::vector<uint32_t> indices_;    // reference's layer buffers
};

// ---- PointCloud (point_cloud.rs:29-63) ----------------------------------------

class PointCloud {
 public:
  PointCloud(size_t size, Context& ctx) : size_(size), ctx_(ctx) {
    if (size < 2) throw Error(ISO_ERR_INVALID_ARGUMENT, "size must be at least 2");
  }

  // PointCloud::extract: one vertex per mixed cell (its centre), cell order,
  // no faces.  Program sources only here (grids: Context + the C ABI).
  template <class S>
  void extract(const S& source, Extractor& extractor) {
    static_assert(is_lowerable<S>::value, "PointCloud::extract takes a lowerable source");
    Program prog;
    source.lower(prog);
    ctx_.check(iso_program_upload(ctx_.get(), prog.data(), static_cast<uint32_t>(prog.size())),
               "iso_program_upload");
    iso_grid_desc desc;
    ctx_.check(iso_grid_desc_reference(static_cast<int64_t>(size_), ISO_SOURCE_PROGRAM, &desc),
               "iso_grid_desc_reference");
    iso_mesh_out out{};
    out.memory = ISO_MEMORY_HOST;
    out.vertices = points_.data();
    out.vertex_capacity = points_.size() / 3;
    int st = iso_extract_point_cloud(ctx_.get(), &desc, &out);
    if (st == ISO_ERR_CAPACITY) {
      points_.resize(3 * out.vertex_count);
      out.vertices = points_.data();
      out.vertex_capacity = out.vertex_count;
      st = iso_mesh_fetch(ctx_.get(), &out);
    }
    ctx_.check(st, "iso_extract_point_cloud");
    extractor.extract_vertices(points_.data(), out.vertex_count);
  }

 private:
  size_t size_;
  Context& ctx_;
  std::vector<float> points_;
};

// ---- HermiteSource and the normal sinks (source.rs:52-94, extractor.rs) -------

// CentralDifference::new_with_epsilon(source, epsilon): lowers like its
// source; sample_normal runs on the GPU.
template <class S>
struct CentralDifference {
  S source;
  float epsilon = 0.000001f;
  void lower(Program& p) const { source.lower(p); }
};

template <class T>
struct is_central_difference : std::false_type {};
template <class S>
struct is_central_difference<CentralDifference<S>> : std::true_type {};

// ---- ExtendedMarchingCubes (extended_marching_cubes.rs:43-186) ----------------

class ExtendedMarchingCubes {
 public:
  ExtendedMarchingCubes(size_t size, Context& ctx) : size_(size), ctx_(ctx) {
    if (size < 2) throw Error(ISO_ERR_INVALID_ARGUMENT, "size must be at least 2");
  }

  // ExtendedMarchingCubes::extract: the source's Directed values, and its
  // HermiteSource normals — the shapes' own (a lowerable source of shapes,
  // unions and translations) or CentralDifference<S>'s.
  template <class S>
  void extract(const S& source, Extractor& extractor) {
    static_assert(is_lowerable<S>::value, "ExtendedMarchingCubes::extract takes a lowerable source");
    float epsilon = 0.0f;
    if constexpr (is_central_difference<S>::value) epsilon = source.epsilon;
    Program prog;
    source.lower(prog);
    ctx_.check(iso_program_upload(ctx_.get(), prog.data(), static_cast<uint32_t>(prog.size())),
               "iso_program_upload");
    iso_grid_desc desc;
    ctx_.check(iso_grid_desc_reference(static_cast<int64_t>(size_), ISO_SOURCE_PROGRAM, &desc),
               "iso_grid_desc_reference");
    iso_mesh_out out{};
    out.memory = ISO_MEMORY_HOST;
    out.vertices = vertices_.data();
    out.indices = indices_.data();
    out.vertex_capacity = vertices_.size() / 3;
    out.index_capacity = indices_.size();
    int st = iso_extract_emc(ctx_.get(), &desc, epsilon, &out);
    if (st == ISO_ERR_CAPACITY) {
      vertices_.resize(3 * out.vertex_count);
      indices_.resize(out.index_count);
      out.vertices = vertices_.data();
      out.indices = indices_.data();
      out.vertex_capacity = out.vertex_count;
      out.index_capacity = out.index_count;
      st = iso_mesh_fetch(ctx_.get(), &out);
    }
    ctx_.check(st, "iso_extract_emc");
    extractor.extract_vertices(vertices_.data(), out.vertex_count);
    for (uint64_t i = 0; i < out.index_count; ++i) extractor.extract_index(indices_[i]);
  }

 private:
  size_t size_;
  Context& ctx_;
  std::vector<float> vertices_;
  std::vector<uint32_t> indices_;
};

namespace detail {
// Runs an extraction into host vectors reused across calls (grown through
// iso_mesh_fetch on ISO_ERR_CAPACITY) and replays it into the extractor.
template <class Call>
void extract_host(Context& ctx, Call call, std::vector<float>& v, std::vector<uint32_t>& i,
                  Extractor& extractor, const char* what) {
  iso_mesh_out out{};
  out.memory = ISO_MEMORY_HOST;
  out.vertices = v.data();
  out.indices = i.data();
  out.vertex_capacity = v.size() / 3;
  out.index_capacity = i.size();
  int st = call(&out);
  if (st == ISO_ERR_CAPACITY) {
    v.resize(3 * out.vertex_count);
    i.resize(out.index_count);
    out.vertices = v.data();
    out.indices = i.data();
    out.vertex_capacity = out.vertex_count;
    out.index_capacity = out.index_count;
    st = iso_mesh_fetch(ctx.get(), &out);
  }
  ctx.check(st, what);
  extractor.extract_vertices(v.data(), out.vertex_count);
  for (uint64_t k = 0; k < out.index_count; ++k) extractor.extract_index(i[k]);
}

inline void upload(Context& ctx, const Program& prog) {
  ctx.check(iso_program_upload(ctx.get(), prog.data(), static_cast<uint32_t>(prog.size())),
            "iso_program_upload");
}
}  // namespace detail

// ---- DualContouring (dual_contouring.rs:38-105) --------------------------------

// MinimiseQEF (feature/qef.rs:24-68): the feature placement the GPU runs.
struct MinimiseQEF {};

class DualContouring {
 public:
  // DualContouring::new(size, MinimiseQEF {}).
  DualContouring(size_t size, MinimiseQEF, Context& ctx) : size_(size), ctx_(ctx) {
    if (size < 2) throw Error(ISO_ERR_INVALID_ARGUMENT, "size must be at least 2");
  }

  // DualContouring::extract over a HermiteSource: CentralDifference<S>
  // (CentralDifference::new_with_epsilon(source, e)) or a lowerable source of
  // shapes, unions and translations, whose own analytic normals are used.
  template <class S>
  void extract(const S& source, Extractor& extractor) {
    static_assert(is_lowerable<S>::value, "DualContouring::extract takes a lowerable source");
    float epsilon = 0.0f;
    if constexpr (is_central_difference<S>::value) epsilon = source.epsilon;
    Program prog;
    source.lower(prog);
    detail::upload(ctx_, prog);
    iso_grid_desc desc;
    ctx_.check(iso_grid_desc_reference(static_cast<int64_t>(size_), ISO_SOURCE_PROGRAM, &desc),
               "iso_grid_desc_reference");
    detail::extract_host(
        ctx_, [&](iso_mesh_out* out) { return iso_extract_dc(ctx_.get(), &desc, epsilon, out); },
        vertices_, indices_, extractor, "iso_extract_dc");
  }

 private:
  size_t size_;
  Context& ctx_;
  std::vector<float> vertices_;
  std::vector<uint32_t> indices_;
};

// ---- LinearHashedMarchingCubes (linear_hashed_marching_cubes.rs:38-84) --------

class LinearHashedMarchingCubes {
 public:
  // LinearHashedMarchingCubes::new(max_depth), max_depth in [0, 20].
  LinearHashedMarchingCubes(int max_depth, Context& ctx) : depth_(max_depth), ctx_(ctx) {
    if (max_depth < 0 || max_depth > 20)
      throw Error(ISO_ERR_INVALID_ARGUMENT, "max_depth must be in [0, 20]");
  }

  template <class S>
  void extract(const S& source, Extractor& extractor) {
    static_assert(is_lowerable<S>::value, "LinearHashedMarchingCubes::extract takes a lowerable source");
    Program prog;
    source.lower(prog);
    detail::upload(ctx_, prog);
    detail::extract_host(
        ctx_, [&](iso_mesh_out* out) { return iso_extract_lhmc(ctx_.get(), depth_, out); },
        vertices_, indices_, extractor, "iso_extract_lhmc");
  }

 private:
  int depth_;
  Context& ctx_;
  std::vector<float> vertices_;
  std::vector<uint32_t> indices_;
};

// IndexedInterleavedNormals (extractor.rs:97-127): xyz then the source's
// normal per vertex, and the indices.  OnlyInterleavedNormals
// (extractor.rs:46-69) is the same without indices (keep_indices = false).
template <class S>
class IndexedInterleavedNormals : public Extractor {
 public:
  IndexedInterleavedNormals(std::vector<float>& vertices, std::vector<uint32_t>* indices,
                            const CentralDifference<S>& source, Context& ctx)
      : vertices_(vertices), indices_(indices), source_(source), ctx_(ctx) {}
  void extract_vertex(Vec3 v) override {
    const float xyz[3] = {v.x, v.y, v.z};
    extract_vertices(xyz, 1);
  }
  void extract_index(size_t index) override {
    if (indices_) indices_->push_back(static_cast<uint32_t>(index));
  }
  void extract_vertices(const float* xyz, uint64_t n) override {
    if (n == 0) return;
    Program prog;
    source_.lower(prog);
    ctx_.check(iso_program_upload(ctx_.get(), prog.data(), static_cast<uint32_t>(prog.size())),
               "iso_program_upload");
    std::vector<float> normals(3 * n);
    ctx_.check(iso_sample_normals(ctx_.get(), xyz, n, source_.epsilon, normals.data(),
                                  ISO_MEMORY_HOST),
               "iso_sample_normals");
    for (uint64_t v = 0; v < n; ++v) {
      for (int a = 0; a < 3; ++a) vertices_.push_back(xyz[3 * v + a]);
      for (int a = 0; a < 3; ++a) vertices_.push_back(normals[3 * v + a]);
    }
  }

 private:
  std::vector<float>& vertices_;
  std::vector<uint32_t>* indices_;
  CentralDifference<S> source_;
  Context& ctx_;
};

}  // namespace isosurface
