// Drives the C++ mirror the way the crate's callers drive the Rust API:
//   MarchingCubes::new(65).extract(&DemoSource(Sphere(0.35)), &mut IndexedVertices)
// and, for a source the GPU cannot evaluate, the same sphere written as a
// host lambda (sampled on the host, uploaded, extracted on the GPU).
// Writes the meshes as raw little-endian arrays for the test suite:
//   <out>.program.f32 / .program.u32 and <out>.host.f32 / .host.u32, and the
//   MarchingCubes::<Directed> meshes <out>.directed.* / .directed_host.*
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>
#include <type_traits>
#include <vector>

#include "isosurface.hpp"

namespace {

struct HostSphere {  // a user source: Sphere(0.35) behind DemoSource, on the host
  isosurface::Signed sample_scalar(isosurface::Vec3 p) const {
    const float x = p.x - 0.5f, y = p.y - 0.5f, z = p.z - 0.5f;
    return {std::sqrt((x * x + y * y) + z * z) - 0.35f};
  }
};

struct HostDirectedSphere {  // the same sphere's VectorSource form (sphere.rs:41-55)
  isosurface::Directed sample_vector(isosurface::Vec3 p) const {
    const float ax = std::fabs(p.x - 0.5f), ay = std::fabs(p.y - 0.5f), az = std::fabs(p.z - 0.5f);
    const float r2 = 0.35f * 0.35f;
    const float l_yz = r2 - (ay * ay + az * az), l_xz = r2 - (ax * ax + az * az),
                l_xy = r2 - (ax * ax + ay * ay);
    const float big = 3.40282347e+38f;
    return {{l_yz < 0.0f ? big : ax - std::sqrt(l_yz), l_xz < 0.0f ? big : ay - std::sqrt(l_xz),
             l_xy < 0.0f ? big : az - std::sqrt(l_xy)}};
  }
};

template <class T>
void dump(const std::string& path, const std::vector<T>& v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(T)));
}

}  // namespace

int main(int argc, char** argv) {
  const std::string out = argc > 1 ? argv[1] : "example_mc";
  try {
    isosurface::Context ctx(0);
    isosurface::MarchingCubes mc(65, ctx);

    std::vector<float> v;
    std::vector<uint32_t> i;
    isosurface::IndexedVertices sink(v, i);
    mc.extract(isosurface::DemoSource(isosurface::Sphere{0.35f}), sink);
    dump(out + ".program.f32", v);
    dump(out + ".program.u32", i);
    std::printf("program source: %zu vertices, %zu triangles\n", v.size() / 3, i.size() / 3);

    std::vector<float> hv;
    std::vector<uint32_t> hi;
    isosurface::IndexedVertices hsink(hv, hi);
    mc.extract(HostSphere{}, hsink);
    dump(out + ".host.f32", hv);
    dump(out + ".host.u32", hi);
    std::printf("host source: %zu vertices, %zu triangles\n", hv.size() / 3, hi.size() / 3);

    // PointCloud::new(65).extract(&source, &mut OnlyVertices) and the normal
    // sink IndexedInterleavedNormals over CentralDifference::new_with_epsilon.
    std::vector<float> pts;
    isosurface::OnlyVertices psink(pts);
    isosurface::PointCloud pc(65, ctx);
    pc.extract(isosurface::DemoSource(isosurface::Sphere{0.35f}), psink);
    dump(out + ".points.f32", pts);
    std::printf("point cloud: %zu points\n", pts.size() / 3);

    const auto demo = isosurface::DemoSource(isosurface::Sphere{0.35f});
    isosurface::CentralDifference<std::remove_cv_t<decltype(demo)>> cd{demo, 1e-4f};
    std::vector<float> nv;
    std::vector<uint32_t> ni;
    isosurface::IndexedInterleavedNormals<std::remove_cv_t<decltype(demo)>> nsink(nv, &ni, cd, ctx);
    mc.extract(demo, nsink);
    dump(out + ".normals.f32", nv);
    dump(out + ".normals.u32", ni);
    std::printf("normal sink: %zu vertices\n", nv.size() / 6);

    // MarchingCubes::<Directed>::new(65), for the lowered and the host source.
    isosurface::MarchingCubes<isosurface::Directed> dmc(65, ctx);
    std::vector<float> dv;
    std::vector<uint32_t> di;
    isosurface::IndexedVertices dsink(dv, di);
    dmc.extract(isosurface::DemoSource(isosurface::Sphere{0.35f}), dsink);
    dump(out + ".directed.f32", dv);
    dump(out + ".directed.u32", di);
    std::vector<float> dhv;
    std::vector<uint32_t> dhi;
    isosurface::IndexedVertices dhsink(dhv, dhi);
    dmc.extract(HostDirectedSphere{}, dhsink);
    dump(out + ".directed_host.f32", dhv);
    dump(out + ".directed_host.u32", dhi);
    std::printf("directed: %zu vertices, %zu triangles\n", dv.size() / 3, di.size() / 3);

    // ExtendedMarchingCubes::new(33) of a cube (its own normals) and of the
    // cube behind CentralDifference.
    const auto cube = isosurface::DemoSource(isosurface::Translate<isosurface::RectangularPrism>{
        {0.013f, 0.021f, -0.017f}, isosurface::RectangularPrism{{0.2f, 0.2f, 0.2f}}});
    isosurface::ExtendedMarchingCubes emc(33, ctx);
    std::vector<float> ev;
    std::vector<uint32_t> ei;
    isosurface::IndexedVertices esink(ev, ei);
    emc.extract(cube, esink);
    dump(out + ".emc.f32", ev);
    dump(out + ".emc.u32", ei);
    std::vector<float> ecv;
    std::vector<uint32_t> eci;
    isosurface::IndexedVertices ecsink(ecv, eci);
    emc.extract(isosurface::CentralDifference<std::remove_cv_t<decltype(cube)>>{cube, 1e-4f}, ecsink);
    dump(out + ".emc_cd.f32", ecv);
    dump(out + ".emc_cd.u32", eci);
    std::printf("extended: %zu vertices, %zu triangles\n", ev.size() / 3, ei.size() / 3);

    // DualContouring::new(33, MinimiseQEF) over CentralDifference(cube, 1e-6),
    // and LinearHashedMarchingCubes::new(6) of the cube.
    isosurface::DualContouring dc(33, isosurface::MinimiseQEF{}, ctx);
    std::vector<float> dcv;
    std::vector<uint32_t> dci;
    isosurface::IndexedVertices dcsink(dcv, dci);
    dc.extract(isosurface::CentralDifference<std::remove_cv_t<decltype(cube)>>{cube, 1e-6f}, dcsink);
    dump(out + ".dc.f32", dcv);
    dump(out + ".dc.u32", dci);
    // ... and over the cube's own (HermiteSource) normals.
    std::vector<float> dcav;
    std::vector<uint32_t> dcai;
    isosurface::IndexedVertices dcasink(dcav, dcai);
    dc.extract(cube, dcasink);
    dump(out + ".dc_hermite.f32", dcav);
    dump(out + ".dc_hermite.u32", dcai);
    isosurface::LinearHashedMarchingCubes lh(6, ctx);
    std::vector<float> lhv;
    std::vector<uint32_t> lhi;
    isosurface::IndexedVertices lhsink(lhv, lhi);
    lh.extract(cube, lhsink);
    dump(out + ".lhmc.f32", lhv);
    dump(out + ".lhmc.u32", lhi);
    std::printf("dual contouring: %zu vertices; octree: %zu vertices\n", dcv.size() / 3,
                lhv.size() / 3);

    // Error channel: a malformed program surfaces as isosurface::Error.
    try {
      isosurface::Program bad{isosurface::detail::node(ISO_OP_UNION)};
      ctx.check(iso_program_upload(ctx.get(), bad.data(), 1), "iso_program_upload");
      std::printf("error path: missing\n");
      return 1;
    } catch (const isosurface::Error& e) {
      std::printf("error path: status %d\n", e.status());
    }
  } catch (const isosurface::Error& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 2;
  }
  return 0;
}
