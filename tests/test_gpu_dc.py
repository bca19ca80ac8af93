"""GPU dual contouring against the oracle (restated reference CPU path).

DualContouring::new(size, MinimiseQEF {}).extract over a CentralDifference
source with IndexedVertices: the GPU output must equal the oracle's bit for
bit — vertex order, f32 vertex bits (feature points from the f64 SVD
included) and the u32 index list."""
import ctypes as C
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from isosurface_amd import (CentralDifference, Cylinder, DualContouring, IndexedVertices,
                            Intersection, MinimiseQEF, RectangularPrism, Sphere, Torus, Translate,
                            Union, Difference, DemoSource, _abi as A)
from isosurface_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DIGESTS = json.loads((Path(__file__).resolve().parent / "golden" / "mesh_digests.json").read_text())


def assert_same_mesh(gpu, ref):
    assert len(gpu.vertices) == len(ref.vertices)
    assert len(gpu.indices) == len(ref.indices)
    assert np.array_equal(gpu.indices, ref.indices)
    if gpu.vertices.tobytes() != ref.vertices.tobytes():
        diff = np.nonzero(np.any(gpu.vertices.view(np.uint32) != ref.vertices.view(np.uint32),
                                 axis=1))[0]
        raise AssertionError(f"{len(diff)} vertices differ, first {diff[:5]}: "
                             f"gpu {gpu.vertices[diff[:3]]} ref {ref.vertices[diff[:3]]}")


def dc(ctx, src, size, eps=1e-6):
    return DualContouring(size, MinimiseQEF(), ctx).extract_mesh(CentralDifference(src, eps))


SOURCES = {
    "sphere": lambda: W.sphere_source(),
    "box": lambda: DemoSource(RectangularPrism((0.21, 0.17, 0.29))),
    "rotated_feature_box_sphere": lambda: DemoSource(
        Intersection(RectangularPrism((0.3, 0.3, 0.3)), Sphere(0.38))),
    "cylinder": lambda: DemoSource(Cylinder(0.23, 0.3)),
    "torus": lambda: DemoSource(Torus(0.28, 0.09)),
    "box_minus_sphere": lambda: DemoSource(Difference(Sphere(0.3),
                                                      RectangularPrism((0.33, 0.2, 0.25)))),
    "boxes32": lambda: W.boxes_union(32, seed=4),
}


@pytest.mark.parametrize("name", sorted(SOURCES))
@pytest.mark.parametrize("size", [4, 9, 24, 33])
def test_dc_matches_oracle(ctx, name, size):
    src = SOURCES[name]()
    gpu = dc(ctx, src, size)
    ref = O.extract_dc(src.program(), O.reference_desc(size), 1e-6)
    print(f"{name} size={size} V={len(ref.vertices)} T={len(ref.indices) // 3}", flush=True)
    assert_same_mesh(gpu, ref)


@pytest.mark.parametrize("size", [2, 3])
def test_dc_tiny_lattices(ctx, size):
    # size 2 has one dual point per axis in x and y: no dual cube, empty mesh.
    src = SOURCES["sphere"]()
    gpu = dc(ctx, src, size)
    ref = O.extract_dc(src.program(), O.reference_desc(size), 1e-6)
    assert_same_mesh(gpu, ref)


@pytest.mark.parametrize("eps", [1e-6, 1e-4, 1e-2])
def test_dc_epsilon(ctx, eps):
    src = SOURCES["rotated_feature_box_sphere"]()
    gpu = dc(ctx, src, 20, eps)
    ref = O.extract_dc(src.program(), O.reference_desc(20), eps)
    assert_same_mesh(gpu, ref)


def test_dc_offset_lattice(ctx):
    # A lattice that is not the reference's unit cube: origin and spacing.
    src = Union(Translate((0.1, -0.2, 0.05), RectangularPrism((0.4, 0.3, 0.5))), Sphere(0.45))
    d = O.lattice_desc(23, 19, 27, origin=(-1.0, -0.9, -1.1), spacing=(0.09, 0.1, 0.085), source=0)
    ref = O.extract_dc(src.program(), d, 1e-6)
    ctx.upload_program(src.program())
    gd = A.IsoGridDesc()
    gd.nx, gd.ny, gd.nz = 23, 19, 27
    for i, (o, s) in enumerate(zip((-1.0, -0.9, -1.1), (0.09, 0.1, 0.085))):
        gd.origin[i] = o
        gd.spacing[i] = s
    gd.source = A.ISO_SOURCE_PROGRAM
    gpu = ctx.extract_dc(gd, 1e-6)
    assert len(ref.vertices) > 0
    assert_same_mesh(gpu, ref)


@pytest.mark.parametrize("seed", range(6))
def test_dc_random_csg(ctx, seed):
    src = DemoSource(W.random_csg(W.SplitMix64(100 + seed), 4 + seed))
    gpu = dc(ctx, src, 21)
    ref = O.extract_dc(src.program(), O.reference_desc(21), 1e-6)
    assert_same_mesh(gpu, ref)


@pytest.mark.parametrize("name", ["dc_sphere_r0.35_size33", "dc_boxes32_size65",
                                  "dc_boxes32_size128", "dc_spheres64_minus_torus_size65"])
def test_dc_matches_golden_digests(ctx, name):
    g = DIGESTS[name]
    if "boxes32" in name:
        src = W.boxes_union(32, seed=4)
    elif "sphere_r0.35" in name:
        src = W.sphere_source()
    else:
        src = W.spheres_minus_torus()[0]
    mesh = dc(ctx, src, g["size"])
    assert len(mesh.vertices) == g["vertices"] and len(mesh.indices) == 3 * g["triangles"]
    assert hashlib.sha256(mesh.vertices.astype("<f4").tobytes()).hexdigest() == g["vertices_sha256"]
    assert hashlib.sha256(mesh.indices.astype("<u4").tobytes()).hexdigest() == g["indices_sha256"]


def test_dc_through_the_reference_api(ctx):
    vertices, indices = [], []
    src = W.boxes_union(32, seed=4)
    DualContouring(33, MinimiseQEF(), ctx).extract(CentralDifference(src),
                                                   IndexedVertices(vertices, indices))
    ref = O.extract_dc(src.program(), O.reference_desc(33), 1e-6)
    assert np.array(vertices, np.float32).tobytes() == ref.vertices.tobytes()
    assert np.array_equal(np.array(indices, np.uint32), ref.indices)


def test_dc_then_mc_share_a_context(ctx):
    # DC and MC reuse the context's scratch; interleaving must not disturb either.
    from isosurface_amd import MarchingCubes
    src = W.boxes_union(32, seed=4)
    a = dc(ctx, src, 40)
    m = MarchingCubes(40, ctx).extract_mesh(src)
    b = dc(ctx, src, 40)
    assert a.vertices.tobytes() == b.vertices.tobytes()
    assert np.array_equal(a.indices, b.indices)
    ref = O.extract_mc(src.program(), desc=O.reference_desc(40))
    assert m.vertices.tobytes() == ref.vertices.tobytes()


def test_dc_error_paths(ctx):
    lib = ctx.lib
    src = W.sphere_source()
    ctx.upload_program(src.program())
    d = ctx.reference_desc(9, A.ISO_SOURCE_GRID)
    out = A.IsoMeshOut(None, None, 0, 0, 0, 0, A.ISO_MEMORY_HOST, 0)
    assert lib.iso_extract_dc(ctx.handle, C.byref(d), C.c_float(1e-6), C.byref(out)) == \
        A.ISO_ERR_INVALID_ARGUMENT
    d = ctx.reference_desc(9, A.ISO_SOURCE_PROGRAM)
    for bad in (-1e-6, float("nan"), float("inf")):  # 0: the program's own normals
        assert lib.iso_extract_dc(ctx.handle, C.byref(d), C.c_float(bad), C.byref(out)) == \
            A.ISO_ERR_INVALID_ARGUMENT
    # Host output buffers too small: the mesh is kept for iso_mesh_fetch.
    st = lib.iso_extract_dc(ctx.handle, C.byref(d), C.c_float(1e-6), C.byref(out))
    assert st == A.ISO_ERR_CAPACITY
    nv, ni = int(out.vertex_count), int(out.index_count)
    ref = O.extract_dc(src.program(), O.reference_desc(9), 1e-6)
    assert (nv, ni) == (len(ref.vertices), len(ref.indices))
    with pytest.raises(A.IsoError):  # Difference has no analytic normals
        DualContouring(9, MinimiseQEF(), ctx).extract_mesh(SOURCES["box_minus_sphere"]())


HERMITE = ["sphere", "box", "cylinder", "torus", "boxes32"]


@pytest.mark.parametrize("name", HERMITE)
@pytest.mark.parametrize("size", [9, 24, 33])
def test_dc_hermite_normals_match_oracle(ctx, name, size):
    # DualContouring over a source that is its own HermiteSource (epsilon 0:
    # the program's analytic normals, dual_contouring.rs:65-69, source.rs:43-48).
    src = SOURCES[name]()
    gpu = DualContouring(size, MinimiseQEF(), ctx).extract_mesh(src)
    ref = O.extract_dc(src.program(), O.reference_desc(size), 0.0)
    print(f"{name} size={size} V={len(ref.vertices)} T={len(ref.indices) // 3}", flush=True)
    assert len(ref.vertices) > 0
    assert_same_mesh(gpu, ref)


def test_dc_hermite_needs_analytic_normals(ctx):
    src = SOURCES["box_minus_sphere"]()  # Difference has no HermiteSource form
    ctx.upload_program(src.program())
    desc = ctx.reference_desc(9, A.ISO_SOURCE_PROGRAM)
    out = A.IsoMeshOut(None, None, 0, 0, 0, 0, A.ISO_MEMORY_HOST, 0)
    assert ctx.lib.iso_extract_dc(ctx.handle, C.byref(desc), C.c_float(0.0), C.byref(out)) == \
        A.ISO_ERR_PROGRAM
    assert ctx.lib.iso_extract_dc(ctx.handle, C.byref(desc), C.c_float(-1.0), C.byref(out)) == \
        A.ISO_ERR_INVALID_ARGUMENT
    with pytest.raises(TypeError):
        DualContouring(9, MinimiseQEF(), ctx).extract_mesh(object())


@pytest.mark.parametrize("name", ["dc_hermite_sphere_r0.35_size33", "dc_hermite_boxes32_size128"])
def test_dc_hermite_matches_golden_digests(ctx, name):
    g = DIGESTS[name]
    src = W.boxes_union(32, seed=4) if "boxes32" in name else W.sphere_source()
    mesh = DualContouring(g["size"], MinimiseQEF(), ctx).extract_mesh(src)
    assert len(mesh.vertices) == g["vertices"] and len(mesh.indices) == 3 * g["triangles"]
    assert hashlib.sha256(mesh.vertices.astype("<f4").tobytes()).hexdigest() == g["vertices_sha256"]
    assert hashlib.sha256(mesh.indices.astype("<u4").tobytes()).hexdigest() == g["indices_sha256"]
