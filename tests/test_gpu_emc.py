"""ExtendedMarchingCubes on the GPU against the oracle, through the C ABI.

iso_extract_emc must reproduce the oracle's extraction bit for bit: vertex
order (feature vertices first in their cells) and f32 bits (crossings and
QEF feature points), and the index lists after the feature-edge flips —
including the degenerate faces the reference's flips make where a cell's
two loops share a diagonal.  Sources with the shapes' own normals and with
CentralDifference normals, several sizes, host and device outputs."""
import ctypes as C

import numpy as np
import pytest

from isosurface_amd import (CentralDifference, DemoSource, ExtendedMarchingCubes, IndexedVertices,
                            RectangularPrism, SmoothUnion, Sphere, Torus, Translate, Union,
                            Cylinder, _abi as A)
from isosurface_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def cube():
    return DemoSource(Translate((0.013, 0.021, -0.017), RectangularPrism((0.2, 0.2, 0.2))))


def analytic_union():
    # shapes and unions only: every one has a HermiteSource
    return DemoSource(Union(Union(Translate((0.1, 0.0, 0.05), RectangularPrism((0.15, 0.15, 0.15))),
                                  Translate((-0.12, 0.05, 0.0), Sphere(0.18))),
                            Union(Translate((0.0, -0.1, -0.1), Torus(0.2, 0.06)),
                                  Translate((0.0, 0.15, 0.1), Cylinder(0.08, 0.2)))))


CASES = {
    "sphere": (lambda: W.sphere_source(), 0.0),
    "cube": (cube, 0.0),
    "analytic_union": (analytic_union, 0.0),
    "boxes_cd": (lambda: W.boxes_union(), 1e-4),
    "spheres_minus_torus_cd": (lambda: W.spheres_minus_torus()[0], 1e-4),
    "random_csg_cd": (lambda: W.random_csg(W.SplitMix64(3), 6), 1e-4),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("size", [2, 5, 17, 33, 65])
def test_matches_oracle(ctx, name, size):
    make, eps = CASES[name]
    src = make()
    ctx.upload_program(src.program())
    desc = ctx.reference_desc(size, A.ISO_SOURCE_PROGRAM)
    mesh = ctx.extract_emc(desc, eps)
    ref = O.extract_emc(src.program(), O.reference_desc(size), epsilon=eps)
    assert len(mesh.vertices) == len(ref.vertices) and len(mesh.indices) == len(ref.indices)
    assert mesh.vertices.tobytes() == ref.vertices.tobytes()
    assert np.array_equal(mesh.indices, ref.indices)


def test_flips_and_degenerate_faces_are_reproduced(ctx):
    # The CSG workload at 65 has feature cells, flipped fan edges and the
    # reference's degenerate faces from loops that share a diagonal.
    src = W.spheres_minus_torus()[0]
    ctx.upload_program(src.program())
    mesh = ctx.extract_emc(ctx.reference_desc(65, A.ISO_SOURCE_PROGRAM), 1e-4)
    ref = O.extract_emc(src.program(), O.reference_desc(65), epsilon=1e-4)
    f = ref.indices.reshape(-1, 3)
    degenerate = (f[:, 0] == f[:, 1]) | (f[:, 1] == f[:, 2]) | (f[:, 0] == f[:, 2])
    feature = set(np.nonzero(ref.keys[:, 3] == 3)[0].tolist())
    joined = sum(1 for t in f if sum(int(v) in feature for v in t) >= 2)
    assert degenerate.sum() > 0 and joined > 0
    assert np.array_equal(mesh.indices, ref.indices)
    assert mesh.vertices.tobytes() == ref.vertices.tobytes()


def test_reference_api(ctx):
    v, i = [], []
    ExtendedMarchingCubes(33, ctx).extract(CentralDifference(W.boxes_union(), 1e-4),
                                           IndexedVertices(v, i))
    ref = O.extract_emc(W.boxes_union().program(), O.reference_desc(33), epsilon=1e-4)
    assert np.array(v, np.float32).tobytes() == ref.vertices.tobytes()
    assert np.array_equal(np.array(i, np.uint32), ref.indices)
    v, i = [], []
    ExtendedMarchingCubes(33, ctx).extract(cube(), IndexedVertices(v, i))
    ref = O.extract_emc(cube().program(), O.reference_desc(33))
    assert np.array(v, np.float32).tobytes() == ref.vertices.tobytes()
    assert np.array_equal(np.array(i, np.uint32), ref.indices)


def test_device_output_and_capacity(ctx):
    src = cube()
    ctx.upload_program(src.program())
    desc = ctx.reference_desc(65, A.ISO_SOURCE_PROGRAM)
    ref = O.extract_emc(src.program(), O.reference_desc(65))
    nv, ni = len(ref.vertices), len(ref.indices)
    dv, di = C.c_void_p(), C.c_void_p()
    ctx.check(ctx.lib.iso_device_alloc(ctx.handle, 12 * nv, C.byref(dv)), "alloc")
    ctx.check(ctx.lib.iso_device_alloc(ctx.handle, 4 * ni, C.byref(di)), "alloc")
    try:
        small = A.IsoMeshOut(dv, di, 10, 30, 0, 0, A.ISO_MEMORY_DEVICE, 0)
        assert ctx.lib.iso_extract_emc(ctx.handle, C.byref(desc), C.c_float(0.0),
                                       C.byref(small)) == A.ISO_ERR_CAPACITY
        assert small.vertex_count == nv and small.index_count == ni
        out = A.IsoMeshOut(dv, di, nv, ni, 0, 0, A.ISO_MEMORY_DEVICE, 0)
        ctx.check(ctx.lib.iso_extract_emc(ctx.handle, C.byref(desc), C.c_float(0.0),
                                          C.byref(out)), "extract")
        verts = np.empty((nv, 3), np.float32)
        idx = np.empty(ni, np.uint32)
        ctx.lib.iso_memcpy_to_host(ctx.handle, verts.ctypes.data, dv, verts.nbytes)
        ctx.lib.iso_memcpy_to_host(ctx.handle, idx.ctypes.data, di, idx.nbytes)
    finally:
        ctx.lib.iso_device_free(ctx.handle, dv)
        ctx.lib.iso_device_free(ctx.handle, di)
    assert verts.tobytes() == ref.vertices.tobytes()
    assert np.array_equal(idx, ref.indices)


def test_error_paths(ctx):
    from isosurface_amd import Context, Difference
    with Context(0) as fresh:
        out = A.IsoMeshOut(None, None, 0, 0, 0, 0, A.ISO_MEMORY_HOST, 0)
        pdesc = fresh.reference_desc(9, A.ISO_SOURCE_PROGRAM)
        gdesc = fresh.reference_desc(9, A.ISO_SOURCE_GRID)
        assert fresh.lib.iso_extract_emc(fresh.handle, C.byref(pdesc), C.c_float(0.0),
                                         C.byref(out)) == A.ISO_ERR_NO_SOURCE
        fresh.upload_program(W.sphere_source().program())
        assert fresh.lib.iso_extract_emc(fresh.handle, C.byref(gdesc), C.c_float(0.0),
                                         C.byref(out)) == A.ISO_ERR_INVALID_ARGUMENT
        for bad in (-1.0, float("nan"), float("inf")):
            assert fresh.lib.iso_extract_emc(fresh.handle, C.byref(pdesc), C.c_float(bad),
                                             C.byref(out)) == A.ISO_ERR_INVALID_ARGUMENT
        # a difference has no analytic normals, but central differences work
        fresh.upload_program(DemoSource(Difference(Sphere(0.1), Sphere(0.3))).program())
        assert fresh.lib.iso_extract_emc(fresh.handle, C.byref(pdesc), C.c_float(0.0),
                                         C.byref(out)) == A.ISO_ERR_PROGRAM
        fresh.extract_emc(pdesc, 1e-4)
        fresh.upload_program(SmoothUnion(Sphere(0.2), Sphere(0.3), 0.05).program())
        assert fresh.lib.iso_extract_emc(fresh.handle, C.byref(pdesc), C.c_float(1e-4),
                                         C.byref(out)) == A.ISO_ERR_PROGRAM
