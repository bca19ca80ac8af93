"""MarchingCubes::<Directed> on the GPU against the oracle, through the C ABI.

iso_program_sample_vector must reproduce the reference's own directed
known-answer tests and the oracle's f32 bits at random points;
iso_extract_mc_directed must reproduce the oracle's directed extraction
bit for bit (vertex order and f32 bits, index lists) for program sources,
directed grids, host and device outputs and a degenerate lattice."""
import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest

from isosurface_amd import (Cylinder, DemoSource, Difference, Directed, IndexedVertices,
                            Intersection, MarchingCubes, RectangularPrism, SmoothUnion, Sphere,
                            Torus, Translate, Union, _abi as A)
from isosurface_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

VKATS = json.loads((Path(__file__).resolve().parent / "golden" /
                    "reference_vector_kats.json").read_text())
FMAX = np.finfo(np.float32).max


def every_operator_source():
    """Every shape and every CSG operator, with translations."""
    a = Translate((0.1, 0.0, 0.0), Sphere(0.25))
    b = Translate((-0.1, 0.05, 0.0), Torus(0.2, 0.05))
    c = Cylinder(0.1, 0.3)
    d = RectangularPrism((0.3, 0.15, 0.2))
    e = Translate((0.0, 0.1, 0.1), Sphere(0.2))
    return DemoSource(Union(Difference(c, Union(a, b)), Intersection(d, e)))


SOURCES = {
    "sphere": lambda: W.sphere_source(),
    "spheres_minus_torus": lambda: W.spheres_minus_torus()[0],
    "boxes": lambda: W.boxes_union(),
    "every_operator": every_operator_source,
}


def same_bits_or_both_nan(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    nan = np.isnan(b)
    return np.array_equal(np.isnan(a), nan) and a[~nan].tobytes() == b[~nan].tobytes()


def sample_vector_program(ctx, program, pts):
    ctx.upload_program(program)
    pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    out = np.empty_like(pts)
    ctx.check(ctx.lib.iso_program_sample_vector(ctx.handle, pts.ctypes.data, len(pts),
                                                out.ctypes.data), "iso_program_sample_vector")
    return out


def test_sample_vector_reference_kats(ctx):
    for k in VKATS:
        got = sample_vector_program(ctx, k["program"], np.array([k["point"]], np.float32))[0]
        want = np.array([FMAX if e == "MAX" else e for e in k["expected"]], np.float32)
        assert np.all(got == want), (k["where"], got, want)


@pytest.mark.parametrize("name", sorted(SOURCES))
def test_sample_vector_matches_oracle(ctx, name):
    src = SOURCES[name]()
    prog = src.program()
    rng = np.random.default_rng(7)
    pts = rng.uniform(-0.1, 1.1, size=(4000, 3)).astype(np.float32)
    # points on the axes through the shapes' centres (the torus's a.x == 0
    # and a.y == 0 branches) and on lattice coordinates
    pts[:200, 0] = 0.5
    pts[200:400, 1] = 0.5
    pts[400:600] = (np.arange(600)[400:600, None] % 17 / np.float32(16)).astype(np.float32)
    got = ctx.sample_vector(src, pts)
    want = np.stack([O.sample_vector(prog, *p) for p in pts])
    assert same_bits_or_both_nan(got, want)


@pytest.mark.parametrize("name", sorted(SOURCES))
@pytest.mark.parametrize("size", [2, 3, 5, 17, 65, 66])
def test_program_extraction_matches_oracle(ctx, name, size):
    src = SOURCES[name]()
    ctx.upload_program(src.program())
    desc = ctx.reference_desc(size, A.ISO_SOURCE_PROGRAM)
    mesh = ctx.extract_mc(desc, directed=True)
    ref = O.extract_mc(src.program(), desc=O.reference_desc(size), directed=True)
    assert mesh.vertices.tobytes() == ref.vertices.tobytes()
    assert np.array_equal(mesh.indices, ref.indices)


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_grid_extraction_matches_oracle(ctx, seed):
    # Random directed values with the special ones the shapes produce:
    # f32::MAX, -f32::MAX, zeros of both signs, NaN.
    rng = np.random.default_rng(100 + seed)
    nz, ny, nx = 23, 37, 70
    g3 = rng.normal(size=(nz, ny, nx, 3)).astype(np.float32)
    r = rng.random(g3.shape)
    g3[r < 0.05] = FMAX
    g3[(r >= 0.05) & (r < 0.08)] = -FMAX
    g3[(r >= 0.08) & (r < 0.10)] = 0.0
    g3[(r >= 0.10) & (r < 0.11)] = -0.0
    g3[(r >= 0.11) & (r < 0.115)] = np.nan
    d = A.IsoGridDesc()
    d.nx, d.ny, d.nz = nx, ny, nz
    origin, spacing = (0.25, -1.0, 0.5), (0.5, 0.25, 1.0)
    for i in range(3):
        d.origin[i], d.spacing[i] = origin[i], spacing[i]
    d.source = A.ISO_SOURCE_GRID
    ctx.upload_grid_directed(g3)
    mesh = ctx.extract_mc(d, directed=True)
    ref = O.extract_mc(grid=g3, desc=O.lattice_desc(nx, ny, nz, origin, spacing), directed=True)
    assert same_bits_or_both_nan(mesh.vertices, ref.vertices)
    assert np.array_equal(mesh.indices, ref.indices)


def test_degenerate_lattice_uses_the_reference_axis_rule(ctx):
    # Along x the coordinates 1e8 + i (i < 4) all round to 1e8, so an x
    # edge's endpoints coincide and (p_a - p_b).abs().max_component_index()
    # picks z: the crossing interpolates the z components.
    rng = np.random.default_rng(5)
    nz, ny, nx = 4, 4, 4
    g3 = rng.normal(size=(nz, ny, nx, 3)).astype(np.float32)
    origin, spacing = (1e8, 0.0, 0.0), (1.0, 0.5, 0.5)
    d = A.IsoGridDesc()
    d.nx, d.ny, d.nz = nx, ny, nz
    for i in range(3):
        d.origin[i], d.spacing[i] = origin[i], spacing[i]
    d.source = A.ISO_SOURCE_GRID
    ctx.upload_grid_directed(g3)
    mesh = ctx.extract_mc(d, directed=True)
    ref = O.extract_mc(grid=g3, desc=O.lattice_desc(nx, ny, nz, origin, spacing), directed=True)
    assert len(ref.indices) > 0
    assert same_bits_or_both_nan(mesh.vertices, ref.vertices)
    assert np.array_equal(mesh.indices, ref.indices)


def test_reference_api_and_host_vector_source(ctx):
    # MarchingCubes::<Directed>::new(65).extract(&DemoSource(Sphere(0.35)), ...)
    ref = O.extract_mc(W.sphere_source().program(), desc=O.reference_desc(65), directed=True)
    v, i = [], []
    MarchingCubes(65, ctx, distance=Directed).extract(W.sphere_source(), IndexedVertices(v, i))
    assert np.array(v, np.float32).tobytes() == ref.vertices.tobytes()
    assert np.array_equal(np.array(i, np.uint32), ref.indices)

    class HostSphere:  # a user VectorSource sampled on the host (sphere.rs:41-55)
        def sample_vector(self, x, y, z):
            a = [np.abs(c - np.float32(0.5)).astype(np.float32) for c in (x, y, z)]
            r2 = np.float32(0.35) * np.float32(0.35)
            out = []
            for k in range(3):
                o1, o2 = [a[j] for j in range(3) if j != k]
                lk = (r2 - (o1 * o1 + o2 * o2)).astype(np.float32)
                with np.errstate(invalid="ignore"):
                    out.append(np.where(lk < 0, FMAX, a[k] - np.sqrt(lk)).astype(np.float32))
            return out

    v, i = [], []
    MarchingCubes(65, ctx, distance=Directed).extract(HostSphere(), IndexedVertices(v, i))
    assert np.array(v, np.float32).tobytes() == ref.vertices.tobytes()
    assert np.array_equal(np.array(i, np.uint32), ref.indices)


def test_device_output_and_capacity(ctx):
    src = SOURCES["every_operator"]()
    ctx.upload_program(src.program())
    desc = ctx.reference_desc(65, A.ISO_SOURCE_PROGRAM)
    ref = O.extract_mc(src.program(), desc=O.reference_desc(65), directed=True)
    nv, ni = len(ref.vertices), len(ref.indices)
    dv, di = C.c_void_p(), C.c_void_p()
    ctx.check(ctx.lib.iso_device_alloc(ctx.handle, 12 * nv, C.byref(dv)), "alloc")
    ctx.check(ctx.lib.iso_device_alloc(ctx.handle, 4 * ni, C.byref(di)), "alloc")
    try:
        small = A.IsoMeshOut(dv, di, 10, 30, 0, 0, A.ISO_MEMORY_DEVICE, 0)
        assert ctx.lib.iso_extract_mc_directed(ctx.handle, C.byref(desc), C.byref(small)) == \
            A.ISO_ERR_CAPACITY
        assert small.vertex_count == nv and small.index_count == ni
        out = A.IsoMeshOut(dv, di, nv, ni, 0, 0, A.ISO_MEMORY_DEVICE, A.ISO_MESH_GRAPH)
        ctx.check(ctx.lib.iso_extract_mc_directed(ctx.handle, C.byref(desc), C.byref(out)),
                  "extract")
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
    from isosurface_amd import Context
    with Context(0) as fresh:
        out = A.IsoMeshOut(None, None, 0, 0, 0, 0, A.ISO_MEMORY_HOST, 0)
        pdesc = fresh.reference_desc(9, A.ISO_SOURCE_PROGRAM)
        gdesc = fresh.reference_desc(9, A.ISO_SOURCE_GRID)
        assert fresh.lib.iso_extract_mc_directed(fresh.handle, C.byref(pdesc), C.byref(out)) == \
            A.ISO_ERR_NO_SOURCE
        assert fresh.lib.iso_extract_mc_directed(fresh.handle, C.byref(gdesc), C.byref(out)) == \
            A.ISO_ERR_NO_SOURCE
        # a scalar grid is not a directed one
        fresh.upload_grid(np.ones((10, 9, 9), np.float32))
        assert fresh.lib.iso_extract_mc_directed(fresh.handle, C.byref(gdesc), C.byref(out)) == \
            A.ISO_ERR_NO_SOURCE
        fresh.upload_grid_directed(np.ones((4, 4, 4, 3), np.float32))
        assert fresh.lib.iso_extract_mc_directed(fresh.handle, C.byref(gdesc), C.byref(out)) == \
            A.ISO_ERR_INVALID_ARGUMENT
        fresh.upload_program(SmoothUnion(Sphere(0.2), Sphere(0.3), 0.05).program())
        assert fresh.lib.iso_extract_mc_directed(fresh.handle, C.byref(pdesc), C.byref(out)) == \
            A.ISO_ERR_PROGRAM
        pts = np.zeros((1, 3), np.float32)
        vals = np.zeros((1, 3), np.float32)
        assert fresh.lib.iso_program_sample_vector(fresh.handle, pts.ctypes.data, 1,
                                                   vals.ctypes.data) == A.ISO_ERR_PROGRAM
        # the scalar path still accepts the smooth union
        fresh.extract_mc(pdesc)
