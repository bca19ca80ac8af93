"""The single-pass grid extraction (one launch: signs, count, look-back,
emit) against the oracle and against the three-pass path, over lattice
shapes that exercise its tiling: one tile per cell layer, several tiles per
layer (wide rows), one cell row per tile (rows of 65,536 cells), thin and
degenerate lattices, zeros and NaN, capacity-limited outputs."""
import ctypes as C

import numpy as np
import pytest

from isosurface_amd import _abi as A
from isosurface_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def desc_of(nx, ny, nz, spacing=1.0, origin=0.0):
    d = A.IsoGridDesc()
    d.nx, d.ny, d.nz = nx, ny, nz
    for i in range(3):
        d.origin[i] = origin
        d.spacing[i] = spacing
    d.source = A.ISO_SOURCE_GRID
    return d


def same(a, b) -> bool:
    return (len(a.indices) == len(b.indices) and np.array_equal(a.indices, b.indices)
            and a.vertices.tobytes() == b.vertices.tobytes())


def check(ctx, g, spacing=1.0, oracle=True):
    nz, ny, nx = g.shape
    g = np.ascontiguousarray(g, np.float32)
    ctx.upload_grid(g)
    d = desc_of(nx, ny, nz, spacing)
    one = ctx.extract_mc(d, flags=A.ISO_MESH_SINGLE_PASS)
    three = ctx.extract_mc(d)
    print(f"{nx}x{ny}x{nz}: V={len(one.vertices)} T={len(one.indices) // 3}", flush=True)
    assert same(one, three)
    if oracle:
        ref = O.extract_mc(grid=g, desc=O.lattice_desc(nx, ny, nz, spacing=(spacing,) * 3))
        assert one.vertices.tobytes() == ref.vertices.tobytes()
        assert np.array_equal(one.indices, ref.indices)
    return one


def blob_grid(nx, ny, nz, seed, scale=None):
    """A smooth random field: sums of a few waves, every sign pattern."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    f = np.zeros((nz, ny, nx), np.float64)
    for _ in range(6):
        k = rng.normal(0, 0.35 if scale is None else scale, 3)
        f += np.sin(k[0] * x + k[1] * y + k[2] * z + rng.uniform(0, 6.3))
    return (f - 0.3).astype(np.float32)


@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 2, 2), (2, 5, 3), (65, 65, 66), (66, 9, 7),
                                   (129, 33, 12), (200, 130, 9), (257, 40, 6), (1025, 70, 4),
                                   (4097, 20, 3), (300, 2, 40), (2, 300, 40)])
def test_shapes_match_oracle(ctx, shape):
    nx, ny, nz = shape
    check(ctx, blob_grid(nx, ny, nz, seed=nx * 7 + ny))


def test_rows_of_16499_cells(ctx):
    # 258 segments per row: one cell row per tile, ny - 1 tiles per layer
    g = blob_grid(16500, 5, 3, seed=3, scale=0.05)
    check(ctx, g, oracle=True)


def test_random_signs_every_case(ctx):
    # independent random signs: every case, every owner configuration
    rng = np.random.default_rng(5)
    for shape in [(70, 20, 9), (130, 3, 4), (257, 5, 3)]:
        g = rng.normal(size=shape[::-1]).astype(np.float32)
        check(ctx, g)


def test_zeros_and_nan_are_inside(ctx):
    rng = np.random.default_rng(8)
    g = rng.normal(size=(7, 30, 90)).astype(np.float32)
    g[rng.random(g.shape) < 0.1] = 0.0
    g[rng.random(g.shape) < 0.05] = -0.0
    g[rng.random(g.shape) < 0.05] = np.nan
    nz, ny, nx = g.shape
    ctx.upload_grid(g)
    d = desc_of(nx, ny, nz)
    one = ctx.extract_mc(d, flags=A.ISO_MESH_SINGLE_PASS)
    three = ctx.extract_mc(d)
    assert np.array_equal(one.indices, three.indices)
    assert np.array_equal(np.isnan(one.vertices), np.isnan(three.vertices))
    ok = ~np.isnan(one.vertices)
    assert one.vertices[ok].tobytes() == three.vertices[ok].tobytes()


def test_csg_grid_size257_matches_golden_and_repeats(ctx):
    import hashlib
    import json
    from pathlib import Path
    g = json.loads((Path(__file__).resolve().parent / "golden" /
                    "mesh_digests.json").read_text())["spheres64_minus_torus_size257"]
    src = W.spheres_minus_torus()[0]
    ctx.upload_program(src.program())
    d = ctx.reference_desc(257, A.ISO_SOURCE_PROGRAM)
    grid = np.empty((d.nz, d.ny, d.nx), np.float32)
    ctx.check(ctx.lib.iso_program_sample_lattice(ctx.handle, C.byref(d), grid.ctypes.data,
                                                 A.ISO_MEMORY_HOST), "sample_lattice")
    ctx.upload_grid(grid)
    d.source = A.ISO_SOURCE_GRID
    for _ in range(3):  # epochs advance; every repeat is the same mesh
        mesh = ctx.extract_mc(d, flags=A.ISO_MESH_SINGLE_PASS)
        assert len(mesh.vertices) == g["vertices"]
        assert hashlib.sha256(mesh.vertices.astype("<f4").tobytes()).hexdigest() == \
            g["vertices_sha256"]
        assert hashlib.sha256(mesh.indices.astype("<u4").tobytes()).hexdigest() == \
            g["indices_sha256"]


def test_device_output_capacity(ctx):
    g = blob_grid(90, 50, 20, seed=11)
    nz, ny, nx = g.shape
    ctx.upload_grid(g)
    d = desc_of(nx, ny, nz)
    full = ctx.extract_mc(d)
    nv, ni = len(full.vertices), len(full.indices)
    sp = A.ISO_MESH_SINGLE_PASS
    lib = ctx.lib
    vbuf = C.c_void_p()
    ibuf = C.c_void_p()
    ctx.check(lib.iso_device_alloc(ctx.handle, max(nv, 1) * 12, C.byref(vbuf)), "alloc")
    ctx.check(lib.iso_device_alloc(ctx.handle, max(ni, 1) * 4, C.byref(ibuf)), "alloc")
    try:
        # too small: counts reported, nothing written past the capacity
        out = A.IsoMeshOut(vbuf.value, ibuf.value, nv // 2, ni // 2, 0, 0, A.ISO_MEMORY_DEVICE, sp)
        assert lib.iso_extract_mc(ctx.handle, C.byref(d), C.byref(out)) == A.ISO_ERR_CAPACITY
        assert (out.vertex_count, out.index_count) == (nv, ni)
        out = A.IsoMeshOut(vbuf.value, ibuf.value, nv, ni, 0, 0, A.ISO_MEMORY_DEVICE, sp)
        ctx.check(lib.iso_extract_mc(ctx.handle, C.byref(d), C.byref(out)), "extract")
        v = np.empty((nv, 3), np.float32)
        i = np.empty(ni, np.uint32)
        ctx.check(lib.iso_memcpy_to_host(ctx.handle, v.ctypes.data, vbuf.value, nv * 12), "copy")
        ctx.check(lib.iso_memcpy_to_host(ctx.handle, i.ctypes.data, ibuf.value, ni * 4), "copy")
        assert v.tobytes() == full.vertices.tobytes()
        assert np.array_equal(i, full.indices)
    finally:
        lib.iso_device_free(ctx.handle, vbuf)
        lib.iso_device_free(ctx.handle, ibuf)
