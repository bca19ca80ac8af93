"""PointCloud and the normal sinks on the GPU against the oracle
(point_cloud.rs:50-63, extractor.rs:46-69,97-127, source.rs:82-94): the
same points in the same order and the same normal bits."""
import ctypes as C

import numpy as np
import pytest

from isosurface_amd import (CentralDifference, IndexedInterleavedNormals, IndexedVertices,
                            MarchingCubes, OnlyInterleavedNormals, PointCloud, OnlyVertices,
                            _abi as A)
from isosurface_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("size,which", [(17, "sphere"), (65, "sphere"), (33, "csg"), (129, "csg"),
                                        (200, "csg")])
def test_point_cloud_matches_oracle(ctx, size, which):
    src = W.sphere_source() if which == "sphere" else W.spheres_minus_torus()[0]
    got = PointCloud(size, ctx).extract_points(src)
    want = O.extract_point_cloud(src.program(), desc=O.reference_desc(size, 0))
    print(size, which, len(got), flush=True)
    assert len(got) == len(want) > 0
    assert got.tobytes() == want.tobytes()


def test_point_cloud_from_a_host_sampled_grid(ctx):
    src = W.spheres_minus_torus()[0]
    g = O.sample_lattice(src.program(), O.reference_desc(48, 0))
    desc = ctx.reference_desc(48, A.ISO_SOURCE_GRID)
    ctx.upload_grid(g)
    got = ctx.extract_point_cloud(desc)
    want = O.extract_point_cloud(grid=g, desc=O.reference_desc(48, 1))
    assert got.tobytes() == want.tobytes()


def test_point_cloud_extractor_protocol_and_capacity(ctx):
    src = W.sphere_source()
    pc = PointCloud(40, ctx)
    sink = OnlyVertices()
    pc.extract(src, sink)  # first call: capacity 1, grown through iso_mesh_fetch
    want = O.extract_point_cloud(src.program(), desc=O.reference_desc(40, 0))
    assert np.array_equal(np.asarray(sink.vertices, np.float32).reshape(-1, 3), want)
    # device output too small: the count is reported with ISO_ERR_CAPACITY
    ctx.upload_program(src.program())
    desc = ctx.reference_desc(40, A.ISO_SOURCE_PROGRAM)
    buf = C.c_void_p()
    ctx.check(ctx.lib.iso_device_alloc(ctx.handle, 12, C.byref(buf)), "alloc")
    out = A.IsoMeshOut(buf, None, 1, 0, 0, 0, A.ISO_MEMORY_DEVICE, 0)
    assert ctx.lib.iso_extract_point_cloud(ctx.handle, C.byref(desc), C.byref(out)) == \
        A.ISO_ERR_CAPACITY
    assert out.vertex_count == len(want) and out.index_count == 0
    ctx.lib.iso_device_free(ctx.handle, buf)


@pytest.mark.parametrize("eps", [1e-6, 1e-3])
def test_normals_match_oracle_at_mesh_vertices(ctx, eps):
    src = W.spheres_minus_torus()[0]
    mesh = MarchingCubes(65, ctx).extract_mesh(src)
    cd = CentralDifference(src, eps)
    got = ctx.sample_normals(cd, mesh.vertices)
    want = O.sample_normals(src.program(), mesh.vertices, eps)
    assert got.tobytes() == want.tobytes()


def test_interleaved_normal_sinks(ctx):
    src = W.sphere_source()
    cd = CentralDifference(src, 1e-4)
    ref = IndexedVertices()
    MarchingCubes(33, ctx).extract(src, ref)
    v = np.asarray(ref.vertices, np.float32).reshape(-1, 3)
    sink = IndexedInterleavedNormals(source=cd, context=ctx)
    MarchingCubes(33, ctx).extract(src, sink)
    out = np.asarray(sink.vertices, np.float32).reshape(-1, 6)
    assert out[:, :3].tobytes() == v.tobytes()
    assert out[:, 3:].tobytes() == O.sample_normals(src.program(), v, 1e-4).tobytes()
    assert sink.indices == ref.indices
    only = OnlyInterleavedNormals(source=cd, context=ctx)
    MarchingCubes(33, ctx).extract(src, only)
    assert np.asarray(only.vertices, np.float32).tobytes() == out.tobytes()
    with pytest.raises(TypeError):
        IndexedInterleavedNormals(source=src)


def test_normals_error_paths(ctx):
    lib = ctx.lib
    ctx.upload_program(W.sphere_source().program())
    p = np.zeros((4, 3), np.float32)
    n = np.zeros((4, 3), np.float32)
    assert lib.iso_sample_normals(ctx.handle, p.ctypes.data, 4, C.c_float(0.0), n.ctypes.data,
                                  A.ISO_MEMORY_HOST) == A.ISO_ERR_INVALID_ARGUMENT
    assert lib.iso_sample_normals(ctx.handle, p.ctypes.data, 4, C.c_float(1e-3), n.ctypes.data,
                                  7) == A.ISO_ERR_INVALID_ARGUMENT
    assert lib.iso_sample_normals(ctx.handle, None, 4, C.c_float(1e-3), n.ctypes.data,
                                  A.ISO_MEMORY_HOST) == A.ISO_ERR_INVALID_ARGUMENT
    assert lib.iso_sample_normals(ctx.handle, p.ctypes.data, 0, C.c_float(1e-3), n.ctypes.data,
                                  A.ISO_MEMORY_HOST) == A.ISO_OK
