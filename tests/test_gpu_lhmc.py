"""GPU linear-hashed octree marching cubes against the oracle.

LinearHashedMarchingCubes::new(max_depth).extract with IndexedVertices.  The
reference visits primal vertices in std HashMap order (random per process);
the GPU visits them in ascending Morton order, and must then equal the
oracle run in that order bit for bit.  Against other visiting orders the
comparison is canonical: the same MortonKey edges, the same triangles, and
positions within the few ulps that the first-encountering cube's edge
orientation can move them."""
import ctypes as C
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from isosurface_amd import (Context, IndexedVertices, LinearHashedMarchingCubes, _abi as A)
from isosurface_amd import workloads as W
from isosurface_amd.sources import (Cylinder, DemoSource, RectangularPrism, Sphere, Torus,
                                    Translate, Union)
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DIGESTS = json.loads((Path(__file__).resolve().parent / "golden" / "mesh_digests.json").read_text())

SOURCES = {
    "sphere": lambda: W.sphere_source(),
    "box": lambda: DemoSource(RectangularPrism((0.21, 0.17, 0.29))),
    "torus": lambda: DemoSource(Torus(0.28, 0.09)),
    "cyl_union": lambda: DemoSource(Union(Cylinder(0.2, 0.3),
                                          Translate((0.2, 0.1, -0.1), Sphere(0.15)))),
    "boxes32": lambda: W.boxes_union(32, seed=4),
    "corner_sphere": lambda: Sphere(0.3),
}


def assert_same_mesh(gpu, ref):
    assert len(gpu.vertices) == len(ref.vertices)
    assert len(gpu.indices) == len(ref.indices)
    assert np.array_equal(gpu.indices, ref.indices)
    assert gpu.vertices.tobytes() == ref.vertices.tobytes()


@pytest.mark.parametrize("name", sorted(SOURCES))
@pytest.mark.parametrize("depth", [0, 1, 2, 3, 4, 5, 6])
def test_lhmc_matches_oracle(ctx, name, depth):
    src = SOURCES[name]()
    print(f"{name} depth={depth}", flush=True)
    gpu = LinearHashedMarchingCubes(depth, ctx).extract_mesh(src)
    ref = O.extract_lhmc(src.program(), depth, 0)
    assert_same_mesh(gpu, ref)


@pytest.mark.parametrize("name", ["lhmc_boxes32_depth7", "lhmc_sphere_r0.35_depth8"])
def test_lhmc_matches_golden_digests(ctx, name):
    g = DIGESTS[name]
    src = W.boxes_union(32, seed=4) if "boxes32" in name else W.sphere_source()
    mesh = LinearHashedMarchingCubes(g["size"], ctx).extract_mesh(src)
    assert len(mesh.vertices) == g["vertices"] and len(mesh.indices) == 3 * g["triangles"]
    assert hashlib.sha256(mesh.vertices.astype("<f4").tobytes()).hexdigest() == g["vertices_sha256"]
    assert hashlib.sha256(mesh.indices.astype("<u4").tobytes()).hexdigest() == g["indices_sha256"]


def canonical(vertices, indices, keys):
    order = np.lexsort((keys[:, 1], keys[:, 0]))
    rank = np.empty(len(order), np.int64)
    rank[order] = np.arange(len(order))
    tris = rank[indices.astype(np.int64)].reshape(-1, 3)
    tris = tris[np.lexsort(tris.T[::-1])]
    return vertices[order], keys[order], tris


@pytest.mark.parametrize("seed", [7, 11])
def test_lhmc_equals_other_visiting_orders_canonically(ctx, seed):
    src = W.boxes_union(32, seed=4)
    gpu = LinearHashedMarchingCubes(6, ctx).extract_mesh(src)
    ref0 = O.extract_lhmc(src.program(), 6, 0)
    refs = O.extract_lhmc(src.program(), 6, seed)
    # the GPU mesh carries the sorted-order oracle's MortonKeys vertex for vertex
    va, ka, ta = canonical(gpu.vertices, gpu.indices, ref0.keys)
    vb, kb, tb = canonical(refs.vertices, refs.indices, refs.keys)
    assert np.array_equal(ka, kb) and np.array_equal(ta, tb)
    assert np.allclose(va, vb, rtol=1e-6, atol=1e-7)


def test_lhmc_through_the_reference_api(ctx):
    vertices, indices = [], []
    src = SOURCES["torus"]()
    LinearHashedMarchingCubes(5, ctx).extract(src, IndexedVertices(vertices, indices))
    ref = O.extract_lhmc(src.program(), 5, 0)
    assert np.array(vertices, np.float32).tobytes() == ref.vertices.tobytes()
    assert np.array_equal(np.array(indices, np.uint32), ref.indices)


def test_lhmc_deepest_supported_depth_is_closed(ctx):
    mesh = LinearHashedMarchingCubes(9, ctx).extract_mesh(W.sphere_source())
    t = mesh.indices.reshape(-1, 3).astype(np.int64)
    e = np.sort(np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]), axis=1)
    _, counts = np.unique(e, axis=0, return_counts=True)
    assert np.all(counts == 2)
    assert len(mesh.vertices) - len(counts) + len(t) == 2
    r = np.linalg.norm(mesh.vertices.astype(np.float64) - 0.5, axis=1)
    assert np.all(np.abs(r - 0.35) < 1.0 / 512)


def test_lhmc_error_paths(ctx):
    lib = ctx.lib
    out = A.IsoMeshOut(None, None, 0, 0, 0, 0, A.ISO_MEMORY_HOST, 0)
    ctx.upload_program(W.sphere_source().program())
    for bad in (-1, 21, 64):
        assert lib.iso_extract_lhmc(ctx.handle, bad, C.byref(out)) == A.ISO_ERR_INVALID_ARGUMENT
    assert lib.iso_extract_lhmc(ctx.handle, 4, C.byref(out)) == A.ISO_ERR_CAPACITY
    ref = O.extract_lhmc(W.sphere_source().program(), 4, 0)
    assert (int(out.vertex_count), int(out.index_count)) == (len(ref.vertices), len(ref.indices))
    fresh = Context(0)
    try:
        assert lib.iso_extract_lhmc(fresh.handle, 4, C.byref(out)) == A.ISO_ERR_NO_SOURCE
    finally:
        fresh.close()
    with pytest.raises(ValueError):
        LinearHashedMarchingCubes(21, ctx)


SQRT3 = float(np.sqrt(3.0))
DEEP = {
    # a sphere through the domain's origin corner: the octree refines down
    # the corner nodes, whose keys lie at each depth's level split (the first
    # keys of depth d have level d - 1: 1 key below depth 11, then 2, 11, 93,
    # ... 58185 at depth 16)
    "origin_sphere": [(0.05, 9), (0.02, 10), (0.01, 11), (0.004, 12), (0.002, 13),
                      (0.0005, 16)],
    # a small sphere away from the corner, refined to the deepest depth
    "small_sphere": [(1e-3, 14), (1e-4, 17), (2e-5, 20)],
}


@pytest.mark.parametrize("name,scale,depth", [(n, s, d) for n, v in DEEP.items() for s, d in v])
def test_lhmc_deep_trees_match_oracle(ctx, name, scale, depth):
    # max_depth 9 and beyond: primal vertices by the sorted path, levels by
    # the split table; bit for bit against the oracle (libm log level()).
    if name == "origin_sphere":
        src = Translate((scale, scale, scale), Sphere(scale * SQRT3))
    else:
        src = Translate((0.3, 0.4, 0.6), Sphere(scale))
    gpu = LinearHashedMarchingCubes(depth, ctx).extract_mesh(src)
    ref = O.extract_lhmc(src.program(), depth, 0)
    print(f"{name} r={scale} depth={depth} V={len(ref.vertices)} T={len(ref.indices) // 3}",
          flush=True)
    assert len(ref.vertices) > 0
    assert_same_mesh(gpu, ref)


def test_lhmc_dense_and_sorted_primal_paths_agree(ctx):
    # max_depth 8 lists primal vertices on a dense lattice and 9 by sorting;
    # the same source through both equals the oracle at each depth
    src = SOURCES["cyl_union"]()
    for depth in (8, 9):
        gpu = LinearHashedMarchingCubes(depth, ctx).extract_mesh(src)
        ref = O.extract_lhmc(src.program(), depth, 0)
        assert_same_mesh(gpu, ref)
