"""Batched chunks on the GPU: every chunk's mesh equals the oracle's
MarchingCubes over that chunk's lattice and program, bit for bit, whether
extracted eagerly, through a replayed HIP graph, or after autotuning."""
import ctypes as C
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from isosurface_amd import BatchedMarchingCubes, Context, IndexedVertices, chunk_desc, _abi as A
from isosurface_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DIGESTS = json.loads((Path(__file__).resolve().parent / "golden" / "mesh_digests.json").read_text())


def small_batch(n=40, seed=21):
    rng = W.SplitMix64(seed)
    sizes = [2, 3, 5, 16, 17, 31, 33, 48, 64, 70]
    chunks = []
    for i in range(n):
        size = sizes[rng.next_u64() % len(sizes)]
        src = W.random_csg(rng, 1 + int(rng.next_u64() % 8))
        d = chunk_desc(size)
        if i % 5 == 3:  # a few chunks with their own origin and anisotropic spacing
            d.origin[0], d.origin[1], d.origin[2] = 0.1, -0.05, 0.2
            d.spacing[2] = d.spacing[2] * 0.75
        chunks.append((src, d))
    return chunks


def oracle_meshes(chunks):
    return [O.extract_mc(src.program(), desc=d) for src, d in chunks]


def assert_same(meshes, refs):
    assert len(meshes) == len(refs)
    for i, (m, r) in enumerate(zip(meshes, refs)):
        assert len(m.vertices) == len(r.vertices), i
        assert np.array_equal(m.indices, r.indices), i
        assert m.vertices.tobytes() == r.vertices.tobytes(), i


def test_batch_matches_oracle_per_chunk(ctx):
    chunks = small_batch()
    meshes = BatchedMarchingCubes(ctx).extract_meshes(chunks)
    assert_same(meshes, oracle_meshes(chunks))


def test_batch_chunk_equals_single_extraction(ctx):
    chunks = small_batch(12, seed=5)
    meshes = BatchedMarchingCubes(ctx).extract_meshes(chunks)
    for (src, d), m in zip(chunks, meshes):
        ctx.upload_program(src.program())
        single = ctx.extract_mc(d)
        assert m.vertices.tobytes() == single.vertices.tobytes()
        assert np.array_equal(m.indices, single.indices)


def device_batch(ctx, n_chunks, nv, ni, flags):
    v = C.c_void_p()
    i = C.c_void_p()
    ctx.check(ctx.lib.iso_device_alloc(ctx.handle, 12 * max(nv, 1), C.byref(v)), "alloc")
    ctx.check(ctx.lib.iso_device_alloc(ctx.handle, 4 * max(ni, 3), C.byref(i)), "alloc")
    voff = np.zeros(n_chunks + 1, np.uint64)
    ioff = np.zeros(n_chunks + 1, np.uint64)
    out = A.IsoBatchOut(v, i, max(nv, 1), max(ni, 3), voff.ctypes.data, ioff.ctypes.data, 0, 0,
                        A.ISO_MEMORY_DEVICE, flags)
    return v, i, voff, ioff, out


def fetch(ctx, v, i, voff, ioff, out):
    nv, ni = int(out.vertex_count), int(out.index_count)
    vv = np.empty((nv, 3), np.float32)
    ii = np.empty(ni, np.uint32)
    ctx.lib.iso_memcpy_to_host(ctx.handle, vv.ctypes.data, v, vv.nbytes)
    ctx.lib.iso_memcpy_to_host(ctx.handle, ii.ctypes.data, i, ii.nbytes)
    n = len(voff) - 1
    return [(vv[int(voff[c]):int(voff[c + 1])], ii[int(ioff[c]):int(ioff[c + 1])])
            for c in range(n)]


@pytest.mark.parametrize("flags", [0, A.ISO_BATCH_GRAPH, A.ISO_BATCH_GRAPH | A.ISO_BATCH_AUTOTUNE])
def test_batch_device_graph_replays(ctx, flags):
    chunks = small_batch(30, seed=9)
    refs = oracle_meshes(chunks)
    nv = sum(len(r.vertices) for r in refs)
    ni = sum(len(r.indices) for r in refs)
    ctx.upload_batch(chunks)
    v, i, voff, ioff, out = device_batch(ctx, len(chunks), nv, ni, flags)
    try:
        for rep in range(4):
            out.vertex_count = out.index_count = 0
            ctx.check(ctx.lib.iso_extract_batch(ctx.handle, C.byref(out)), "iso_extract_batch")
            got = fetch(ctx, v, i, voff, ioff, out)
            for c, (r, (gv, gi)) in enumerate(zip(refs, got)):
                assert gv.tobytes() == r.vertices.tobytes(), (rep, c)
                assert np.array_equal(gi, r.indices), (rep, c)
            # eager work between replays: a single extraction on the same context
            ctx.upload_program(W.sphere_source().program())
            ctx.extract_mc(ctx.reference_desc(17, A.ISO_SOURCE_PROGRAM))
    finally:
        ctx.lib.iso_device_free(ctx.handle, v)
        ctx.lib.iso_device_free(ctx.handle, i)


def test_full_batch_workload_matches_golden(ctx):
    g = DIGESTS["batch4096_seed6"]
    chunks = [(src, chunk_desc(size)) for src, size in W.batch_chunks(4096, seed=6)]
    ctx.upload_batch(chunks)
    meshes = ctx.extract_batch(g["vertices"], 3 * g["triangles"],
                               flags=A.ISO_BATCH_AUTOTUNE)
    v = np.concatenate([m.vertices for m in meshes])
    i = np.concatenate([m.indices for m in meshes])
    assert len(v) == g["vertices"] and len(i) == 3 * g["triangles"]
    assert hashlib.sha256(v.astype("<f4").tobytes()).hexdigest() == g["vertices_sha256"]
    assert hashlib.sha256(i.astype("<u4").tobytes()).hexdigest() == g["indices_sha256"]
    # spot-check chunks directly against the oracle
    for c in (0, 1, 777, 2048, 4095):
        ref = O.extract_mc(chunks[c][0].program(), desc=chunks[c][1])
        assert meshes[c].vertices.tobytes() == ref.vertices.tobytes()
        assert np.array_equal(meshes[c].indices, ref.indices)


def test_batch_through_extractors(ctx):
    chunks = small_batch(6, seed=2)
    exs = [IndexedVertices([], []) for _ in chunks]
    BatchedMarchingCubes(ctx).extract(chunks, exs)
    for ex, r in zip(exs, oracle_meshes(chunks)):
        assert np.array(ex.vertices, np.float32).tobytes() == r.vertices.tobytes()
        assert np.array_equal(np.array(ex.indices, np.uint32), r.indices)


def test_batch_error_paths(ctx):
    lib = ctx.lib
    fresh = Context(0)
    try:
        voff = np.zeros(2, np.uint64)
        out = A.IsoBatchOut(None, None, 0, 0, voff.ctypes.data, voff.ctypes.data, 0, 0,
                            A.ISO_MEMORY_HOST, 0)
        assert lib.iso_extract_batch(fresh.handle, C.byref(out)) == A.ISO_ERR_NO_SOURCE
    finally:
        fresh.close()
    prog = W.sphere_source().program()
    nodes = __import__("isosurface_amd").to_nodes(prog)
    descs = (A.IsoChunkDesc * 1)()
    descs[0].grid = chunk_desc(8)
    descs[0].node_offset = 0
    descs[0].node_count = len(prog) + 1  # beyond the node array
    assert lib.iso_batch_upload(ctx.handle, C.cast(nodes, C.c_void_p), len(prog), descs, 1) == \
        A.ISO_ERR_INVALID_ARGUMENT
    descs[0].node_count = len(prog)
    descs[0].grid.source = A.ISO_SOURCE_GRID
    assert lib.iso_batch_upload(ctx.handle, C.cast(nodes, C.c_void_p), len(prog), descs, 1) == \
        A.ISO_ERR_INVALID_ARGUMENT
    descs[0].grid.source = A.ISO_SOURCE_PROGRAM
    assert lib.iso_batch_upload(ctx.handle, C.cast(nodes, C.c_void_p), len(prog), descs, 1) == 0
    voff = np.zeros(2, np.uint64)
    ioff = np.zeros(2, np.uint64)
    out = A.IsoBatchOut(None, None, 0, 0, voff.ctypes.data, ioff.ctypes.data, 0, 0,
                        A.ISO_MEMORY_HOST, 0)
    assert lib.iso_extract_batch(ctx.handle, C.byref(out)) == A.ISO_ERR_CAPACITY
    ref = O.extract_mc(prog, desc=chunk_desc(8))
    assert int(out.vertex_count) == len(ref.vertices) and int(voff[1]) == len(ref.vertices)
    assert int(ioff[1]) == len(ref.indices)
    out.vertex_offsets = None
    assert lib.iso_extract_batch(ctx.handle, C.byref(out)) == A.ISO_ERR_INVALID_ARGUMENT
