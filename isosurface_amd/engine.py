"""GPU contexts and the extractor front doors (Python mirror of the crate API).

  MarchingCubes(size).extract(source, extractor)   marching_cubes.rs:46-82
  DualContouring(size, MinimiseQEF()).extract(CentralDifference(source), extractor)
                                                   dual_contouring.rs:48-105
  LinearHashedMarchingCubes(max_depth).extract(source, extractor)
                                                   linear_hashed_marching_cubes.rs:51-84

The extraction runs on the GPU through the C ABI; the result is then
replayed into the extractor with the reference's sink protocol: every
extract_vertex first (in first-encounter order, which the GPU reproduces),
then every extract_index (mesh.rs:91-100, extractor.rs:17-20).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _abi as A
from .sources import CentralDifference, Source, to_nodes


class Extractor:
    """Sink of mesh data (extractor.rs:17-20)."""

    def extract_vertex(self, vertex) -> None:  # pragma: no cover - interface
        raise NotImplementedError

    def extract_index(self, index: int) -> None:  # pragma: no cover - interface
        raise NotImplementedError

    # Bulk forms; sinks may override them to avoid per-element calls.
    def extract_vertices(self, vertices: np.ndarray) -> None:
        for v in vertices:
            self.extract_vertex(v)

    def extract_indices(self, indices: np.ndarray) -> None:
        for i in indices:
            self.extract_index(int(i))


class IndexedVertices(Extractor):
    """xyz f32 per vertex and u32 indices (extractor.rs:72-94)."""

    def __init__(self, vertices: Optional[list] = None, indices: Optional[list] = None):
        self.vertices = vertices if vertices is not None else []
        self.indices = indices if indices is not None else []

    def extract_vertex(self, v):
        self.vertices.extend(float(c) for c in v)

    def extract_index(self, index):
        self.indices.append(int(index) & 0xFFFFFFFF)

    def extract_vertices(self, vertices):
        self.vertices.extend(np.asarray(vertices, dtype=np.float32).reshape(-1).tolist())

    def extract_indices(self, indices):
        self.indices.extend(np.asarray(indices, dtype=np.uint32).tolist())


class OnlyVertices(Extractor):
    """Vertices only; indices are discarded (extractor.rs:24-42)."""

    def __init__(self, vertices: Optional[list] = None):
        self.vertices = vertices if vertices is not None else []

    def extract_vertex(self, v):
        self.vertices.extend(float(c) for c in v)

    def extract_index(self, index):
        pass

    def extract_vertices(self, vertices):
        self.vertices.extend(np.asarray(vertices, dtype=np.float32).reshape(-1).tolist())

    def extract_indices(self, indices):
        pass


class IndexedInterleavedNormals(Extractor):
    """xyz and normal per vertex (6 f32) and u32 indices (extractor.rs:97-127);
    the normal is source.sample_normal(v) of a CentralDifference source,
    evaluated on the GPU (iso_sample_normals)."""

    def __init__(self, vertices: Optional[list] = None, indices: Optional[list] = None,
                 source: Optional[CentralDifference] = None, context: Optional["Context"] = None):
        if not isinstance(source, CentralDifference):
            raise TypeError("the normal sinks need a HermiteSource: a CentralDifference source")
        self.vertices = vertices if vertices is not None else []
        self.indices = indices if indices is not None else []
        self.source = source
        self.ctx = context

    def extract_vertex(self, v):
        self.extract_vertices(np.asarray([v], dtype=np.float32))

    def extract_index(self, index):
        self.indices.append(int(index) & 0xFFFFFFFF)

    def extract_vertices(self, vertices):
        v = np.asarray(vertices, dtype=np.float32).reshape(-1, 3)
        if not len(v):
            return
        ctx = self.ctx or default_context()
        n = ctx.sample_normals(self.source, v)
        self.vertices.extend(np.concatenate([v, n], axis=1).reshape(-1).tolist())

    def extract_indices(self, indices):
        self.indices.extend(np.asarray(indices, dtype=np.uint32).tolist())


class OnlyInterleavedNormals(IndexedInterleavedNormals):
    """xyz and normal per vertex, faces discarded (extractor.rs:46-69)."""

    def __init__(self, vertices: Optional[list] = None,
                 source: Optional[CentralDifference] = None, context: Optional["Context"] = None):
        super().__init__(vertices, None, source, context)

    def extract_index(self, index):
        pass

    def extract_indices(self, indices):
        pass


@dataclass
class Mesh:
    vertices: np.ndarray  # (V, 3) float32
    indices: np.ndarray   # (3T,) uint32


class Context:
    """One device context (iso_ctx): owns the GPU scratch buffers."""

    def __init__(self, device: int = 0):
        self.lib = A.load()
        h = C.c_void_p()
        st = self.lib.iso_ctx_create(device, None, C.byref(h))
        if st != A.ISO_OK:
            raise A.IsoError(st, "iso_ctx_create: " + self.lib.iso_status_string(st).decode())
        self.handle = h
        self._program_key = None

    def close(self):
        if self.handle:
            self.lib.iso_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- plumbing
    def check(self, st: int, what: str) -> int:
        if st != A.ISO_OK:
            msg = self.lib.iso_ctx_last_error(self.handle).decode()
            raise A.IsoError(st, f"{what}: {msg}")
        return st

    def upload_program(self, program) -> None:
        key = tuple((op, tuple(p)) for op, p in program)
        if key == self._program_key:
            return
        nodes = to_nodes(program)
        self.check(self.lib.iso_program_upload(self.handle, nodes, len(program)),
                   "iso_program_upload")
        self._program_key = key

    def upload_grid(self, values: np.ndarray) -> None:
        values = np.ascontiguousarray(values, dtype=np.float32)
        nz, ny, nx = values.shape
        self.check(self.lib.iso_grid_upload(self.handle, values.ctypes.data, nx, ny, nz),
                   "iso_grid_upload")
        self._program_key = self._program_key  # grids and programs are independent

    def sample(self, source: Source, points: np.ndarray) -> np.ndarray:
        """ScalarSource::sample_scalar at each point, on the GPU."""
        self.upload_program(source.program())
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        out = np.empty(len(pts), dtype=np.float32)
        self.check(self.lib.iso_program_sample(self.handle, pts.ctypes.data, len(pts),
                                               out.ctypes.data), "iso_program_sample")
        return out

    def sample_vector(self, source: Source, points: np.ndarray) -> np.ndarray:
        """VectorSource::sample_vector at each point, on the GPU: (n, 3) float32."""
        self.upload_program(source.program())
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        out = np.empty_like(pts)
        self.check(self.lib.iso_program_sample_vector(self.handle, pts.ctypes.data, len(pts),
                                                      out.ctypes.data), "iso_program_sample_vector")
        return out

    def upload_grid_directed(self, values: np.ndarray) -> None:
        """Directed values (nz, ny, nx, 3) for iso_extract_mc_directed of a grid."""
        values = np.ascontiguousarray(values, dtype=np.float32)
        nz, ny, nx, three = values.shape
        if three != 3:
            raise ValueError("directed grids hold 3 floats per point")
        self.check(self.lib.iso_grid_upload_directed(self.handle, values.ctypes.data, nx, ny, nz),
                   "iso_grid_upload_directed")

    def sample_normals(self, source: CentralDifference, points: np.ndarray) -> np.ndarray:
        """CentralDifference::sample_normal at each point (not normalised), on the GPU."""
        self.upload_program(source.program())
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        out = np.empty_like(pts)
        self.check(self.lib.iso_sample_normals(self.handle, pts.ctypes.data, len(pts),
                                               C.c_float(source.epsilon), out.ctypes.data,
                                               A.ISO_MEMORY_HOST), "iso_sample_normals")
        return out

    def extract_point_cloud(self, desc: A.IsoGridDesc, capacity: int = 0) -> np.ndarray:
        """iso_extract_point_cloud into a host array (grown on ISO_ERR_CAPACITY)."""
        cap = max(int(capacity), 1)
        pts = np.empty((cap, 3), dtype=np.float32)
        out = A.IsoMeshOut(pts.ctypes.data, None, cap, 0, 0, 0, A.ISO_MEMORY_HOST, 0)
        st = self.lib.iso_extract_point_cloud(self.handle, C.byref(desc), C.byref(out))
        if st == A.ISO_ERR_CAPACITY:
            n = int(out.vertex_count)
            pts = np.empty((max(n, 1), 3), dtype=np.float32)
            out = A.IsoMeshOut(pts.ctypes.data, None, max(n, 1), 0, 0, 0, A.ISO_MEMORY_HOST, 0)
            st = self.lib.iso_mesh_fetch(self.handle, C.byref(out))
        self.check(st, "iso_extract_point_cloud")
        return pts[:int(out.vertex_count)].copy()

    def reference_desc(self, size: int, source_kind: int) -> A.IsoGridDesc:
        d = A.IsoGridDesc()
        self.check(self.lib.iso_grid_desc_reference(size, source_kind, C.byref(d)),
                   "iso_grid_desc_reference")
        return d

    def extract_mc(self, desc: A.IsoGridDesc, vertex_capacity: int = 0,
                   index_capacity: int = 0, directed: bool = False, flags: int = 0) -> Mesh:
        """iso_extract_mc (iso_extract_mc_directed with directed=True) into
        host arrays (grown on ISO_ERR_CAPACITY); `flags` are ISO_MESH_*
        (A.ISO_MESH_SINGLE_PASS: one launch over a grid source instead of
        the classify / count / emit launches)."""
        vcap = max(int(vertex_capacity), 1)
        icap = max(int(index_capacity), 3)
        verts = np.empty((vcap, 3), dtype=np.float32)
        idx = np.empty(icap, dtype=np.uint32)
        out = A.IsoMeshOut(verts.ctypes.data, idx.ctypes.data, vcap, icap, 0, 0,
                           A.ISO_MEMORY_HOST, flags)
        fn = self.lib.iso_extract_mc_directed if directed else self.lib.iso_extract_mc
        st = fn(self.handle, C.byref(desc), C.byref(out))
        if st == A.ISO_ERR_CAPACITY:
            nv, ni = int(out.vertex_count), int(out.index_count)
            verts = np.empty((max(nv, 1), 3), dtype=np.float32)
            idx = np.empty(max(ni, 3), dtype=np.uint32)
            out = A.IsoMeshOut(verts.ctypes.data, idx.ctypes.data, max(nv, 1), max(ni, 3), 0, 0,
                               A.ISO_MEMORY_HOST, 0)
            st = self.lib.iso_mesh_fetch(self.handle, C.byref(out))
        self.check(st, "iso_extract_mc")
        nv, ni = int(out.vertex_count), int(out.index_count)
        return Mesh(verts[:nv].copy(), idx[:ni].copy())

    def extract_dc(self, desc: A.IsoGridDesc, epsilon: float, vertex_capacity: int = 0,
                   index_capacity: int = 0) -> Mesh:
        """iso_extract_dc into host arrays (grown on ISO_ERR_CAPACITY)."""
        vcap = max(int(vertex_capacity), 1)
        icap = max(int(index_capacity), 3)
        verts = np.empty((vcap, 3), dtype=np.float32)
        idx = np.empty(icap, dtype=np.uint32)
        out = A.IsoMeshOut(verts.ctypes.data, idx.ctypes.data, vcap, icap, 0, 0,
                           A.ISO_MEMORY_HOST, 0)
        st = self.lib.iso_extract_dc(self.handle, C.byref(desc), C.c_float(epsilon),
                                     C.byref(out))
        if st == A.ISO_ERR_CAPACITY:
            nv, ni = int(out.vertex_count), int(out.index_count)
            verts = np.empty((max(nv, 1), 3), dtype=np.float32)
            idx = np.empty(max(ni, 3), dtype=np.uint32)
            out = A.IsoMeshOut(verts.ctypes.data, idx.ctypes.data, max(nv, 1), max(ni, 3), 0, 0,
                               A.ISO_MEMORY_HOST, 0)
            st = self.lib.iso_mesh_fetch(self.handle, C.byref(out))
        self.check(st, "iso_extract_dc")
        nv, ni = int(out.vertex_count), int(out.index_count)
        return Mesh(verts[:nv].copy(), idx[:ni].copy())

    def extract_emc(self, desc: A.IsoGridDesc, epsilon: float, vertex_capacity: int = 0,
                    index_capacity: int = 0) -> Mesh:
        """iso_extract_emc into host arrays (grown on ISO_ERR_CAPACITY)."""
        vcap = max(int(vertex_capacity), 1)
        icap = max(int(index_capacity), 3)
        verts = np.empty((vcap, 3), dtype=np.float32)
        idx = np.empty(icap, dtype=np.uint32)
        out = A.IsoMeshOut(verts.ctypes.data, idx.ctypes.data, vcap, icap, 0, 0,
                           A.ISO_MEMORY_HOST, 0)
        st = self.lib.iso_extract_emc(self.handle, C.byref(desc), C.c_float(epsilon),
                                      C.byref(out))
        if st == A.ISO_ERR_CAPACITY:
            nv, ni = int(out.vertex_count), int(out.index_count)
            verts = np.empty((max(nv, 1), 3), dtype=np.float32)
            idx = np.empty(max(ni, 3), dtype=np.uint32)
            out = A.IsoMeshOut(verts.ctypes.data, idx.ctypes.data, max(nv, 1), max(ni, 3), 0, 0,
                               A.ISO_MEMORY_HOST, 0)
            st = self.lib.iso_mesh_fetch(self.handle, C.byref(out))
        self.check(st, "iso_extract_emc")
        nv, ni = int(out.vertex_count), int(out.index_count)
        return Mesh(verts[:nv].copy(), idx[:ni].copy())

    def extract_lhmc(self, max_depth: int, vertex_capacity: int = 0,
                     index_capacity: int = 0) -> Mesh:
        """iso_extract_lhmc into host arrays (grown on ISO_ERR_CAPACITY)."""
        vcap = max(int(vertex_capacity), 1)
        icap = max(int(index_capacity), 3)
        verts = np.empty((vcap, 3), dtype=np.float32)
        idx = np.empty(icap, dtype=np.uint32)
        out = A.IsoMeshOut(verts.ctypes.data, idx.ctypes.data, vcap, icap, 0, 0,
                           A.ISO_MEMORY_HOST, 0)
        st = self.lib.iso_extract_lhmc(self.handle, int(max_depth), C.byref(out))
        if st == A.ISO_ERR_CAPACITY:
            nv, ni = int(out.vertex_count), int(out.index_count)
            verts = np.empty((max(nv, 1), 3), dtype=np.float32)
            idx = np.empty(max(ni, 3), dtype=np.uint32)
            out = A.IsoMeshOut(verts.ctypes.data, idx.ctypes.data, max(nv, 1), max(ni, 3), 0, 0,
                               A.ISO_MEMORY_HOST, 0)
            st = self.lib.iso_mesh_fetch(self.handle, C.byref(out))
        self.check(st, "iso_extract_lhmc")
        nv, ni = int(out.vertex_count), int(out.index_count)
        return Mesh(verts[:nv].copy(), idx[:ni].copy())

    def upload_batch(self, chunks) -> None:
        """iso_batch_upload: chunks = [(Source or node list, IsoGridDesc), ...]."""
        nodes, descs = [], (A.IsoChunkDesc * len(chunks))()
        for i, (src, desc) in enumerate(chunks):
            prog = src.program() if isinstance(src, Source) else list(src)
            descs[i].grid = desc
            descs[i].grid.source = A.ISO_SOURCE_PROGRAM
            descs[i].node_offset = len(nodes)
            descs[i].node_count = len(prog)
            nodes.extend(prog)
        arr = to_nodes(nodes)
        self.check(self.lib.iso_batch_upload(self.handle, C.cast(arr, C.c_void_p), len(nodes),
                                             descs, len(chunks)), "iso_batch_upload")
        self._batch_chunks = len(chunks)

    def extract_batch(self, vertex_capacity: int = 0, index_capacity: int = 0,
                      flags: int = 0) -> List[Mesh]:
        """iso_extract_batch into host arrays; one Mesh per chunk."""
        n = self._batch_chunks
        voff = np.zeros(n + 1, np.uint64)
        ioff = np.zeros(n + 1, np.uint64)
        vcap = max(int(vertex_capacity), 1)
        icap = max(int(index_capacity), 3)
        verts = np.empty((vcap, 3), dtype=np.float32)
        idx = np.empty(icap, dtype=np.uint32)
        out = A.IsoBatchOut(verts.ctypes.data, idx.ctypes.data, vcap, icap, voff.ctypes.data,
                            ioff.ctypes.data, 0, 0, A.ISO_MEMORY_HOST, flags)
        st = self.lib.iso_extract_batch(self.handle, C.byref(out))
        if st == A.ISO_ERR_CAPACITY:
            nv, ni = int(out.vertex_count), int(out.index_count)
            verts = np.empty((max(nv, 1), 3), dtype=np.float32)
            idx = np.empty(max(ni, 3), dtype=np.uint32)
            mo = A.IsoMeshOut(verts.ctypes.data, idx.ctypes.data, max(nv, 1), max(ni, 3), 0, 0,
                              A.ISO_MEMORY_HOST, 0)
            st = self.lib.iso_mesh_fetch(self.handle, C.byref(mo))
        self.check(st, "iso_extract_batch")
        return [Mesh(verts[int(voff[i]):int(voff[i + 1])].copy(),
                     idx[int(ioff[i]):int(ioff[i + 1])].copy()) for i in range(n)]

    # ---- timing
    def set_timing(self, on: bool) -> None:
        self.check(self.lib.iso_ctx_set_timing(self.handle, 1 if on else 0), "set_timing")

    def kernel_stats(self) -> dict:
        arr = (A.IsoKernelStat * 64)()
        n = C.c_int()
        self.check(self.lib.iso_ctx_kernel_stats(self.handle, arr, 64, C.byref(n)),
                   "kernel_stats")
        return {arr[i].name.decode(): (int(arr[i].launches), float(arr[i].total_ms))
                for i in range(min(n.value, 64))}

    def reset_kernel_stats(self) -> None:
        self.check(self.lib.iso_ctx_reset_kernel_stats(self.handle), "reset_kernel_stats")

    def synchronize(self) -> None:
        self.check(self.lib.iso_ctx_synchronize(self.handle), "synchronize")


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _replay(mesh: Mesh, extractor: Extractor) -> None:
    extractor.extract_vertices(mesh.vertices)
    extractor.extract_indices(mesh.indices)


def sample_lattice_host(source, desc: A.IsoGridDesc) -> np.ndarray:
    """Samples an arbitrary host source over the lattice (x fastest).

    The coordinates are the reference's: i as f32 * one_over_size
    (primal_grid.rs:50, 63-67)."""
    def axis(n, o, s):
        return (np.float32(o) + np.arange(n, dtype=np.float32) * np.float32(s)).astype(np.float32)

    xs = axis(desc.nx, desc.origin[0], desc.spacing[0])
    ys = axis(desc.ny, desc.origin[1], desc.spacing[1])
    zs = axis(desc.nz, desc.origin[2], desc.spacing[2])
    z, y, x = np.meshgrid(zs, ys, xs, indexing="ij")
    v = source.sample_scalar(x, y, z)
    return np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(desc.nz, desc.ny,
                                                                         desc.nx))


def sample_lattice_host_vector(source, desc: A.IsoGridDesc) -> np.ndarray:
    """Samples an arbitrary host VectorSource (`sample_vector(x, y, z)` over
    numpy arrays, returning the three components) over the lattice:
    (nz, ny, nx, 3) float32, the reference's coordinates."""
    def axis(n, o, s):
        return (np.float32(o) + np.arange(n, dtype=np.float32) * np.float32(s)).astype(np.float32)

    xs = axis(desc.nx, desc.origin[0], desc.spacing[0])
    ys = axis(desc.ny, desc.origin[1], desc.spacing[1])
    zs = axis(desc.nz, desc.origin[2], desc.spacing[2])
    z, y, x = np.meshgrid(zs, ys, xs, indexing="ij")
    vx, vy, vz = source.sample_vector(x, y, z)
    v = np.stack([np.broadcast_to(np.asarray(c, dtype=np.float32), x.shape) for c in (vx, vy, vz)],
                 axis=-1)
    return np.ascontiguousarray(v.reshape(desc.nz, desc.ny, desc.nx, 3))


class Signed:
    """The scalar distance (distance.rs:40-70): MarchingCubes(size) default."""


class Directed:
    """The per-axis distance (distance.rs:43-103): MarchingCubes(size,
    distance=Directed) samples sources with sample_vector."""


class MarchingCubes:
    """MarchingCubes::<D>::new(size) (marching_cubes.rs:38-50) on the GPU,
    D = Signed (default) or Directed."""

    def __init__(self, size: int, context: Optional[Context] = None, distance=Signed):
        if size < 2:
            raise ValueError("size must be at least 2")
        if distance not in (Signed, Directed):
            raise TypeError("distance must be Signed or Directed")
        self.size = int(size)
        self.ctx = context or default_context()
        self.directed = distance is Directed
        self._last_counts = (0, 0)

    def extract_mesh(self, source) -> Mesh:
        if isinstance(source, Source):
            self.ctx.upload_program(source.program())
            desc = self.ctx.reference_desc(self.size, A.ISO_SOURCE_PROGRAM)
        else:
            desc = self.ctx.reference_desc(self.size, A.ISO_SOURCE_GRID)
            if self.directed:
                self.ctx.upload_grid_directed(sample_lattice_host_vector(source, desc))
            else:
                self.ctx.upload_grid(sample_lattice_host(source, desc))
        mesh = self.ctx.extract_mc(desc, *self._last_counts, directed=self.directed)
        self._last_counts = (len(mesh.vertices), len(mesh.indices))
        return mesh

    def extract(self, source, extractor: Extractor) -> None:
        """MarchingCubes::extract (marching_cubes.rs:59-82)."""
        _replay(self.extract_mesh(source), extractor)


class PointCloud:
    """PointCloud::<Signed>::new(size) (point_cloud.rs:29-42) on the GPU: one
    vertex per mixed cell, its centre, in cell order; no faces."""

    def __init__(self, size: int, context: Optional[Context] = None):
        if size < 2:
            raise ValueError("size must be at least 2")
        self.size = int(size)
        self.ctx = context or default_context()
        self._last_count = 0

    def extract_points(self, source) -> np.ndarray:
        if isinstance(source, Source):
            self.ctx.upload_program(source.program())
            desc = self.ctx.reference_desc(self.size, A.ISO_SOURCE_PROGRAM)
        else:
            desc = self.ctx.reference_desc(self.size, A.ISO_SOURCE_GRID)
            self.ctx.upload_grid(sample_lattice_host(source, desc))
        pts = self.ctx.extract_point_cloud(desc, self._last_count)
        self._last_count = len(pts)
        return pts

    def extract(self, source, extractor: Extractor) -> None:
        """PointCloud::extract (point_cloud.rs:50-63): vertices only."""
        extractor.extract_vertices(self.extract_points(source))


class MinimiseQEF:
    """MinimiseQEF (feature/qef.rs:24-68): the feature point minimising the
    quadratic error of the cell corners' tangent planes, by the reference's
    f64 SVD with singular values below 0.1 dropped (and the smallest one on
    an edge feature).  Evaluated on the GPU."""


class DualContouring:
    """DualContouring::new(size, MinimiseQEF {}) (dual_contouring.rs:48-54)
    on the GPU."""

    def __init__(self, size: int, place_feature: Optional[MinimiseQEF] = None,
                 context: Optional[Context] = None):
        if size < 2:
            raise ValueError("size must be at least 2")
        if place_feature is not None and not isinstance(place_feature, MinimiseQEF):
            raise TypeError("only MinimiseQEF feature placement runs on the GPU")
        self.size = int(size)
        self.place_feature = place_feature or MinimiseQEF()
        self.ctx = context or default_context()
        self._last_counts = (0, 0)

    def extract_mesh(self, source) -> Mesh:
        """The source is CentralDifference(Source) (normals by central
        differences) or a program Source whose HermiteSource normals are
        analytic (shapes, unions and translations)."""
        if isinstance(source, CentralDifference):
            epsilon = float(source.epsilon)
        elif isinstance(source, Source):
            epsilon = 0.0
        else:
            raise TypeError("DualContouring needs a HermiteSource: a program Source (analytic "
                            "normals) or CentralDifference(Source)")
        self.ctx.upload_program(source.program())
        desc = self.ctx.reference_desc(self.size, A.ISO_SOURCE_PROGRAM)
        mesh = self.ctx.extract_dc(desc, epsilon, *self._last_counts)
        self._last_counts = (len(mesh.vertices), len(mesh.indices))
        return mesh

    def extract(self, source, extractor: Extractor) -> None:
        """DualContouring::extract (dual_contouring.rs:65-105)."""
        _replay(self.extract_mesh(source), extractor)


class ExtendedMarchingCubes:
    """ExtendedMarchingCubes::new(size) (extended_marching_cubes.rs:43-57) on
    the GPU.  The source is a program Source (its HermiteSource normals:
    shapes, unions and translations) or CentralDifference(Source) (normals by
    central differences); its values are the Directed distances."""

    def __init__(self, size: int, context: Optional[Context] = None):
        if size < 2:
            raise ValueError("size must be at least 2")
        self.size = int(size)
        self.ctx = context or default_context()
        self._last_counts = (0, 0)

    def extract_mesh(self, source) -> Mesh:
        if isinstance(source, CentralDifference):
            epsilon = float(source.epsilon)
        elif isinstance(source, Source):
            epsilon = 0.0
        else:
            raise TypeError("ExtendedMarchingCubes needs a program Source (analytic normals) "
                            "or CentralDifference(Source)")
        self.ctx.upload_program(source.program())
        desc = self.ctx.reference_desc(self.size, A.ISO_SOURCE_PROGRAM)
        mesh = self.ctx.extract_emc(desc, epsilon, *self._last_counts)
        self._last_counts = (len(mesh.vertices), len(mesh.indices))
        return mesh

    def extract(self, source, extractor: Extractor) -> None:
        """ExtendedMarchingCubes::extract (extended_marching_cubes.rs:59-90)."""
        _replay(self.extract_mesh(source), extractor)


class LinearHashedMarchingCubes:
    """LinearHashedMarchingCubes::new(max_depth)
    (linear_hashed_marching_cubes.rs:51-56) on the GPU.  The source is
    sampled over the unit cube at octree node centres; max_depth <= 20."""

    def __init__(self, max_depth: int, context: Optional[Context] = None):
        if not 0 <= int(max_depth) <= 20:
            raise ValueError("max_depth must be in [0, 20]")
        self.max_depth = int(max_depth)
        self.ctx = context or default_context()
        self._last_counts = (0, 0)

    def extract_mesh(self, source: Source) -> Mesh:
        if not isinstance(source, Source):
            raise TypeError("LinearHashedMarchingCubes evaluates program sources on the GPU")
        self.ctx.upload_program(source.program())
        mesh = self.ctx.extract_lhmc(self.max_depth, *self._last_counts)
        self._last_counts = (len(mesh.vertices), len(mesh.indices))
        return mesh

    def extract(self, source: Source, extractor: Extractor) -> None:
        """LinearHashedMarchingCubes::extract (linear_hashed_marching_cubes.rs:62-84)."""
        _replay(self.extract_mesh(source), extractor)


def chunk_desc(size: int) -> A.IsoGridDesc:
    """A chunk of size^3 cells: (size+1)^3 samples, spacing 1/size, over [0, 1]^3."""
    d = A.IsoGridDesc()
    d.nx = d.ny = d.nz = size + 1
    for i in range(3):
        d.origin[i] = 0.0
        d.spacing[i] = float(np.float32(1.0) / np.float32(size))
    d.source = A.ISO_SOURCE_PROGRAM
    return d


class BatchedMarchingCubes:
    """Many chunks, each MarchingCubes over its own lattice and program,
    extracted in one pass (iso_batch_upload + iso_extract_batch).  extract
    gives one mesh per chunk, each identical to extracting that chunk alone."""

    def __init__(self, context: Optional[Context] = None, graph: bool = False,
                 autotune: bool = False):
        self.ctx = context or default_context()
        self.flags = (A.ISO_BATCH_GRAPH if graph else 0) | (A.ISO_BATCH_AUTOTUNE if autotune else 0)
        self._last_counts = (0, 0)

    def extract_meshes(self, chunks) -> List[Mesh]:
        """chunks = [(Source, IsoGridDesc), ...]."""
        self.ctx.upload_batch(chunks)
        meshes = self.ctx.extract_batch(*self._last_counts, flags=self.flags)
        self._last_counts = (sum(len(m.vertices) for m in meshes),
                             sum(len(m.indices) for m in meshes))
        return meshes

    def extract(self, chunks, extractors: Sequence[Extractor]) -> None:
        for mesh, ex in zip(self.extract_meshes(chunks), extractors):
            _replay(mesh, ex)
