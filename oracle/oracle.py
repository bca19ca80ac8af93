"""ctypes wrapper of the CPU oracle — TEST INFRASTRUCTURE ONLY.

Only the test suite, the smoke check and the benchmark's CPU-baseline leg
import this module, and only as the checker.  The product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# ISO_ORACLE_LIB selects another build of the same sources (the sanitizer
# build, oracle/_san/liboracle.so).
LIB_PATH = Path(os.environ.get("ISO_ORACLE_LIB", str(HERE / "liboracle.so")))


class IsoNode(C.Structure):
    _fields_ = [("op", C.c_uint32), ("reserved", C.c_uint32), ("p", C.c_float * 6)]


class IsoGridDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
        ("origin", C.c_float * 3), ("spacing", C.c_float * 3),
        ("source", C.c_int32), ("flags", C.c_int32),
    ]


class OrMesh(C.Structure):
    _fields_ = [
        ("vertices", C.POINTER(C.c_float)),
        ("indices", C.POINTER(C.c_uint32)),
        ("edge_keys", C.POINTER(C.c_int64)),
        ("n_vertices", C.c_uint64),
        ("n_indices", C.c_uint64),
        ("cases", C.POINTER(C.c_uint8)),
        ("n_cases", C.c_uint64),
    ]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        lib = C.CDLL(str(LIB_PATH))
        lib.or_sample.restype = C.c_float
        lib.or_sample.argtypes = [C.POINTER(IsoNode), C.c_uint32, C.c_float, C.c_float, C.c_float]
        lib.or_sample_f64.restype = C.c_double
        lib.or_sample_f64.argtypes = [C.POINTER(IsoNode), C.c_uint32, C.c_double, C.c_double,
                                      C.c_double]
        lib.or_program_check.restype = C.c_int
        lib.or_program_check.argtypes = [C.POINTER(IsoNode), C.c_uint32]
        lib.or_sample_lattice.restype = C.c_int
        lib.or_sample_lattice.argtypes = [C.POINTER(IsoNode), C.c_uint32, C.POINTER(IsoGridDesc),
                                          C.c_void_p]
        lib.or_sample_lattice_layers.restype = C.c_int
        lib.or_sample_lattice_layers.argtypes = [C.POINTER(IsoNode), C.c_uint32,
                                                 C.POINTER(IsoGridDesc), C.c_int64, C.c_int64,
                                                 C.c_void_p]
        lib.or_extract_mc.restype = C.c_int
        lib.or_extract_mc.argtypes = [C.POINTER(IsoNode), C.c_uint32, C.c_void_p,
                                      C.POINTER(IsoGridDesc), C.c_int, C.POINTER(OrMesh)]
        lib.or_extract_point_cloud.restype = C.c_int
        lib.or_extract_mc_directed.argtypes = lib.or_extract_mc.argtypes
        lib.or_loop_table.argtypes = [C.c_void_p]
        lib.or_sample_vector.argtypes = [C.POINTER(IsoNode), C.c_uint32, C.c_float, C.c_float,
                                         C.c_float, C.c_void_p]
        lib.or_extract_point_cloud.argtypes = [C.POINTER(IsoNode), C.c_uint32, C.c_void_p,
                                               C.POINTER(IsoGridDesc), C.POINTER(OrMesh)]
        lib.or_sample_normals.restype = C.c_int
        lib.or_sample_normals.argtypes = [C.POINTER(IsoNode), C.c_uint32, C.c_void_p, C.c_uint64,
                                          C.c_float, C.c_void_p]
        lib.or_edge_points_f64.restype = C.c_int
        lib.or_edge_points_f64.argtypes = [C.POINTER(IsoNode), C.c_uint32, C.POINTER(IsoGridDesc),
                                           C.c_void_p, C.c_uint64, C.c_void_p]
        lib.or_mesh_free.restype = None
        lib.or_mesh_free.argtypes = [C.POINTER(OrMesh)]
        lib.or_tri_table.restype = None
        lib.or_tri_table.argtypes = [C.c_void_p]
        lib.or_edge_mask_table.restype = None
        lib.or_edge_mask_table.argtypes = [C.c_void_p]
        _lib = lib
    return _lib


def nodes(program):
    arr = (IsoNode * len(program))()
    for i, (op, params) in enumerate(program):
        arr[i].op = op
        for j, v in enumerate(params):
            arr[i].p[j] = v
    return arr


def reference_desc(size: int, source: int = 0) -> IsoGridDesc:
    """The lattice of MarchingCubes::new(size) (primal_grid.rs:44-69)."""
    d = IsoGridDesc()
    d.nx, d.ny, d.nz = size, size, size + 1
    ovs = np.float32(1.0) / np.float32(size - 1)
    for i in range(3):
        d.origin[i] = 0.0
        d.spacing[i] = float(ovs)
    d.source = source
    return d


def lattice_desc(nx, ny, nz, origin=(0.0, 0.0, 0.0), spacing=(1.0, 1.0, 1.0), source=1):
    d = IsoGridDesc()
    d.nx, d.ny, d.nz = nx, ny, nz
    for i in range(3):
        d.origin[i] = origin[i]
        d.spacing[i] = spacing[i]
    d.source = source
    return d


def as_desc(d) -> IsoGridDesc:
    """The oracle's descriptor type from any struct with the same fields."""
    if isinstance(d, IsoGridDesc):
        return d
    o = IsoGridDesc()
    o.nx, o.ny, o.nz = d.nx, d.ny, d.nz
    for i in range(3):
        o.origin[i] = d.origin[i]
        o.spacing[i] = d.spacing[i]
    o.source, o.flags = d.source, d.flags
    return o


def sample(program, x, y, z) -> float:
    lib = load()
    arr = nodes(program)
    return float(np.float32(lib.or_sample(arr, len(program), x, y, z)))


def sample_f64(program, x, y, z) -> float:
    lib = load()
    arr = nodes(program)
    return float(lib.or_sample_f64(arr, len(program), x, y, z))


def sample_lattice(program, desc: IsoGridDesc) -> np.ndarray:
    lib = load()
    arr = nodes(program)
    out = np.empty((desc.nz, desc.ny, desc.nx), dtype=np.float32)
    st = lib.or_sample_lattice(arr, len(program), C.byref(desc), out.ctypes.data)
    assert st == 0, st
    return out


def sample_lattice_threads(program, desc: IsoGridDesc, threads: int = 8) -> np.ndarray:
    """sample_lattice on several threads (disjoint layer ranges; ctypes
    releases the interpreter lock during each call): the same values."""
    from concurrent.futures import ThreadPoolExecutor

    lib = load()
    arr = nodes(program)
    desc = as_desc(desc)
    out = np.empty((desc.nz, desc.ny, desc.nx), dtype=np.float32)
    bounds = np.linspace(0, desc.nz, max(1, threads) * 4 + 1).astype(np.int64)

    def run(k):
        z0, z1 = int(bounds[k]), int(bounds[k + 1])
        if z1 > z0:
            st = lib.or_sample_lattice_layers(arr, len(program), C.byref(desc), z0, z1,
                                              out[z0:z1].ctypes.data)
            assert st == 0, st

    with ThreadPoolExecutor(max(1, threads)) as ex:
        list(ex.map(run, range(len(bounds) - 1)))
    return out


class OracleMesh:
    def __init__(self, vertices, indices, keys, cases):
        self.vertices = vertices  # (V,3) f32
        self.indices = indices    # (3T,) u32
        self.keys = keys          # (V,4) i64: lattice point + axis
        self.cases = cases        # per-cell cube index (or None)


def extract_mc(program=None, grid=None, desc: IsoGridDesc = None, want_cases=False,
               directed=False) -> OracleMesh:
    """MarchingCubes::extract + IndexedVertices, restated on the CPU; with
    directed=True, MarchingCubes::<Directed> (grid: 3 floats per point)."""
    lib = load()
    arr = nodes(program) if program is not None else None
    n = len(program) if program is not None else 0
    gptr = None
    if grid is not None:
        grid = np.ascontiguousarray(grid, dtype=np.float32)
        gptr = grid.ctypes.data
    m = OrMesh()
    desc = as_desc(desc)
    fn = lib.or_extract_mc_directed if directed else lib.or_extract_mc
    st = fn(arr, n, gptr, C.byref(desc), 1 if want_cases else 0, C.byref(m))
    if st != 0:
        raise ValueError(f"oracle extraction failed: status {st}")
    try:
        nv, ni = int(m.n_vertices), int(m.n_indices)
        v = np.ctypeslib.as_array(m.vertices, shape=(nv * 3,)).reshape(nv, 3).copy() if nv else \
            np.zeros((0, 3), np.float32)
        i = np.ctypeslib.as_array(m.indices, shape=(ni,)).copy() if ni else np.zeros(0, np.uint32)
        k = np.ctypeslib.as_array(m.edge_keys, shape=(nv * 4,)).reshape(nv, 4).copy() if nv else \
            np.zeros((0, 4), np.int64)
        c = None
        if want_cases and m.n_cases:
            c = np.ctypeslib.as_array(m.cases, shape=(int(m.n_cases),)).copy()
    finally:
        lib.or_mesh_free(C.byref(m))
    return OracleMesh(v, i, k, c)


def sample_vector(program, x, y, z) -> np.ndarray:
    """VectorSource::sample_vector of a program at one point: 3 float32."""
    out = np.zeros(3, np.float32)
    st = load().or_sample_vector(nodes(program), len(program), x, y, z, out.ctypes.data)
    if st != 0:
        raise ValueError("program has no directed form")
    return out


def hermite_normal(program, x, y, z) -> np.ndarray:
    """HermiteSource::sample_normal of a program at one point (not normalised)."""
    lib = load()
    lib.or_sample_hermite_normal.argtypes = [C.POINTER(IsoNode), C.c_uint32, C.c_float,
                                             C.c_float, C.c_float, C.c_void_p]
    out = np.zeros(3, np.float32)
    if lib.or_sample_hermite_normal(nodes(program), len(program), x, y, z, out.ctypes.data):
        raise ValueError("program has no HermiteSource form")
    return out


def extract_emc(program, desc: IsoGridDesc, epsilon: float = 0.0, want_features=False):
    """ExtendedMarchingCubes::extract + IndexedVertices, restated on the CPU.
    Returns an OracleMesh; with want_features, .cases holds one byte per cell
    (1: the cell placed a feature vertex)."""
    lib = load()
    lib.or_extract_emc.argtypes = [C.POINTER(IsoNode), C.c_uint32, C.c_float,
                                   C.POINTER(IsoGridDesc), C.c_int, C.POINTER(OrMesh)]
    m = OrMesh()
    desc = as_desc(desc)
    st = lib.or_extract_emc(nodes(program), len(program), epsilon, C.byref(desc),
                            1 if want_features else 0, C.byref(m))
    if st != 0:
        raise ValueError(f"oracle extraction failed: status {st}")
    try:
        nv, ni = int(m.n_vertices), int(m.n_indices)
        v = np.ctypeslib.as_array(m.vertices, shape=(nv * 3,)).reshape(nv, 3).copy() if nv else \
            np.zeros((0, 3), np.float32)
        i = np.ctypeslib.as_array(m.indices, shape=(ni,)).copy() if ni else np.zeros(0, np.uint32)
        k = np.ctypeslib.as_array(m.edge_keys, shape=(nv * 4,)).reshape(nv, 4).copy() if nv else \
            np.zeros((0, 4), np.int64)
        c = None
        if want_features and m.n_cases:
            c = np.ctypeslib.as_array(m.cases, shape=(int(m.n_cases),)).copy()
    finally:
        lib.or_mesh_free(C.byref(m))
    return OracleMesh(v, i, k, c)


def extract_point_cloud(program=None, grid=None, desc: IsoGridDesc = None) -> np.ndarray:
    """PointCloud::extract restated on the CPU: (n, 3) float32 cell centres."""
    lib = load()
    arr = nodes(program) if program is not None else None
    n = len(program) if program is not None else 0
    gptr = None
    if grid is not None:
        grid = np.ascontiguousarray(grid, dtype=np.float32)
        gptr = grid.ctypes.data
    m = OrMesh()
    desc = as_desc(desc)
    st = lib.or_extract_point_cloud(arr, n, gptr, C.byref(desc), C.byref(m))
    assert st == 0, st
    try:
        nv = int(m.n_vertices)
        v = np.ctypeslib.as_array(m.vertices, shape=(nv * 3,)).reshape(nv, 3).copy() if nv else \
            np.zeros((0, 3), np.float32)
    finally:
        lib.or_mesh_free(C.byref(m))
    return v


def sample_normals(program, points: np.ndarray, epsilon: float) -> np.ndarray:
    """CentralDifference::sample_normal (not normalised) at each point."""
    lib = load()
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    out = np.empty_like(pts)
    st = lib.or_sample_normals(nodes(program), len(program), pts.ctypes.data, len(pts),
                               C.c_float(epsilon), out.ctypes.data)
    assert st == 0, st
    return out


def edge_points_f64(program, desc: IsoGridDesc, keys: np.ndarray) -> np.ndarray:
    lib = load()
    arr = nodes(program)
    keys = np.ascontiguousarray(keys, dtype=np.int64)
    out = np.empty((len(keys), 3), dtype=np.float64)
    st = lib.or_edge_points_f64(arr, len(program), C.byref(desc), keys.ctypes.data, len(keys),
                                out.ctypes.data)
    assert st == 0
    return out


def tri_table() -> np.ndarray:
    out = np.empty(256 * 16, dtype=np.int8)
    load().or_tri_table(out.ctypes.data)
    return out.reshape(256, 16)


def loop_table() -> np.ndarray:
    out = np.empty(256 * 17, dtype=np.int8)
    load().or_loop_table(out.ctypes.data)
    return out.reshape(256, 17)


def edge_mask_table() -> np.ndarray:
    out = np.empty(256, dtype=np.uint16)
    load().or_edge_mask_table(out.ctypes.data)
    return out


def _dc_bind(lib):
    if getattr(lib, "_dc_bound", False):
        return
    lib.or_extract_dc.restype = C.c_int
    lib.or_extract_dc.argtypes = [C.POINTER(IsoNode), C.c_uint32, C.POINTER(IsoGridDesc),
                                  C.c_float, C.POINTER(OrMesh)]
    lib.or_svd_solve.restype = C.c_int
    lib.or_svd_solve.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                 C.c_void_p]
    lib.or_place_feature.restype = None
    lib.or_place_feature.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    lib._dc_bound = True


def extract_dc(program, desc: IsoGridDesc, epsilon: float = 1e-6) -> OracleMesh:
    """DualContouring::extract (MinimiseQEF), restated: CentralDifference
    normals, or with epsilon 0 the program's own HermiteSource normals."""
    lib = load()
    _dc_bind(lib)
    arr = nodes(program)
    m = OrMesh()
    st = lib.or_extract_dc(arr, len(program), C.byref(desc), epsilon, C.byref(m))
    assert st == 0, st
    try:
        nv, ni = int(m.n_vertices), int(m.n_indices)
        v = np.ctypeslib.as_array(m.vertices, shape=(nv * 3,)).reshape(nv, 3).copy() if nv else \
            np.zeros((0, 3), np.float32)
        i = np.ctypeslib.as_array(m.indices, shape=(ni,)).copy() if ni else np.zeros(0, np.uint32)
        k = np.ctypeslib.as_array(m.edge_keys, shape=(nv * 4,)).reshape(nv, 4).copy() if nv else \
            np.zeros((0, 4), np.int64)
    finally:
        lib.or_mesh_free(C.byref(m))
    return OracleMesh(v, i, k, None)


def svd_solve(a: np.ndarray, b: np.ndarray, drop_min: bool = False):
    """SVD::new + (optional Edge rank drop) + SVD::solve on an n x 3 system."""
    lib = load()
    _dc_bind(lib)
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.empty(3, np.float64)
    d = np.empty(3, np.float64)
    st = lib.or_svd_solve(a.ctypes.data, b.ctypes.data, len(a), 1 if drop_min else 0,
                          x.ctypes.data, d.ctypes.data)
    assert st == 0
    return x, d


def place_feature(corners: np.ndarray, normals: np.ndarray) -> np.ndarray:
    lib = load()
    _dc_bind(lib)
    c = np.ascontiguousarray(corners, dtype=np.float32)
    n = np.ascontiguousarray(normals, dtype=np.float32)
    out = np.empty(3, np.float32)
    lib.or_place_feature(c.ctypes.data, n.ctypes.data, out.ctypes.data)
    return out


def _lhmc_bind(lib):
    if getattr(lib, "_lhmc_bound", False):
        return
    lib.or_extract_lhmc.restype = C.c_int
    lib.or_extract_lhmc.argtypes = [C.POINTER(IsoNode), C.c_uint32, C.c_int, C.c_uint64,
                                    C.POINTER(OrMesh)]
    lib.or_lhmc_octree.restype = C.c_int
    lib.or_lhmc_octree.argtypes = [C.POINTER(IsoNode), C.c_uint32, C.c_int, C.c_uint64,
                                   C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64)]
    lib.or_morton_level.restype = C.c_int
    lib.or_morton_level.argtypes = [C.c_uint64]
    lib.or_morton_center.restype = None
    lib.or_morton_center.argtypes = [C.c_uint64, C.c_void_p]
    lib.or_morton_size.restype = C.c_float
    lib.or_morton_size.argtypes = [C.c_uint64]
    for name in ("or_morton_add", "or_morton_sub"):
        getattr(lib, name).restype = C.c_uint64
        getattr(lib, name).argtypes = [C.c_uint64, C.c_uint64]
    for name in ("or_morton_primal_vertex", "or_morton_dual_vertex"):
        getattr(lib, name).restype = C.c_uint64
        getattr(lib, name).argtypes = [C.c_uint64, C.c_int, C.c_uint]
    lib._lhmc_bound = True


def morton():
    """The oracle's Morton helpers (morton.rs)."""
    lib = load()
    _lhmc_bind(lib)
    return lib


def morton_center(k: int) -> np.ndarray:
    lib = morton()
    out = np.empty(3, np.float32)
    lib.or_morton_center(k, out.ctypes.data)
    return out


def extract_lhmc(program, max_depth: int, order_seed: int = 0) -> OracleMesh:
    """LinearHashedMarchingCubes::new(max_depth).extract, restated; primal
    vertices visited in ascending Morton order (seed 0) or a seeded shuffle."""
    lib = load()
    _lhmc_bind(lib)
    arr = nodes(program)
    m = OrMesh()
    st = lib.or_extract_lhmc(arr, len(program), max_depth, order_seed, C.byref(m))
    assert st == 0, st
    try:
        nv, ni = int(m.n_vertices), int(m.n_indices)
        v = np.ctypeslib.as_array(m.vertices, shape=(nv * 3,)).reshape(nv, 3).copy() if nv else \
            np.zeros((0, 3), np.float32)
        i = np.ctypeslib.as_array(m.indices, shape=(ni,)).copy() if ni else np.zeros(0, np.uint32)
        k = np.ctypeslib.as_array(m.edge_keys, shape=(nv * 4,)).reshape(nv, 4).copy() if nv else \
            np.zeros((0, 4), np.int64)
    finally:
        lib.or_mesh_free(C.byref(m))
    return OracleMesh(v, i, k[:, :2].astype(np.uint64), None)


def lhmc_octree(program, max_depth: int):
    """(keys in BFS order, values, leaf flags) of the implicit octree."""
    lib = load()
    _lhmc_bind(lib)
    arr = nodes(program)
    n = C.c_uint64(0)
    cap = 1
    while True:
        keys = np.empty(cap, np.uint64)
        vals = np.empty(cap, np.float32)
        leaf = np.empty(cap, np.uint8)
        st = lib.or_lhmc_octree(arr, len(program), max_depth, cap, keys.ctypes.data,
                                vals.ctypes.data, leaf.ctypes.data, C.byref(n))
        if st == 0:
            k = int(n.value)
            return keys[:k], vals[:k], leaf[:k].astype(bool)
        assert st == 2, st  # ISO_ERR_CAPACITY
        cap = int(n.value)


def extract_batch(chunks):
    """Per chunk (program, desc): MarchingCubes over the chunk's lattice —
    the batched extraction is, by definition, these independent meshes."""
    return [extract_mc(program, desc=desc) for program, desc in chunks]
