"""ctypes declarations of the C ABI in include/isosurface.h."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libisosurface_amd.so"

ISO_OK = 0
ISO_ERR_INVALID_ARGUMENT = 1
ISO_ERR_CAPACITY = 2
ISO_ERR_DEVICE = 3
ISO_ERR_OUT_OF_MEMORY = 4
ISO_ERR_PROGRAM = 5
ISO_ERR_NO_DEVICE = 6
ISO_ERR_NO_SOURCE = 7
ISO_ERR_OVERFLOW = 8

ISO_OP_SPHERE = 1
ISO_OP_TORUS = 2
ISO_OP_BOX = 3
ISO_OP_CYLINDER = 4
ISO_OP_UNION = 16
ISO_OP_INTERSECTION = 17
ISO_OP_DIFFERENCE = 18
ISO_OP_SMOOTH_UNION = 19
ISO_OP_TRANSLATE = 32
ISO_OP_END_TRANSLATE = 33

ISO_SOURCE_PROGRAM = 0
ISO_SOURCE_GRID = 1
ISO_MEMORY_HOST = 0
ISO_MEMORY_DEVICE = 1
ISO_BATCH_GRAPH = 1
ISO_BATCH_AUTOTUNE = 2


class IsoNode(C.Structure):
    _fields_ = [("op", C.c_uint32), ("reserved", C.c_uint32), ("p", C.c_float * 6)]


class IsoGridDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int64),
        ("ny", C.c_int64),
        ("nz", C.c_int64),
        ("origin", C.c_float * 3),
        ("spacing", C.c_float * 3),
        ("source", C.c_int32),
        ("flags", C.c_int32),
    ]


class IsoMeshOut(C.Structure):
    _fields_ = [
        ("vertices", C.c_void_p),
        ("indices", C.c_void_p),
        ("vertex_capacity", C.c_uint64),
        ("index_capacity", C.c_uint64),
        ("vertex_count", C.c_uint64),
        ("index_count", C.c_uint64),
        ("memory", C.c_int32),
        ("flags", C.c_int32),
    ]


class IsoChunkDesc(C.Structure):
    _fields_ = [("grid", IsoGridDesc), ("node_offset", C.c_uint32), ("node_count", C.c_uint32)]


class IsoBatchOut(C.Structure):
    _fields_ = [
        ("vertices", C.c_void_p),
        ("indices", C.c_void_p),
        ("vertex_capacity", C.c_uint64),
        ("index_capacity", C.c_uint64),
        ("vertex_offsets", C.c_void_p),
        ("index_offsets", C.c_void_p),
        ("vertex_count", C.c_uint64),
        ("index_count", C.c_uint64),
        ("memory", C.c_int32),
        ("flags", C.c_uint32),
    ]


ISO_MESH_GRAPH = 1
ISO_MESH_SINGLE_PASS = 2
ISO_BAND_PARAMS = 15617
ISO_BAND_CERTIFY = 1
ISO_BAND_NO_NETWORK = 2
ISO_BAND_BRICK8 = 4


class IsoBandTrainParams(C.Structure):
    _fields_ = [("steps", C.c_uint32), ("batch", C.c_uint32), ("lr", C.c_float),
                ("lr_min", C.c_float), ("clamp", C.c_float), ("seed", C.c_uint32),
                ("overshoot", C.c_float)]


class IsoBandStats(C.Structure):
    _fields_ = [("bricks", C.c_uint64), ("band_bricks", C.c_uint64),
                ("exact_samples", C.c_uint64), ("centre_samples", C.c_uint64),
                ("overturned_culls", C.c_uint64)]


class IsoKernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_uint64), ("total_ms", C.c_double)]


# (name, restype, argtypes) of every exported entry point.
SIGNATURES = [
    ("iso_abi_version", C.c_int, []),
    ("iso_status_string", C.c_char_p, [C.c_int]),
    ("iso_ctx_create", C.c_int, [C.c_int, C.c_void_p, C.POINTER(C.c_void_p)]),
    ("iso_ctx_destroy", C.c_int, [C.c_void_p]),
    ("iso_ctx_synchronize", C.c_int, [C.c_void_p]),
    ("iso_ctx_last_error", C.c_char_p, [C.c_void_p]),
    ("iso_program_validate", C.c_int, [C.POINTER(IsoNode), C.c_uint32, C.c_char_p, C.c_size_t]),
    ("iso_program_upload", C.c_int, [C.c_void_p, C.POINTER(IsoNode), C.c_uint32]),
    ("iso_program_sample", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    ("iso_sample_normals", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_float, C.c_void_p,
                                     C.c_int32]),
    ("iso_extract_point_cloud", C.c_int, [C.c_void_p, C.POINTER(IsoGridDesc),
                                          C.POINTER(IsoMeshOut)]),
    ("iso_program_sample_lattice", C.c_int, [C.c_void_p, C.POINTER(IsoGridDesc), C.c_void_p,
                                             C.c_int32]),
    ("iso_grid_upload", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_int64]),
    ("iso_grid_bind_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_int64]),
    ("iso_grid_desc_reference", C.c_int, [C.c_int64, C.c_int32, C.POINTER(IsoGridDesc)]),
    ("iso_extract_mc", C.c_int, [C.c_void_p, C.POINTER(IsoGridDesc), C.POINTER(IsoMeshOut)]),
    ("iso_extract_mc_noalloc", C.c_int, [C.c_void_p, C.POINTER(IsoGridDesc), C.POINTER(IsoMeshOut)]),
    ("iso_extract_mc_directed", C.c_int, [C.c_void_p, C.POINTER(IsoGridDesc),
                                          C.POINTER(IsoMeshOut)]),
    ("iso_extract_emc", C.c_int, [C.c_void_p, C.POINTER(IsoGridDesc), C.c_float,
                                  C.POINTER(IsoMeshOut)]),
    ("iso_program_sample_vector", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    ("iso_grid_upload_directed", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64,
                                           C.c_int64]),
    ("iso_extract_dc", C.c_int, [C.c_void_p, C.POINTER(IsoGridDesc), C.c_float,
                                 C.POINTER(IsoMeshOut)]),
    ("iso_extract_lhmc", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(IsoMeshOut)]),
    ("iso_batch_upload", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(IsoChunkDesc),
                                   C.c_uint32]),
    ("iso_extract_batch", C.c_int, [C.c_void_p, C.POINTER(IsoBatchOut)]),
    ("iso_band_init", C.c_int, [C.c_void_p, C.c_uint32]),
    ("iso_band_set_params", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    ("iso_band_get_params", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    ("iso_band_train", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                 C.c_int32, C.POINTER(IsoBandTrainParams),
                                 C.POINTER(C.c_float)]),
    ("iso_band_gradient", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                    C.c_float, C.c_float, C.c_void_p, C.POINTER(C.c_float)]),
    ("iso_band_brick_size", C.c_int, [C.c_uint32]),
    ("iso_band_predict", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_int32]),
    ("iso_extract_mc_band", C.c_int, [C.c_void_p, C.POINTER(IsoGridDesc), C.c_float, C.c_uint32,
                                      C.POINTER(IsoMeshOut), C.POINTER(IsoBandStats)]),
    ("iso_band_missed_cells", C.c_int, [C.c_void_p, C.POINTER(IsoGridDesc),
                                        C.POINTER(C.c_uint64)]),
    ("iso_mesh_fetch", C.c_int, [C.c_void_p, C.POINTER(IsoMeshOut)]),
    ("iso_device_alloc", C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    ("iso_device_free", C.c_int, [C.c_void_p, C.c_void_p]),
    ("iso_memcpy_to_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    ("iso_memcpy_to_host", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    ("iso_memcpy_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    ("iso_ctx_set_timing", C.c_int, [C.c_void_p, C.c_int]),
    ("iso_ctx_kernel_stats", C.c_int, [C.c_void_p, C.POINTER(IsoKernelStat), C.c_int,
                                       C.POINTER(C.c_int)]),
    ("iso_ctx_reset_kernel_stats", C.c_int, [C.c_void_p]),
]

_lib = None


def load() -> C.CDLL:
    """Loads libisosurface_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m isosurface_amd.build` "
                "(there is no CPU fallback)")
        lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | C.RTLD_GLOBAL)
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class IsoError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{message} (status {status})")
        self.status = status
