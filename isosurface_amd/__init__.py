"""isosurface_amd: MI355X (gfx950) isosurface extraction.

The hot path of the swiftcoder/isosurface crate (SDF sampling, marching-cubes
classification, edge interpolation, vertex dedup and triangle emission) runs
in hand-written HIP kernels behind the C ABI of include/isosurface.h.  This
package is the Python view of that ABI, mirroring the crate's
Source / Extractor / MarchingCubes / DualContouring /
LinearHashedMarchingCubes API; the C++ mirror for compiled hosts is
isosurface_amd/host/isosurface.hpp.
"""
from .engine import (BatchedMarchingCubes, Context, Directed, DualContouring,
                     ExtendedMarchingCubes, Extractor,
                     IndexedInterleavedNormals, IndexedVertices, LinearHashedMarchingCubes,
                     MarchingCubes, Mesh, MinimiseQEF, OnlyInterleavedNormals, OnlyVertices,
                     PointCloud, Signed, chunk_desc, default_context, sample_lattice_host,
                     sample_lattice_host_vector)
from .sources import (CentralDifference, Cylinder, DemoSource, Difference, Intersection,
                      RectangularPrism,
                      SmoothUnion, Source, Sphere, Torus, Translate, Union, smooth_union_all,
                      to_nodes, union_all)

__all__ = [
    "BatchedMarchingCubes", "chunk_desc", "CentralDifference", "DualContouring", "LinearHashedMarchingCubes", "MinimiseQEF",
    "Context", "Extractor", "IndexedVertices", "MarchingCubes", "Mesh", "OnlyVertices",
    "IndexedInterleavedNormals", "OnlyInterleavedNormals", "PointCloud", "Signed", "Directed",
    "ExtendedMarchingCubes",
    "sample_lattice_host_vector",
    "default_context", "sample_lattice_host", "Cylinder", "DemoSource", "Difference",
    "Intersection", "RectangularPrism", "SmoothUnion", "Source", "Sphere", "Torus",
    "Translate", "Union", "smooth_union_all", "to_nodes", "union_all",
]
