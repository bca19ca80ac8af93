"""GPU SDF evaluation: bit-identical to the oracle's restatement of the
crate's sample_scalar (implicit/*.rs) at random points."""
import json
from pathlib import Path

import numpy as np
import pytest

from isosurface_amd import workloads as W
from isosurface_amd.sources import (Cylinder, DemoSource, Difference, Intersection,
                                    RectangularPrism, SmoothUnion, Sphere, Torus, Translate, Union,
                                    union_all)
from oracle import oracle as O

pytestmark = pytest.mark.gpu

KATS = json.loads((Path(__file__).resolve().parent / "golden" / "reference_kats.json").read_text())

SOURCES = {
    "sphere": Sphere(0.3),
    "torus": Torus(0.3, 0.05),
    "box": RectangularPrism((0.1, 0.2, 0.3)),
    "cylinder": Cylinder(0.2, 0.3),
    "union": Union(Sphere(0.2), Translate((0.1, 0.0, 0.0), Sphere(0.2))),
    "intersection": Intersection(RectangularPrism((0.2, 0.2, 0.2)), Sphere(0.25)),
    "difference": Difference(Sphere(0.2), RectangularPrism((0.3, 0.3, 0.3))),
    "difference_rev": Difference(Union(Sphere(0.1), Translate((0.2, 0, 0), Sphere(0.1))),
                                 Sphere(0.3)),
    "smooth": SmoothUnion(Sphere(0.2), Translate((0.25, 0.0, 0.0), Sphere(0.2)), 0.05),
    "nested_translate": Translate((0.1, 0.2, 0.3), Union(
        Translate((-0.1, 0.0, 0.0), Torus(0.2, 0.05)),
        Translate((0.05, 0.05, 0.0), Intersection(Sphere(0.2), Cylinder(0.1, 0.3))))),
    "spheres64_minus_torus": W.spheres_minus_torus()[0],
    "boxes32": W.boxes_union(),
}


@pytest.mark.parametrize("name", sorted(SOURCES))
def test_program_sampling_is_bit_identical(ctx, name):
    src = SOURCES[name]
    prog = src.program()
    rng = np.random.RandomState(abs(hash(name)) % (2 ** 31))
    pts = (rng.rand(4096, 3) * 1.2 - 0.6).astype(np.float32)
    got = ctx.sample(src, pts)
    want = np.array([O.sample(prog, *map(float, p)) for p in pts], dtype=np.float32)
    assert got.tobytes() == want.tobytes()


@pytest.mark.parametrize("kat", KATS, ids=[k["where"] for k in KATS])
def test_reference_known_answers_on_gpu(ctx, kat):
    from isosurface_amd.sources import Source

    class Raw(Source):
        def program(self):
            return [(op, tuple(p)) for op, p in kat["program"]]

    got = ctx.sample(Raw(), np.array([kat["point"]], np.float32))
    assert float(got[0]) == kat["expected"]


def test_sqrt_and_division_are_correctly_rounded(ctx):
    # Sphere(0) is sqrt((x*x + y*y) + z*z): compare with numpy's IEEE f32.
    rng = np.random.RandomState(3)
    pts = (rng.rand(200000, 3) * 2 - 1).astype(np.float32)
    pts[:1000] *= np.float32(1e-20)
    pts[1000:2000] *= np.float32(1e15)
    got = ctx.sample(Sphere(0.0), pts)
    x, y, z = pts[:, 0], pts[:, 1], pts[:, 2]
    want = np.sqrt((x * x + y * y) + z * z)
    assert got.tobytes() == want.tobytes()
    # smooth-min divides by k: exercise the division path too
    src = SmoothUnion(Sphere(0.2), Translate((0.01, 0.0, 0.0), Sphere(0.2)), 0.37)
    prog = src.program()
    got = ctx.sample(src, pts[:3000])
    want = np.array([O.sample(prog, *map(float, p)) for p in pts[:3000]], np.float32)
    assert got.tobytes() == want.tobytes()


def test_lattice_sampling_matches_oracle(ctx):
    src, _ = W.spheres_minus_torus()
    prog = src.program()
    ctx.upload_program(prog)
    desc = ctx.reference_desc(33, 0)
    out = np.empty((desc.nz, desc.ny, desc.nx), np.float32)
    from isosurface_amd import _abi as A
    import ctypes as C
    ctx.check(ctx.lib.iso_program_sample_lattice(ctx.handle, C.byref(desc), out.ctypes.data,
                                                 A.ISO_MEMORY_HOST), "sample_lattice")
    want = O.sample_lattice(prog, O.reference_desc(33))
    assert out.tobytes() == want.tobytes()
