"""The oracle's MarchingCubes::<Directed> restatement, pinned without the GPU.

* sample_vector: the 31 directed known-answer tests of the reference's own
  unit tests (sphere.rs, torus.rs, cylinder.rs, rectangular_prism.rs,
  csg.rs), compared as the reference's assert_eq! compares (==, f32::MAX).
* Extraction: a directed grid whose three components equal a scalar grid
  must give exactly the Signed extraction of that grid (classification by
  "any component > 0" and interpolation of the edge-axis component both
  reduce to Signed's); hand-computed crossings of one cell, where each axis
  has its own zero; and, for a sphere, the directed crossings lie on the
  exact surface (each component is linear along its axis there), which a
  wrong component choice would miss by far.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from isosurface_amd import workloads as W
from isosurface_amd.sources import DemoSource, SmoothUnion, Sphere
from oracle import oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden"
VKATS = json.loads((GOLDEN / "reference_vector_kats.json").read_text())
FMAX = np.finfo(np.float32).max


def expected(kat):
    return np.array([FMAX if e == "MAX" else e for e in kat["expected"]], np.float32)


@pytest.mark.parametrize("kat", VKATS, ids=[k["where"] for k in VKATS])
def test_sample_vector_reference_kats(kat):
    got = O.sample_vector(kat["program"], *kat["point"])
    assert np.all(got == expected(kat)), (got, kat["expected"])


def test_vector_kats_cover_every_shape():
    files = {k["where"].split(":")[0] for k in VKATS}
    assert len(VKATS) == 31
    assert files == {f"src/implicit/{n}.rs" for n in
                     ("sphere", "torus", "cylinder", "rectangular_prism", "csg")}


def test_smooth_union_has_no_directed_form():
    prog = SmoothUnion(Sphere(0.2), Sphere(0.3), 0.05).program()
    with pytest.raises(ValueError):
        O.sample_vector(prog, 0.0, 0.0, 0.0)
    with pytest.raises(ValueError):
        O.extract_mc(prog, desc=O.reference_desc(5), directed=True)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_equal_components_reduce_to_signed(seed):
    rng = np.random.default_rng(seed)
    nz, ny, nx = 9, 8, 7
    g = rng.normal(size=(nz, ny, nx)).astype(np.float32)
    g[rng.random(g.shape) < 0.05] = 0.0
    g3 = np.repeat(g[..., None], 3, axis=-1)
    desc = O.lattice_desc(nx, ny, nz, origin=(0.25, -1.0, 0.5), spacing=(0.5, 0.25, 1.0))
    a = O.extract_mc(grid=g, desc=desc, want_cases=True)
    b = O.extract_mc(grid=g3, desc=desc, want_cases=True, directed=True)
    assert np.array_equal(a.cases, b.cases)
    assert a.vertices.tobytes() == b.vertices.tobytes()
    assert np.array_equal(a.indices, b.indices)


def test_one_cell_crossings_follow_each_axis_component():
    # Corner 0 inside with a different zero along each axis; every other
    # point outside (1, 1, 1).  Along x the x components go -0.25 -> 1 (zero
    # at 0.2), along y -0.5 -> 1 (1/3), along z -0.75 -> 1 (3/7).
    g3 = np.ones((2, 2, 2, 3), np.float32)
    g3[0, 0, 0] = (-0.25, -0.5, -0.75)
    m = O.extract_mc(grid=g3, desc=O.lattice_desc(2, 2, 2), directed=True)
    assert len(m.indices) == 3
    want = {(0.2, 0.0, 0.0), (0.0, 1.0 / 3.0, 0.0), (0.0, 0.0, 3.0 / 7.0)}
    got = {tuple(v) for v in m.vertices.astype(np.float64)}
    for w in want:
        assert min(np.abs(np.array(g) - np.array(w)).max() for g in got) < 1e-6, (w, got)


def test_outside_if_any_component_positive():
    # One point whose x component is positive and y, z negative is outside:
    # the cell of a lattice whose only non-positive point is that one has no
    # crossings; flipping x to -0.0 (not positive) makes it inside.
    g3 = np.ones((2, 2, 2, 3), np.float32)
    g3[0, 0, 0] = (0.5, -1.0, -1.0)
    m = O.extract_mc(grid=g3, desc=O.lattice_desc(2, 2, 2), want_cases=True, directed=True)
    assert m.cases[0] == 0 and len(m.indices) == 0
    g3[0, 0, 0] = (-0.0, -1.0, -1.0)
    m = O.extract_mc(grid=g3, desc=O.lattice_desc(2, 2, 2), want_cases=True, directed=True)
    assert m.cases[0] == 1 and len(m.indices) == 3


def test_sphere_crossings_lie_on_the_surface():
    r = 0.35
    src = DemoSource(Sphere(r))
    m = O.extract_mc(src.program(), desc=O.reference_desc(33), directed=True)
    s = O.extract_mc(src.program(), desc=O.reference_desc(33))
    d = np.linalg.norm(m.vertices.astype(np.float64) - 0.5, axis=1) - r
    ds = np.linalg.norm(s.vertices.astype(np.float64) - 0.5, axis=1) - r
    assert np.abs(d).max() < 2e-6
    assert np.abs(ds).max() > 1e-4  # the scalar interpolation is visibly off
    # The same cells are mixed: classification agrees with the scalar one.
    assert len(m.vertices) == len(s.vertices) and len(m.indices) == len(s.indices)


def test_csg_sources_extract():
    src, _ = W.spheres_minus_torus(8, seed=3)
    m = O.extract_mc(src.program(), desc=O.reference_desc(17), directed=True)
    assert len(m.indices) > 0 and len(m.indices) % 3 == 0
    assert m.indices.max() < len(m.vertices)
