"""The oracle's ExtendedMarchingCubes restatement, pinned without the GPU.

* EDGE_LOOPS: the decoded loop table against SHA-256 of the reference's
  table text (with and without the three rows' unread trailing entries);
  every case's loops use exactly its crossed edges.
* HermiteSource normals: the reference's own 10 normal known-answer tests;
  the torus normal at points of its surface has the tube radius as length.
* Extraction: a smooth sphere has no feature cell and gives exactly the
  directed marching-cubes mesh; an axis-aligned box's feature vertices lie
  on the box's edges and corners; the flipped box mesh is a closed, oriented
  2-manifold (every directed edge once, its reverse once, chi = 2) with
  faces joining two feature vertices — a wrong rotation winding breaks the
  orientation.
"""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from isosurface_amd import workloads as W
from isosurface_amd.sources import (DemoSource, Difference, RectangularPrism, SmoothUnion, Sphere,
                                    Torus, Translate)
from oracle import oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden"
NKATS = json.loads((GOLDEN / "reference_normal_kats.json").read_text())

# SHA-256 of EDGE_LOOPS as int8[256][17] row-major, computed from the text of
# the reference's marching_cubes_tables.rs:337-594: as written, and with the
# entries after each row's last loop (which extended_marching_cubes.rs never
# reads) replaced by -1.
LOOPS_SHA_RAW = "ee9140b6c1aca4bfed861ed12573d652338750055279058011c3240552caa6ee"
LOOPS_SHA_READ = "2ed4921dc1dc2e3aa9c290a9b6466f899b0ad0e0bdb5a289dcc02fe6e8a41bae"
UNREAD_TRAILING = {167: [6], 181: [2, 10], 229: [8, 3]}


def test_loop_table_matches_the_reference_text():
    t = O.loop_table()
    assert hashlib.sha256(t.tobytes()).hexdigest() == LOOPS_SHA_READ
    raw = t.copy()
    for c, extra in UNREAD_TRAILING.items():
        n = 1 + raw[c, 0] + raw[c, 1:1 + raw[c, 0]].sum()
        raw[c, n:n + len(extra)] = extra
    assert hashlib.sha256(raw.tobytes()).hexdigest() == LOOPS_SHA_RAW


def test_loops_use_exactly_the_crossed_edges():
    t = O.loop_table()
    mask = O.edge_mask_table()
    for c in range(256):
        n = t[c, 0]
        sizes = t[c, 1:1 + n]
        edges = t[c, 1 + n:1 + n + sizes.sum()]
        assert all(s >= 3 for s in sizes)
        assert sum(1 << int(e) for e in set(edges.tolist())) == mask[c]


@pytest.mark.parametrize("kat", NKATS, ids=[k["where"] for k in NKATS])
def test_normal_reference_kats(kat):
    n = O.hermite_normal(kat["program"], *kat["point"]).astype(np.float32)
    unit = n / np.float32(np.sqrt((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]))
    assert np.all(unit == np.array(kat["expected"], np.float32))


def test_torus_normal_has_tube_length_on_the_surface():
    R, r = 0.3, 0.05
    prog = Torus(R, r).program()
    rng = np.random.default_rng(1)
    for u, v in rng.uniform(0, 2 * np.pi, size=(50, 2)):
        p = np.array([(R + r * np.cos(v)) * np.cos(u), (R + r * np.cos(v)) * np.sin(u),
                      r * np.sin(v)])
        n = O.hermite_normal(prog, *p.astype(np.float32)).astype(np.float64)
        assert abs(np.linalg.norm(n) - r) < 1e-6
        want = np.array([np.cos(v) * np.cos(u), np.cos(v) * np.sin(u), np.sin(v)])
        assert np.abs(n / np.linalg.norm(n) - want).max() < 1e-4


def test_operators_without_hermite_form_are_rejected():
    with pytest.raises(ValueError):
        O.hermite_normal(Difference(Sphere(0.1), Sphere(0.2)).program(), 0.0, 0.0, 0.0)
    with pytest.raises(ValueError):
        O.extract_emc(Difference(Sphere(0.1), Sphere(0.2)).program(), O.reference_desc(5))
    with pytest.raises(ValueError):
        O.extract_emc(SmoothUnion(Sphere(0.1), Sphere(0.2), 0.05).program(), O.reference_desc(5),
                      epsilon=1e-4)
    # with central differences a difference is fine
    O.extract_emc(DemoSource(Difference(Sphere(0.1), Sphere(0.3))).program(),
                  O.reference_desc(9), epsilon=1e-4)


def test_smooth_sphere_is_directed_marching_cubes():
    prog = W.sphere_source().program()
    m = O.extract_emc(prog, O.reference_desc(33), want_features=True)
    d = O.extract_mc(prog, desc=O.reference_desc(33), directed=True)
    assert m.cases.sum() == 0
    assert m.vertices.tobytes() == d.vertices.tobytes()
    assert np.array_equal(m.indices, d.indices)


# A cube: RectangularPrism's normal is the dominant axis of p
# (clamp_to_cardinal_axis), the face normal only when the half extents agree.
HALF = np.array([0.2, 0.2, 0.2])
SHIFT = np.array([0.013, 0.021, -0.017])


def box_mesh(size=33):
    box = DemoSource(Translate(tuple(SHIFT), RectangularPrism(tuple(HALF))))
    return O.extract_emc(box.program(), O.reference_desc(size), want_features=True)


def test_box_feature_vertices_lie_on_edges_and_corners():
    m = box_mesh()
    feat = m.keys[:, 3] == 3
    assert feat.sum() == m.cases.sum() > 0
    q = np.abs(m.vertices[feat].astype(np.float64) - 0.5 - SHIFT)
    on_face = np.abs(q - HALF) < 1e-5           # per axis: on the face plane
    assert np.all(on_face.sum(axis=1) >= 2)     # an edge (two planes) or a corner (three)
    assert np.all(q <= HALF + 1e-5)


def test_box_mesh_is_a_closed_oriented_manifold_after_flips():
    m = box_mesh()
    f = m.indices.reshape(-1, 3).astype(np.int64)
    directed = {}
    for a, b, c in f:
        for u, v in ((a, b), (b, c), (c, a)):
            directed[(u, v)] = directed.get((u, v), 0) + 1
    assert all(n == 1 for n in directed.values())
    assert all((v, u) in directed for (u, v) in directed)
    V, E, F = len(m.vertices), len(directed) // 2, len(f)
    assert V - E + F == 2
    feat = set(np.nonzero(m.keys[:, 3] == 3)[0].tolist())
    assert sum(1 for t in f if sum(int(v) in feat for v in t) == 2) > 0
