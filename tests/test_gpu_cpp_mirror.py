"""The C++ host mirror (isosurface_amd/host/isosurface.hpp) end to end: the
reference-style extract(source, extractor) call, for a GPU-lowered source and
for a host-sampled user source, must produce the oracle's mesh."""
import hashlib
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
EXAMPLE = ROOT / "isosurface_amd" / "host" / "example_mc"
GOLD = json.loads((ROOT / "tests" / "golden" / "mesh_digests.json").read_text())["sphere_r0.35_size65"]


def test_cpp_mirror_matches_oracle(tmp_path):
    assert EXAMPLE.exists(), "build the library first (python -m isosurface_amd.build)"
    out = tmp_path / "mesh"
    r = subprocess.run([str(EXAMPLE), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert "error path: status 5" in r.stdout
    for kind in ("program", "host"):
        v = np.fromfile(f"{out}.{kind}.f32", dtype="<f4")
        i = np.fromfile(f"{out}.{kind}.u32", dtype="<u4")
        assert len(v) == 3 * GOLD["vertices"] and len(i) == 3 * GOLD["triangles"]
        assert hashlib.sha256(v.tobytes()).hexdigest() == GOLD["vertices_sha256"], kind
        assert hashlib.sha256(i.tobytes()).hexdigest() == GOLD["indices_sha256"], kind
    # PointCloud and the interleaved-normal sink through the same mirror
    from isosurface_amd import workloads as W
    from oracle import oracle as O
    src = W.sphere_source(0.35)
    pts = np.fromfile(f"{out}.points.f32", dtype="<f4").reshape(-1, 3)
    want = O.extract_point_cloud(src.program(), desc=O.reference_desc(65, 0))
    assert pts.tobytes() == want.tobytes()
    nv = np.fromfile(f"{out}.normals.f32", dtype="<f4").reshape(-1, 6)
    ni = np.fromfile(f"{out}.normals.u32", dtype="<u4")
    v = np.fromfile(f"{out}.program.f32", dtype="<f4").reshape(-1, 3)
    assert nv[:, :3].tobytes() == v.tobytes()
    assert ni.tobytes() == np.fromfile(f"{out}.program.u32", dtype="<u4").tobytes()
    assert nv[:, 3:].tobytes() == O.sample_normals(src.program(), v, 1e-4).tobytes()
    # MarchingCubes::<Directed>: the lowered sphere and the host VectorSource
    ref = O.extract_mc(src.program(), desc=O.reference_desc(65, 0), directed=True)
    for kind in ("directed", "directed_host"):
        dv = np.fromfile(f"{out}.{kind}.f32", dtype="<f4")
        di = np.fromfile(f"{out}.{kind}.u32", dtype="<u4")
        assert dv.tobytes() == ref.vertices.tobytes(), kind
        assert np.array_equal(di, ref.indices), kind
    # ExtendedMarchingCubes: a cube with its own normals, and behind CentralDifference
    from isosurface_amd.sources import DemoSource, RectangularPrism, Translate
    cube = DemoSource(Translate((0.013, 0.021, -0.017), RectangularPrism((0.2, 0.2, 0.2))))
    for kind, eps in (("emc", 0.0), ("emc_cd", 1e-4)):
        ref = O.extract_emc(cube.program(), O.reference_desc(33, 0), epsilon=eps)
        ev = np.fromfile(f"{out}.{kind}.f32", dtype="<f4")
        ei = np.fromfile(f"{out}.{kind}.u32", dtype="<u4")
        assert ev.tobytes() == ref.vertices.tobytes(), kind
        assert np.array_equal(ei, ref.indices), kind
    # DualContouring and LinearHashedMarchingCubes of the same cube
    ref = O.extract_dc(cube.program(), O.reference_desc(33, 0), 1e-6)
    assert np.fromfile(f"{out}.dc.f32", dtype="<f4").tobytes() == ref.vertices.tobytes()
    assert np.array_equal(np.fromfile(f"{out}.dc.u32", dtype="<u4"), ref.indices)
    ref = O.extract_dc(cube.program(), O.reference_desc(33, 0), 0.0)
    assert np.fromfile(f"{out}.dc_hermite.f32", dtype="<f4").tobytes() == ref.vertices.tobytes()
    assert np.array_equal(np.fromfile(f"{out}.dc_hermite.u32", dtype="<u4"), ref.indices)
    ref = O.extract_lhmc(cube.program(), 6, 0)
    assert np.fromfile(f"{out}.lhmc.f32", dtype="<f4").tobytes() == ref.vertices.tobytes()
    assert np.array_equal(np.fromfile(f"{out}.lhmc.u32", dtype="<u4"), ref.indices)
