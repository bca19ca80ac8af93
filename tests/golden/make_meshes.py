"""Writes mesh_digests.json: counts and SHA-256 digests of the oracle's
meshes (marching cubes Signed and Directed, dual contouring, octree MC,
extended MC, batched chunks), point cloud and vertex normals for the
benchmark workloads.

The oracle (oracle/, a restatement of the reference's CPU extraction) is run
here once; the GPU tests compare their output digests with these values, so
full-size parity needs no CPU re-run on the GPU host.

usage: python tests/golden/make_meshes.py [--only NAME[,NAME...]]
(--only recomputes the named entries and keeps the others as they are)
"""
import hashlib
import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from isosurface_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402


def digest(a) -> str:
    return hashlib.sha256(a.tobytes()).hexdigest()


SMOOTH_SIZES = (129, 257)


def smooth_name(size: int) -> str:
    return f"smooth_spheres1024_k0.02_size{size}"


def smooth_cases(names=None):
    """Marching cubes of the smooth-min (k=0.02) blend of 1024 seeded spheres
    (the learned-band field), sampled by the oracle on 8 threads."""
    prog = W.smooth_spheres(1024, 0.02, 5).program()
    out = {}
    for size in SMOOTH_SIZES:
        if names is not None and smooth_name(size) not in names:
            continue
        d = O.reference_desc(size)
        grid = O.sample_lattice_threads(prog, d, threads=8)
        out[smooth_name(size)] = (size, O.extract_mc(grid=grid, desc=d))
    return out


def mesh_entry(size, m) -> dict:
    return {
        "size": size,
        "vertices": int(len(m.vertices)),
        "triangles": int(len(m.indices) // 3),
        "vertices_sha256": digest(m.vertices.astype("<f4")),
        "indices_sha256": digest(m.indices.astype("<u4")),
    }


# DualContouring over the sources' own HermiteSource normals (epsilon 0).
DC_HERMITE = {
    "dc_hermite_sphere_r0.35_size33": (lambda: W.sphere_source(), 33),
    "dc_hermite_boxes32_size128": (lambda: W.boxes_union(32, seed=4), 128),
}


def dc_hermite_cases(names=None):
    return {name: (size, O.extract_dc(make().program(), O.reference_desc(size), 0.0))
            for name, (make, size) in DC_HERMITE.items() if names is None or name in names}


def main() -> None:
    path = Path(__file__).with_name("mesh_digests.json")
    if len(sys.argv) > 2 and sys.argv[1] == "--only":
        names = set(sys.argv[2].split(","))
        known = {smooth_name(size) for size in SMOOTH_SIZES} | set(DC_HERMITE)
        if names - known:
            raise SystemExit(f"--only supports {sorted(known)}; unknown: {sorted(names - known)}")
        out = json.loads(path.read_text())
        fresh = {k: mesh_entry(*v) for k, v in smooth_cases(names).items()}
        fresh.update({k: mesh_entry(*v) for k, v in dc_hermite_cases(names).items()})
        out.update(fresh)
        for k, v in fresh.items():
            print(k, v["vertices"], v["triangles"], flush=True)
        path.write_text(json.dumps(out, indent=1) + "\n")
        return
    cases = {
        "sphere_r0.35_size65": (W.sphere_source(), 65),
        "spheres64_minus_torus_size65": (W.spheres_minus_torus()[0], 65),
        "spheres64_minus_torus_size129": (W.spheres_minus_torus()[0], 129),
        "spheres64_minus_torus_size257": (W.spheres_minus_torus()[0], 257),
    }
    dc_cases = {
        "dc_sphere_r0.35_size33": (W.sphere_source(), 33),
        "dc_boxes32_size65": (W.boxes_union(32, seed=4), 65),
        "dc_boxes32_size128": (W.boxes_union(32, seed=4), 128),
        "dc_spheres64_minus_torus_size65": (W.spheres_minus_torus()[0], 65),
    }
    out = {}
    meshes = [(name, size, O.extract_mc(src.program(), desc=O.reference_desc(size)))
              for name, (src, size) in cases.items()]
    meshes += [(name, size, O.extract_dc(src.program(), O.reference_desc(size), 1e-6))
               for name, (src, size) in dc_cases.items()]
    lhmc_cases = {
        "lhmc_boxes32_depth7": (W.boxes_union(32, seed=4), 7),
        "lhmc_sphere_r0.35_depth8": (W.sphere_source(), 8),
    }
    meshes += [(name, depth, O.extract_lhmc(src.program(), depth, 0))
               for name, (src, depth) in lhmc_cases.items()]
    boxes = W.boxes_union(32, seed=4)
    meshes.append(("directed_boxes32_size128", 128,
                   O.extract_mc(boxes.program(), desc=O.reference_desc(128), directed=True)))
    meshes.append(("emc_boxes32_size128", 128, O.extract_emc(boxes.program(), O.reference_desc(128))))
    meshes += [(name, size, m) for name, (size, m) in smooth_cases().items()]
    meshes += [(name, size, m) for name, (size, m) in dc_hermite_cases().items()]
    for name, size, m in meshes:
        out[name] = mesh_entry(size, m)
        print(name, out[name]["vertices"], out[name]["triangles"], flush=True)
    # The batched-chunk workload: per-chunk oracle meshes concatenated in chunk order.
    from isosurface_amd.engine import chunk_desc  # noqa: E402
    vs, ix = [], []
    for src, size in W.batch_chunks(4096, seed=6):
        m = O.extract_mc(src.program(), desc=chunk_desc(size))
        vs.append(m.vertices)
        ix.append(m.indices)
    import numpy as np  # noqa: E402
    v = np.concatenate(vs)
    i = np.concatenate(ix)
    out["batch4096_seed6"] = {
        "size": 0,
        "vertices": int(len(v)),
        "triangles": int(len(i) // 3),
        "vertices_sha256": digest(v.astype("<f4")),
        "indices_sha256": digest(i.astype("<u4")),
    }
    print("batch4096_seed6", out["batch4096_seed6"]["vertices"],
          out["batch4096_seed6"]["triangles"], flush=True)
    # PointCloud and CentralDifference(1e-4) vertex normals of the CSG at 257.
    csg = W.spheres_minus_torus()[0]
    pts = O.extract_point_cloud(csg.program(), desc=O.reference_desc(257))
    out["pointcloud_spheres64_minus_torus_size257"] = {
        "size": 257, "points": int(len(pts)), "points_sha256": digest(pts.astype("<f4"))}
    mv = O.extract_mc(csg.program(), desc=O.reference_desc(257)).vertices
    nrm = O.sample_normals(csg.program(), mv, 1e-4)
    out["normals_spheres64_minus_torus_size257_eps1e-4"] = {
        "size": 257, "points": int(len(nrm)), "normals_sha256": digest(nrm.astype("<f4"))}
    print("pointcloud", len(pts), "normals", len(nrm), flush=True)
    path.write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
