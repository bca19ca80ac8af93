"""Writes mesh_digests.json: counts and SHA-256 digests of the oracle's
meshes (marching cubes Signed and Directed, dual contouring, octree MC,
extended MC, batched chunks), point cloud and vertex normals for the
benchmark workloads.

The oracle (oracle/, a restatement of the reference's CPU extraction) is run
here once; the GPU tests compare their output digests with these values, so
full-size parity needs no CPU re-run on the GPU host.

usage: python tests/golden/make_meshes.py
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


def main() -> None:
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
    for name, size, m in meshes:
        out[name] = {
            "size": size,
            "vertices": int(len(m.vertices)),
            "triangles": int(len(m.indices) // 3),
            "vertices_sha256": digest(m.vertices.astype("<f4")),
            "indices_sha256": digest(m.indices.astype("<u4")),
        }
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
    path = Path(__file__).with_name("mesh_digests.json")
    path.write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
