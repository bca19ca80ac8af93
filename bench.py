#!/usr/bin/env python3
"""Benchmark: marching cubes at 256^3 cells on one MI355X.

Workload (one "step"): MarchingCubes over a 257 x 257 x 258 FP32 sample
lattice already resident in HBM — the reference's MarchingCubes::new(257)
layout, 256 x 256 x 257 = 16,842,752 cells — sampled from the union of 64
seeded spheres minus a torus.  A step is one full extraction: sign
classification, per-segment counts, scan, vertex and triangle emission into
device buffers, and the read-back of the mesh sizes.  The output is
bit-identical to the reference's CPU path (checked here against the golden
digests of the oracle's mesh).

Printed JSON line:
  value         cells/s of the timed steps (wall clock, synchronised)
  roofline      the classify + interpolate + emit stage of the headline path
                (the single-pass kernel, or the three-pass path's kernels):
                algorithmic bytes (4 B per sample + 12 B per vertex + 12 B
                per triangle) / the sum of the kernels' mean durations over
                the timed steps (HIP events attached to each kernel's
                dispatch, so they span the kernel alone), against MI355X's 8 TB/s HBM
                peak; traffic from the committed counter summary
  kernels       each kernel's own algorithmic bytes, duration and fraction
  device_copy   a 1 GiB device-to-device copy: the measured bandwidth the
                stage's fraction is also quoted against
  large_grid    MarchingCubes::new(513) over a 541 MB resident grid (above the
                256 MB Infinity Cache) by both paths, meshes compared
  program_path  the same extraction sampling the CSG program on the GPU
                instead of reading the grid (VALU-bound, reported aside)
  point_cloud, vertex_normals
                PointCloud over the same resident grid, and the
                CentralDifference(1e-4) normals of the mesh's vertices (the
                normal sinks' pass), checked against the oracle's digests
  batched_chunks
                4096 chunks of 16^3..64^3 cells, each its own seeded 1-8
                primitive CSG program, extracted by one captured HIP graph
                per batch after the autotuner picked the launch shapes
  learned_band  1024^3 cells of the smooth-min (k=0.02) blend of 1024 seeded
                spheres: the Fourier-feature MLP is trained on the GPU on 2M
                samples the plain path generates, its margin calibrated on
                held-out samples; band extraction is timed against the plain
                dense GPU extraction of the same program, which checks it
                (identical mesh, missed surface cells).  The network's own
                culls are statistical (a held-out calibration, no proof), so
                its missed_cells can be nonzero; with the certificate a brick
                leaves the band only when a super-brick or brick centre
                sample proves it (the field is 1-Lipschitz), and the network
                culls that no proof allows are counted (overturned_culls)
  dual_contouring, octree_mc
                the sharp-feature workload (union of 32 seeded boxes):
                DualContouring at 128^3 with central-difference normals and
                the f64 QEF, LinearHashedMarchingCubes at max_depth 7,
                MarchingCubes::<Directed> at 128^3 and ExtendedMarchingCubes
                at 128^3; each checked against the oracle's golden digests
  cpu_baseline  the oracle (a single-threaded restatement of the reference
                CPU path) on the same lattice, median of a few runs
"""
from __future__ import annotations

import argparse
import ctypes as C
import hashlib
import json
import os
import platform
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from isosurface_amd import Context, _abi as A  # noqa: E402
from isosurface_amd import workloads as W  # noqa: E402

METRIC = "MC cells/sec @256³ + % HBM roofline; learned-band speedup at 0 missed cells"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


class DeviceMesh:
    def __init__(self, ctx: Context, nv: int, ni: int):
        self.ctx = ctx
        self.v = C.c_void_p()
        self.i = C.c_void_p()
        self.nv, self.ni = nv, ni
        ctx.check(ctx.lib.iso_device_alloc(ctx.handle, 12 * nv, C.byref(self.v)), "alloc")
        ctx.check(ctx.lib.iso_device_alloc(ctx.handle, 4 * ni, C.byref(self.i)), "alloc")

    def out(self, flags: int = 0) -> A.IsoMeshOut:
        return A.IsoMeshOut(self.v, self.i, self.nv, self.ni, 0, 0, A.ISO_MEMORY_DEVICE, flags)

    def free(self):
        self.ctx.lib.iso_device_free(self.ctx.handle, self.v)
        self.ctx.lib.iso_device_free(self.ctx.handle, self.i)


def extract(ctx: Context, desc, mesh: DeviceMesh, graph: bool = False,
            flags: int = 0) -> A.IsoMeshOut:
    out = mesh.out(flags | (A.ISO_MESH_GRAPH if graph else 0))
    ctx.check(ctx.lib.iso_extract_mc(ctx.handle, C.byref(desc), C.byref(out)), "iso_extract_mc")
    return out


def timed_steps(ctx, desc, mesh, steps, graph=False, flags=0):
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        extract(ctx, desc, mesh, graph, flags)
    ctx.synchronize()
    return time.perf_counter() - t0


def kernel_pass(ctx, desc, mesh, steps, flags=0):
    """Per-kernel mean durations (HIP events taken by each kernel's dispatch)."""
    ctx.reset_kernel_stats()
    ctx.set_timing(True)
    for _ in range(steps):
        extract(ctx, desc, mesh, flags=flags)
    ctx.synchronize()
    ctx.set_timing(False)
    stats = ctx.kernel_stats()
    ctx.reset_kernel_stats()
    return {k: {"launches": n, "mean_us": 1e3 * ms / max(n, 1)} for k, (n, ms) in stats.items()}


def measured_traffic():
    """HBM bytes per launch of the stage's kernels, as measured by the
    committed FETCH_SIZE / WRITE_SIZE counter passes (scaled by a 64 MiB
    copy; see the summary's header), or None when the summary is absent."""
    path = ROOT / "profiles" / "mc_grid_257_hbm_traffic.md"
    if not path.exists():
        return None
    out = {"kernels": {}}
    for line in path.read_text().splitlines():
        parts = line.strip().split()
        if len(parts) == 2 and parts[0].startswith("traffic_bytes."):
            name = parts[0][len("traffic_bytes."):].rstrip(":")
            if name == "stage":
                out["stage"] = int(parts[1])
            else:
                out["kernels"][name] = int(parts[1])
    return out if out["kernels"] else None


# VALU issue peak: 256 CUs x 4 SIMDs at 2.4 GHz, one wave64 FP32 VALU
# instruction per SIMD every 2 clocks (32 lanes a clock: the rate behind the
# 157.3 TFLOP/s FP32 vector figure).
VALU_PEAK_GINSTR = 256 * 4 * 2.4e9 / 2 / 1e9
FP32_PEAK_TFLOPS = 157.3


def valu_counters():
    """VALU wave-instructions and FP32 FLOPs per launch of the VALU-bound
    kernels, from the committed SQ counter passes over `bench.py
    --valu-probe`: {kernel: {"instr": n, "flop": n}}, or None."""
    path = ROOT / "profiles" / "valu_counters.md"
    if not path.exists():
        return None
    out = {}
    for line in path.read_text().splitlines():
        parts = line.strip().split()
        if len(parts) != 2:
            continue
        for key, field in (("valu_instr.", "instr"), ("valu_f32_flop.", "flop")):
            if parts[0].startswith(key):
                out.setdefault(parts[0][len(key):].rstrip(":"), {})[field] = int(parts[1])
    return out or None


def valu_roofline(durations_us: dict) -> dict:
    """Achieved VALU issue rate of each kernel: counted instructions per
    launch over its live mean duration."""
    counts = valu_counters()
    if not counts:
        return None
    res = {}
    for k, us in durations_us.items():
        if k in counts and us and "instr" in counts[k]:
            rate = counts[k]["instr"] / (us * 1e-6) / 1e9
            flops = counts[k].get("flop", 0) / (us * 1e-6) / 1e12
            res[k] = {"bound": "valu", "valu_wave_instr": counts[k]["instr"], "mean_us": us,
                      "achieved": rate, "peak": VALU_PEAK_GINSTR, "unit": "G wave-instr/s",
                      "frac": rate / VALU_PEAK_GINSTR,
                      "f32_tflops": flops, "f32_frac": flops / FP32_PEAK_TFLOPS}
    return res


def valu_probe(ctx: Context, size: int) -> None:
    """A few launches of each VALU-bound kernel, for SQ counter passes: the
    program-path signs at size^3, DualContouring::new(128) over
    CentralDifference on the 32-box union (corner normals and feature
    points), and the certificate-only narrow
    band at 1025^3 on the smooth spheres (the bench's configurations)."""
    src, _ = W.spheres_minus_torus()
    ctx.upload_program(src.program())
    pdesc = ctx.reference_desc(size, A.ISO_SOURCE_PROGRAM)
    mesh = DeviceMesh(ctx, 1 << 20, 3 << 21)
    for _ in range(3):
        ctx.check(ctx.lib.iso_extract_mc(ctx.handle, C.byref(pdesc), C.byref(mesh.out())), "p")
    mesh.free()
    boxes = W.boxes_union(32, seed=4)
    ctx.upload_program(boxes.program())
    ddesc = ctx.reference_desc(128, A.ISO_SOURCE_PROGRAM)
    mesh = DeviceMesh(ctx, 1 << 20, 3 << 21)
    for _ in range(3):
        ctx.check(ctx.lib.iso_extract_dc(ctx.handle, C.byref(ddesc), C.c_float(1e-6),
                                         C.byref(mesh.out())), "dc")
    mesh.free()
    smooth = W.smooth_spheres(1024, k=0.02, seed=5)
    ctx.upload_program(smooth.program())
    bdesc = ctx.reference_desc(1025, A.ISO_SOURCE_PROGRAM)
    mesh = DeviceMesh(ctx, 1 << 23, 3 << 24)
    for _ in range(2):
        st = ctx.lib.iso_extract_mc_band(ctx.handle, C.byref(bdesc), C.c_float(0.0),
                                         A.ISO_BAND_NO_NETWORK, C.byref(mesh.out()), None)
        if st not in (A.ISO_OK, A.ISO_ERR_CAPACITY):
            ctx.check(st, "band")
    mesh.free()
    ctx.synchronize()


def digest(ctx, mesh: DeviceMesh, out) -> tuple:
    nv, ni = int(out.vertex_count), int(out.index_count)
    v = np.empty((nv, 3), np.float32)
    i = np.empty(ni, np.uint32)
    ctx.lib.iso_memcpy_to_host(ctx.handle, v.ctypes.data, mesh.v, v.nbytes)
    ctx.lib.iso_memcpy_to_host(ctx.handle, i.ctypes.data, mesh.i, i.nbytes)
    return (hashlib.sha256(v.astype("<f4").tobytes()).hexdigest(),
            hashlib.sha256(i.astype("<u4").tobytes()).hexdigest())


def host_cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), or the platform's name."""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.lower().startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(grid: np.ndarray, size: int, budget_s: float) -> dict:
    """The oracle (restated reference CPU path, one thread) on the same lattice."""
    from oracle import oracle as O

    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
        pinned = True
    except (AttributeError, OSError):
        pinned = False
    desc = O.reference_desc(size, source=1)
    times = []
    t_start = time.perf_counter()
    while len(times) < 5 and (time.perf_counter() - t_start) < budget_s:
        t0 = time.perf_counter()
        O.extract_mc(grid=grid, desc=desc)
        times.append(time.perf_counter() - t0)
    cells = (size - 1) * (size - 1) * size
    med = statistics.median(times)
    return {
        "value": cells / med,
        "unit": "cells/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle MarchingCubes over the same resident {size}x{size}x{size + 1} grid, "
                   f"{len(times)} runs, median {med:.3f} s, pinned={pinned}, "
                   f"host CPU {host_cpu_model()}, {os.cpu_count()} logical cpus"),
        "host_cpu": host_cpu_model(),
    }


def side_workloads(ctx: Context, steps: int, with_cpu: bool) -> dict:
    """Dual contouring at 128^3 and octree MC at depth 7 on the 32-box union."""
    golden = json.loads((ROOT / "tests" / "golden" / "mesh_digests.json").read_text())
    src = W.boxes_union(32, seed=4)
    ctx.upload_program(src.program())
    res = {}

    def run(name, call, key, cells, oracle_call):
        probe = DeviceMesh(ctx, 1, 3)
        out = probe.out()
        st = call(out)
        if st not in (A.ISO_OK, A.ISO_ERR_CAPACITY):
            ctx.check(st, name)
        nv, ni = int(out.vertex_count), int(out.index_count)
        probe.free()
        mesh = DeviceMesh(ctx, max(nv, 1), max(ni, 3))
        out = mesh.out()
        ctx.check(call(out), name)
        dv, di = digest(ctx, mesh, out)
        parity = dv == golden[key]["vertices_sha256"] and di == golden[key]["indices_sha256"]
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.check(call(mesh.out()), name)
        ctx.synchronize()
        dt = (time.perf_counter() - t0) / steps
        ctx.reset_kernel_stats()
        ctx.set_timing(True)
        ctx.check(call(mesh.out()), name)
        ctx.synchronize()
        ctx.set_timing(False)
        ks = {k: round(1e3 * ms / max(n, 1), 2) for k, (n, ms) in ctx.kernel_stats().items()}
        ctx.reset_kernel_stats()
        mesh.free()
        entry = {"ms_per_extraction": 1e3 * dt, "cells_per_s": cells / dt, "vertices": nv,
                 "triangles": ni // 3, "parity_vs_oracle": parity, "kernels_us": ks}
        if with_cpu:
            t0 = time.perf_counter()
            oracle_call()
            entry["cpu_oracle_s"] = time.perf_counter() - t0
        log(f"{name}: {1e3 * dt:.3f} ms, parity {parity}")
        return entry

    from_oracle = None
    if with_cpu:
        from oracle import oracle as O
        from_oracle = O
    dc_desc = ctx.reference_desc(128, A.ISO_SOURCE_PROGRAM)
    res["dual_contouring"] = run(
        "iso_extract_dc",
        lambda out: ctx.lib.iso_extract_dc(ctx.handle, C.byref(dc_desc), C.c_float(1e-6),
                                           C.byref(out)),
        "dc_boxes32_size128", 127 * 127 * 128,
        lambda: from_oracle.extract_dc(src.program(), from_oracle.reference_desc(128), 1e-6))
    res["dual_contouring"]["workload"] = (
        "DualContouring::new(128, MinimiseQEF) on CentralDifference(union of 32 seeded boxes), "
        "128x128x129 primal samples, f64 QEF per mixed cell")
    res["dual_contouring_hermite"] = run(
        "iso_extract_dc (analytic normals)",
        lambda out: ctx.lib.iso_extract_dc(ctx.handle, C.byref(dc_desc), C.c_float(0.0),
                                           C.byref(out)),
        "dc_hermite_boxes32_size128", 127 * 127 * 128,
        lambda: from_oracle.extract_dc(src.program(), from_oracle.reference_desc(128), 0.0))
    res["dual_contouring_hermite"]["workload"] = (
        "DualContouring::new(128, MinimiseQEF) on the union of 32 seeded boxes with their own "
        "(HermiteSource) normals, 128x128x129 primal samples")
    res["octree_mc"] = run(
        "iso_extract_lhmc",
        lambda out: ctx.lib.iso_extract_lhmc(ctx.handle, 7, C.byref(out)),
        "lhmc_boxes32_depth7", 128 ** 3,
        lambda: from_oracle.extract_lhmc(src.program(), 7, 0))
    res["octree_mc"]["workload"] = ("LinearHashedMarchingCubes::new(7) on the union of 32 "
                                    "seeded boxes (cells/s counted against the 128^3 grid)")
    res["directed_mc"] = run(
        "iso_extract_mc_directed",
        lambda out: ctx.lib.iso_extract_mc_directed(ctx.handle, C.byref(dc_desc), C.byref(out)),
        "directed_boxes32_size128", 127 * 127 * 128,
        lambda: from_oracle.extract_mc(src.program(), desc=from_oracle.reference_desc(128),
                                       directed=True))
    res["directed_mc"]["workload"] = ("MarchingCubes::<Directed>::new(128) on the union of 32 "
                                      "seeded boxes (sample_vector evaluated on the GPU)")
    res["extended_mc"] = run(
        "iso_extract_emc",
        lambda out: ctx.lib.iso_extract_emc(ctx.handle, C.byref(dc_desc), C.c_float(0.0),
                                            C.byref(out)),
        "emc_boxes32_size128", 127 * 127 * 128,
        lambda: from_oracle.extract_emc(src.program(), from_oracle.reference_desc(128)))
    res["extended_mc"]["workload"] = ("ExtendedMarchingCubes::new(128) on the union of 32 seeded "
                                      "boxes with their own (HermiteSource) normals")
    return res


def points_and_normals(ctx: Context, desc, mesh: "DeviceMesh", nv: int, steps: int) -> dict:
    """PointCloud over the resident 257^3 grid and the CentralDifference(1e-4)
    normals of the MarchingCubes::new(257) vertices (the normal sinks'
    per-vertex pass), both into device memory, checked against the oracle's
    digests."""
    golden = json.loads((ROOT / "tests" / "golden" / "mesh_digests.json").read_text())
    gp = golden["pointcloud_spheres64_minus_torus_size257"]
    gn = golden["normals_spheres64_minus_torus_size257_eps1e-4"]
    res = {}
    cap = gp["points"]
    pbuf = C.c_void_p()
    ctx.check(ctx.lib.iso_device_alloc(ctx.handle, 12 * cap, C.byref(pbuf)), "alloc")

    def cloud():
        out = A.IsoMeshOut(pbuf, None, cap, 0, 0, 0, A.ISO_MEMORY_DEVICE, 0)
        ctx.check(ctx.lib.iso_extract_point_cloud(ctx.handle, C.byref(desc), C.byref(out)),
                  "iso_extract_point_cloud")
        return int(out.vertex_count)

    n = cloud()
    pts = np.empty((n, 3), np.float32)
    ctx.lib.iso_memcpy_to_host(ctx.handle, pts.ctypes.data, pbuf, pts.nbytes)
    parity = n == gp["points"] and hashlib.sha256(pts.astype("<f4").tobytes()).hexdigest() == \
        gp["points_sha256"]
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        cloud()
    dt = (time.perf_counter() - t0) / steps
    ctx.lib.iso_device_free(ctx.handle, pbuf)
    cells = (desc.nx - 1) * (desc.ny - 1) * (desc.nz - 1)
    res["point_cloud"] = {
        "workload": "PointCloud::new(257) over the resident 257x257x258 FP32 grid (device output)",
        "ms_per_extraction": 1e3 * dt, "cells_per_s": cells / dt, "points": n,
        "parity_vs_oracle": parity}
    log(f"iso_extract_point_cloud: {1e3 * dt:.3f} ms, parity {parity}")

    nbuf = C.c_void_p()
    ctx.check(ctx.lib.iso_device_alloc(ctx.handle, 12 * nv, C.byref(nbuf)), "alloc")

    def normals():
        ctx.check(ctx.lib.iso_sample_normals(ctx.handle, mesh.v, nv, C.c_float(1e-4), nbuf,
                                             A.ISO_MEMORY_DEVICE), "iso_sample_normals")

    normals()
    nrm = np.empty((nv, 3), np.float32)
    ctx.lib.iso_memcpy_to_host(ctx.handle, nrm.ctypes.data, nbuf, nrm.nbytes)
    parity = nv == gn["points"] and hashlib.sha256(nrm.astype("<f4").tobytes()).hexdigest() == \
        gn["normals_sha256"]
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        normals()
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / steps
    ctx.lib.iso_device_free(ctx.handle, nbuf)
    res["vertex_normals"] = {
        "workload": ("CentralDifference::new_with_epsilon(CSG, 1e-4).sample_normal at the "
                     f"{nv} vertices of MarchingCubes::new(257) (the normal sinks' pass, "
                     "device memory)"),
        "ms": 1e3 * dt, "normals_per_s": nv / dt, "parity_vs_oracle": parity}
    log(f"iso_sample_normals: {1e3 * dt:.3f} ms for {nv} vertices, parity {parity}")
    return res


def batched_chunks(ctx: Context, steps: int) -> dict:
    """The batched-chunk workload: 4096 chunks, one graph replay per batch."""
    from isosurface_amd.engine import chunk_desc
    golden = json.loads((ROOT / "tests" / "golden" / "mesh_digests.json").read_text())
    g = golden["batch4096_seed6"]
    work = W.batch_chunks(4096, seed=6)
    cells = sum(size ** 3 for _, size in work)
    ctx.upload_batch([(src, chunk_desc(size)) for src, size in work])
    n = len(work)
    nv, ni = g["vertices"], 3 * g["triangles"]
    v, i = C.c_void_p(), C.c_void_p()
    ctx.check(ctx.lib.iso_device_alloc(ctx.handle, 12 * nv, C.byref(v)), "alloc")
    ctx.check(ctx.lib.iso_device_alloc(ctx.handle, 4 * ni, C.byref(i)), "alloc")
    voff = np.zeros(n + 1, np.uint64)
    ioff = np.zeros(n + 1, np.uint64)
    flags = A.ISO_BATCH_GRAPH | A.ISO_BATCH_AUTOTUNE

    def call():
        out = A.IsoBatchOut(v, i, nv, ni, voff.ctypes.data, ioff.ctypes.data, 0, 0,
                            A.ISO_MEMORY_DEVICE, flags)
        ctx.check(ctx.lib.iso_extract_batch(ctx.handle, C.byref(out)), "iso_extract_batch")
        return out

    t0 = time.perf_counter()
    out = call()  # eager pass, autotuning, graph capture
    first = time.perf_counter() - t0
    call()  # first replay
    vv = np.empty((int(out.vertex_count), 3), np.float32)
    ii = np.empty(int(out.index_count), np.uint32)
    ctx.lib.iso_memcpy_to_host(ctx.handle, vv.ctypes.data, v, vv.nbytes)
    ctx.lib.iso_memcpy_to_host(ctx.handle, ii.ctypes.data, i, ii.nbytes)
    parity = (hashlib.sha256(vv.astype("<f4").tobytes()).hexdigest() == g["vertices_sha256"] and
              hashlib.sha256(ii.astype("<u4").tobytes()).hexdigest() == g["indices_sha256"])
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        call()
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / steps
    # One eager pass (same tuned shapes) with per-kernel events; graph
    # replays record none.
    flags = A.ISO_BATCH_AUTOTUNE
    ctx.reset_kernel_stats()
    ctx.set_timing(True)
    call()
    ctx.synchronize()
    ctx.set_timing(False)
    kus = {k: round(1e3 * ms / max(c, 1), 2) for k, (c, ms) in ctx.kernel_stats().items()}
    ctx.reset_kernel_stats()
    ctx.lib.iso_device_free(ctx.handle, v)
    ctx.lib.iso_device_free(ctx.handle, i)
    nvo, nio = int(out.vertex_count), int(out.index_count)
    # emit's algorithmic bytes: the mesh written, and the two stored corner
    # samples each vertex interpolates between
    emit_bytes = 12 * nvo + 4 * nio + 8 * nvo
    emit_us = kus.get("batch_emit")
    log(f"batched chunks: {1e3 * dt:.3f} ms per batch, parity {parity}, kernels {kus}")
    return {"workload": ("4096 chunks, S^3 cells with S uniform in {16,32,48,64}, each a seeded "
                         "1-8 sphere/box/torus CSG program (seed 6); one HIP graph replay per "
                         "batch, launch shapes autotuned on the first extraction"),
            "cells": cells, "ms_per_batch": 1e3 * dt, "cells_per_s": cells / dt,
            "first_call_ms": 1e3 * first, "vertices": nvo,
            "triangles": nio // 3, "parity_vs_oracle": parity, "kernels_us": kus,
            "emit_algorithmic_bytes": emit_bytes,
            "emit_GBps": emit_bytes / (emit_us * 1e-6) / 1e9 if emit_us else None}


def brick_centres(size: int) -> np.ndarray:
    """Centres of the band bricks of MarchingCubes::new(size) (x fastest),
    as the band kernels place them."""
    from isosurface_amd.band import brick_size
    bs = brick_size()
    sp = np.float32(1.0) / np.float32(size - 1)
    axes = []
    for n in (size, size, size + 1):
        lo = np.arange(0, n - 1, bs)
        hi = np.minimum(lo + bs, n - 1)
        axes.append(((0.5 * (lo + hi)).astype(np.float32) * sp).astype(np.float32))
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return np.ascontiguousarray(np.stack([x, y, z], -1).reshape(-1, 3), np.float32)


def copy_peak(ctx: Context, mib: int = 1024, reps: int = 10) -> dict:
    """Device-to-device copy bandwidth on this GPU (read + write bytes per
    second), the measured ceiling the stage's fraction is also quoted
    against."""
    nb = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    ctx.check(ctx.lib.iso_device_alloc(ctx.handle, nb, C.byref(a)), "alloc")
    ctx.check(ctx.lib.iso_device_alloc(ctx.handle, nb, C.byref(b)), "alloc")
    ctx.check(ctx.lib.iso_memcpy_device(ctx.handle, b, a, nb), "copy")
    times = []
    for _ in range(reps):
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.check(ctx.lib.iso_memcpy_device(ctx.handle, b, a, nb), "copy")
        ctx.synchronize()
        times.append(time.perf_counter() - t0)
    ctx.lib.iso_device_free(ctx.handle, a)
    ctx.lib.iso_device_free(ctx.handle, b)
    best = min(times)
    gbs = 2 * nb / best / 1e9
    log(f"device copy: {mib} MiB in {1e3 * best:.3f} ms, {gbs:.0f} GB/s (read + write)")
    return {"bytes": nb, "ms": 1e3 * best, "GBps": gbs,
            "note": "hipMemcpy device to device, best of runs, read + write bytes"}


def large_grid(ctx: Context, size: int, steps: int, copy_gbs: float) -> dict:
    """MarchingCubes::new(size) over a resident grid larger than the 256 MB
    Infinity Cache (size 513: 513 x 513 x 514 FP32 samples, 541 MB), by both
    paths; the two meshes must be identical."""
    src, _ = W.spheres_minus_torus()
    ctx.upload_program(src.program())
    desc = ctx.reference_desc(size, A.ISO_SOURCE_GRID)
    n = desc.nx * desc.ny * desc.nz
    cells = (desc.nx - 1) * (desc.ny - 1) * (desc.nz - 1)
    dgrid = C.c_void_p()
    ctx.check(ctx.lib.iso_device_alloc(ctx.handle, 4 * n, C.byref(dgrid)), "alloc grid")
    ctx.check(ctx.lib.iso_program_sample_lattice(ctx.handle, C.byref(desc), dgrid,
                                                 A.ISO_MEMORY_DEVICE), "sample_lattice")
    ctx.check(ctx.lib.iso_grid_bind_device(ctx.handle, dgrid, desc.nx, desc.ny, desc.nz), "bind")
    probe = DeviceMesh(ctx, 1, 3)
    out = probe.out()
    st = ctx.lib.iso_extract_mc(ctx.handle, C.byref(desc), C.byref(out))
    if st not in (A.ISO_OK, A.ISO_ERR_CAPACITY):
        ctx.check(st, "iso_extract_mc")
    nv, ni = int(out.vertex_count), int(out.index_count)
    probe.free()
    mesh = DeviceMesh(ctx, max(nv, 1), max(ni, 3))
    single = A.ISO_MESH_SINGLE_PASS
    res = {"workload": (f"MarchingCubes::new({size}) over a resident FP32 "
                        f"{desc.nx}x{desc.ny}x{desc.nz} grid ({4 * n / 1e6:.0f} MB, above the "
                        "256 MB Infinity Cache) of the union of 64 seeded spheres minus a torus"),
           "cells": cells, "samples": n, "vertices": nv, "triangles": ni // 3}
    digests = {}
    stage_bytes = 4 * n + 12 * nv + 4 * ni
    for name, flags, stage in (("three_pass", 0, ["mc_signs_grid", "mc_count", "mc_emit"]),
                               ("single_pass", single, ["mc_fused"])):
        o = extract(ctx, desc, mesh, flags=flags)
        digests[name] = digest(ctx, mesh, o)
        for _ in range(2):
            extract(ctx, desc, mesh, flags=flags)
        el = timed_steps(ctx, desc, mesh, steps, flags=flags)
        ks = kernel_pass(ctx, desc, mesh, steps, flags=flags)
        us = sum(ks[k]["mean_us"] for k in stage)
        achieved = stage_bytes / (us * 1e-6) / 1e9
        res[name] = {"ms_per_step": 1e3 * el / steps, "cells_per_s": cells * steps / el,
                     "kernels_us": {k: round(v["mean_us"], 2) for k, v in ks.items()},
                     "stage_us": us, "achieved_GBps": achieved,
                     "frac_of_8TBps": achieved / HBM_PEAK_GBS,
                     "frac_of_measured_copy": achieved / copy_gbs if copy_gbs else None}
        log(f"{size}^3 {name}: {1e3 * el / steps:.3f} ms/step, stage {us:.1f} us, "
            f"{achieved:.0f} GB/s = {achieved / HBM_PEAK_GBS:.3f} of 8 TB/s")
    res["algorithmic_bytes"] = stage_bytes
    res["paths_identical"] = digests["three_pass"] == digests["single_pass"]
    mesh.free()
    ctx.lib.iso_device_free(ctx.handle, dgrid)
    return res


def learned_band(ctx: Context, size: int, reps: int, steps: int, seeds=(13, 14, 15)) -> dict:
    """The learned narrow band against the dense path at size^3 cells, with
    4x4x4 and 8x8x8-cell bricks; the network is trained once per seed."""
    from isosurface_amd.band import NarrowBand, brick_half_diagonal, training_samples
    src = W.smooth_spheres(1024, k=0.02, seed=5)
    upper = (1.0, 1.0, (size + 1) / size)  # the lattice's extra z layer
    bricks = {"4": 0, "8": A.ISO_BAND_BRICK8}
    hd = {b: brick_half_diagonal(size + 1, f) for b, f in bricks.items()}
    t0 = time.perf_counter()
    pts, sdf = training_samples(ctx, src, 1_000_000, 1_000_000, seed=11, coarse=257, upper=upper)
    cal_p, cal_s = training_samples(ctx, src, 500_000, 500_000, seed=12, coarse=257, upper=upper)
    t_samples = time.perf_counter() - t0
    band = NarrowBand(ctx)
    centres = brick_centres(size + 1)
    # The same network before training: its calibrated margin keeps every brick.
    band.init(7)
    m0 = band.calibrate(cal_p, cal_s, hd["4"])
    untrained_fraction = float(np.mean(~(np.abs(band.predict(centres)) > hd["4"] + m0)))
    ctx.upload_program(src.program())
    desc = ctx.reference_desc(size + 1, A.ISO_SOURCE_PROGRAM)
    cells = size * size * (size + 1)
    probe = DeviceMesh(ctx, 1, 3)
    out = probe.out()
    st = ctx.lib.iso_extract_mc(ctx.handle, C.byref(desc), C.byref(out))
    if st not in (A.ISO_OK, A.ISO_ERR_CAPACITY):
        ctx.check(st, "iso_extract_mc")
    nv, ni = int(out.vertex_count), int(out.index_count)
    probe.free()
    # Room for a band mesh that differs from the dense one.
    mesh = DeviceMesh(ctx, 2 * max(nv, 1), 2 * max(ni, 3))
    times = []
    for _ in range(max(1, reps)):
        ctx.synchronize()
        t0 = time.perf_counter()
        out = mesh.out()
        ctx.check(ctx.lib.iso_extract_mc(ctx.handle, C.byref(desc), C.byref(out)), "dense")
        ctx.synchronize()
        times.append(time.perf_counter() - t0)
    t_dense = min(times)
    dd = digest(ctx, mesh, out)

    def run(flags: int, mode: str, margin: float) -> dict:
        stats = A.IsoBandStats()
        times = []
        for _ in range(max(1, reps)):
            ctx.synchronize()
            t0 = time.perf_counter()
            outb = mesh.out()
            st = ctx.lib.iso_extract_mc_band(ctx.handle, C.byref(desc), margin, flags,
                                             C.byref(outb), C.byref(stats))
            ctx.synchronize()
            times.append(time.perf_counter() - t0)
            if st not in (A.ISO_OK, A.ISO_ERR_CAPACITY):
                ctx.check(st, "band")
        t_band = min(times)
        same = st == A.ISO_OK and \
            (int(outb.vertex_count), int(outb.index_count)) == (nv, ni) and \
            digest(ctx, mesh, outb) == dd
        missed = C.c_uint64(0)
        ctx.check(ctx.lib.iso_band_missed_cells(ctx.handle, C.byref(desc), C.byref(missed)),
                  "missed")
        # one more run with per-kernel events
        ctx.reset_kernel_stats()
        ctx.set_timing(True)
        st2 = ctx.lib.iso_extract_mc_band(ctx.handle, C.byref(desc), margin, flags,
                                          C.byref(mesh.out()), None)
        ctx.synchronize()
        ctx.set_timing(False)
        if st2 not in (A.ISO_OK, A.ISO_ERR_CAPACITY):
            ctx.check(st2, "band")
        kus = {k: round(1e3 * ms / max(n, 1), 2) for k, (n, ms) in ctx.kernel_stats().items()}
        ctx.reset_kernel_stats()
        r = {"band_ms": 1e3 * t_band, "kernels_us": kus, "speedup": t_dense / t_band,
             "band_cells_per_s": cells / t_band, "bricks": int(stats.bricks),
             "band_bricks": int(stats.band_bricks),
             "band_fraction": int(stats.band_bricks) / max(int(stats.bricks), 1),
             "exact_samples": int(stats.exact_samples),
             "centre_samples": int(stats.centre_samples),
             "overturned_culls": int(stats.overturned_culls),
             "missed_cells": int(missed.value), "identical_to_dense": same, "margin": margin}
        log(f"band [{mode}]: {1e3 * t_band:.1f} ms, x{t_dense / t_band:.2f}, band "
            f"{r['band_fraction']:.3f}, missed {missed.value}, identical {same}")
        return r

    overshoot, lr = 8.0, 2e-3
    weight = (1.0 + 9.0 * np.exp(-np.abs(sdf) / 0.01)).astype(np.float32)
    per_seed, training = [], []
    first = {}
    for i, seed in enumerate(seeds):
        band.init(7)
        t0 = time.perf_counter()
        loss = band.train(pts, sdf, weight, steps=steps, batch=16384, lr=lr, lr_min=5e-5,
                          clamp=0.05, seed=seed, overshoot=overshoot)
        training.append({"seed": seed, "final_loss": loss,
                         "train_ms": 1e3 * (time.perf_counter() - t0)})
        margins = {b: band.calibrate(cal_p, cal_s, hd[b]) for b in bricks}
        row = {"seed": seed}
        for b, f in bricks.items():
            row[f"brick{b}"] = run(f, f"learned, {b}^3 bricks, seed {seed}", margins[b])
            if i == 0:
                first[b] = {
                    "learned": row[f"brick{b}"],
                    "learned_with_certificate": run(
                        f | A.ISO_BAND_CERTIFY, f"learned+certificate, {b}^3 bricks", margins[b]),
                    "certificate_without_network": run(
                        f | A.ISO_BAND_NO_NETWORK, f"certificate only, {b}^3 bricks", margins[b]),
                }
        per_seed.append(row)
    mesh.free()
    res = {
        "workload": (f"MarchingCubes::new({size + 1}) ({cells} cells) on the smooth-min (k=0.02) "
                     "of 1024 seeded spheres, radii U[0.01,0.05], program source"),
        "dense_ms": 1e3 * t_dense, "dense_cells_per_s": cells / t_dense,
        "bricks": int(centres.shape[0]), "half_diagonal": hd["4"],
        "vertices": nv, "triangles": ni // 3,
        **first["4"]["learned"],
        "untrained_network_band_fraction": untrained_fraction,
        "learned_with_certificate": first["4"]["learned_with_certificate"],
        "certificate_without_network": first["4"]["certificate_without_network"],
        "brick8": {"half_diagonal": hd["8"], **first["8"]},
        "learned_only_over_seeds": per_seed,
        "training": {"samples": len(sdf), "steps": steps, "batch": 16384, "lr": lr,
                     "overshoot": overshoot, "runs": training,
                     "sample_generation_ms": 1e3 * t_samples},
    }
    for b in bricks:
        sp = [r[f"brick{b}"]["speedup"] for r in per_seed]
        ms = [r[f"brick{b}"]["missed_cells"] for r in per_seed]
        log(f"learned band, {b}^3 bricks: dense {1e3 * t_dense:.1f} ms; learned x"
            f"{', x'.join(f'{v:.2f}' for v in sp)} over seeds (missed {ms}); with certificate x"
            f"{first[b]['learned_with_certificate']['speedup']:.2f}; certificate alone x"
            f"{first[b]['certificate_without_network']['speedup']:.2f}")
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=257)
    ap.add_argument("--program-steps", type=int, default=5)
    ap.add_argument("--cpu-budget", type=float, default=25.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-side", action="store_true",
                    help="skip the batched, dual-contouring and octree workloads")
    ap.add_argument("--band-steps", type=int, default=100000,
                    help="training steps of the learned-band network")
    ap.add_argument("--band-size", type=int, default=1024,
                    help="cells per axis of the learned-band workload (0: skip it)")
    ap.add_argument("--large-size", type=int, default=513,
                    help="size of the above-Infinity-Cache grid workload (0: skip it)")
    ap.add_argument("--valu-probe", action="store_true",
                    help="only a few launches of the VALU-bound kernels (program-path "
                         "signs, DC features, band exact samples), for SQ counter passes")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only a few grid extractions and three 64 MiB device copies (the "
                         "scale reference), for hardware-counter passes")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the extraction path has no multi-GPU mode "
              "(sharding is a design note only); run with --gpus 1", file=sys.stderr)
        return 2

    size = args.size
    ctx = Context(0)
    if args.valu_probe:
        valu_probe(ctx, size)
        print(json.dumps({"valu_probe": True, "size": size}), flush=True)
        ctx.close()
        return 0
    src, _ = W.spheres_minus_torus()
    ctx.upload_program(src.program())
    desc = ctx.reference_desc(size, A.ISO_SOURCE_GRID)
    n = desc.nx * desc.ny * desc.nz
    cells = (desc.nx - 1) * (desc.ny - 1) * (desc.nz - 1)
    log(f"lattice {desc.nx}x{desc.ny}x{desc.nz}: {n} samples, {cells} cells")

    # Resident FP32 input grid, sampled from the CSG program on the device.
    dgrid = C.c_void_p()
    ctx.check(ctx.lib.iso_device_alloc(ctx.handle, 4 * n, C.byref(dgrid)), "alloc grid")
    ctx.check(ctx.lib.iso_program_sample_lattice(ctx.handle, C.byref(desc), dgrid,
                                                 A.ISO_MEMORY_DEVICE), "sample_lattice")
    ctx.check(ctx.lib.iso_grid_bind_device(ctx.handle, dgrid, desc.nx, desc.ny, desc.nz), "bind")

    # Size device outputs from a first extraction.
    mesh = DeviceMesh(ctx, 1, 3)
    out = mesh.out()
    st = ctx.lib.iso_extract_mc(ctx.handle, C.byref(desc), C.byref(out))
    if st not in (A.ISO_OK, A.ISO_ERR_CAPACITY):
        ctx.check(st, "iso_extract_mc")
    nv, ni = int(out.vertex_count), int(out.index_count)
    mesh.free()
    mesh = DeviceMesh(ctx, max(nv, 1), max(ni, 3))
    out = extract(ctx, desc, mesh)
    log(f"mesh: {nv} vertices, {ni // 3} triangles")

    parity = None
    golden = json.loads((ROOT / "tests" / "golden" / "mesh_digests.json").read_text())
    key = f"spheres64_minus_torus_size{size}"
    if key in golden:
        dv, di = digest(ctx, mesh, out)
        parity = dv == golden[key]["vertices_sha256"] and di == golden[key]["indices_sha256"]
        log(f"parity vs oracle digests: {parity}")

    if args.traffic_probe:
        # Counter passes read per-dispatch FETCH_SIZE / WRITE_SIZE: a few
        # single-pass and three-pass extractions, then copies of exactly 64 MiB.
        for _ in range(3):
            extract(ctx, desc, mesh)
        for _ in range(3):
            extract(ctx, desc, mesh, flags=A.ISO_MESH_SINGLE_PASS)
        nb = 64 << 20
        a, b = C.c_void_p(), C.c_void_p()
        ctx.check(ctx.lib.iso_device_alloc(ctx.handle, nb, C.byref(a)), "alloc")
        ctx.check(ctx.lib.iso_device_alloc(ctx.handle, nb, C.byref(b)), "alloc")
        for _ in range(3):
            ctx.check(ctx.lib.iso_memcpy_device(ctx.handle, b, a, nb), "copy")
        ctx.lib.iso_device_free(ctx.handle, a)
        ctx.lib.iso_device_free(ctx.handle, b)
        print(json.dumps({"traffic_probe": True, "copy_bytes": nb, "samples": n}), flush=True)
        return 0

    # Each step is one iso_extract_mc call that returns the mesh counts to
    # the host, by the three-pass path (classify, count and emit launches,
    # the library's default) and by the single pass (one launch,
    # ISO_MESH_SINGLE_PASS); the headline is the faster of the two, the
    # other and the three-pass graph replay are reported beside it.
    single = A.ISO_MESH_SINGLE_PASS
    for _ in range(args.warmup):
        extract(ctx, desc, mesh, graph=True)
    graph_elapsed = timed_steps(ctx, desc, mesh, args.steps, graph=True)
    for _ in range(args.warmup):
        extract(ctx, desc, mesh, flags=single)
    single_elapsed = timed_steps(ctx, desc, mesh, args.steps, flags=single)
    for _ in range(args.warmup):
        extract(ctx, desc, mesh)
    three_elapsed = timed_steps(ctx, desc, mesh, args.steps)
    # Clear the mesh so that the checks below read what each path wrote.
    zeros = np.zeros(max(3 * mesh.nv, mesh.ni), np.uint32)
    for flags, graph in ((single, False), (0, True)):
        ctx.lib.iso_memcpy_to_device(ctx.handle, mesh.v, zeros.ctypes.data, 12 * mesh.nv)
        ctx.lib.iso_memcpy_to_device(ctx.handle, mesh.i, zeros.ctypes.data, 4 * mesh.ni)
        o = extract(ctx, desc, mesh, graph=graph, flags=flags)
        if parity is not None:
            dv, di = digest(ctx, mesh, o)
            parity = parity and dv == golden[key]["vertices_sha256"] and \
                di == golden[key]["indices_sha256"]
    log(f"single-pass and graph replay parity vs oracle digests: {parity}")
    use_single = single_elapsed < three_elapsed
    elapsed = single_elapsed if use_single else three_elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    value = cells * args.steps / elapsed
    log(f"grid path: {1e3 * three_elapsed / args.steps:.4f} ms/step three-pass, "
        f"{1e3 * single_elapsed / args.steps:.4f} single pass, "
        f"{1e3 * graph_elapsed / args.steps:.4f} three-pass graph replay; "
        f"headline {value / 1e9:.2f} G cells/s")

    three_kernels = kernel_pass(ctx, desc, mesh, args.steps)
    single_kernels = kernel_pass(ctx, desc, mesh, args.steps, flags=single)
    # The classify + interpolate + emit stage is the unit of the roofline:
    # algorithmic bytes 4 per sample read + 12 per vertex + 12 per triangle
    # written, over the headline path's kernel time (the single kernel's mean
    # duration, or the three-pass path's three kernels summed).
    stage_bytes = 4 * n + 12 * nv + 4 * ni
    three_stage = ["mc_signs_grid", "mc_count", "mc_emit"]
    three_us = sum(three_kernels[k]["mean_us"] for k in three_stage)
    single_us = single_kernels["mc_fused"]["mean_us"]
    stage_us = single_us if use_single else three_us
    achieved = stage_bytes / (stage_us * 1e-6) / 1e9
    segs = (desc.nx - 1 + 63) // 64 * (desc.ny - 1) * (desc.nz - 1)
    kernel_bytes = {
        "mc_fused": stage_bytes,                   # every sample once, mesh out
        "mc_signs_grid": 4 * n + n // 8,           # every sample once, one bit out
        "mc_count": n // 8 + 8 * segs,              # sign bits in, first ids out
        "mc_emit": 8 * nv + 12 * nv + 4 * ni,       # endpoint samples in, mesh out
    }
    all_kernels = {**three_kernels, **single_kernels}
    per_kernel = {
        k: {"mean_us": round(all_kernels[k]["mean_us"], 3), "algorithmic_bytes": b,
            "achieved_GBps": b / (all_kernels[k]["mean_us"] * 1e-6) / 1e9,
            "frac": b / (all_kernels[k]["mean_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS}
        for k, b in kernel_bytes.items()}
    traffic = measured_traffic()
    stage_traffic = None
    if traffic:
        if use_single:
            stage_traffic = traffic["kernels"].get("mc_fused")
        elif all(k in traffic["kernels"] for k in three_stage):
            stage_traffic = sum(traffic["kernels"][k] for k in three_stage)
    log(f"kernels: three-pass {json.dumps(three_kernels)} single pass "
        f"{json.dumps(single_kernels)}")

    copy = copy_peak(ctx)
    # Program path: the CSG sampled by the extraction itself (no grid).
    pdesc = ctx.reference_desc(size, A.ISO_SOURCE_PROGRAM)
    extract(ctx, pdesc, mesh)
    p_elapsed = timed_steps(ctx, pdesc, mesh, args.program_steps)
    p_kernels = kernel_pass(ctx, pdesc, mesh, max(1, args.program_steps // 2))
    program_path = {
        "value": cells * args.program_steps / p_elapsed,
        "unit": "cells/s",
        "ms_per_step": 1e3 * p_elapsed / args.program_steps,
        "kernels_us": {k: round(v["mean_us"], 2) for k, v in p_kernels.items()},
        "note": "64-sphere-minus-torus SDF evaluated on the GPU (VALU-bound, not an HBM stream)",
    }
    log(f"program path: {program_path['ms_per_step']:.3f} ms/step")
    # The PointCloud and normal-sink rows on the same lattice and mesh.
    out = extract(ctx, desc, mesh)
    widened = points_and_normals(ctx, desc, mesh, nv, max(1, args.program_steps))

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "path": "single pass (one launch)" if use_single else "three-pass (classify, count, emit)",
        "three_pass_ms_per_step": 1e3 * three_elapsed / args.steps,
        "single_pass_ms_per_step": 1e3 * single_elapsed / args.steps,
        "three_pass_graph_replay_ms_per_step": 1e3 * graph_elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"MarchingCubes::new({size}) over a resident FP32 "
                         f"{desc.nx}x{desc.ny}x{desc.nz} grid of the union of 64 seeded spheres "
                         "minus a torus (R=0.3, r=0.05); IndexedVertices output"),
            "cells": cells,
            "samples": n,
            "vertices": nv,
            "triangles": ni // 3,
            "parity_vs_oracle": parity,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": ("k_mc_fused (classify, count, look-back scan, interpolate and emit: "
                       "one extraction, one launch)") if use_single else
                      ("k_signs_grid + k_count + k_emit (classify, count and scan, "
                       "interpolate and emit: one extraction)"),
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": stage_traffic,
            "algorithmic_bytes": stage_bytes,
            "mean_us": stage_us,
            "three_pass_mean_us": three_us,
            "single_pass_mean_us": single_us,
            "measured_copy_GBps": copy["GBps"],
            "frac_of_measured_copy": achieved / copy["GBps"],
        },
        "device_copy": copy,
        "kernels": per_kernel,
        "traffic_per_kernel": traffic.get("kernels") if traffic else None,
        "program_path": program_path,
        **widened,
    }

    if args.large_size > 0:
        result["large_grid"] = large_grid(ctx, args.large_size, max(5, args.steps // 5),
                                          copy["GBps"])
        ctx.upload_program(src.program())

    if not args.no_side:
        if args.band_size > 0:
            result["learned_band"] = learned_band(ctx, args.band_size, 2, args.band_steps)
        result["batched_chunks"] = batched_chunks(ctx, max(2, args.program_steps))
        result.update(side_workloads(ctx, max(1, args.program_steps), not args.no_cpu_baseline))

    # VALU-bound kernels against the VALU issue peak: counted instructions
    # per launch (committed SQ counter pass) over this run's durations.
    durations = {"mc_signs_program": program_path["kernels_us"].get("mc_signs_program")}
    if "dual_contouring" in result:
        for k in ("dc_normals", "dc_feature"):
            durations[k] = result["dual_contouring"]["kernels_us"].get(k)
    if "learned_band" in result:
        durations["band_exact"] = result["learned_band"]["certificate_without_network"][
            "kernels_us"].get("band_exact")
    result["valu_roofline"] = valu_roofline(durations)

    if not args.no_cpu_baseline:
        host_grid = np.empty((desc.nz, desc.ny, desc.nx), np.float32)
        ctx.lib.iso_memcpy_to_host(ctx.handle, host_grid.ctypes.data, dgrid, host_grid.nbytes)
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline(host_grid, size, args.cpu_budget)
    else:
        result["cpu_baseline"] = None

    mesh.free()
    ctx.lib.iso_device_free(ctx.handle, dgrid)
    ctx.close()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
