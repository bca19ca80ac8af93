"""The learned narrow band (no reference counterpart; see include/isosurface.h).

    band = NarrowBand(ctx)
    band.init(seed)
    pts, sdf = training_samples(ctx, source, ...)       # made by the plain path
    band.train(pts, sdf, steps=..., batch=16384)
    margin = band.calibrate(pts2, sdf2, half_diag)     # held-out cull condition
    mesh, stats = band.extract_mesh(desc, margin)        # exact SDF only in the band
    band.missed_cells(desc)                              # checked against the dense path
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Tuple

import numpy as np

from . import _abi as A
from .engine import Context, MarchingCubes, Mesh, default_context
from .sources import Source

N_PARAMS = A.ISO_BAND_PARAMS


class NarrowBand:
    def __init__(self, context: Optional[Context] = None):
        self.ctx = context or default_context()
        self.clamp = 0.05

    @property
    def lib(self):
        return self.ctx.lib

    def init(self, seed: int = 0) -> None:
        self.ctx.check(self.lib.iso_band_init(self.ctx.handle, seed), "iso_band_init")

    def set_params(self, theta: np.ndarray) -> None:
        theta = np.ascontiguousarray(theta, np.float32)
        self.ctx.check(self.lib.iso_band_set_params(self.ctx.handle, theta.ctypes.data,
                                                    len(theta)), "iso_band_set_params")

    def get_params(self) -> np.ndarray:
        out = np.empty(N_PARAMS, np.float32)
        self.ctx.check(self.lib.iso_band_get_params(self.ctx.handle, out.ctypes.data, N_PARAMS),
                       "iso_band_get_params")
        return out

    def train(self, points: np.ndarray, sdf: np.ndarray, weight: Optional[np.ndarray] = None,
              steps: int = 2000, batch: int = 16384, lr: float = 3e-3, lr_min: float = 1e-4,
              clamp: float = 0.05, seed: int = 1, overshoot: float = 1.0) -> float:
        """Adam on (points, sdf); `overshoot` > 1 weighs the part of each error
        that lies beyond the interval between 0 and the target more, so the
        network errs toward the band (see iso_band_train)."""
        points = np.ascontiguousarray(points, np.float32)
        sdf = np.ascontiguousarray(sdf, np.float32)
        wp = None
        if weight is not None:
            weight = np.ascontiguousarray(weight, np.float32)
            wp = weight.ctypes.data
        self.clamp = clamp
        prm = A.IsoBandTrainParams(steps, batch, lr, lr_min, clamp, seed, overshoot)
        loss = C.c_float(0.0)
        self.ctx.check(self.lib.iso_band_train(self.ctx.handle, points.ctypes.data, sdf.ctypes.data,
                                               wp, len(sdf), A.ISO_MEMORY_HOST, C.byref(prm),
                                               C.byref(loss)), "iso_band_train")
        return float(loss.value)

    def gradient(self, points, sdf, weight=None, clamp: float = 0.05,
                 overshoot: float = 1.0) -> Tuple[float, np.ndarray]:
        points = np.ascontiguousarray(points, np.float32)
        sdf = np.ascontiguousarray(sdf, np.float32)
        wp = None
        if weight is not None:
            weight = np.ascontiguousarray(weight, np.float32)
            wp = weight.ctypes.data
        grad = np.empty(N_PARAMS, np.float32)
        loss = C.c_float(0.0)
        self.ctx.check(self.lib.iso_band_gradient(self.ctx.handle, points.ctypes.data,
                                                  sdf.ctypes.data, wp, len(sdf), clamp,
                                                  overshoot, grad.ctypes.data, C.byref(loss)),
                       "iso_band_gradient")
        return float(loss.value), grad

    def predict(self, points: np.ndarray) -> np.ndarray:
        points = np.ascontiguousarray(points, np.float32)
        out = np.empty(len(points), np.float32)
        self.ctx.check(self.lib.iso_band_predict(self.ctx.handle, points.ctypes.data, len(points),
                                                 out.ctypes.data, A.ISO_MEMORY_HOST),
                       "iso_band_predict")
        return out

    def calibrate(self, points: np.ndarray, sdf: np.ndarray, half_diag: float,
                  safety: float = 1.1, slack: float = 2e-4) -> float:
        """The smallest margin m under which, on held-out samples, the cull
        test |f_hat| > half_diag + m keeps every sample that may be surface
        (|f| <= half_diag) and never culls a sample whose predicted sign is
        wrong; times `safety`, plus `slack`.  (A brick is culled from its
        centre's prediction, so the samples stand in for brick centres.)"""
        f_hat = self.predict(points).astype(np.float64)
        f = np.asarray(sdf, np.float64)
        must_keep = (np.abs(f) <= half_diag) | ((f_hat > 0) != (f > 0))
        if not np.any(must_keep):
            return slack
        m = max(0.0, float(np.max(np.abs(f_hat[must_keep]) - half_diag)))
        return m * safety + slack

    def extract_mesh(self, desc: A.IsoGridDesc, margin: float, flags: int = 0,
                     vertex_capacity: int = 0, index_capacity: int = 0):
        """Band extraction; flags: A.ISO_BAND_CERTIFY (a brick leaves the
        band only when an exact super-brick or brick centre sample proves
        it, whatever the network predicted; stats.overturned_culls counts the
        network culls no proof allowed) or A.ISO_BAND_NO_NETWORK (the same
        proofs, no network), with A.ISO_BAND_BRICK8 for 8x8x8-cell bricks."""
        vcap = max(int(vertex_capacity), 1)
        icap = max(int(index_capacity), 3)
        verts = np.empty((vcap, 3), np.float32)
        idx = np.empty(icap, np.uint32)
        out = A.IsoMeshOut(verts.ctypes.data, idx.ctypes.data, vcap, icap, 0, 0,
                           A.ISO_MEMORY_HOST, 0)
        stats = A.IsoBandStats()
        st = self.lib.iso_extract_mc_band(self.ctx.handle, C.byref(desc), margin, flags,
                                          C.byref(out), C.byref(stats))
        if st == A.ISO_ERR_CAPACITY:
            nv, ni = int(out.vertex_count), int(out.index_count)
            verts = np.empty((max(nv, 1), 3), np.float32)
            idx = np.empty(max(ni, 3), np.uint32)
            mo = A.IsoMeshOut(verts.ctypes.data, idx.ctypes.data, max(nv, 1), max(ni, 3), 0, 0,
                              A.ISO_MEMORY_HOST, 0)
            st = self.lib.iso_mesh_fetch(self.ctx.handle, C.byref(mo))
        self.ctx.check(st, "iso_extract_mc_band")
        nv, ni = int(out.vertex_count), int(out.index_count)
        return Mesh(verts[:nv].copy(), idx[:ni].copy()), stats

    def missed_cells(self, desc: A.IsoGridDesc) -> int:
        n = C.c_uint64(0)
        self.ctx.check(self.lib.iso_band_missed_cells(self.ctx.handle, C.byref(desc), C.byref(n)),
                       "iso_band_missed_cells")
        return int(n.value)


def brick_size(flags: int = 0) -> int:
    """Cells per edge of a narrow-band brick under iso_extract_mc_band's
    flags (iso_band_brick_size): 4, or 8 with A.ISO_BAND_BRICK8."""
    return int(A.load().iso_band_brick_size(flags))


def brick_half_diagonal(size: int, flags: int = 0) -> float:
    """Half-diagonal of a band brick of MarchingCubes::new(size)."""
    return (0.5 * brick_size(flags) * float(np.sqrt(3.0)) *
            float(np.float32(1.0) / np.float32(size - 1)))


def training_samples(ctx: Context, source: Source, n_uniform: int, n_surface: int,
                     seed: int = 0, jitter: float = 0.01, coarse: int = 129,
                     upper=(1.0, 1.0, 1.0)):
    """(points, sdf) generated by the plain path: uniform points in
    [0, upper] plus vertices of a coarse plain GPU extraction, jittered by
    N(0, jitter); targets are the program's exact values (GPU)."""
    rng = np.random.default_rng(seed)
    pts = [(rng.random((n_uniform, 3), dtype=np.float32) *
            np.asarray(upper, np.float32)).astype(np.float32)]
    if n_surface:
        mesh = MarchingCubes(coarse, ctx).extract_mesh(source)
        pick = rng.integers(0, len(mesh.vertices), n_surface)
        near = mesh.vertices[pick] + rng.normal(0.0, jitter, (n_surface, 3)).astype(np.float32)
        pts.append(np.clip(near, 0.0, 1.0).astype(np.float32))
    points = np.ascontiguousarray(np.concatenate(pts), np.float32)
    sdf = ctx.sample(source, points)
    return points, sdf
