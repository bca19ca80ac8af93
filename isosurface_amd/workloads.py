"""Seeded synthetic sources of the benchmark workloads.

The reference has no random generator; these workloads are defined here,
with SplitMix64 (Steele, Lea & Flood 2014) as the generator and
u = (x >> 40) * 2^-24 as the uniform in [0, 1).  All shapes live in the
[-0.5, 0.5]^3 SDF domain and are sampled over [0, 1]^3 through DemoSource's
p - 0.5 shift (examples/common/sources.rs:38-43).
"""
from __future__ import annotations

from typing import List, Tuple

from .sources import (DemoSource, Difference, RectangularPrism, Source, Sphere, Torus, Translate,
                      Union, smooth_union_all, union_all)

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo: float, hi: float) -> float:
        u = (self.next_u64() >> 40) * (1.0 / (1 << 24))
        return lo + (hi - lo) * u


def sphere_source(radius: float = 0.35) -> Source:
    """64^3-cell CPU case: a sphere of radius 0.35 at the domain centre."""
    return DemoSource(Sphere(radius))


def spheres_minus_torus(n_spheres: int = 64, seed: int = 2) -> Tuple[Source, List[tuple]]:
    """Union of n seeded spheres (centres U[-0.4,0.4]^3, radii U[0.03,0.12])
    minus a torus (R=0.3, r=0.05): Difference::new(torus, union) = max(union, -torus)."""
    rng = SplitMix64(seed)
    spheres = []
    for _ in range(n_spheres):
        c = (rng.uniform(-0.4, 0.4), rng.uniform(-0.4, 0.4), rng.uniform(-0.4, 0.4))
        r = rng.uniform(0.03, 0.12)
        spheres.append((c, r))
    u = union_all(Translate(c, Sphere(r)) for c, r in spheres)
    return DemoSource(Difference(Torus(0.3, 0.05), u)), spheres


def boxes_union(n_boxes: int = 32, seed: int = 4) -> Source:
    """Union of n seeded axis-aligned boxes (centres U[-0.3,0.3]^3,
    half-extents U[0.05,0.2]^3): the sharp-feature dual-contouring source."""
    rng = SplitMix64(seed)
    boxes = []
    for _ in range(n_boxes):
        c = (rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3))
        h = (rng.uniform(0.05, 0.2), rng.uniform(0.05, 0.2), rng.uniform(0.05, 0.2))
        boxes.append(Translate(c, RectangularPrism(h)))
    return DemoSource(union_all(boxes))


def smooth_spheres(n_spheres: int = 1024, k: float = 0.02, seed: int = 5) -> Source:
    """Smooth-min (k) left fold of n seeded spheres, radii U[0.01,0.05],
    centres U[-0.45,0.45]^3."""
    rng = SplitMix64(seed)
    parts = []
    for _ in range(n_spheres):
        c = (rng.uniform(-0.45, 0.45), rng.uniform(-0.45, 0.45), rng.uniform(-0.45, 0.45))
        r = rng.uniform(0.01, 0.05)
        parts.append(Translate(c, Sphere(r)))
    return DemoSource(smooth_union_all(parts, k))


def random_csg(rng: SplitMix64, n_prims: int) -> Source:
    """A small random CSG tree of spheres, boxes and tori (batched chunks)."""
    prims = []
    for _ in range(n_prims):
        kind = rng.next_u64() % 3
        c = (rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3))
        if kind == 0:
            p: Source = Sphere(rng.uniform(0.05, 0.25))
        elif kind == 1:
            p = RectangularPrism((rng.uniform(0.05, 0.2), rng.uniform(0.05, 0.2),
                                  rng.uniform(0.05, 0.2)))
        else:
            p = Torus(rng.uniform(0.1, 0.25), rng.uniform(0.02, 0.06))
        prims.append(Translate(c, p))
    acc = prims[0]
    for p in prims[1:]:
        op = rng.next_u64() % 4
        acc = Difference(p, acc) if op == 3 else Union(acc, p)
    return DemoSource(acc)


CHUNK_SIZES = (16, 32, 48, 64)


def batch_chunks(n_chunks: int = 4096, seed: int = 6) -> List[Tuple[Source, int]]:
    """The batched-extraction workload: n chunks, each (source, S) with S
    drawn uniformly from {16, 32, 48, 64} and a seeded CSG of 1-8 spheres,
    boxes and tori.  Chunk i is S^3 cells: a lattice of (S+1)^3 samples with
    spacing 1/S over [0, 1]^3 (chunk_desc)."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(n_chunks):
        size = CHUNK_SIZES[rng.next_u64() % 4]
        n_prims = 1 + int(rng.next_u64() % 8)
        out.append((random_csg(rng, n_prims), size))
    return out
