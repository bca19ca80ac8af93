"""Why the band certificate bounds the rounding of positions and values.

A brick is proven to lie on one side of the surface when its centre value
is farther from zero than its half-diagonal plus the f32 error.  A fixed
allowance (1e-4 relative + 2e-5 absolute) is not enough once coordinates
are large: on a lattice at origin 4000.3 (unit in the last place 2.4e-4) a
box whose corners lie on lattice points gives bricks whose centre value
clears that allowance while a sample of the brick's closed box has the
other sign.  This CPU test restates the fixed-allowance test with the
oracle and finds those bricks, so the GPU test of the same lattice
(test_gpu_band.py::test_certificate_at_a_far_origin) exercises them."""
import numpy as np

from isosurface_amd.sources import RectangularPrism, Translate
from oracle import oracle as O

F = np.float32


def far_box(origin=4000.3, lo=44, hi=92, n=130):
    """A cube with faces on lattice points lo and hi of an n^3 lattice of
    spacing 1/(n-1) at `origin` (every axis); returns (program, desc)."""
    s = F(1.0) / F(n - 1)
    o = F(origin)
    a = F(o + F(lo) * s)
    b = F(o + F(hi) * s)
    c = float(F(F(0.5) * F(a + b)))
    h = float(F(F(0.5) * F(b - a)))
    prog = Translate((c, c, c), RectangularPrism((h, h, h))).program()
    return prog, O.lattice_desc(n, n, n, (float(o),) * 3, (float(s),) * 3, source=0)


def fixed_allowance_wrong_proofs(prog, d, brick=4) -> int:
    n = d.nx
    s = F(d.spacing[0])
    o = F(d.origin[0])
    inside = ~(O.sample_lattice(prog, d) > 0)
    nb = (n - 1 + brick - 1) // brick
    ext = F(brick) * s
    d2 = F(0)
    for _ in range(3):
        d2 = F(d2 + F(ext * ext))
    hd = F(F(0.5) * np.sqrt(d2))
    bad = 0
    for bz in range(nb):
        z0, z1 = brick * bz, min(brick * bz + brick, n - 1)
        for by in range(nb):
            y0, y1 = brick * by, min(brick * by + brick, n - 1)
            for bx in range(nb):
                x0, x1 = brick * bx, min(brick * bx + brick, n - 1)
                box = inside[z0:z1 + 1, y0:y1 + 1, x0:x1 + 1]
                if box.all() or not box.any():
                    continue  # one sign throughout: any proof is right
                p = [F(o + F(0.5 * (q0 + q1)) * s) for q0, q1 in ((x0, x1), (y0, y1), (z0, z1))]
                v = F(O.sample(prog, *map(float, p)))
                if abs(v) > hd * F(1 + 1e-4) + F(2e-5):
                    bad += 1
    return bad


def test_fixed_allowance_fails_at_a_far_origin():
    prog, d = far_box()
    assert fixed_allowance_wrong_proofs(prog, d) > 0


def test_fixed_allowance_holds_at_the_unit_lattice():
    prog, d = far_box(origin=0.0)
    assert fixed_allowance_wrong_proofs(prog, d) == 0
