"""The GPU's closed-form vertex numbering, checked against the oracle.

The emit kernel never hashes: it derives every vertex id from the lattice
position of its edge.  A lattice edge's vertex is owned by the lowest cell
(z-major) containing the edge; owners own e5, e6, e10 always and the
low-face edges only on the lattice boundary; a vertex's id is the owner's
running count plus the rank of the edge among the owner's owned edges in
first-appearance order of its triangle list.  This file restates that rule
in plain Python and checks it reproduces the oracle's first-encounter
numbering (mesh.rs:240-251) exactly, including lattice boundaries and zero
samples."""
import numpy as np
import pytest

from oracle import oracle as O

CORNER = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
EDGES = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6),
         (3, 7)]
TRI = O.tri_table()
MASK = [int(v) for v in O.edge_mask_table()]


def _before():
    out = [[0] * 12 for _ in range(256)]
    for c in range(256):
        seen = []
        for e in map(int, TRI[c]):
            if e < 0:
                break
            if e not in seen:
                seen.append(e)
        acc = 0
        for e in seen:
            out[c][e] = acc
            acc |= 1 << e
    return out


BEFORE = _before()
GEOM = []
OF_AXIS_MIN = {}
for _e, (_u, _v) in enumerate(EDGES):
    _a, _b = CORNER[_u], CORNER[_v]
    _axis = [d for d in range(3) if _a[d] != _b[d]][0]
    _mn = tuple(min(_a[d], _b[d]) for d in range(3))
    GEOM.append((_axis, _mn))
    OF_AXIS_MIN[(_axis, _mn)] = _e


def owned(x, y, z):
    m = (1 << 5) | (1 << 6) | (1 << 10)
    if y == 0:
        m |= (1 << 4) | (1 << 9)
    if z == 0:
        m |= (1 << 2) | (1 << 1)
    if x == 0:
        m |= (1 << 7) | (1 << 11)
    if y == 0 and z == 0:
        m |= 1
    if x == 0 and z == 0:
        m |= 1 << 3
    if x == 0 and y == 0:
        m |= 1 << 8
    return m


def closed_form_indices(g):
    nz, ny, nx = g.shape
    s = ~(g > 0)

    def case(x, y, z):
        return sum(1 << i for i, (dx, dy, dz) in enumerate(CORNER) if s[z + dz, y + dy, x + dx])

    cells = [(x, y, z) for z in range(nz - 1) for y in range(ny - 1) for x in range(nx - 1)]
    base, acc = {}, 0
    for c in cells:
        base[c] = acc
        acc += bin(owned(*c) & MASK[case(*c)]).count("1")
    idx = []
    for (x, y, z) in cells:
        c = case(x, y, z)
        for e in map(int, TRI[c]):
            if e < 0:
                break
            axis, mn = GEOM[e]
            cc = (x, y, z)
            r, rel = [0, 0, 0], [0, 0, 0]
            for d in range(3):
                if d != axis:
                    if mn[d] == 0 and cc[d] >= 1:
                        r[d], rel[d] = -1, 1
                    else:
                        rel[d] = mn[d]
            o = (x + r[0], y + r[1], z + r[2])
            eo = OF_AXIS_MIN[(axis, tuple(rel))]
            oc = case(*o)
            oown = owned(*o) & MASK[oc]
            idx.append(base[o] + bin(oown & BEFORE[oc][eo]).count("1"))
    return acc, np.array(idx, np.uint32)


@pytest.mark.parametrize("seed", range(120))
def test_closed_form_numbering_matches_first_encounter(seed):
    rng = np.random.RandomState(1000 + seed)
    nx, ny, nz = rng.randint(2, 8, size=3)
    g = rng.randn(nz, ny, nx).astype(np.float32)
    if seed % 3 == 0:
        g[rng.rand(nz, ny, nx) < 0.3] = 0.0
    m = O.extract_mc(grid=g, desc=O.lattice_desc(nx, ny, nz))
    nv, idx = closed_form_indices(g)
    assert nv == len(m.vertices)
    assert np.array_equal(idx, m.indices)


def test_owned_edges_partition_lattice_edges():
    # Every crossed lattice edge is owned by exactly one cell.
    rng = np.random.RandomState(7)
    g = rng.randn(5, 6, 7).astype(np.float32)
    nz, ny, nx = g.shape
    seen = {}
    for z in range(nz - 1):
        for y in range(ny - 1):
            for x in range(nx - 1):
                for e in range(12):
                    if owned(x, y, z) >> e & 1:
                        axis, mn = GEOM[e]
                        key = (x + mn[0], y + mn[1], z + mn[2], axis)
                        assert key not in seen
                        seen[key] = (x, y, z)
    n_edges = (nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1)
    assert len(seen) == n_edges


M64 = (1 << 64) - 1


def _popc(v):
    return bin(v).count("1")


def owned_count_planes(c, m, y0, z0, first):
    """The kernels' bit-sliced owned-vertex count over the cells in mask m of
    one 64-cell segment (c = the eight corner bit-planes)."""
    n = _popc(m & (c[6] ^ c[7])) + _popc(m & (c[5] ^ c[6])) + _popc(m & (c[2] ^ c[6]))
    if y0:
        n += _popc(m & (c[4] ^ c[5])) + _popc(m & (c[1] ^ c[5]))
    if z0:
        n += _popc(m & (c[2] ^ c[3])) + _popc(m & (c[1] ^ c[2]))
    if y0 and z0:
        n += _popc(m & (c[0] ^ c[1]))
    if first and (m & 1):
        n += ((c[4] ^ c[7]) & 1) + ((c[3] ^ c[7]) & 1)
        if z0:
            n += (c[0] ^ c[3]) & 1
        if y0:
            n += (c[0] ^ c[4]) & 1
    return n


def owned_planes(c, valid, y0, z0, first):
    """The count kernel's per-cell owned-vertex counts, bit-sliced in four
    planes by a ripple-carry adder over the owned-edge masks."""
    # e6, e5, e10 of every cell: a full adder of the three masks
    a, b, d = valid & (c[6] ^ c[7]), valid & (c[5] ^ c[6]), valid & (c[2] ^ c[6])
    p = [a ^ b ^ d, (a & b) | (d & (a ^ b)), 0, 0]

    def add(m):
        for k in range(4):
            carry = p[k] & m
            p[k] ^= m
            m = carry

    if y0:
        add(valid & (c[4] ^ c[5]))
        add(valid & (c[1] ^ c[5]))
    if z0:
        add(valid & (c[2] ^ c[3]))
        add(valid & (c[1] ^ c[2]))
    if y0 and z0:
        add(valid & (c[0] ^ c[1]))
    if first:
        v0 = valid & 1
        add(v0 & (c[4] ^ c[7]))
        add(v0 & (c[3] ^ c[7]))
        if z0:
            add(v0 & (c[0] ^ c[3]))
        if y0:
            add(v0 & (c[0] ^ c[4]))
    return p


def planes_count(p, m):
    return sum((1 << k) * _popc(p[k] & m) for k in range(4))


@pytest.mark.parametrize("seed", range(31))
def test_bit_sliced_owned_counts_match_per_cell_rule(seed):
    rng = np.random.RandomState(500 + seed)
    nx, ny, nz = int(rng.randint(2, 200)), int(rng.randint(2, 4)), int(rng.randint(2, 4))
    s = ~(rng.randn(nz, ny, nx).astype(np.float32) > 0)
    if seed == 30:
        # a 3-D checkerboard crosses every edge: the corner cell owns all 12
        z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        s = ((x + y + z) % 2) == 0

    def rowbits(y, z, x0):
        v = 0
        for b in range(65):
            if x0 + b < nx and s[z, y, x0 + b]:
                v |= 1 << b
        return v & M64, (v >> 64) & 1

    def hi(a, n):
        return ((a >> 1) | (n << 63)) & M64

    for cz in range(nz - 1):
        for cy in range(ny - 1):
            for w in range((nx - 1 + 63) // 64):
                x0 = 64 * w
                w00, n00 = rowbits(cy, cz, x0)
                w10, n10 = rowbits(cy + 1, cz, x0)
                w01, n01 = rowbits(cy, cz + 1, x0)
                w11, n11 = rowbits(cy + 1, cz + 1, x0)
                c = [w00, hi(w00, n00), hi(w10, n10), w10, w01, hi(w01, n01), hi(w11, n11), w11]
                left = nx - 1 - x0
                valid = M64 if left >= 64 else (1 << left) - 1
                for b in (0, 1, 13, 63, 64):
                    lm = M64 if b == 64 else (1 << b) - 1
                    want = 0
                    for x in range(x0, min(x0 + b, nx - 1)):
                        case = sum(1 << i for i, (dx, dy, dz) in enumerate(CORNER)
                                   if s[cz + dz, cy + dy, x + dx])
                        want += _popc(owned(x, cy, cz) & MASK[case])
                    assert owned_count_planes(c, valid & lm, cy == 0, cz == 0, w == 0) == want
                    planes = owned_planes(c, valid, cy == 0, cz == 0, w == 0)
                    assert planes_count(planes, lm) == want
