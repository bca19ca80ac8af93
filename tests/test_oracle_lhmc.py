"""Pins the oracle's linear-hashed octree marching cubes.

Morton arithmetic is checked against closed forms on decoded integer
coordinates; the whole extraction is checked against a second, independent
restatement below that works on (level, x, y, z) coordinates instead of
dilated integers (pure Python, small depths only).  The level() quirk
(morton.rs:35, 54-59: floor(ln k * 0.48089834696), one level low on exact
powers of 8) is pinned by its integer closed form, valid below 8^11."""
import math

import numpy as np
import pytest

from isosurface_amd import workloads as W
from isosurface_amd.sources import (Cylinder, DemoSource, RectangularPrism, Sphere, Torus,
                                    Translate, Union)
from oracle import oracle as O

f32 = np.float32
MAX_LEVEL = 21
REMAP_CUBE = [2, 3, 1, 0, 6, 7, 5, 4]
CORNER = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
EDGE = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6),
        (3, 7)]


def encode(level, x, y, z):
    k = 1
    for b in range(level - 1, -1, -1):
        k = (k << 3) | ((z >> b) & 1) << 2 | ((y >> b) & 1) << 1 | ((x >> b) & 1)
    return k


def decode(k):
    level = (k.bit_length() - 1) // 3
    x = y = z = 0
    for b in range(level):
        x |= ((k >> (3 * b)) & 1) << b
        y |= ((k >> (3 * b + 1)) & 1) << b
        z |= ((k >> (3 * b + 2)) & 1) << b
    return level, x, y, z


def quirk_level(k):
    """level() in closed form below 8^11: the true depth, one less on 8^L, L >= 1."""
    if k == 0:
        return 0
    L = (k.bit_length() - 1) // 3
    return L - 1 if (L >= 1 and k == 1 << (3 * L)) else L


def test_level_quirk_closed_form():
    lib = O.morton()
    keys = [0, 1]
    for L in range(1, 11):
        p = 8 ** L
        keys += [p - 1, p, p + 1, p + 7, 2 * p - 1, 2 * p, 7 * p + 3, 8 * p - 1]
    rng = np.random.default_rng(0)
    keys += [int(v) for v in rng.integers(1, 8 ** 11, 2000, dtype=np.int64)]
    for k in keys:
        assert lib.or_morton_level(k) == quirk_level(k), k
    # the quirk itself: every exact power of 8 above the root is one level low
    assert [lib.or_morton_level(8 ** L) for L in range(1, 8)] == list(range(0, 7))


def test_center_and_size_brute_force():
    lib = O.morton()
    for L in range(0, 5):
        for x in range(2 ** L):
            for y in range(2 ** L):
                for z in range(2 ** L):
                    k = encode(L, x, y, z)
                    ql = quirk_level(k)
                    s = f32(1.0) / f32(2 << ql)
                    assert lib.or_morton_size(k) == s
                    # coordinates masked to the (quirky) level's bit count
                    m = (1 << ql) - 1
                    want = np.array([f32(c & m) * (s * f32(2)) + s for c in (x, y, z)], f32)
                    assert np.array_equal(O.morton_center(k), want), (L, x, y, z)
    # the origin-corner node of depth L is centred at 1/2^L, not 1/2^(L+1)
    for L in range(1, 6):
        assert np.all(O.morton_center(8 ** L) == f32(1.0) / f32(2 ** L))


def test_dilated_add_sub_match_coordinate_arithmetic():
    lib = O.morton()
    rng = np.random.default_rng(1)
    for _ in range(3000):
        L = int(rng.integers(1, 9))
        x, y, z = (int(v) for v in rng.integers(0, 2 ** L, 3))
        wx, wy, wz = (int(v) for v in rng.integers(0, 2, 3))
        k = encode(L, x, y, z)
        w = wx | wy << 1 | wz << 2
        # add: the sentinel sits in the x lane, so x carries run into it
        s = lib.or_morton_add(k, w)
        X = x + (1 << L) + wx
        want = 0
        for b in range(L + 2):
            want |= ((X >> b) & 1) << (3 * b) | (((y + wy) >> b) & 1) << (3 * b + 1) | \
                (((z + wz) >> b) & 1) << (3 * b + 2)
        assert s == want
        if x >= wx and y >= wy and z >= wz:
            assert lib.or_morton_sub(k, w) == encode(L, x - wx, y - wy, z - wz)


def test_primal_and_dual_vertex_closed_form():
    lib = O.morton()
    for L in range(1, 5):
        for x in range(2 ** L):
            for y in range(2 ** L):
                for z in range(2 ** L):
                    k = encode(L, x, y, z)
                    lv = quirk_level(k)
                    for w in range(8):
                        got = lib.or_morton_primal_vertex(k, lv, w)
                        if lv != L:  # origin-corner nodes never yield a vertex
                            assert got == 0
                            continue
                        c = (x + (w & 1), y + (w >> 1 & 1), z + (w >> 2 & 1))
                        if min(c) == 0 or max(c) >= 2 ** L:
                            assert got == 0
                        else:
                            sh = MAX_LEVEL - L
                            assert got == encode(MAX_LEVEL, *(v << sh for v in c))
                            # dual_vertex undoes it at the same level
                            assert lib.or_morton_dual_vertex(got, L, w) == k


# ---------------------------------------------------------------------------
# Independent restatement on coordinates
# ---------------------------------------------------------------------------
def py_center(k):
    L, x, y, z = decode(k)
    ql = quirk_level(k)
    s = f32(1.0) / f32(2 << ql)
    m = (1 << ql) - 1
    return tuple(f32(c & m) * (s * f32(2)) + s for c in (x, y, z))


def py_lhmc(program, max_depth):
    sample = lambda p: f32(O.sample(program, float(p[0]), float(p[1]), float(p[2])))
    # octree, depth by depth (the BFS visits depths in order)
    nodes, leaves = {}, []
    frontier = [1]
    while frontier:
        nxt = []
        for k in frontier:
            v = sample(py_center(k))
            ql = quirk_level(k)
            s = f32(1.0) / f32(2 << ql)
            if ql < 2 or (ql < max_depth and abs(v) < s * f32(1.73205080757)):
                nxt += [(k << 3) | i for i in range(8)]
            else:
                leaves.append(k)
            nodes[k] = v
        frontier = nxt
    # primal vertices as level-21 coordinates with their deepest leaf level
    primal = {}
    for k in leaves:
        L, x, y, z = decode(k)
        if quirk_level(k) != L:
            continue
        for w in range(8):
            c = (x + (w & 1), y + (w >> 1 & 1), z + (w >> 2 & 1))
            if min(c) == 0 or max(c) >= 2 ** L:
                continue
            key = tuple(v << (MAX_LEVEL - L) for v in c)
            primal[key] = max(primal.get(key, -1), L)
    order = sorted(primal, key=lambda c: encode(MAX_LEVEL, *c))
    keys = [1] * 8
    corners = [(f32(0), f32(0), f32(0))] * 8
    values = [f32(0)] * 8
    cache, verts, faces = {}, [], []
    tri = O.tri_table()
    for c in order:
        lv = primal[c]
        for i in range(8):
            w = REMAP_CUBE[i]
            d = (c[0] >> (MAX_LEVEL - lv)) - (w & 1), (c[1] >> (MAX_LEVEL - lv)) - (w >> 1 & 1), \
                (c[2] >> (MAX_LEVEL - lv)) - (w >> 2 & 1)
            L = lv
            while True:
                m = encode(L, *d)
                if m <= 1:
                    break
                if m in nodes:
                    keys[i], corners[i], values[i] = m, py_center(m), nodes[m]
                    break
                L -= 1
                d = (d[0] >> 1, d[1] >> 1, d[2] >> 1)
        cube = sum(1 << i for i in range(8) if not values[i] > 0)
        for t in range(5):
            te = tri[cube][3 * t:3 * t + 3]
            if te[0] < 0:
                break
            ids = []
            for e in te:
                u, v = EDGE[e]
                ek = (min(keys[u], keys[v]), max(keys[u], keys[v]))
                if ek not in cache:
                    a, b = values[u], values[v]
                    dl = b - a
                    tt = f32(0.5) if dl == 0 else -a / dl
                    ss = f32(1) - tt
                    cache[ek] = len(verts)
                    verts.append([corners[u][j] * ss + corners[v][j] * tt for j in range(3)])
                ids.append(cache[ek])
            faces += ids
    return np.array(verts, f32).reshape(-1, 3), np.array(faces, np.uint32)


SMALL = {
    "sphere": lambda: W.sphere_source(),
    "box": lambda: DemoSource(RectangularPrism((0.21, 0.17, 0.29))),
    "torus": lambda: DemoSource(Torus(0.28, 0.09)),
    "cyl_union": lambda: DemoSource(Union(Cylinder(0.2, 0.3),
                                          Translate((0.2, 0.1, -0.1), Sphere(0.15)))),
    "boxes32": lambda: W.boxes_union(32, seed=4),
    "corner_sphere": lambda: Sphere(0.3),  # surface through the origin-corner nodes
}


@pytest.mark.parametrize("name", sorted(SMALL))
@pytest.mark.parametrize("depth", [2, 3, 4])
def test_lhmc_matches_coordinate_restatement(name, depth):
    prog = SMALL[name]().program()
    v, i = py_lhmc(prog, depth)
    m = O.extract_lhmc(prog, depth)
    assert np.array_equal(m.indices, i)
    assert m.vertices.tobytes() == v.tobytes()


def test_octree_node_set():
    prog = W.boxes_union(32, seed=4).program()
    keys, vals, leaf = O.lhmc_octree(prog, 5)
    assert np.all(np.diff(keys.astype(np.float64)) > 0)  # BFS order is ascending key order
    internal = set(int(k) for k, l in zip(keys, leaf) if not l)
    allk = set(int(x) for x in keys)
    for k in keys[1:]:
        assert int(k) >> 3 in internal  # every non-root node's parent was refined
    for k in internal:
        assert all((k << 3 | c) in allk for c in range(8))  # refined nodes have 8 children
    assert all(quirk_level(int(k)) <= 5 for k in keys)


def canonical(vertices, indices, keys):
    """Vertex order by MortonKey, index triples renumbered, triangles sorted."""
    order = np.lexsort((keys[:, 1], keys[:, 0]))
    rank = np.empty(len(order), np.int64)
    rank[order] = np.arange(len(order))
    tris = rank[indices.astype(np.int64)].reshape(-1, 3)
    tris = tris[np.lexsort(tris.T[::-1])]
    return vertices[order], keys[order], tris


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_lhmc_order_independence(seed):
    prog = W.boxes_union(32, seed=4).program()
    a = O.extract_lhmc(prog, 5, 0)
    b = O.extract_lhmc(prog, 5, seed)
    va, ka, ta = canonical(a.vertices, a.indices, a.keys)
    vb, kb, tb = canonical(b.vertices, b.indices, b.keys)
    assert np.array_equal(ka, kb) and np.array_equal(ta, tb)
    # positions depend on which dual cube met the edge first: a few ulps apart
    assert np.allclose(va, vb, rtol=1e-6, atol=1e-7)


def test_lhmc_sphere_is_closed():
    m = O.extract_lhmc(W.sphere_source().program(), 6)
    t = m.indices.reshape(-1, 3).astype(np.int64)
    e = np.sort(np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]), axis=1)
    _, counts = np.unique(e, axis=0, return_counts=True)
    assert np.all(counts == 2)
    assert len(m.vertices) - len(counts) + len(t) == 2
    r = np.linalg.norm(m.vertices.astype(np.float64) - 0.5, axis=1)
    assert np.all(np.abs(r - 0.35) < 1.0 / 64)


def oracle_split(lib, d):
    """The first offset k of depth d's keys [8^d, 8^(d+1)) at which the
    oracle's level() reaches d (bisection; level() is monotone)."""
    base = 8 ** d
    lo, hi = 0, (2 ** 64 - 1 - base) if d == 21 else 7 * base - 1
    while lo < hi:
        mid = (lo + hi) // 2
        if lib.or_morton_level(base + mid) >= d:
            hi = mid
        else:
            lo = mid + 1
    return lo


def test_level_split_points_follow_the_truncated_constant():
    # level(a) = floor(ln(a) c) with c = 0.48089834696 < 1/(3 ln 2): the keys of
    # depth d below exp(d / c) = 8^d exp(d (1/c - 3 ln 2)) sit one level low.
    # The oracle's split points (with libm's f64 log) must sit at that bound:
    # exactly up to depth 16, within the f64 rounding of ln(a) c beyond.
    from decimal import ROUND_CEILING, Decimal, getcontext
    getcontext().prec = 60
    c = Decimal("0.48089834696")
    lib = O.morton()
    splits = []
    for d in range(1, 22):
        base = 8 ** d
        bound = int(((Decimal(d) / c).exp() - base).to_integral_value(rounding=ROUND_CEILING))
        got = oracle_split(lib, d)
        splits.append(got)
        assert lib.or_morton_level(base + got) == d
        assert lib.or_morton_level(base + got - 1) == d - 1
        assert lib.or_morton_level(8 * base - 1 if d < 21 else 2 ** 64 - 1) == d
        if d <= 16:
            assert got == bound, d
        else:
            assert abs(got - bound) <= base >> 40, d
    # one key (the power of 8) per depth below 8^11, then growing runs
    assert splits[:10] == [1] * 10 and splits[10:13] == [2, 11, 93]
