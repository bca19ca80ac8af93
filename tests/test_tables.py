"""Marching-cubes tables: pinned to the reference's table text, plus the
invariants that make them a valid triangulation."""
import hashlib
import struct

import numpy as np

from oracle import oracle as O

CORNER = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
EDGES = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6),
         (3, 7)]


def test_triangle_table_sha256():
    # SHA-256 of TRIANGLE_CONNECTION (marching_cubes_tables.rs:74-331) as
    # int8[256][16] row-major, computed from the reference's source text.
    t = O.tri_table()
    assert hashlib.sha256(t.astype(np.int8).tobytes()).hexdigest() == (
        "19bf7699e214903d72c94c296546f2e31337d637a1e4b118c3108a0f428e809b")


def test_edge_mask_sha256():
    # SHA-256 of EDGE_CROSSING_MASK (marching_cubes_tables.rs:49-70) as u16le[256].
    m = O.edge_mask_table()
    assert hashlib.sha256(b"".join(struct.pack("<H", int(v)) for v in m)).hexdigest() == (
        "2572f4b459bce46d930e57b2c261ef8143b6d43157272c698c64d70dd295b596")


def test_mask_is_sign_change_set():
    m = O.edge_mask_table()
    for case in range(256):
        want = 0
        for e, (u, v) in enumerate(EDGES):
            if ((case >> u) & 1) != ((case >> v) & 1):
                want |= 1 << e
        assert int(m[case]) == want, case
        assert m[case] == m[255 - case]


def test_triangles_use_exactly_the_crossed_edges_and_pad_with_minus_one():
    t = O.tri_table()
    m = O.edge_mask_table()
    hist = {}
    for case in range(256):
        row = [int(v) for v in t[case]]
        n = 0
        while n < 16 and row[n] >= 0:
            n += 1
        assert n % 3 == 0 and all(v == -1 for v in row[n:])
        used = 0
        for e in row[:n]:
            used |= 1 << e
        assert used == int(m[case])
        hist[n // 3] = hist.get(n // 3, 0) + 1
    assert hist == {0: 2, 1: 16, 2: 50, 3: 80, 4: 76, 5: 32}


def test_triangles_are_nondegenerate():
    t = O.tri_table()
    for case in range(256):
        row = [int(v) for v in t[case] if v >= 0]
        for k in range(0, len(row), 3):
            assert len(set(row[k:k + 3])) == 3


def _faces_and_antipodes():
    faces = [sum(1 << k for k in range(8) if CORNER[k][a] == v) for a in range(3) for v in (0, 1)]
    anti = {k: CORNER.index(tuple(1 - c for c in CORNER[k])) for k in range(8)}
    return faces, anti


def test_triangle_count_closed_form():
    # The count and single-pass kernels take a cell's triangle count from
    # its corners, V - 2F - 4H (V corners inside, F faces with all four
    # corners inside, H = 1 when the two corners outside are antipodal),
    # instead of the table: it must equal the table's count in every mixed case.
    t = O.tri_table()
    faces, anti = _faces_and_antipodes()
    for case in range(1, 255):
        n = int(np.count_nonzero(t[case] >= 0)) // 3
        v = bin(case).count("1")
        f = sum(1 for m in faces if case & m == m)
        out = [k for k in range(8) if not (case >> k) & 1]
        h = int(len(out) == 2 and anti[out[0]] == out[1])
        assert n == v - 2 * f - 4 * h, case


def test_triangle_count_gate_sequence():
    # The kernels' bit-sliced evaluation of V - 2F - 4H (mcd::tri_planes),
    # run here with one bit per case instead of one per cell.
    full = (1 << 256) - 1
    c = [sum(((case >> k) & 1) << case for case in range(256)) for k in range(8)]

    def full_add(a, b, d):
        x = a ^ b
        return x ^ d, (a & b) | (x & d)

    s1, k1 = full_add(c[0], c[1], c[2])
    s2, k2 = full_add(c[3], c[4], c[5])
    v0, k4 = full_add(s1, s2, c[6] ^ c[7])
    twos, k5 = full_add(k1, k2, c[6] & c[7])
    v1, v2 = twos ^ k4, k5 ^ (twos & k4)
    c04, c37, c15, c26 = c[0] & c[4], c[3] & c[7], c[1] & c[5], c[2] & c[6]
    sa, ka = full_add(c04 & c37, c15 & c26, c04 & c15)
    sb, kb = full_add(c26 & c37, c[0] & c[1] & c[2] & c[3], c[4] & c[5] & c[6] & c[7])
    f0, f1 = sa ^ sb, ka ^ kb ^ (sa & sb)
    inv = lambda x: full & ~x  # noqa: E731
    h = v2 & v1 & inv(v0) & (inv(c[0] | c[6]) | inv(c[1] | c[7]) | inv(c[2] | c[4]) |
                             inv(c[3] | c[5]))
    d0, d1 = v1 ^ f0, v2 ^ f1 ^ (inv(v1) & f0)
    planes = [v0, d0, d1 ^ h]
    t = O.tri_table()
    for case in range(1, 255):
        got = sum(((planes[i] >> case) & 1) << i for i in range(3))
        assert got == int(np.count_nonzero(t[case] >= 0)) // 3, case
