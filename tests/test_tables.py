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
