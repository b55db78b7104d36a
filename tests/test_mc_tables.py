"""The packed marching-cubes case table against the reference's own tables.

oracle/_ref/libref_mctables.so is compiled from the reference's
GPU/MarchingCubesHelper.cpp (oracle/Makefile, target `ref`); the comparison is
entry by entry.  The derivation of the edge flags and vertex counts from the
packed table (which is how the kernels use it) is checked against the
reference's stored edgeFlags[] and numVertices[].
"""
import numpy as np
import pytest

import oracle

EDGES = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5),
         (2, 6), (3, 7)]
CORNERS = [(0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0), (1, 0, 0), (1, 0, 1), (1, 1, 1),
           (1, 1, 0)]


def unpack(v):
    out = []
    for k in range(16):
        e = (int(v) >> (4 * k)) & 0xF
        if e == 0xF:
            break
        out.append(e)
    return out


@pytest.fixture(scope="module")
def ref():
    t = oracle.reference_mc_tables()
    if t is None:
        pytest.skip("oracle/_ref not built (reference tree absent when building)")
    return t


def test_packed_table_equals_reference(ref):
    packed = oracle.mc_tri_packed()
    for c in range(256):
        row = [int(x) for x in ref["tri"][c] if x >= 0]
        assert unpack(packed[c]) == row, c
        # entries after the first -1 are all -1
        k = len(row)
        assert all(x < 0 for x in ref["tri"][c][k:])


def test_vertex_counts_derive_from_table(ref):
    packed = oracle.mc_tri_packed()
    for c in range(256):
        assert len(unpack(packed[c])) == ref["num_vertices"][c]


def test_edge_flags_derive_from_corner_crossings(ref):
    assert [tuple(e) for e in ref["edge_vertices"].tolist()] == EDGES
    assert [tuple(o) for o in ref["corner_offsets"].tolist()] == CORNERS
    for c in range(256):
        m = 0
        for e, (a, b) in enumerate(EDGES):
            if ((c >> a) & 1) != ((c >> b) & 1):
                m |= 1 << e
        assert m == ref["edge_flags"][c], c


def test_table_uses_only_crossed_edges():
    packed = oracle.mc_tri_packed()
    for c in range(256):
        for e in unpack(packed[c]):
            a, b = EDGES[e]
            assert ((c >> a) & 1) != ((c >> b) & 1), (c, e)
        assert len(unpack(packed[c])) % 3 == 0
    assert len(unpack(packed[0])) == 0 and len(unpack(packed[255])) == 0
