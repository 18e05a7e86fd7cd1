"""The committed lookup tables: identical to the reference's data, and
self-consistent in the ways the extraction relies on."""
import re

import numpy as np
import pytest

from binarymeshfitting_amd.tables import TRI_TABLE, DMC_POLYS, EDGE_CORNERS, LOCAL_EDGE
from tests.conftest import REFERENCE


def _ref_rows(path, name):
    text = path.read_text()
    start = text.index(name)
    body = text[text.index("{", start):]
    depth = 0
    for i, ch in enumerate(body):
        depth += ch == "{"
        depth -= ch == "}"
        if depth == 0:
            body = body[1:i]
            break
    return [[int(v) for v in re.findall(r"-?\d+", r)] for r in re.findall(r"\{([^{}]*)\}", body)]


@pytest.mark.skipif(not (REFERENCE / "MCTable.h").exists(), reason="reference sources not present")
def test_tri_table_matches_reference():
    ref = _ref_rows(REFERENCE / "MCTable.h", "tri_table")
    assert len(ref) == 256
    for m in range(256):
        row = ref[m]
        n = row.index(-1) if -1 in row else 16
        assert TRI_TABLE[m] == row[:n], m


@pytest.mark.skipif(not (REFERENCE / "MCTable.h").exists(), reason="reference sources not present")
def test_edge_map_matches_reference():
    """The reference's edge_map (MCTable.h:5-23) low 12 bits = edges whose
    corners differ; derived here from EDGE_CORNERS."""
    text = (REFERENCE / "MCTable.h").read_text()
    body = text[text.index("edge_map"):]
    body = body[body.index("{") + 1: body.index("};")]
    ref = [int(v, 16) for v in re.findall(r"0x([0-9A-Fa-f]+)", body)]
    assert len(ref) == 256
    for m in range(256):
        crossing = sum(1 << e for e, (a, b) in enumerate(EDGE_CORNERS) if ((m >> a) & 1) != ((m >> b) & 1))
        assert ref[m] & 0xFFF == crossing, m


@pytest.mark.skipif(not (REFERENCE / "Tables.hpp").exists(), reason="reference sources not present")
def test_dmc_table_matches_reference():
    ref = _ref_rows(REFERENCE / "Tables.hpp", "EdgeTable")
    for m in range(256):
        row = ref[m]
        polys, cur = [], []
        for e in row:
            if e == -2:
                break
            if e == -1:
                polys.append(cur)
                cur = []
            else:
                cur.append(e)
        if cur:
            polys.append(cur)
        assert DMC_POLYS[m] == polys, m


def test_tri_table_uses_exactly_the_crossing_edges():
    for m in range(256):
        crossing = {e for e, (a, b) in enumerate(EDGE_CORNERS) if ((m >> a) & 1) != ((m >> b) & 1)}
        used = set(TRI_TABLE[m])
        assert used == crossing, m
        assert len(TRI_TABLE[m]) % 3 == 0 and len(TRI_TABLE[m]) <= 15


def test_triangles_have_distinct_edges():
    for m in range(256):
        r = TRI_TABLE[m]
        for t in range(0, len(r), 3):
            assert len({r[t], r[t + 1], r[t + 2]}) == 3, (m, t)


def test_local_edge_geometry():
    # local edge endpoints agree with the (offset, axis) form
    for e, (dx, dy, dz, a) in enumerate(LOCAL_EDGE):
        c0, c1 = EDGE_CORNERS[e]
        assert c0 == 4 * dx + 2 * dy + dz
        assert c1 - c0 == (4, 2, 1)[a]


def test_cell_surfaces_are_closed_within_the_cube():
    """Per mask, every triangle side that is not shared by two triangles of
    the same cell lies on a cube face (both of its edges on that face)."""
    faces = []
    for axis in range(3):
        for side in (0, 1):
            fe = set()
            for e, (dx, dy, dz, a) in enumerate(LOCAL_EDGE):
                d = (dx, dy, dz)
                if a != axis and d[axis] == side:
                    fe.add(e)
            faces.append(fe)
    for m in range(1, 255):
        r = TRI_TABLE[m]
        sides = {}
        for t in range(0, len(r), 3):
            tri = r[t:t + 3]
            for k in range(3):
                key = tuple(sorted((tri[k], tri[(k + 1) % 3])))
                sides[key] = sides.get(key, 0) + 1
        for (a, b), n in sides.items():
            assert n in (1, 2), (m, a, b, n)
            if n == 1:
                assert any(a in f and b in f for f in faces), (m, a, b)
