"""Lookup tables of the extraction path, decoded from include/bmf_mc_tables.h.

The header is the single committed copy of the marching-cubes triangle table
(the reference's MCTable.h data) and of the dual-marching-cubes polygon table
(Tables.hpp); both the HIP kernels and this module read it.
"""
import re
from pathlib import Path

_HEADER = Path(__file__).resolve().parent.parent / "include" / "bmf_mc_tables.h"

# Local edge -> (dx, dy, dz, axis) of the lattice edge relative to the cell
# origin, in the order of DMCChunk::polygonize_cell (DMCChunk.cpp:543-565).
LOCAL_EDGE = [
    (0, 0, 0, 0), (0, 0, 1, 0), (0, 1, 0, 0), (0, 1, 1, 0),
    (0, 0, 0, 1), (0, 0, 1, 1), (1, 0, 0, 1), (1, 0, 1, 1),
    (0, 0, 0, 2), (0, 1, 0, 2), (1, 0, 0, 2), (1, 1, 0, 2),
]


def _corner(dx, dy, dz):
    return 4 * dx + 2 * dy + dz


# Local edge -> the two cell-corner bits it joins.
EDGE_CORNERS = []
for (dx, dy, dz, a) in LOCAL_EDGE:
    d = [dx, dy, dz]
    e = list(d)
    e[a] += 1
    EDGE_CORNERS.append((_corner(*d), _corner(*e)))


def _parse(name):
    text = _HEADER.read_text()
    body = text[text.index("#define " + name):]
    body = body[body.index("{") + 1: body.index("}")]
    return [int(v, 16) for v in re.findall(r"0x([0-9a-fA-F]+)ull", body)]


def _decode_tri(v):
    n = v >> 60
    return [(v >> (4 * s)) & 15 for s in range(n)]


def _decode_dmc(v):
    n = v >> 60
    nib = [(v >> (4 * s)) & 15 for s in range(n)]
    polys, cur = [], []
    for e in nib:
        if e == 0xF:
            polys.append(cur)
            cur = []
        else:
            cur.append(e)
    if cur:
        polys.append(cur)
    return polys


TRI_PACKED = _parse("BMF_TRI_TABLE_INIT")
DMC_PACKED = _parse("BMF_DMC_POLY_TABLE_INIT")
assert len(TRI_PACKED) == 256 and len(DMC_PACKED) == 256
TRI_TABLE = [_decode_tri(v) for v in TRI_PACKED]
DMC_POLYS = [_decode_dmc(v) for v in DMC_PACKED]
