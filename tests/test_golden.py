"""The oracle against the committed regression fixture (tests/golden).

The fixture freezes the oracle's densities and meshes on nine small chunks
(terrain at 5 and 7 octaves, overlap and boundary processing, a single cell,
solid and air chunks, an implicit sphere, a CSG union); tests/golden/make_golden.py
wrote it. A change that alters one bit of the oracle's output fails here.
"""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

FIXTURE = Path(__file__).resolve().parent / "golden" / "regression_meshes.json"
CASES = json.loads(FIXTURE.read_text())


def digest(a, dtype):
    return hashlib.sha256(np.ascontiguousarray(a, dtype).tobytes()).hexdigest()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_matches_fixture(oracle, case):
    c, e = case["chunk"], case["expect"]
    spec = oracle.SamplerSpec(**case["sampler"])
    chunk = oracle.make_chunk(c["pos"], c["size"], c["dim"], c["overlap"], c["iters"], c["boundary_processing"])
    ref = oracle.extract_chunk(spec, chunk, want_density=True)
    assert (ref["V"], ref["nI"], ref["contains_mesh"]) == (e["V"], e["nI"], e["contains_mesh"])
    dens = ref["density"]
    assert [int(v) for v in dens.ravel()[:8].view(np.uint32)] == e["density_head"]
    assert digest(dens, "<f4") == e["density_sha256"]
    assert digest(ref["pos"], "<f4") == e["pos_sha256"]
    assert digest(ref["boundary"], "u1") == e["boundary_sha256"]
    assert digest(ref["idx"], "<u4") == e["idx_sha256"]


def test_fixture_covers_the_edge_cases():
    """Empty meshes (solid, air, single cell) and meshes with smoothing,
    overlap and boundary processing are all present."""
    by = {c["name"]: c["expect"] for c in CASES}
    assert by["terrain-16-solid"]["V"] == 0 and not by["terrain-16-solid"]["contains_mesh"]
    assert by["terrain-air"]["V"] == 0 and by["terrain-air"]["contains_mesh"]  # air, dim % 32 != 0
    assert all(by[n]["V"] > 0 for n in ("terrain-16", "terrain-64", "sphere-40", "csg-32"))
    assert any(c["chunk"]["boundary_processing"] and c["chunk"]["overlap"] > 0 for c in CASES)
