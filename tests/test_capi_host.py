"""C-ABI checks that need no GPU: the library loads, exports every symbol
include/bmf.h declares, and validates descriptors on the host."""
import re
from pathlib import Path

import pytest

from binarymeshfitting_amd import capi, workloads

ROOT = Path(__file__).resolve().parent.parent


def test_exports_match_header():
    decl = set(re.findall(r"\b(bmf_[a-z_0-9]+)\s*\(", (ROOT / "include" / "bmf.h").read_text()))
    L = capi.lib()
    for name in sorted(decl):
        assert hasattr(L, name), name
    assert set(capi.EXPORTS) <= decl


def test_version_and_strerror():
    L = capi.lib()
    assert b"gfx950" in L.bmf_version()
    assert L.bmf_strerror(capi.ECAPACITY) == b"arena capacity exceeded"
    assert L.bmf_strerror(99) == b"unknown error"


@pytest.mark.parametrize("bad", [
    dict(dim=1), dict(dim=257), dict(size=0.0), dict(size=float("nan")), dict(overlap=-0.1),
    dict(iters=-1), dict(iters=65), dict(boundary_processing=2), dict(pos=(float("inf"), 0.0, 0.0)),
])
def test_validate_rejects(bad):
    args = dict(pos=(0.0, 0.0, 0.0), size=64.0, dim=65, overlap=0.0, iters=4, boundary_processing=0)
    args.update(bad)
    assert capi.validate_chunks(workloads.to_capi([workloads.Chunk(**args)])) == capi.EINVAL


def test_validate_accepts_configs():
    for wl in (workloads.single_chunk(), workloads.uniform_grid(), workloads.mixed_lod()):
        assert capi.validate_chunks(workloads.to_capi(wl.chunks)) == capi.OK
    lim = capi.Limits(8, 1 << 20, 1 << 20, 1 << 20)
    assert capi.validate_chunks(workloads.to_capi(workloads.uniform_grid(n_side=3).chunks), lim) == capi.EINVAL
    assert capi.validate_chunks([], lim) == capi.OK


def test_mixed_lod_workload_shape():
    wl = workloads.mixed_lod()
    dims = [c.dim for c in wl.chunks]
    assert len(dims) == 1536
    assert dims.count(33) == 922 and dims.count(65) == 461 and dims.count(129) == 153
    for c in wl.chunks:
        ext = (c.dim - 1) * {33: 1, 65: 2, 129: 4}[c.dim]
        assert c.size == ext and all(0 <= p <= 4096 - ext for p in c.pos)
        assert 0 <= c.iters <= 8
    assert workloads.mixed_lod().chunks == wl.chunks  # deterministic


def test_splitmix64_known_values():
    # reference outputs of splitmix64 seeded with 0 (Vigna's published sequence)
    r = workloads.SplitMix64(0)
    assert r.next() == 0xE220A8397B1DCDAF
    assert r.next() == 0x6E789E6AA1B965F4
    assert r.next() == 0x06C45D188009454F
