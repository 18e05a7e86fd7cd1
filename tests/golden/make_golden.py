"""Writes tests/golden/regression_meshes.json from the CPU oracle.

The reference holds no golden vectors for the extraction path (its only
result-pinning data are the lookup tables, checked in tests/test_tables.py),
so this fixture is self-generated: it freezes the oracle's outputs on a few
small chunks (densities, mesh sizes and SHA-256 digests of the vertex
positions, boundary flags and triangle indices) so that any later change to
the oracle or to the HIP path that alters a single bit fails
tests/test_golden.py (oracle) and tests/test_gpu_golden.py (HIP path). It
guards against regressions; the oracle itself is pinned by the independent
formulations in tests/test_oracle_*.py.

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from binarymeshfitting_amd import workloads  # noqa: E402
from oracle import pyoracle  # noqa: E402

OUT = Path(__file__).resolve().parent / "regression_meshes.json"


def digest(a, dtype):
    return hashlib.sha256(np.ascontiguousarray(a, dtype).tobytes()).hexdigest()


def _csg_small():
    wl = workloads.csg_scene(n_side=1, c=32, seed=42, n_spheres=24, n_boxes=12)
    s = dict(wl.sampler)
    s["spheres"] = np.asarray(s["spheres"], np.float32).tolist()
    s["boxes"] = np.asarray(s["boxes"], np.float32).tolist()
    return s


def cases():
    terrain = workloads.terrain_sampler()
    lod = workloads.terrain_sampler(workloads.TerrainParams(octaves=7, freq=0.005))
    sphere = dict(kind=1, world_size=40.0)
    csg = _csg_small()
    C = workloads.Chunk
    return [
        ("terrain-16-solid", terrain, C((0.0, 0.0, 0.0), 16.0, 17, 0.0, 3, 0)),
        ("terrain-16", terrain, C((3.0, 20.0, 1.0), 16.0, 17, 0.0, 3, 0)),
        ("terrain-1cell", terrain, C((5.0, 30.0, 9.0), 1.0, 2, 0.0, 1, 0)),
        ("terrain-32-overlap-bp", terrain, C((10.0, 7.5, 3.0), 32.0, 33, 0.055, 4, 1)),
        ("terrain-64", terrain, C((0.0, 0.0, 0.0), 64.0, 65, 0.0, 4, 0)),
        ("terrain-7oct-32-spacing2", lod, C((256.0, 0.0, 128.0), 64.0, 33, 0.0, 2, 0)),
        ("sphere-40", sphere, C((-20.0, -19.0, -21.5), 40.0, 41, 0.0, 1, 0)),
        ("csg-32", csg, C((0.0, 0.0, 0.0), 32.0, 33, 0.0, 0, 0)),
        ("terrain-air", terrain, C((0.0, 512.0, 0.0), 16.0, 17, 0.0, 2, 0)),
    ]


def record(oracle, sampler, c):
    spec = oracle.SamplerSpec(**sampler)
    chunk = oracle.make_chunk(c.pos, c.size, c.dim, c.overlap, c.iters, c.boundary_processing)
    ref = oracle.extract_chunk(spec, chunk, want_density=True)
    dens = ref["density"]
    return dict(
        V=int(ref["V"]), nI=int(ref["nI"]), contains_mesh=bool(ref["contains_mesh"]),
        density_sha256=digest(dens, "<f4"),
        density_head=[int(v) for v in dens.ravel()[:8].view(np.uint32)],
        pos_sha256=digest(ref["pos"], "<f4"),
        boundary_sha256=digest(ref["boundary"], "u1"),
        idx_sha256=digest(ref["idx"], "<u4"),
    )


def main():
    pyoracle.lib()
    out = []
    for name, sampler, c in cases():
        r = record(pyoracle, sampler, c)
        out.append(dict(name=name, sampler=sampler, chunk=dict(pos=list(c.pos), size=c.size, dim=c.dim,
                                                               overlap=c.overlap, iters=c.iters,
                                                               boundary_processing=c.boundary_processing),
                        expect=r))
        print(name, r["V"], r["nI"], r["contains_mesh"])
    OUT.write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
