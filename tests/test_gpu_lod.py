"""bmf_lod_queue (the watcher's LOD pass on the GPU) and bmf_extract_queued
against the oracle's sequential restatement, through the C ABI.

Bar: node flags, the dirty list, the generate batch's descriptors and items
bit-exact with the oracle on seeded worlds (focus moved after the build, flag
perturbations that reach every branch, several max_gen values); the queued
extraction gives the meshes of bmf_extract on the oracle's descriptors.
"""
from dataclasses import replace

import numpy as np
import pytest

from binarymeshfitting_amd import capi, workloads as wl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    with capi.Context(max_chunks=4096, max_samples=600 * 33 ** 3, max_vertices=1 << 21, max_indices=1 << 23) as c:
        c.set_sampler(**wl.terrain_sampler())
        yield c


CASES = [
    # (world size, build focus, run focus, perturbation, seed, max_gen, max_level)
    (4096.0, (0.0, 0.0, 0.0), (300.0, 20.0, -150.0), 0.0, 1, 400, 7),
    (4096.0, (0.0, 0.0, 0.0), (2000.0, 100.0, -500.0), 0.0, 2, 400, 7),
    (4096.0, (100.0, 0.0, 100.0), (-900.0, 50.0, 700.0), 0.1, 3, 400, 7),
    (8192.0, (0.0, 0.0, 0.0), (1500.0, 0.0, 1500.0), 0.2, 4, 1 << 20, 9),
    (8192.0, (500.0, 0.0, -500.0), (-3000.0, 10.0, 0.0), 0.3, 5, 64, 9),
]


@pytest.mark.parametrize("case", CASES, ids=[f"world{i}" for i in range(len(CASES))])
def test_lod_queue_matches_oracle(ctx, oracle, case):
    size, build, run, perturb, seed, max_gen, max_level = case
    base = wl.LodParams(max_level=max_level, chunk_resolution=33, process_iters=2, overlap=0.035)
    nodes, leaves = wl.lod_world(world_size=size, build_focus=build, params=base, seed=seed, perturb=perturb)
    # the renderables list: the leaves, then a seeded sample of inner nodes, in a seeded order
    rng = np.random.default_rng(seed)
    inner = np.flatnonzero(nodes["world_leaf"] == 0)
    extra = rng.choice(inner, size=min(len(inner), 16), replace=False) if len(inner) else np.zeros(0, np.int64)
    rlist = np.concatenate([leaves, extra]).astype(np.int32)
    rlist = rlist[rng.permutation(len(rlist))] if seed % 2 else rlist
    p = replace(base, focus=run, max_gen=max_gen)
    rc, flags, dirty, items, nc = ctx.lod_queue(nodes, rlist, p)
    capi.check(rc, "bmf_lod_queue")
    oflags, odirty, oitems, odescs, onc = oracle.lod_queue(nodes, rlist, p)
    print(f"nodes {len(nodes)}, list {len(rlist)}: dirty {len(odirty)}, chunks {onc}")
    assert nc == onc and onc > 0
    assert np.array_equal(flags, oflags)
    assert np.array_equal(dirty, odirty)
    assert np.array_equal(items, oitems)
    # the queued descriptors extract as the oracle's descriptors do
    ctx.extract_queued(capi.MODE_BATCHED)
    rc, res_q, _, _ = ctx.results(nc)
    capi.check(rc, "bmf_result")
    meshes_q = [ctx.mesh(j) if res_q[j].vertex_count else None for j in range(nc)]
    cd = [capi.chunk_desc(tuple(float(v) for v in d["pos"]), float(d["size"]), int(d["dim"]), float(d["overlap"]),
                          int(d["iters"]), int(d["boundary_processing"])) for d in odescs]
    ctx.extract(cd, capi.MODE_BATCHED)
    rc, res_h, _, _ = ctx.results(nc)
    capi.check(rc, "bmf_result")
    for j in range(nc):
        a, b = res_q[j], res_h[j]
        assert (a.vertex_count, a.index_count, a.contains_mesh, a.delta) == \
            (b.vertex_count, b.index_count, b.contains_mesh, b.delta)
        assert tuple(a.overlap_pos) == tuple(b.overlap_pos)
        if a.vertex_count:
            for x, y in zip(meshes_q[j], ctx.mesh(j)):
                assert np.array_equal(x, y)


def test_queued_plain_mode_and_state(ctx, oracle):
    nodes, leaves = wl.lod_world(world_size=4096.0, build_focus=(0.0, 0.0, 0.0))
    p = wl.LodParams(focus=(64.0, 30.0, 64.0), chunk_resolution=33, process_iters=1)
    rc, _, _, _, nc = ctx.lod_queue(nodes, leaves, p)
    capi.check(rc, "bmf_lod_queue")
    ctx.extract_queued(capi.MODE_PLAIN)
    rc, res_p, _, _ = ctx.results(nc)
    capi.check(rc, "bmf_result")
    pos_p = [ctx.mesh(j)[0] for j in range(nc) if res_p[j].vertex_count]
    ctx.extract_queued(capi.MODE_BATCHED)   # the descriptors stay queued across extractions
    rc, res_b, _, _ = ctx.results(nc)
    pos_b = [ctx.mesh(j)[0] for j in range(nc) if res_b[j].vertex_count]
    assert len(pos_p) == len(pos_b) and all(np.array_equal(a, b) for a, b in zip(pos_p, pos_b))
    # a host upload replaces them
    ctx.extract([capi.chunk_desc((0.0, 0.0, 0.0), 32.0, 33, 0.0, 0, 0)], capi.MODE_BATCHED)
    with pytest.raises(capi.BmfError) as e:
        ctx.extract_queued(capi.MODE_BATCHED)
    assert e.value.code == capi.ESTATE
    # capacity: the generate batch does not fit cap_items
    rc, _, _, _, need = ctx.lod_queue(nodes, leaves, p, cap_items=3)
    assert rc == capi.ECAPACITY and need == nc
