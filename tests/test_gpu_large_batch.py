"""Batches of more than 1,024 descriptors: up to 4,096 the plans run in one
block with several descriptors per thread, beyond that as several passes
over many blocks.  Bar: every chunk's mesh bit-exact with the oracle and
with the same chunks extracted in batches of 1,000 (one descriptor per
thread); overflow of the sample arena reported for every sampled chunk of a
large batch."""
import numpy as np
import pytest

from binarymeshfitting_amd import capi, workloads

pytestmark = pytest.mark.gpu

TERRAIN = workloads.terrain_sampler()


def _batch(n, seed=5):
    """n small chunks (5..17 samples per axis) scattered over and off the
    terrain band, K in 0..3, some with boundary processing and overlap."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        dim = int(rng.choice([5, 9, 12, 17]))
        size = float(rng.choice([4.0, 8.0, 16.0]))
        y = float(rng.uniform(-30.0, 90.0))
        out.append(workloads.Chunk((float(rng.uniform(-500, 500)), y, float(rng.uniform(-500, 500))), size, dim,
                                   float(rng.choice([0.0, 0.035])), int(rng.integers(0, 4)), int(rng.integers(0, 2))))
    return out


def _ochunk(oracle, c):
    return oracle.make_chunk(c.pos, c.size, c.dim, c.overlap, c.iters, c.boundary_processing)


@pytest.mark.parametrize("n", [1025, 3000, 5000])
def test_large_batch_matches_oracle_and_small_batches(oracle, n):
    cs = _batch(n)
    spec = oracle.SamplerSpec(**TERRAIN)
    with capi.Context(max_chunks=8192, max_samples=n * 17 ** 3, max_vertices=1 << 22, max_indices=1 << 24) as ctx:
        ctx.set_sampler(**TERRAIN)
        ctx.extract(workloads.to_capi(cs), capi.MODE_BATCHED)
        rc, res, _, _ = ctx.results(n)
        capi.check(rc, "bmf_result")
        big = {j: ctx.mesh(j) for j in range(n) if res[j].vertex_count}
        counts = [(r.vertex_count, r.index_count, r.contains_mesh, r.sampled) for r in res]
        # the same chunks in batches of at most 1,024 (one plan block)
        small = {}
        for s0 in range(0, n, 1000):
            part = cs[s0:s0 + 1000]
            ctx.extract(workloads.to_capi(part), capi.MODE_BATCHED)
            rc, r2, _, _ = ctx.results(len(part))
            capi.check(rc, "bmf_result")
            for q, r in enumerate(r2):
                assert (r.vertex_count, r.index_count, r.contains_mesh, r.sampled) == counts[s0 + q]
                if r.vertex_count:
                    small[s0 + q] = ctx.mesh(q)
    smoothed = 0
    for j, c in enumerate(cs):
        ref = oracle.extract_chunk(spec, _ochunk(oracle, c))
        V, I, cm, _ = counts[j]
        assert V == ref["V"] and I == ref["nI"] and bool(cm) == ref["contains_mesh"], (j, c)
        if V:
            pos, idx, bnd = big[j]
            assert np.array_equal(idx, ref["idx"]) and np.array_equal(bnd, ref["boundary"]), (j, c)
            assert np.array_equal(pos, ref["pos"]), (j, c)
            for a, b in zip(big[j], small[j]):
                assert np.array_equal(a, b)
            smoothed += c.iters > 0
    print(f"{n} chunks, {len(big)} with a mesh, {smoothed} smoothed")
    assert smoothed > 100


def test_large_batch_sample_overflow():
    cs = _batch(5000, seed=6)
    with capi.Context(max_chunks=8192, max_samples=100 * 17 ** 3, max_vertices=1 << 22, max_indices=1 << 24) as ctx:
        ctx.set_sampler(**TERRAIN)
        ctx.extract(workloads.to_capi(cs), capi.MODE_BATCHED)
        rc, res, _, _ = ctx.results(len(cs))
        assert rc == capi.ECAPACITY
        sampled = [r for r in res if r.status == capi.ECAPACITY]
        assert len(sampled) > 100 and all(r.vertex_count == 0 for r in sampled)
