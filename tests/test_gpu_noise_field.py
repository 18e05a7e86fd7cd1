"""The NOISE sampler (the reference's NoiseSamplers block shapes with the
project's own noise) on the GPU against the oracle, through the C ABI.

Bar: densities of sample_block and every mesh of batched and plain
extraction bit-exact with the oracle for each preset (the 2D fields go
through the sampling kernel's column cache, the 3D ones through the direct
evaluation); chunks that the plan's bound classifies without sampling carry
the oracle's flags; the smallest and largest chunk sizes of the column
cache; the differentiable forward with a NOISE sampler equal to the
extraction.
"""
import numpy as np
import pytest

from binarymeshfitting_amd import capi, workloads

pytestmark = pytest.mark.gpu

# (preset, chunk size, y origins of the chunk rows) so that the rows cross the surface band
PRESETS = {
    "terrain2d": (capi.PRESET_TERRAIN2D, 64.0, (-64.0, 0.0, 200.0)),
    "terrain2d_pert": (capi.PRESET_TERRAIN2D_PERT, 128.0, (-128.0, 0.0, 900.0)),
    "world": (capi.PRESET_WORLD, 128.0, (-128.0, 0.0, -1300.0)),
    "terrain3d": (capi.PRESET_TERRAIN3D, 16.0, (-16.0, 0.0, 40.0)),
    "terrain3d_pert": (capi.PRESET_TERRAIN3D_PERT, 128.0, (-128.0, 0.0, 700.0)),
    "windy3d": (capi.PRESET_WINDY3D, 256.0, (-256.0, 0.0, 1100.0)),
}


@pytest.fixture(scope="module")
def ctx():
    with capi.Context(max_chunks=256, max_samples=64 * 65 ** 3, max_vertices=1 << 22, max_indices=1 << 24) as c:
        yield c


def _chunks(size, ys, dim=33, iters=2):
    return [workloads.Chunk((x * size, y, z * size), size, dim, 0.0, iters, 0)
            for x in (-1, 0) for z in (0, 1) for y in ys]


def _ochunk(oracle, c):
    return oracle.make_chunk(c.pos, c.size, c.dim, c.overlap, c.iters, c.boundary_processing)


@pytest.mark.parametrize("name", list(PRESETS))
def test_sample_block_bit_exact(ctx, oracle, name):
    which, size, ys = PRESETS[name]
    sd = capi.noise_preset(which)
    ctx.set_sampler(**sd)
    c = workloads.Chunk((size / 3, ys[0] / 2, -size / 5), size, 41, 0.0, 0, 0)
    d = ctx.sample_block(workloads.to_capi([c])[0])
    delta, opos = oracle.geometry(_ochunk(oracle, c))
    ref = oracle.sample_block(oracle.SamplerSpec(**sd), c.dim, opos, delta)
    assert np.array_equal(d, ref), np.abs(d.astype(np.float64) - ref).max()


@pytest.mark.parametrize("name", list(PRESETS))
@pytest.mark.parametrize("mode", [capi.MODE_BATCHED, capi.MODE_PLAIN], ids=["batched", "plain"])
def test_extraction_bit_exact(ctx, oracle, name, mode):
    which, size, ys = PRESETS[name]
    sd = capi.noise_preset(which)
    spec = oracle.SamplerSpec(**sd)
    ctx.set_sampler(**sd)
    cs = _chunks(size, ys)
    ctx.extract(workloads.to_capi(cs), mode)
    rc, res, _, _ = ctx.results(len(cs))
    capi.check(rc, "bmf_result")
    total = culled = 0
    for j, c in enumerate(cs):
        ref = oracle.extract_chunk(spec, _ochunk(oracle, c))
        assert bool(res[j].contains_mesh) == ref["contains_mesh"], (name, c)
        assert res[j].vertex_count == ref["V"] and res[j].index_count == ref["nI"], (name, c)
        culled += res[j].sampled == 0
        if ref["V"]:
            pos, idx, bnd = ctx.mesh(j)
            assert np.array_equal(idx, ref["idx"]) and np.array_equal(bnd, ref["boundary"]), (name, c)
            assert np.array_equal(pos, ref["pos"]), (name, c)
            total += ref["V"]
    print(f"{name}: {total} vertices, {culled} of {len(cs)} chunks classified without sampling")
    assert total > 0
    if mode == capi.MODE_BATCHED:
        assert culled >= 4  # the far row lies outside the field's bound


@pytest.mark.parametrize("dim", [2, 3, 5, 129])
def test_column_cache_edge_sizes(ctx, oracle, dim):
    sd = capi.noise_preset(capi.PRESET_TERRAIN2D)
    ctx.set_sampler(**sd)
    c = workloads.Chunk((5.0, -20.0, 7.0), 48.0, dim, 0.0, 0, 0)
    d = ctx.sample_block(workloads.to_capi([c])[0])
    delta, opos = oracle.geometry(_ochunk(oracle, c))
    ref = oracle.sample_block(oracle.SamplerSpec(**sd), c.dim, opos, delta)
    assert np.array_equal(d, ref)


def test_forward_diff_with_a_noise_field(ctx, oracle):
    sd = capi.noise_preset(capi.PRESET_TERRAIN2D_PERT)
    ctx.set_sampler(**sd)
    c = workloads.Chunk((0.0, -64.0, 0.0), 128.0, 65, 0.0, 4, 0)
    with capi.Tape(ctx, max_dim=65, max_vertices=1 << 18, max_indices=1 << 20, checkpoint_every=2) as t:
        t.forward(capi.chunk_desc(c.pos, c.size, c.dim, c.overlap, c.iters, c.boundary_processing))
        pos, idx, bnd = t.mesh()
    ref = oracle.extract_chunk(oracle.SamplerSpec(**sd), _ochunk(oracle, c))
    assert ref["V"] > 0
    assert np.array_equal(idx, ref["idx"]) and np.array_equal(pos, ref["pos"])
