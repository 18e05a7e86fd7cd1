"""Vertex normals of the triangle path on the GPU against the CPU oracle.

Bar: normals bit-exact with the FP32 oracle (same operation sequence),
positions identical to the run without normals; the FP32 normals' error
against the FP64 twin is reported.
"""
import numpy as np
import pytest

from binarymeshfitting_amd import capi, workloads

pytestmark = pytest.mark.gpu

TERRAIN = workloads.terrain_sampler()


@pytest.fixture(scope="module")
def ctx():
    with capi.Context(max_chunks=64, max_samples=20 * 65 ** 3, max_vertices=1 << 21, max_indices=1 << 23) as c:
        c.set_sampler(**TERRAIN)
        yield c


@pytest.mark.parametrize("mode", [capi.MODE_PLAIN, capi.MODE_BATCHED], ids=["plain", "batched"])
def test_normals_match_oracle(ctx, oracle, mode):
    chunks = [workloads.Chunk((0.0, 0.0, 0.0), 64.0, 65, 0.0, K, bp) for K, bp in ((0, 0), (3, 0), (4, 1), (8, 0))]
    chunks.append(workloads.Chunk((64.0, 0.0, 0.0), 32.0, 33, 0.0, 2, 0))
    cd = workloads.to_capi(chunks)
    ctx.extract(cd, mode)
    capi.check(ctx.results(len(cd))[0], "bmf_result")
    plain_pos = [ctx.mesh(j)[0] for j in range(len(cd))]
    ctx.set_mesh_options(normals=True)
    try:
        ctx.extract(cd, mode)
        rc, res, _, _ = ctx.results(len(cd))
        capi.check(rc, "bmf_result")
        spec = oracle.SamplerSpec(**TERRAIN)
        worst64 = 0.0
        for j, c in enumerate(chunks):
            pos, idx, bnd = ctx.mesh(j)
            nrm = ctx.normals(j)
            assert np.array_equal(pos, plain_pos[j])
            base = oracle.extract_chunk(spec, oracle.make_chunk(c.pos, c.size, c.dim, c.overlap, 0, c.boundary_processing))
            p32, n32 = oracle.smooth_normals(base["pos"], base["boundary"], base["idx"], c.iters, c.boundary_processing)
            assert np.array_equal(pos, p32)
            assert np.array_equal(nrm, n32), (c, np.abs(nrm - n32).max())
            _, n64 = oracle.smooth_normals(base["pos"], base["boundary"], base["idx"], c.iters, c.boundary_processing,
                                           f64=True)
            worst64 = max(worst64, float(np.quantile(np.abs(nrm - n64).max(1), 0.99)))
        print(f"normals: p99 |FP32 - FP64| over chunks (worst) {worst64:.3g}")
    finally:
        ctx.set_mesh_options(normals=False)


def test_normals_rejected_in_quad_mode(ctx):
    with pytest.raises(capi.BmfError) as e:
        ctx.set_mesh_options(mesh=capi.MESH_QUADS, normals=True)
    assert e.value.code == capi.EINVAL
