"""bmf_format (VBO-ready arrays of a batch, formatted on the device) against
the oracle's restatement of WorldOctreeNode::format / GLChunk::format_data /
DMCChunk::copy_verts_and_inds and ColorMapper::generate_colors.

Bar: every array bit-exact with the oracle's formatting of the same meshes
(which the parity tests compare with the CPU extraction), in grid and world
space, white and noise colours, chunk-local and batch-global indices, in
both extraction modes; capacity overflow reported with the totals needed.
"""
import numpy as np
import pytest

from binarymeshfitting_amd import capi, workloads

pytestmark = pytest.mark.gpu

TERRAIN = workloads.terrain_sampler()


@pytest.fixture(scope="module")
def ctx():
    with capi.Context(max_chunks=64, max_samples=40 * 65 ** 3, max_vertices=1 << 21, max_indices=1 << 23) as c:
        c.set_sampler(**TERRAIN)
        yield c


def _chunks():
    # surface chunks, chunks culled above and below the terrain band, a chunk
    # with overlap (non-zero overlap_pos shift) and mixed sizes
    out = [workloads.Chunk((float(64 * i), 0.0, float(64 * k)), 64.0, 65, 0.0, 4, 0) for i in range(2) for k in range(2)]
    out.insert(1, workloads.Chunk((0.0, 128.0, 0.0), 64.0, 65, 0.0, 4, 0))
    out.insert(3, workloads.Chunk((0.0, -96.0, 0.0), 64.0, 65, 0.0, 4, 0))
    out.append(workloads.Chunk((20.0, 10.0, 30.0), 32.0, 33, 0.055, 3, 1))
    out.append(workloads.Chunk((-40.0, 4.0, 7.0), 128.0, 65, 0.0, 0, 0))
    return out


def _oracle_format(oracle, ctx, n, space, colors, glob, normals):
    rc, res, _, _ = ctx.results(n)
    capi.check(rc, "bmf_result")
    chunks = []
    for j in range(n):
        r = res[j]
        d = dict(V=r.vertex_count, I=r.index_count, contains_mesh=bool(r.contains_mesh), delta=r.delta,
                 opos=tuple(r.overlap_pos), pos=np.zeros((0, 3), np.float32), idx=np.zeros(0, np.uint32))
        if r.vertex_count:
            d["pos"], d["idx"], _ = ctx.mesh(j)
        if normals:
            d["nrm"] = ctx.normals(j) if r.vertex_count else np.zeros((0, 3), np.float32)
        chunks.append(d)
    pos, col, nrm, idx = oracle.format_batch(chunks, space, 1 if colors == capi.COLOR_WHITE else 0, glob)
    if colors == capi.COLOR_NOISE:
        col = oracle.vertex_colors(np.concatenate([c["pos"] for c in chunks if c["contains_mesh"]]))
    return pos, col, nrm, idx


@pytest.mark.parametrize("mode", [capi.MODE_PLAIN, capi.MODE_BATCHED], ids=["plain", "batched"])
@pytest.mark.parametrize("space,colors,glob", [(capi.SPACE_GRID, capi.COLOR_WHITE, True),
                                               (capi.SPACE_WORLD, capi.COLOR_NOISE, True),
                                               (capi.SPACE_WORLD, capi.COLOR_NONE, False)])
def test_format_matches_oracle(ctx, oracle, mode, space, colors, glob):
    chunks = _chunks()
    ctx.extract(workloads.to_capi(chunks), mode)
    rc, V, I, pos, col, nrm, idx = ctx.format(space=space, colors=colors, global_indices=glob)
    capi.check(rc, "bmf_format")
    opos, ocol, _, oidx = _oracle_format(oracle, ctx, len(chunks), space, colors, glob, False)
    assert V == len(opos) and I == len(oidx) and V > 0
    assert np.array_equal(pos, opos)
    assert np.array_equal(idx, oidx)
    if colors != capi.COLOR_NONE:
        assert np.array_equal(col, ocol)


def test_format_normals_and_quads(ctx, oracle):
    chunks = _chunks()
    ctx.set_mesh_options(mesh=capi.MESH_QUADS, normals=True)
    try:
        ctx.extract(workloads.to_capi(chunks), capi.MODE_BATCHED)
        rc, V, I, pos, col, nrm, idx = ctx.format(space=capi.SPACE_WORLD, colors=capi.COLOR_WHITE, normals=True)
        capi.check(rc, "bmf_format")
        opos, _, onrm, oidx = _oracle_format(oracle, ctx, len(chunks), capi.SPACE_WORLD, capi.COLOR_WHITE, True, True)
        assert np.array_equal(pos, opos) and np.array_equal(nrm, onrm) and np.array_equal(idx, oidx)
        assert I % 4 == 0
    finally:
        ctx.set_mesh_options()


def test_format_capacity_and_state(ctx):
    chunks = _chunks()
    ctx.extract(workloads.to_capi(chunks), capi.MODE_BATCHED)
    rc, V, I, *_ = ctx.format()
    capi.check(rc, "bmf_format")
    rc, V2, I2, *_ = ctx.format(cap_vertices=V - 1)
    assert rc == capi.ECAPACITY and (V2, I2) == (V, I)
    rc, *_ = ctx.format(normals=True)  # normals were not computed by this extraction
    assert rc == capi.ESTATE
