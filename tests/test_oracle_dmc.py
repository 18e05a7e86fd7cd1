"""Pins of the dual-marching-cubes quad-mode oracle (CPU only).

The reference's quad path has no caller, so the oracle is pinned by an
independent numpy restatement (vertex counts from the polygon table, quads
from a brute-force walk over edges, mass points from the marching-cubes
vertices), by topology (a closed sphere gives a consistently oriented
2-manifold with Euler characteristic 2, wound like the marching-cubes table),
by flush_to_tris, and by the N=3 smoothing reducing to the triangle one.
"""
import numpy as np
import pytest

from binarymeshfitting_amd import workloads
from binarymeshfitting_amd.tables import DMC_POLYS, LOCAL_EDGE

RING = [[(0, -1, -1, 3), (0, -1, 0, 2), (0, 0, -1, 1), (0, 0, 0, 0)],
        [(-1, 0, -1, 7), (-1, 0, 0, 6), (0, 0, -1, 5), (0, 0, 0, 4)],
        [(-1, -1, 0, 11), (-1, 0, 0, 10), (0, -1, 0, 9), (0, 0, 0, 8)]]


def _sphere(oracle, dim=25, r_world=40.0, lo=-12.0):
    spec = oracle.SamplerSpec(kind=oracle.SAMPLER_SPHERE, world_size=r_world)
    d, o = oracle.geometry(oracle.make_chunk((lo, lo, lo), float(dim - 1), dim))
    return oracle.sample_block(spec, dim, o, d)


def _terrain(oracle, dim=17, pos=(0.0, 24.0, 0.0)):
    spec = oracle.SamplerSpec(**workloads.terrain_sampler())
    d, o = oracle.geometry(oracle.make_chunk(pos, float(dim - 1), dim))
    return oracle.sample_block(spec, dim, o, d)


def _masks(dens):
    lab = (dens < 0).astype(np.int64)
    c = dens.shape[0] - 1
    m = np.zeros((c, c, c), np.int64)
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                m |= lab[dx:dx + c, dy:dy + c, dz:dz + c] << (4 * dx + 2 * dy + dz)
    return m


def _numpy_dmc(dens):
    """Independent restatement: vertex ids per (cell, polygon), mass points
    from the edge points, quads by walking every edge."""
    dim = dens.shape[0]
    c = dim - 1
    masks = _masks(dens)
    lab = dens < 0
    base = -np.ones((c, c, c), np.int64)
    pos = []
    for x in range(c):
        for y in range(c):
            for z in range(c):
                m = int(masks[x, y, z])
                if m in (0, 255):
                    continue
                base[x, y, z] = len(pos)
                for poly in DMC_POLYS[m]:
                    acc = np.zeros(3, np.float32)
                    for e in poly:
                        dx, dy, dz, a = LOCAL_EDGE[e]
                        p0 = np.array([x + dx, y + dy, z + dz], np.float32)
                        p1 = p0.copy()
                        p1[a] += 1
                        s0 = dens[tuple(p0.astype(int))]
                        s1 = dens[tuple(p1.astype(int))]
                        mu = np.float32(np.float32(0.0) - s0) / np.float32(s1 - s0)
                        q = (p1 - p0) * mu + p0
                        acc = (acc + q).astype(np.float32)
                    pos.append(acc / np.float32(len(poly)))
    quads = []
    for x in range(dim):
        for y in range(dim):
            for z in range(dim):
                for a in range(3):
                    e1 = [x, y, z]
                    e1[a] += 1
                    if max(e1) >= dim or lab[x, y, z] == lab[tuple(e1)]:
                        continue
                    cells = [(x + r[0], y + r[1], z + r[2], r[3]) for r in RING[a]]
                    if any(min(cc[:3]) < 0 or max(cc[:3]) >= c for cc in cells):
                        continue
                    v = []
                    for cx, cy, cz, le in cells:
                        polys = DMC_POLYS[int(masks[cx, cy, cz])]
                        pi = [k for k, p in enumerate(polys) if le in p][0]
                        v.append(base[cx, cy, cz] + pi)
                    quads.append((v, bool(lab[x, y, z]), a))
    return np.array(pos, np.float32).reshape(-1, 3), quads


@pytest.mark.parametrize("field", ["sphere", "terrain"])
def test_dmc_matches_numpy_restatement(oracle, field):
    dens = _sphere(oracle, dim=13, lo=-6.0) if field == "sphere" else _terrain(oracle, dim=13)
    m = oracle.dmc_mesh(dens)
    pos, quads = _numpy_dmc(dens)
    assert m["V"] == len(pos)
    assert np.array_equal(m["pos"], pos)
    q = m["idx"].reshape(-1, 4)
    assert len(q) == len(quads)
    for got, (v, air_origin, a) in zip(q, quads):
        # same four vertices, cyclic order c0 c1 c3 c2 or its reverse
        fwd, rev = [v[0], v[1], v[3], v[2]], [v[0], v[2], v[3], v[1]]
        assert list(got) in (fwd, rev)


def test_sphere_quads_are_oriented_closed_manifold(oracle):
    dens = _sphere(oracle, dim=25)
    m = oracle.dmc_mesh(dens)
    q = m["idx"].reshape(-1, 4).astype(np.int64)
    directed = {}
    for quad in q:
        for k in range(4):
            e = (quad[k], quad[(k + 1) % 4])
            assert e not in directed          # each directed edge once: consistent orientation
            directed[e] = 1
    for (a, b) in directed:
        assert (b, a) in directed             # closed
    E = len(directed) // 2
    assert m["V"] - E + len(q) == 2           # a sphere
    # right-hand normals point to the solid (the centre), like the MC triangles
    p = m["pos"].astype(np.float64)
    ctr = np.full(3, 12.0)
    n = np.cross(p[q[:, 1]] - p[q[:, 0]], p[q[:, 2]] - p[q[:, 0]]) + np.cross(p[q[:, 2]] - p[q[:, 0]],
                                                                              p[q[:, 3]] - p[q[:, 0]])
    cen = p[q].mean(1)
    live = np.linalg.norm(n, axis=1) > 1e-9       # quads with four collinear vertices have no normal
    assert live.mean() > 0.95
    assert np.all(np.sum(n * (ctr - cen), 1)[live] > 0)
    spec = oracle.SamplerSpec(kind=oracle.SAMPLER_SPHERE, world_size=40.0)
    mc = oracle.extract_chunk(spec, oracle.make_chunk((-12.0, -12.0, -12.0), 24.0, 25))
    t = mc["idx"].reshape(-1, 3).astype(np.int64)
    pt = mc["pos"].astype(np.float64)
    nt = np.cross(pt[t[:, 1]] - pt[t[:, 0]], pt[t[:, 2]] - pt[t[:, 0]])
    live = np.linalg.norm(nt, axis=1) > 1e-9      # samples exactly on the sphere make point triangles
    assert live.mean() > 0.9
    assert np.all(np.sum(nt * (ctr - pt[t].mean(1)), 1)[live] > 0)


def test_vertex_count_is_polygon_count(oracle):
    dens = _terrain(oracle, dim=33)
    m = oracle.dmc_mesh(dens)
    masks = _masks(dens).ravel()
    assert m["V"] == sum(len(DMC_POLYS[int(k)]) for k in masks if k not in (0, 255))
    assert np.all(m["boundary"] == [(min(x, y, z) == 0 or max(x, y, z) == 31)
                                    for x, y, z in zip(m["code"] & 255, (m["code"] >> 8) & 255,
                                                       (m["code"] >> 16) & 255)])


def test_as_triangles_is_flush_to_tris(oracle):
    dens = _terrain(oracle, dim=17)
    q = oracle.dmc_mesh(dens)["idx"].reshape(-1, 4)
    t = oracle.dmc_mesh(dens, as_triangles=True)["idx"].reshape(-1, 6)
    assert np.array_equal(t, q[:, [0, 1, 2, 2, 3, 0]])


def test_smooth_n_reduces_to_triangles(oracle):
    spec = oracle.SamplerSpec(**workloads.terrain_sampler())
    mc = oracle.extract_chunk(spec, oracle.make_chunk((0.0, 16.0, 0.0), 32.0, 33))
    a = oracle.smooth(mc["pos"], mc["boundary"], mc["idx"], 5)
    b = oracle.smooth_n(mc["pos"], mc["boundary"], mc["idx"], 3, 5)
    assert np.array_equal(a, b)


def test_smooth_quads_matches_numpy(oracle):
    dens = _terrain(oracle, dim=17)
    m = oracle.dmc_mesh(dens)
    K = 3
    got = oracle.smooth_n(m["pos"], m["boundary"], m["idx"], 4, K, f64=True)
    p = m["pos"].astype(np.float64).copy()
    q = m["idx"].reshape(-1, 4)
    adj = [[] for _ in range(len(p))]
    for t, quad in enumerate(q):
        for v in quad:
            adj[v].append(t)
    for _ in range(K):
        cen = p[q].sum(1) / 4.0
        for v in range(len(p)):
            if adj[v] and not m["boundary"][v]:
                p[v] = cen[adj[v]].sum(0) / len(adj[v])
    assert np.abs(got - p).max() < 1e-12


def test_qef_placement_on_csg(oracle):
    """QEF vertices of the CSG scene: near the surface (|d| small at the
    vertex), and the FP32 solve within 1e-4 of its FP64 twin on the same
    polygons for most vertices (the truncation band makes rare jumps)."""
    wl = workloads.csg_scene()
    spec = oracle.SamplerSpec(**wl.sampler)
    ch = workloads.Chunk((16.0, 16.0, 16.0), 32.0, 33)
    oc = oracle.make_chunk(ch.pos, ch.size, ch.dim)
    delta, opos = oracle.geometry(oc)
    dens = oracle.sample_block(spec, ch.dim, opos, delta)
    m = oracle.dmc_mesh(dens, spec, opos, delta, placement=oracle.PLACE_QEF, tol=0.1)
    assert m["V"] > 100
    d64 = oracle.sample_block_f64(spec, ch.dim, opos.astype(np.float64), float(delta))
    p64, e64 = oracle.dmc_positions_f64(d64, m["code"], spec, opos, float(delta), oracle.PLACE_QEF, 0.1)
    err = np.abs(m["pos"] - p64).max(1)
    assert np.quantile(err, 0.99) < 1e-3
    mp = oracle.dmc_mesh(dens)
    # QEF vertices move off the mass point
    assert np.abs(m["pos"] - mp["pos"]).max() > 1e-3
