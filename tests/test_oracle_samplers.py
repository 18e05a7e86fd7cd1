"""Pins for the oracle's samplers: the FBM simplex noise (the project's own
definition, no reference counterpart), the reference's implicit samplers and
the CSG field, plus the QEF restatement."""
import math

import numpy as np
import pytest

from tests import refimpl


def test_snoise_equals_all_lattice_point_sum(oracle):
    rng = np.random.default_rng(3)
    L = oracle.lib()
    pts = rng.uniform(-40, 40, size=(400, 3))
    pts[:40] = np.round(pts[:40])  # lattice-aligned coordinates
    for seed in (1337, 7):
        for x, y, z in pts:
            ref = refimpl.snoise_lattice_sum(x, y, z, seed)
            got64 = L.or_snoise3_f64(x, y, z, seed)
            assert abs(got64 - ref) < 1e-12
            got32 = L.or_snoise3_f32(np.float32(x), np.float32(y), np.float32(z), seed)
            assert abs(got32 - refimpl.snoise_lattice_sum(float(np.float32(x)), float(np.float32(y)),
                                                          float(np.float32(z)), seed)) < 2e-5


def test_grad_index_matches_restatement(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(9)
    for i, j, k in rng.integers(-5000, 5000, size=(300, 3)):
        assert L.or_grad_index(int(i), int(j), int(k), 1337) == refimpl.grad_index(int(i), int(j), int(k), 1337)


def test_grad_index_uniform(oracle):
    L = oracle.lib()
    h = [L.or_grad_index(i, j, k, 1337) for i in range(-8, 8) for j in range(-8, 8) for k in range(-8, 8)]
    cnt = np.bincount(h, minlength=16)
    e = len(h) / 16
    assert ((cnt - e) ** 2 / e).sum() < 40  # chi-square, 15 dof


def test_snoise_range_and_continuity(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(5)
    vals = np.array([L.or_snoise3_f64(*rng.uniform(-100, 100, 3), 1) for _ in range(20000)])
    assert np.abs(vals).max() <= 1.0
    assert abs(vals.mean()) < 0.02 and 0.15 < vals.std() < 0.5
    # continuous: small steps give small changes even across simplex faces
    for _ in range(2000):
        p = rng.uniform(-10, 10, 3)
        dp = rng.standard_normal(3) * 1e-7
        assert abs(L.or_snoise3_f64(*(p + dp), 2) - L.or_snoise3_f64(*p, 2)) < 1e-4


def test_terrain_f32_close_to_f64_and_bounded(oracle):
    spec = oracle.SamplerSpec()
    d32 = oracle.sample_block(spec, 17, np.array([3.0, 20.0, -5.0], np.float32), 1.0)
    d64 = oracle.sample_block_f64(spec, 17, np.array([3.0, 20.0, -5.0]), 1.0)
    assert np.abs(d32 - d64).max() < 5e-5
    y = 20.0 + np.arange(17)[None, :, None]
    # |fbm| <= 1 exactly: d within amp of the plane
    assert np.all(np.abs(d32 + (y - 32.0)) <= 24.0)


def test_terrain_seed_and_determinism(oracle):
    a = oracle.sample_block(oracle.SamplerSpec(seed=1337), 9, np.zeros(3, np.float32), 1.0)
    b = oracle.sample_block(oracle.SamplerSpec(seed=1337), 9, np.zeros(3, np.float32), 1.0)
    c = oracle.sample_block(oracle.SamplerSpec(seed=1338), 9, np.zeros(3, np.float32), 1.0)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.abs(a - c).max() > 0.1


def test_terrain_fbm_sums_octaves(oracle):
    """d = amp * clamp(sum_o g^o n(f0 L^o p, seed+o) / sum_o g^o) - (y - y0)."""
    L = oracle.lib()
    spec = oracle.SamplerSpec(octaves=3, freq=0.05, lacunarity=2.0, gain=0.5, amp=10.0, y0=4.0, seed=11)
    for p in [(1.5, 2.25, -3.0), (20.0, 5.0, 7.75)]:
        f0 = float(np.float32(0.05))  # parameters are FP32 fields
        S = sum(0.5 ** o * L.or_snoise3_f64(*(c * f0 * 2 ** o for c in p), 11 + o) for o in range(3))
        exp = 10.0 * max(-1.0, min(1.0, S / 1.75)) - (p[1] - 4.0)
        assert abs(L.or_terrain_value_f64(spec.ref, *p) - exp) < 1e-12
        assert abs(L.or_terrain_value_f32(spec.ref, *p) - exp) < 1e-5


@pytest.mark.parametrize("kind", [1, 2, 3, 4])
def test_implicit_samplers_closed_form(oracle, kind):
    spec = oracle.SamplerSpec(kind=kind, world_size=40.0)
    opos = np.array([-12.0, -11.0, -10.5], np.float32)
    d = oracle.sample_block(spec, 9, opos, 3.0)
    for (i, j, k) in [(0, 0, 0), (4, 4, 4), (8, 3, 1), (2, 7, 5)]:
        x, y, z = (float(np.float32(np.float32(i * 3.0) + opos[0])), float(np.float32(np.float32(j * 3.0) + opos[1])),
                   float(np.float32(np.float32(k * 3.0) + opos[2])))
        if kind == 1:
            e = -(math.sqrt(x * x + y * y + z * z) - 10.0)
        elif kind == 2:
            q = [abs(x) - 5.0, abs(y) - 5.0, abs(z) - 5.0]
            e = -min(max(q), math.sqrt(sum(v * v for v in q)))
        elif kind == 3:
            qx = math.sqrt(x * x + y * y) - 10.0
            e = -(math.sqrt(qx * qx + z * z) - 4.0)
        else:
            e = -y
        assert abs(d[i, j, k] - e) < 1e-5


def test_csg_union_and_gradient(oracle):
    sph = np.array([[0, 0, 0, 5], [9, 0, 0, 4]], np.float32)
    box = np.array([[0, 10, 0, 2, 3, 4]], np.float32)
    spec = oracle.SamplerSpec(kind=oracle.SAMPLER_CSG, spheres=sph, boxes=box)
    L = oracle.lib()
    g = np.zeros(3, np.float64)
    import ctypes
    gp = g.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    # inside sphere 0: value 5 - |p|
    assert abs(L.or_csg_value_f64(spec.ref, 1.0, 0.0, 0.0, gp) - 4.0) < 1e-12
    assert np.allclose(g, [-1, 0, 0])
    # inside the box, nearest face is +y at distance 1 (|py| = 2 of half 3)
    assert abs(L.or_csg_value_f64(spec.ref, 0.0, 12.0, 0.0, gp) - 1.0) < 1e-12
    assert np.allclose(g, [0, -1, 0])
    # far away: floored
    assert L.or_csg_value_f64(spec.ref, 100.0, 100.0, 100.0, gp) == -16.0
    # gradient matches central differences
    rng = np.random.default_rng(1)
    for _ in range(50):
        p = rng.uniform(-8, 14, 3)
        v = L.or_csg_value_f64(spec.ref, *p, gp)
        if v <= -15.9:
            continue
        gd = g.copy()
        h = 1e-6
        fd = [(L.or_csg_value_f64(spec.ref, *(p + h * e), None) - L.or_csg_value_f64(spec.ref, *(p - h * e), None)) /
              (2 * h) for e in np.eye(3)]
        assert np.allclose(gd, fd, atol=1e-4)


# --------------------------------------------------------------- QEF -----
def _qef_textbook(p, n, tol):
    """Mass-point-centred least squares with an eigen pseudo-inverse that
    keeps eigenvalues in [tol, 1/tol] (the reference's truncation band)."""
    A = n.T @ n
    mp = p.mean(0)
    b = (np.einsum("ij,ij->i", p, n)[:, None] * n).sum(0)
    w, Q = np.linalg.eigh(A)
    keep = np.minimum(np.abs(w), 1 / np.where(w == 0, 1e-300, np.abs(w))) >= tol
    inv = np.where(keep, 1 / np.where(w == 0, 1, w), 0)
    x = Q @ (inv * (Q.T @ (b - A @ mp)))
    return x + mp


def test_qef_three_planes_corner(oracle):
    corner = np.array([1.5, -2.0, 0.25])
    n = np.eye(3)
    p = corner + np.array([[0, 1, 2], [3, 0, -1], [0.5, 0.5, 0]])
    for f64 in (True, False):
        x, err = oracle.qef_solve(p, n, 0.1, f64=f64)
        assert np.allclose(x, corner, atol=1e-5)
        # error uses the uncentred ATb with the centred solution (qef_simd.h:434-440, 457):
        # |ATb - ATA (x - masspoint)|^2 = |ATA masspoint|^2 here, not 0
        b = (np.einsum("ij,ij->i", p, n)[:, None] * n).sum(0)
        assert abs(err - np.sum((b - (corner - p.mean(0))) ** 2)) < 1e-4


def test_qef_two_planes_projects_mass_point_on_line(oracle):
    n = np.array([[1.0, 0, 0], [0, 1.0, 0]])
    p = np.array([[2.0, 5, 3], [7, -1.0, -1]])
    x, _ = oracle.qef_solve(p, n, 0.1, f64=True)
    mp = p.mean(0)
    assert np.allclose(x, [2.0, -1.0, mp[2]], atol=1e-12)


def test_qef_one_plane_projects_on_plane(oracle):
    nn = np.array([1.0, 2.0, 2.0]) / 3.0
    p = np.array([[1.0, 1, 1], [3, 0, 0], [0, 3, 0]])
    d = p @ nn
    p = p - (d - 2.0)[:, None] * nn  # all points on the plane n.x = 2
    n = np.tile(nn, (3, 1))
    x, err = oracle.qef_solve(p, n, 0.1, f64=True)
    mp = p.mean(0)
    assert np.allclose(x, mp, atol=1e-9)  # mass point already on the plane
    assert abs(x @ nn - 2.0) < 1e-9


@pytest.mark.parametrize("seed", range(6))
def test_qef_matches_textbook_pinv(oracle, seed):
    rng = np.random.default_rng(seed)
    k = rng.integers(2, 13)
    p = rng.uniform(-1, 1, (k, 3))
    n = rng.standard_normal((k, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    tol = 0.1
    x64, e64 = oracle.qef_solve(p, n, tol, f64=True)
    ref = _qef_textbook(p, n, tol)
    assert np.allclose(x64, ref, atol=1e-6)
    x32, e32 = oracle.qef_solve(p, n, tol)
    assert np.abs(x32 - x64).max() < 1e-4
    # error = |ATb - ATA x| ^ 2 with the uncentred ATb (qef_simd.h:434-440)
    A = n.T @ n
    b = (np.einsum("ij,ij->i", p, n)[:, None] * n).sum(0)
    xc = x64 - p.mean(0)
    assert abs(e64 - np.sum((b - A @ xc) ** 2)) < 1e-9 * max(1.0, e64)


def test_qef_count_limits(oracle):
    x, e = oracle.qef_solve(np.ones((1, 3)), np.ones((1, 3)), 0.1)
    assert np.all(x == 0) and e == 0
    x, e = oracle.qef_solve(np.ones((13, 3)), np.ones((13, 3)), 0.1)
    assert np.all(x == 0) and e == 0
