"""Differentiable extraction on the GPU (bmf_forward_diff / bmf_loss_nearest /
bmf_backward_diff) against the FP64 oracle.

Bar: the tape's forward mesh bit-exact with bmf_extract and the CPU
extraction; recomputed checkpoint states bit-exact with the oracle's FP32
smoothing; the nearest-target assignment exact and the loss within 1e-5
relative of FP64; the FP32 adjoint within 1e-4 relative (vector norm over
the checked samples) of the FP64 oracle gradient and of central finite
differences of the FP64 forward (h = 1e-3, Richardson-combined with h/2).
"""
import numpy as np
import pytest

from binarymeshfitting_amd import capi, workloads

pytestmark = pytest.mark.gpu

TERRAIN = workloads.terrain_sampler()


@pytest.fixture(scope="module")
def ctx():
    with capi.Context(max_chunks=64, max_samples=3 * 129 ** 3, max_vertices=1 << 21, max_indices=1 << 23) as c:
        c.set_sampler(**TERRAIN)
        yield c


@pytest.fixture(scope="module")
def tape(ctx):
    with capi.Tape(ctx, max_dim=129, max_vertices=1 << 20, max_indices=1 << 22, checkpoint_every=2) as t:
        yield t


def _oracle_mesh(oracle, c):
    spec = oracle.SamplerSpec(**TERRAIN)
    ch = oracle.make_chunk(c.pos, c.size, c.dim, c.overlap, c.iters, c.boundary_processing)
    return oracle.extract_chunk(spec, ch, want_density=True)


def _crossing_samples(m, dim):
    st = (dim * dim, dim, 1)
    e = m["edge"].astype(np.int64)
    p, a = e // 3, e % 3
    return np.unique(np.concatenate([p, p + np.take(st, a)]))


@pytest.mark.parametrize("K,bp", [(0, 0), (4, 0), (8, 0), (3, 1)])
def test_tape_forward_matches_extract_and_oracle(ctx, tape, oracle, K, bp):
    c = workloads.Chunk((0.0, 0.0, 0.0), 64.0, 65, 0.0, K, bp)
    cd = capi.chunk_desc(c.pos, c.size, c.dim, c.overlap, c.iters, c.boundary_processing)
    info = tape.forward(cd)
    pos, idx, bnd = tape.mesh()
    ref = _oracle_mesh(oracle, c)
    assert info.vertex_count == ref["V"] and info.index_count == ref["nI"]
    assert np.array_equal(idx, ref["idx"]) and np.array_equal(bnd, ref["boundary"])
    assert np.array_equal(pos, ref["pos"]), np.abs(pos - ref["pos"]).max()
    assert np.array_equal(tape.densities(), ref["density"])
    ctx.extract([cd], capi.MODE_PLAIN)
    rc, res, _, _ = ctx.results(1)
    capi.check(rc, "bmf_result")
    p2, i2, b2 = ctx.mesh(0)
    assert np.array_equal(p2, pos) and np.array_equal(i2, idx) and np.array_equal(b2, bnd)


def test_checkpoints_recompute_every_state(tape, oracle):
    c = workloads.Chunk((0.0, 0.0, 0.0), 64.0, 65, 0.0, 8, 0)
    tape.forward(capi.chunk_desc(c.pos, c.size, c.dim, c.overlap, c.iters, c.boundary_processing))
    ref = _oracle_mesh(oracle, workloads.Chunk(c.pos, c.size, c.dim, 0.0, 0, 0))
    for k in range(0, 9):
        want = oracle.smooth(ref["pos"], ref["boundary"], ref["idx"], k) if k else ref["pos"]
        got = tape.positions(k)
        assert np.array_equal(got, want), (k, np.abs(got - want).max())


def test_forward_from_external_densities(ctx, tape, oracle):
    c = workloads.Chunk((64.0, 0.0, 32.0), 64.0, 65, 0.0, 4, 0)
    ref = _oracle_mesh(oracle, c)
    dev = capi.DeviceArray.from_numpy(ref["density"].astype(np.float32))
    tape.forward(capi.chunk_desc(c.pos, c.size, c.dim, c.overlap, c.iters, c.boundary_processing), dev)
    pos, idx, bnd = tape.mesh()
    assert np.array_equal(idx, ref["idx"]) and np.array_equal(pos, ref["pos"])
    # the context's own sampler is untouched by the external forward
    ctx.extract([capi.chunk_desc(c.pos, c.size, c.dim, c.overlap, c.iters, c.boundary_processing)])
    capi.check(ctx.results(1)[0], "bmf_result")
    assert np.array_equal(ctx.mesh(0)[0], ref["pos"])


@pytest.mark.parametrize("mode", [capi.MODE_PLAIN, capi.MODE_BATCHED])
def test_extract_with_external_sampler(oracle, mode):
    chunks = [workloads.Chunk((0.0, 0.0, 0.0), 32.0, 33, 0.0, 2, 0),
              workloads.Chunk((16.0, 8.0, 0.0), 64.0, 65, 0.0, 0, 0),
              workloads.Chunk((0.0, 20.0, 40.0), 20.0, 21, 0.035, 3, 1)]
    refs = [_oracle_mesh(oracle, c) for c in chunks]
    packed = np.concatenate([r["density"].ravel() for r in refs]).astype(np.float32)
    dev = capi.DeviceArray.from_numpy(packed)
    with capi.Context(max_chunks=8, max_samples=packed.size, max_vertices=1 << 20, max_indices=1 << 22) as c2:
        c2.set_sampler(kind=capi.SAMPLER_EXTERNAL, densities=dev)
        c2.extract(workloads.to_capi(chunks), mode)
        rc, res, _, _ = c2.results(len(chunks))
        capi.check(rc, "bmf_result")
        for j, r in enumerate(refs):
            assert res[j].vertex_count == r["V"] and res[j].index_count == r["nI"]
            if r["V"]:
                pos, idx, bnd = c2.mesh(j)
                assert np.array_equal(idx, r["idx"]) and np.array_equal(pos, r["pos"])


def _config5(ctx, tape, oracle, dim=129, K=8, n_targets=200_000, centre=64.0, radius=40.0):
    c = workloads.Chunk((0.0, 0.0, 0.0), float(dim - 1), dim, 0.0, K, 0)
    info = tape.forward(capi.chunk_desc(c.pos, c.size, c.dim, c.overlap, c.iters, c.boundary_processing))
    targets = workloads.sphere_targets(n=n_targets, seed=3, radius=radius, centre=(centre,) * 3)
    tdev = capi.DeviceArray.from_numpy(targets)
    L, dldp, near = tape.loss_nearest(tdev)
    grad = tape.backward(dldp)
    return c, info, targets, L, dldp.numpy()[: info.vertex_count], near, grad


def test_loss_and_gradient_match_fp64_oracle(ctx, tape, oracle):
    """Config 5 at full size: 129^3 samples, K = 8, 200k targets."""
    c, info, targets, L, dldp, near, grad = _config5(ctx, tape, oracle)
    ref = _oracle_mesh(oracle, c)
    d64 = ref["density"].astype(np.float64)
    pos32, _, _ = tape.mesh()
    kd = oracle.KDTree(targets.astype(np.float64))
    # nearest target of the GPU's own FP32 positions, in FP64
    onear, _ = kd.nearest(pos32.astype(np.float64), threads=16)
    agree = np.mean(onear == near)
    print(f"V={info.vertex_count} nearest agreement with FP64 distances {agree:.6f}")
    assert agree >= 0.999
    # with the kernel's own FP32 distances the assignment is exact
    near32, _ = kd.nearest_f32(pos32, threads=16)
    assert np.array_equal(near, near32)
    L64, _ = oracle.loss_nearest_f64(pos32, targets, near)
    assert abs(L - L64) <= 1e-5 * L64
    # FP64 forward and adjoint, with the loss gradient taken at the GPU's
    # nearest targets (the assignment is checked above; a near-tie resolved
    # differently in FP64 would change dL/dp, not the adjoint under test)
    p64 = oracle.forward_f64(d64, ref, c.iters, 0)
    _, g64 = oracle.loss_nearest_f64(p64, targets, near)
    assert np.abs(dldp - g64).max() <= 1e-4 * np.abs(g64).max()
    grad64 = oracle.adjoint_f64(d64, ref, c.iters, 0, g64)
    g, r = grad.ravel().astype(np.float64), grad64.ravel()
    rel = np.linalg.norm(g - r) / np.linalg.norm(r)
    print(f"FP32 adjoint vs FP64: norm-relative {rel:.3e}, max {np.abs(g - r).max() / np.abs(r).max():.3e}")
    assert rel <= 1e-4
    assert np.abs(g - r).max() <= 1e-3 * np.abs(r).max()
    dim = c.dim
    off = np.ones(d64.size, bool)
    off[_crossing_samples(ref, dim)] = False
    assert np.all(grad.ravel()[off] == 0.0)


def test_nearest_ties_go_to_lower_index(ctx, tape, oracle):
    """Six targets at exactly 0.25 from every vertex (+-0.25 per axis), in a
    shuffled order among random decoys: the search returns the lowest index
    at the minimum FP32 distance, as a brute-force scan does."""
    c = workloads.Chunk((0.0, 16.0, 0.0), 32.0, 33, 0.0, 2, 0)
    info = tape.forward(capi.chunk_desc(c.pos, c.size, c.dim, c.overlap, c.iters, c.boundary_processing))
    pos, _, _ = tape.mesh()
    rng = np.random.default_rng(11)
    offs = np.float32(0.25) * np.concatenate([np.eye(3), -np.eye(3)]).astype(np.float32)
    t = (pos[:, None, :] + offs[None, :, :]).reshape(-1, 3)
    t = np.concatenate([t, rng.uniform(-8, 40, (5000, 3)).astype(np.float32)])
    t = t[rng.permutation(len(t))]
    L, dldp, near = tape.loss_nearest(capi.DeviceArray.from_numpy(t))
    kd = oracle.KDTree(t.astype(np.float64))
    near32, d32 = kd.nearest_f32(pos, threads=16)
    assert np.array_equal(near, near32)
    # the tie rule is exercised: most vertices have two or more of their own
    # six targets at the exact minimum distance
    od = (pos[:, None, :] + offs[None, :, :]) - pos[:, None, :]
    od2 = (od[..., 0] * od[..., 0] + od[..., 1] * od[..., 1]) + od[..., 2] * od[..., 2]
    tied = np.sum(np.sum(od2 == d32[:, None], 1) >= 2)
    print(f"{tied} of {info.vertex_count} vertices with tied nearest targets")
    assert tied >= 0.5 * info.vertex_count
    assert np.array_equal(dldp.numpy()[: info.vertex_count], (pos - t[near]) * (np.float32(2.0) / np.float32(info.vertex_count)))


def test_gradient_matches_finite_differences(ctx, tape, oracle):
    """Config 5: 1,000 samples on sign-changing edges with |d| > 10 h,
    h = 1e-3, central differences of the FP64 forward (nearest targets
    recomputed), Richardson-combined with h/2."""
    c, info, targets, L, dldp, near, grad = _config5(ctx, tape, oracle)
    ref = _oracle_mesh(oracle, c)
    d64 = ref["density"].astype(np.float64)
    kd = oracle.KDTree(targets.astype(np.float64))
    h = 1e-3
    cand = _crossing_samples(ref, c.dim)
    cand = cand[np.abs(d64.ravel()[cand]) > 10 * h]
    rng = np.random.default_rng(1000)
    smp = rng.choice(cand, size=1000, replace=False)
    f1, k1 = oracle.fd_grad_f64(d64, ref, c.iters, 0, kd, smp, h=h, threads=16)
    f2, k2 = oracle.fd_grad_f64(d64, ref, c.iters, 0, kd, smp, h=h / 2, threads=16)
    ok = ~(k1 | k2)
    rich = (4.0 * f2 - f1) / 3.0
    g = grad.ravel()[smp].astype(np.float64)
    rel_plain = np.linalg.norm((g - f1)[ok]) / np.linalg.norm(f1[ok])
    rel_rich = np.linalg.norm((g - rich)[ok]) / np.linalg.norm(rich[ok])
    print(f"samples {ok.sum()} of 1000 without a nearest-target switch; FD h=1e-3 rel {rel_plain:.3e}, "
          f"Richardson rel {rel_rich:.3e}")
    assert ok.mean() >= 0.95
    assert rel_rich <= 1e-4
    assert rel_plain <= 1e-2


def test_backward_boundary_processing_small(ctx, tape, oracle):
    c = workloads.Chunk((0.0, 16.0, 0.0), 32.0, 33, 0.0, 5, 1)
    info = tape.forward(capi.chunk_desc(c.pos, c.size, c.dim, c.overlap, c.iters, c.boundary_processing))
    ref = _oracle_mesh(oracle, c)
    rng = np.random.default_rng(7)
    g = rng.standard_normal((info.vertex_count, 3)).astype(np.float32)
    grad = tape.backward(capi.DeviceArray.from_numpy(g))
    grad64 = oracle.adjoint_f64(ref["density"].astype(np.float64), ref, c.iters, 1, g.astype(np.float64))
    rel = np.linalg.norm(grad.ravel() - grad64.ravel()) / np.linalg.norm(grad64.ravel())
    assert rel <= 1e-5


def test_tape_error_paths(ctx, tape):
    big = capi.chunk_desc((0.0, 0.0, 0.0), 129.0, 130, 0.0, 1, 0)
    with pytest.raises(capi.BmfError) as e:
        tape.forward(big)
    assert e.value.code == capi.ECAPACITY
    with capi.Tape(ctx, max_dim=33, max_vertices=1024, max_indices=4096) as t2:
        with pytest.raises(capi.BmfError) as e:
            t2.info()
        assert e.value.code == capi.ESTATE
        with pytest.raises(capi.BmfError) as e:   # mesh larger than the tape's capacity
            t2.forward(capi.chunk_desc((0.0, 16.0, 0.0), 32.0, 33, 0.0, 2, 0))
        assert e.value.code == capi.ECAPACITY


def test_cached_target_set_matches_per_call_build(ctx, oracle):
    """bmf_tape_set_targets + bmf_tape_loss_nearest give bmf_loss_nearest's
    outputs bit for bit, and the set survives later forward passes."""
    with capi.Tape(ctx, max_dim=65, max_vertices=1 << 18, max_indices=1 << 20, checkpoint_every=2) as t:
        c = workloads.Chunk((0.0, 0.0, 0.0), 64.0, 65, 0.0, 4, 0)
        t.forward(capi.chunk_desc(c.pos, c.size, c.dim, c.overlap, c.iters, c.boundary_processing))
        with pytest.raises(capi.BmfError) as e:
            t.loss_nearest_set()
        assert e.value.code == capi.ESTATE
        tgt = capi.DeviceArray.from_numpy(workloads.sphere_targets(n=50_000, seed=3, radius=20.0,
                                                                   centre=(32.0, 32.0, 32.0)))
        t.set_targets(tgt)
        for K in (4, 2):
            t.forward(capi.chunk_desc(c.pos, c.size, c.dim, c.overlap, K, c.boundary_processing))
            L1, g1, n1 = t.loss_nearest_set()
            L0, g0, n0 = t.loss_nearest(tgt)
            V = t.info().vertex_count
            assert L1 == L0
            assert np.array_equal(n1, n0)
            assert np.array_equal(g1.numpy()[:V], g0.numpy()[:V])


def test_tape_timing_reports_each_call(ctx):
    with capi.Tape(ctx, max_dim=65, max_vertices=1 << 18, max_indices=1 << 20, checkpoint_every=2) as t:
        assert t.timing() == [-1.0, -1.0, -1.0]
        t.forward(capi.chunk_desc((0.0, 0.0, 0.0), 64.0, 65, 0.0, 4, 0))
        t.set_targets(capi.DeviceArray.from_numpy(workloads.sphere_targets(n=10_000, radius=20.0,
                                                                           centre=(32.0, 32.0, 32.0))))
        _, g, _ = t.loss_nearest_set()
        t.backward(g)
        ms = t.timing()
        print("device ms (forward, loss, backward):", ms)
        assert all(0.0 < x < 1000.0 for x in ms)
