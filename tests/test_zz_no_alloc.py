"""bmf_loss_nearest_noalloc (the nearest-target loss with its hierarchy in a
static device array of the library, no allocation) against
bmf_loss_nearest, through the C ABI: identical nearest targets, dL/dp and
loss on two calls in a row, the second with fewer targets."""
import numpy as np
import pytest

from binarymeshfitting_amd import capi, workloads

pytestmark = pytest.mark.gpu


def _noalloc(tape, targets, V):
    g = capi.DeviceArray((max(V, 1), 3), np.float32)
    near = capi.DeviceArray((max(V, 1),), np.int32)
    loss = capi.DeviceArray((1,), np.float64)
    capi.check(capi.lib().bmf_loss_nearest_noalloc(tape.h, targets.ptr, targets.shape[0], g.ptr, near.ptr, loss.ptr),
               "bmf_loss_nearest_noalloc")
    tape.ctx.sync()
    return float(loss.numpy()[0]), g.numpy()[:V], near.numpy()[:V]


def test_noalloc_loss_matches_loss_nearest():
    with capi.Context(max_chunks=4, max_samples=65 ** 3, max_vertices=1 << 19, max_indices=1 << 21) as ctx:
        ctx.set_sampler(**workloads.terrain_sampler())
        with capi.Tape(ctx, max_dim=65, max_vertices=1 << 18, max_indices=1 << 20, checkpoint_every=2) as t:
            c = workloads.Chunk((0.0, 0.0, 0.0), 64.0, 65, 0.0, 4, 0)
            info = t.forward(capi.chunk_desc(c.pos, c.size, c.dim, c.overlap, c.iters, c.boundary_processing))
            V = info.vertex_count
            assert V > 0
            for n, seed, radius in ((200_000, 3, 20.0), (7_000, 5, 12.0)):
                tgt = capi.DeviceArray.from_numpy(workloads.sphere_targets(n=n, seed=seed, radius=radius,
                                                                           centre=(32.0, 32.0, 32.0)))
                Ln, gn, nn = _noalloc(t, tgt, V)
                L, g, near = t.loss_nearest(tgt)
                print(f"n={n}: V={V} loss {L:.6f} (noalloc {Ln:.9f})")
                assert np.array_equal(nn, near)
                assert np.array_equal(gn, g.numpy()[:V])
                assert np.float32(Ln) == np.float32(L)


def test_noalloc_rejects_more_targets_than_its_maximum():
    """The size check happens on the host before any work is enqueued, so
    the claimed count is never read from the (small) buffer."""
    with capi.Context(max_chunks=4, max_samples=33 ** 3, max_vertices=1 << 16, max_indices=1 << 18) as ctx:
        ctx.set_sampler(**workloads.terrain_sampler())
        with capi.Tape(ctx, max_dim=33, max_vertices=1 << 16, max_indices=1 << 18) as t:
            t.forward(capi.chunk_desc((0.0, 16.0, 0.0), 32.0, 33, 0.0, 2, 0))
            tgt = capi.DeviceArray((4, 3), np.float32)
            g = capi.DeviceArray((1 << 16, 3), np.float32)
            rc = capi.lib().bmf_loss_nearest_noalloc(t.h, tgt.ptr, capi.NOALLOC_MAX_TARGETS + 1, g.ptr, None, None)
            assert rc == capi.ECAPACITY
