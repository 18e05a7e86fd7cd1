"""FP32-vs-FP64 error statements for the GPU path's results (test
infrastructure only: imported by tests/ and by the checker section of
bench.py, never by the product package).

Every function takes what the GPU returned for one chunk and recomputes the
same quantity on the same topology with the oracle's FP64 twins: densities
from the FP64 noise / CSG at the FP64 sample positions overlap_pos + i*delta
(NoiseSampler.cpp:19-31, ImplicitSampler.cpp:12-18), vertex placement
mu = (0 - s0) / (s1 - s0) (DMCChunk.cpp:657-674) or the quad mode's mass point
/ QEF (qef_simd.h:392-579), then the K dual/primal smoothing steps
(MeshProcessor.cpp:130-306), normals included, all in FP64.  The error of a
vertex is the largest absolute coordinate difference (grid units for
positions; unit vectors for normals).
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from oracle import pyoracle as po


def stats(per_item, extent=None):
    """max / p99 / median of per-item errors (plus max / extent when given)."""
    e = np.asarray(per_item, np.float64).ravel()
    if e.size == 0:
        out = {"count": 0, "max": 0.0, "p99": 0.0, "median": 0.0}
    else:
        out = {"count": int(e.size), "max": float(e.max()), "p99": float(np.quantile(e, 0.99)),
               "median": float(np.median(e))}
    if extent is not None:
        out["max_over_extent"] = out["max"] / float(extent)
    return out


def merge(parts):
    """Concatenate per-chunk error arrays."""
    parts = [np.asarray(p, np.float64).ravel() for p in parts if len(p)]
    return np.concatenate(parts) if parts else np.zeros(0)


def _chunk(c, iters=None):
    return po.make_chunk(c.pos, c.size, c.dim, c.overlap, c.iters if iters is None else iters,
                         c.boundary_processing)


def triangle_positions(spec, c, gpu_pos):
    """Per-vertex error of the triangle path's final positions."""
    r = po.extract_chunk(spec, _chunk(c), want_f64=True)
    if r["V"] != len(gpu_pos):
        raise ValueError(f"vertex count {len(gpu_pos)} differs from the oracle's {r['V']}")
    if not r["V"]:
        return np.zeros(0)
    return np.abs(np.asarray(gpu_pos, np.float64) - r["pos_f64"]).max(1)


def triangle_normals(spec, c, gpu_pos, gpu_nrm):
    """Per-vertex errors (positions, normals) of the triangle path with
    normals: FP64 placement, then the FP64 smoothing with its normals."""
    base = po.extract_chunk(spec, _chunk(c, iters=0), want_f64=True)
    if base["V"] != len(gpu_pos):
        raise ValueError(f"vertex count {len(gpu_pos)} differs from the oracle's {base['V']}")
    if not base["V"]:
        return np.zeros(0), np.zeros(0)
    p64, n64 = po.smooth_normals(base["pos_f64"], base["boundary"], base["idx"], c.iters, c.boundary_processing,
                                 f64=True)
    ep = np.abs(np.asarray(gpu_pos, np.float64) - p64).max(1)
    en = np.abs(np.asarray(gpu_nrm, np.float64) - n64).max(1)
    return ep, en


def quad_positions(spec, c, gpu_pos, placement=po.PLACE_MASS_POINT, tol=0.1, as_triangles=False, want_normals=False,
                   gpu_nrm=None):
    """Per-vertex error of the quad mode's final positions (and normals when
    want_normals): FP64 densities, FP64 mass point or QEF, FP64 smoothing
    as MeshProcessor<4> (or <3> on the triangles of flush_to_tris)."""
    oc = _chunk(c)
    delta, opos = po.geometry(oc)
    dens = po.sample_block(spec, c.dim, opos, delta)
    m = po.dmc_mesh(dens, spec, opos, delta, placement=placement, tol=tol, as_triangles=as_triangles)
    if m["V"] != len(gpu_pos):
        raise ValueError(f"vertex count {len(gpu_pos)} differs from the oracle's {m['V']}")
    if not m["V"]:
        return (np.zeros(0), np.zeros(0)) if want_normals else np.zeros(0)
    d64 = po.sample_block_f64(spec, c.dim, np.asarray(opos, np.float64), float(delta))
    p64, _ = po.dmc_positions_f64(d64, m["code"], spec, opos, float(delta), placement, tol)
    N = 3 if as_triangles else 4
    if want_normals:
        p64, n64 = po.smooth_normals_n(p64, m["boundary"], m["idx"], N, c.iters, c.boundary_processing, f64=True)
        ep = np.abs(np.asarray(gpu_pos, np.float64) - p64).max(1)
        return ep, np.abs(np.asarray(gpu_nrm, np.float64) - n64).max(1)
    if c.iters > 0 and m["nI"]:
        p64 = po.smooth_n(p64, m["boundary"], m["idx"], N, c.iters, c.boundary_processing, f64=True)
    return np.abs(np.asarray(gpu_pos, np.float64) - p64).max(1)


def over_chunks(fn, items, threads=16):
    """fn(item) over items on a thread pool (the oracle releases the GIL)."""
    with ThreadPoolExecutor(max(1, threads)) as ex:
        return list(ex.map(fn, items))


def gradient(dens32, mesh, K, bp, targets, near, gpu_grad, fd_samples=1000, h=1e-3, seed=1000, threads=16):
    """Error statement of an FP32 adjoint d(loss)/d(density) (loss = mean
    squared distance to the nearest target, the GPU's assignment) against
    (a) the FP64 adjoint of the FP64 forward and (b) central differences of
    the FP64 forward with nearest targets recomputed, Richardson-combined
    over h and h/2, on fd_samples endpoints of sign-changing edges with
    |d| > 10 h; samples whose stencil switches a nearest target (a kink of
    the loss) are excluded and counted."""
    d64 = np.asarray(dens32, np.float64)
    dim = d64.shape[0]
    p64 = po.forward_f64(d64, mesh, K, bp)
    _, g64 = po.loss_nearest_f64(p64, targets, near)
    grad64 = po.adjoint_f64(d64, mesh, K, bp, g64).ravel()
    g = np.asarray(gpu_grad, np.float64).ravel()
    out = {"vs_fp64_adjoint": {"rel_norm": float(np.linalg.norm(g - grad64) / np.linalg.norm(grad64)),
                               "max_rel": float(np.abs(g - grad64).max() / np.abs(grad64).max())}}
    if fd_samples:
        st = (dim * dim, dim, 1)
        e = mesh["edge"].astype(np.int64)
        p, a = e // 3, e % 3
        cand = np.unique(np.concatenate([p, p + np.take(st, a)]))
        cand = cand[np.abs(d64.ravel()[cand]) > 10 * h]
        rng = np.random.default_rng(seed)
        smp = rng.choice(cand, size=min(fd_samples, len(cand)), replace=False)
        kd = po.KDTree(np.asarray(targets, np.float64))
        try:
            f1, k1 = po.fd_grad_f64(d64, mesh, K, bp, kd, smp, h=h, threads=threads)
            f2, k2 = po.fd_grad_f64(d64, mesh, K, bp, kd, smp, h=h / 2, threads=threads)
        finally:
            kd.close()
        ok = ~(k1 | k2)
        rich = (4.0 * f2 - f1) / 3.0
        gs = g[smp]
        out["vs_central_differences"] = {
            "samples": int(len(smp)), "kinks_excluded": int((~ok).sum()), "h": h,
            "rel_norm_richardson": float(np.linalg.norm((gs - rich)[ok]) / np.linalg.norm(rich[ok])),
            "rel_norm_plain_h": float(np.linalg.norm((gs - f1)[ok]) / np.linalg.norm(f1[ok])),
        }
    return out
