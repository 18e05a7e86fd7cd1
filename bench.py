#!/usr/bin/env python3
"""Throughput of the chunk-extraction path on one MI355X.

Workload (default): 512 chunks of 64^3 cells tiled 8x8x8 over the terrain
d = 24*fbm - (y - 32) (FBM of 3D simplex noise, 5 octaves, base frequency
0.02, lacunarity 2, gain 0.5, seed 1337), FP32, marching-cubes triangles
plus 4 dual/primal smoothing iterations; the whole batch is one BATCHED
extraction (persistent kernels, device work queue, hipGraph replay).  A step
is one extraction of the batch; inputs (the descriptors) are uploaded inside
the step, outputs stay in HBM.

Printed JSON line: value = chunks/s of the batched path; beside it the plain
per-chunk GPU path, the CPU oracle (multithreaded, the reference's algorithm)
on a bounded sample, the roofline of the dominant kernel (sampling, FP32
VALU bound) from HIP events on the library stream, and a parity check of
every chunk's mesh against the CPU run.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "64^3-cell chunks meshed/sec, batched vs plain GPU vs CPU, at fixed tolerance"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--profile-reps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-path leg")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = all host cores (at most 16)")
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed passes of the CPU baseline (median)")
    ap.add_argument("--lod-steps", type=int, default=10,
                    help="timed batched steps of the uneven-LOD leg (1,536 chunks); 0 skips it")
    ap.add_argument("--lod-class", type=int, default=0, choices=(0, 1, 2, 3),
                    help="restrict the uneven-LOD leg to the chunks whose label rows take W u64 words "
                         "(W = ceil(dim / 64): 1 for 33^3, 2 for 65^3, 3 for 129^3 samples), so that the frame "
                         "forks no row-word classes onto auxiliary streams; 0 = all chunks")
    ap.add_argument("--world-steps", type=int, default=10,
                    help="timed batched steps of the default-world leg (0 skips it)")
    ap.add_argument("--hermite-steps", type=int, default=5,
                    help="timed batched steps of the CSG quad/QEF leg (64 chunks of 64^3 cells); 0 skips it")
    ap.add_argument("--single-steps", type=int, default=20,
                    help="timed plain extractions of the single 64^3-cell chunk; 0 skips it")
    ap.add_argument("--quads-steps", type=int, default=10,
                    help="timed batched steps of the uniform batch in DMC quad mode (512 chunks); 0 skips it")
    ap.add_argument("--diff-steps", type=int, default=5,
                    help="timed steps of the forward+loss+backward leg (one 128^3-cell chunk, K=8); 0 skips it")
    ap.add_argument("--contexts", type=int, default=3,
                    help="the timed batches of each batched leg alternate over 1 .. this many contexts (each with "
                         "its own arenas and stream, so one batch's sampling overlaps the previous batch's row and "
                         "smoothing kernels); the fastest count is reported, with every count's figure")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the library's default launch shapes for the headline batch")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (3 launches) and exit; its FETCH_SIZE and "
                         "WRITE_SIZE readings scale the counters of the extraction kernels to HBM bytes")
    return ap.parse_args()


def host_cpu():
    """CPU model and core counts of the host (the CPU baseline's hardware)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"model": model, "nproc": os.cpu_count(), "affinity": aff}


def copy_calibration(mib):
    """A plain device-to-device copy of known size, for scaling the L2-to-memory
    byte counters of the other kernels to bytes."""
    import torch

    nbytes = mib << 20
    src = torch.empty(nbytes, dtype=torch.uint8, device="cuda:0").fill_(7)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    assert int(dst[-1]) == 7
    print(json.dumps({"copy_bytes_per_launch": nbytes, "launches": 3}), flush=True)
    return 0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def stream_times(capi, ctx, make_ctx, arr, n, steps, nctx, warmup=2):
    """A stream of `steps` batched extractions of the same batch, timed on
    `ctx` alone (back to back on its stream) and alternating over k = 2 ..
    nctx contexts (`ctx` and further ones from make_ctx(): own arenas, stream
    and captured frame; the same sampler, mesh options and launch shapes), so
    that consecutive batches overlap on the GPU.  Returns the seconds per
    batch for each k, the fastest k (a per-workload choice, like the launch
    shapes) and whether every context's meshes equal the first's bit for
    bit."""
    extra = []
    try:
        for _ in range(nctx - 1):
            cx = make_ctx()
            cx.set_tuning(**ctx.tuning())
            extra.append(cx)
        cs = [ctx] + extra
        for cx in cs:
            for _ in range(max(1, warmup)):
                capi.check(cx.extract_raw(arr, n, capi.MODE_BATCHED), "bmf_extract")
            cx.sync()

        def run(group):
            t0 = time.perf_counter()
            for i in range(steps):
                capi.check(group[i % len(group)].extract_raw(arr, n, capi.MODE_BATCHED), "bmf_extract")
            for cx in group:
                cx.sync()
            return (time.perf_counter() - t0) / steps

        per = {k: run(cs[:k]) for k in range(1, nctx + 1)}
        best = min(per, key=per.get)
        same = True
        if extra:
            rc, res, _, _ = ctx.results(n)
            capi.check(rc, "bmf_result")
            for cx in extra:
                rc2, res2, _, _ = cx.results(n)
                capi.check(rc2, "bmf_result")
                for j in range(n):
                    if (res[j].vertex_count, res[j].index_count) != (res2[j].vertex_count, res2[j].index_count):
                        same = False
                    elif res[j].vertex_count and not all(
                            np.array_equal(u, w) for u, w in zip(ctx.mesh(j), cx.mesh(j))):
                        same = False
    finally:
        for cx in extra:
            cx.close()
    return per, best, same


def stream_fields(n, per, best, same):
    """JSON fields of a leg's stream timings."""
    return {"contexts": best, "meshes_identical_across_contexts": same,
            "one_context": {"chunks_per_s": round(n / per[1], 2), "ms_per_step": round(1e3 * per[1], 4)},
            "stream_over_contexts": {str(k): {"chunks_per_s": round(n / w, 2), "ms_per_step": round(1e3 * w, 4)}
                                     for k, w in per.items()}}


def lod_leg(capi, roofline, workloads, steps, only_class=0, nctx=2):
    """Uneven-LOD batch: 1,536 chunks (922 / 461 / 153 of 32^3 / 64^3 / 128^3
    cells at spacing 1 / 2 / 4, origins from seed 7 in a 4096^3 world), 7-octave
    field, K ~ U{0..8}; shapes from the autotuner, then batched and plain.
    only_class W > 0 keeps the chunks with ceil(dim / 64) == W."""
    wl = workloads.mixed_lod()
    if only_class:
        wl = workloads.Workload(f"{wl.name}-class{only_class}", wl.sampler,
                                [c for c in wl.chunks if (c.dim + 63) // 64 == only_class])
    cd = workloads.to_capi(wl.chunks)
    n = len(cd)
    arr = capi.chunk_array(cd)
    with capi.Context(max_chunks=2048, max_samples=96 * 129 ** 3, max_vertices=1 << 23, max_indices=1 << 25) as ctx:
        ctx.set_sampler(**wl.sampler)
        t0 = time.perf_counter()
        tuned = ctx.autotune(cd, reps=3)
        t_tune = time.perf_counter() - t0

        def make():
            cx = capi.Context(max_chunks=2048, max_samples=96 * 129 ** 3, max_vertices=1 << 23, max_indices=1 << 25)
            cx.set_sampler(**wl.sampler)
            return cx

        per, used, same = stream_times(capi, ctx, make, arr, n, steps, nctx)
        wall, wall1 = per[used], per[1]
        rc, res, _, _ = ctx.results(n)
        capi.check(rc, "bmf_result")
        sampled = sum(r.sampled for r in res)
        nonempty = sum(1 for r in res if r.vertex_count)
        meshes = {j: ctx.mesh(j)[0] for j in range(n) if res[j].vertex_count}
        flops = sum(c.dim ** 3 * roofline.terrain_flops_per_sample(wl.sampler["octaves"])
                    + roofline.smoothing_flops(r.vertex_count, r.index_count // 3, c.iters)
                    for r, c in zip(res, wl.chunks) if r.sampled)
        nbytes = sum(roofline.mesh_bytes(r.vertex_count, r.index_count // 3) for r in res)
        capi.check(ctx.extract_raw(arr, n, capi.MODE_PLAIN), "bmf_extract plain")
        ctx.sync()
        t0 = time.perf_counter()
        capi.check(ctx.extract_raw(arr, n, capi.MODE_PLAIN), "bmf_extract plain")
        ctx.sync()
        plain = time.perf_counter() - t0
    log(f"[bench] uneven LOD: {n / wall:.1f} chunks/s batched ({1e3 * wall:.3f} ms/step over {used} context(s); one "
        f"context {n / wall1:.1f}), plain {n / plain:.1f} chunks/s, sampled {sampled}/{n}, non-empty {nonempty}/{n}, "
        f"tuned {tuned} in {t_tune:.2f} s")
    name = "uneven-lod-1536: 922/461/153 chunks of 32^3/64^3/128^3 cells, spacing 1/2/4, 7 octaves, K~U{0..8}"
    if only_class:
        name += f"; only the {n} chunks with ceil(dim/64) = {only_class}"
    return {"workload": name,
            "chunks_per_s": round(n / wall, 2), "ms_per_step": round(1e3 * wall, 4),
            "cells_per_s": round(wl.cells / wall, 1), "plain_chunks_per_s": round(n / plain, 2),
            "batched_over_plain": round(plain / wall, 2), "sampled_chunks": sampled, "nonempty_chunks": nonempty,
            "nonempty_fraction": round(nonempty / n, 4), "surface_chunks_per_s": round(sampled / wall, 2),
            "tuning": tuned, "autotune_s": round(t_tune, 3),
            **stream_fields(n, per, used, same),
            "roofline": roofline.frame_fractions(flops, nbytes, wall),
            "plain_roofline": roofline.frame_fractions(flops, nbytes, plain),
            "_check": {"workload": wl, "meshes": meshes}}


def world_leg(capi, roofline, workloads, steps, nctx=2):
    """The reference's default world: WorldOctree's sampler (terrain2d_pert_block with the
    world's noise_properties, WorldOctree.cpp:40-54) over the renderable leaves the LOD pass
    leaves around a focus on the surface of a 4096-unit world (max level 7), chunk_resolution
    32, K = 0 (WorldProperties' defaults), overlap 0.035 below the max level
    (ChunkGenerator.cpp:98), as one batch."""
    p = workloads.LodParams(max_level=7, min_level=1)
    focus = (0.0, -140.0, 0.0)
    nodes, leaves = workloads.lod_world(world_size=4096.0, build_focus=focus, params=p)
    chunks = [workloads.Chunk(tuple(float(v) for v in nodes[L]["pos"]), float(nodes[L]["size"]), 32,
                              0.035 if int(nodes[L]["level"]) < p.max_level else 0.0, 0, 0) for L in leaves]
    sd = capi.noise_preset(capi.PRESET_WORLD)
    wl = workloads.Workload("world-default", sd, chunks)
    cd = workloads.to_capi(chunks)
    n = len(cd)
    arr = capi.chunk_array(cd)
    def make():
        cx = capi.Context(max_chunks=1024, max_samples=n * 32 ** 3, max_vertices=1 << 22, max_indices=1 << 24)
        cx.set_sampler(**sd)
        return cx

    with make() as ctx:
        per, used, same = stream_times(capi, ctx, make, arr, n, steps, nctx)
        wall, wall1 = per[used], per[1]
        stages = ctx.profile_stages(cd)
        rc, res, _, _ = ctx.results(n)
        capi.check(rc, "bmf_result")
        sampled = sum(r.sampled for r in res)
        counts = [(r.vertex_count, r.index_count, r.contains_mesh) for r in res]
        meshes = {j: ctx.mesh(j) for j in range(n) if res[j].vertex_count}
        nbytes = sum(roofline.mesh_bytes(r.vertex_count, r.index_count // 3) for r in res)
        capi.check(ctx.extract_raw(arr, n, capi.MODE_PLAIN), "bmf_extract plain")
        ctx.sync()
        t0 = time.perf_counter()
        capi.check(ctx.extract_raw(arr, n, capi.MODE_PLAIN), "bmf_extract plain")
        ctx.sync()
        plain = time.perf_counter() - t0
    V = sum(c[0] for c in counts)
    log(f"[bench] default world: {n / wall:.1f} chunks/s batched ({1e3 * wall:.3f} ms/step over {used} context(s); "
        f"one context {n / wall1:.1f}), plain {n / plain:.1f}"
        f" chunks/s, sampled {sampled}/{n}, non-empty {len(meshes)}/{n}, V={V}, stages "
        + ", ".join(f"{k}={v:.3f}" for k, v in stages.items()))
    return {"workload": f"world-default-{n}: the reference world's terrain2d_pert sampler (own noise; 3 value-noise "
                        f"octaves, 13 gradient-perturbation octaves) on the {n} LOD leaves around {focus} of a "
                        f"4096-unit world, 32^3 samples per chunk, K=0",
            "chunks_per_s": round(n / wall, 2), "ms_per_step": round(1e3 * wall, 4),
            **stream_fields(n, per, used, same),
            "plain_chunks_per_s": round(n / plain, 2), "batched_over_plain": round(plain / wall, 2),
            "sampled_chunks": sampled, "nonempty_chunks": len(meshes), "vertices": V,
            "surface_chunks_per_s": round(sampled / wall, 2),
            "stage_ms": {k: round(v, 4) for k, v in stages.items()},
            "hbm_frac": round(nbytes / wall / (roofline.PEAK_HBM_GBS * 1e9), 6),
            "_check": {"workload": wl, "meshes": meshes, "counts": counts}}


def plan_scaling(capi, workloads, sizes=(512, 4096, 16384, 65536)):
    """Time of the batch prologue (the plan stage: geometry, culling, arena
    offsets) and of the whole frame for growing batches of small chunks
    (9^3 samples, spread over and off the terrain band, K = 1): one block up
    to 4,096 descriptors, three passes over many blocks beyond."""
    rng = np.random.default_rng(3)
    out = []
    with capi.Context(max_chunks=max(sizes), max_samples=max(sizes) * 9 ** 3, max_vertices=1 << 22,
                      max_indices=1 << 24) as ctx:
        ctx.set_sampler(**workloads.terrain_sampler())
        for n in sizes:
            cs = [workloads.Chunk((float(rng.uniform(-4000, 4000)), float(rng.uniform(-40, 100)),
                                   float(rng.uniform(-4000, 4000))), 8.0, 9, 0.0, 1, 0) for _ in range(n)]
            st = ctx.profile_stages(workloads.to_capi(cs))
            out.append({"chunks": n, "plan_ms": round(st["plan"], 4), "frame_ms": round(sum(st.values()), 4)})
    log("[bench] plan scaling: " + ", ".join(f"{r['chunks']}: plan {r['plan_ms']:.3f} / frame {r['frame_ms']:.3f} ms"
                                              for r in out))
    return out


def hermite_leg(capi, roofline, workloads, steps, nctx=2):
    """64 chunks of 64^3 cells tiled 4x4x4 over [0,256)^3 of the CSG union of
    200 spheres and 100 boxes (seed 42); dual-marching-cubes quads with FP32
    QEF placement (tolerance 0.1), batched."""
    wl = workloads.csg_scene()
    cd = workloads.to_capi(wl.chunks)
    n = len(cd)
    arr = capi.chunk_array(cd)
    def make():
        cx = capi.Context(max_chunks=n, max_samples=n * 65 ** 3, max_vertices=1 << 22, max_indices=1 << 24)
        cx.set_sampler(**wl.sampler)
        cx.set_mesh_options(mesh=capi.MESH_QUADS, placement=capi.PLACE_QEF, qef_tolerance=0.1)
        return cx

    with make() as ctx:
        per, used, same = stream_times(capi, ctx, make, arr, n, steps, nctx)
        wall, wall1 = per[used], per[1]
        rc, res, _, _ = ctx.results(n)
        capi.check(rc, "bmf_result")
        V = sum(r.vertex_count for r in res)
        Q = sum(r.index_count for r in res) // 4
        err = np.concatenate([ctx.qef_errors(j) for j in range(n) if res[j].vertex_count] or [np.zeros(1)])
        meshes = {j: ctx.mesh(j)[0] for j in range(n) if res[j].vertex_count}
        stages = ctx.profile_stages(cd)
        prims = len(wl.sampler["spheres"]) * roofline.CSG_SPHERE_FLOPS + \
            len(wl.sampler["boxes"]) * roofline.CSG_BOX_FLOPS
        flops = sum(c.dim ** 3 * prims for r, c in zip(res, wl.chunks) if r.sampled)
        nbytes = sum(roofline.mesh_bytes(r.vertex_count, r.index_count // 4 * 2) for r in res)
    log(f"[bench] hermite/QEF: {n / wall:.1f} chunks/s ({1e3 * wall:.3f} ms/step over {used} context(s); one context "
        f"{n / wall1:.1f}), V={V} quads={Q}, "
        f"stages {', '.join(f'{k}={v:.3f}' for k, v in stages.items())}")
    return {"workload": "csg-64x64: 4x4x4 chunks of 64^3 cells, 200 spheres + 100 boxes (seed 42), DMC quads, QEF tol 0.1",
            "chunks_per_s": round(n / wall, 2), "ms_per_step": round(1e3 * wall, 4), "vertices": V, "quads": Q,
            **stream_fields(n, per, used, same),
            "qef_residual": {"definition": "uncentred |ATb - ATA x|^2 of each vertex's QEF (qef_simd.h:434-440), "
                                           "not an error against FP64",
                             "median": float(np.median(err)), "max": float(err.max())},
            "stage_ms": {k: round(v, 4) for k, v in stages.items()},
            "roofline": roofline.frame_fractions(flops, nbytes, wall,
                                                 note="FLOP: every sample against all 200 spheres and 100 boxes "
                                                      "(the uncull union), counted from csrc/samplers.h"),
            "_check": {"workload": wl, "meshes": meshes}}


def single_leg(capi, workloads, steps, with_cpu):
    """One 64^3-cell chunk (origin 0, unit spacing, K=4) through the plain
    path, against one CPU thread of the oracle."""
    wl = workloads.single_chunk(iters=4)
    cd = workloads.to_capi(wl.chunks)
    arr = capi.chunk_array(cd)
    with capi.Context(max_chunks=1, max_samples=65 ** 3, max_vertices=1 << 20, max_indices=1 << 22) as ctx:
        ctx.set_sampler(**wl.sampler)
        for _ in range(3):
            capi.check(ctx.extract_raw(arr, 1, capi.MODE_PLAIN), "bmf_extract")
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            capi.check(ctx.extract_raw(arr, 1, capi.MODE_PLAIN), "bmf_extract")
            ctx.sync()
        gpu_ms = 1e3 * (time.perf_counter() - t0) / steps
        pos = ctx.mesh(0)[0]
        # the same chunk with vertex normals (the reference's smooth_normals branch)
        ctx.set_mesh_options(normals=True)
        capi.check(ctx.extract_raw(arr, 1, capi.MODE_PLAIN), "bmf_extract normals")
        capi.check(ctx.results(1)[0], "bmf_result")
        npos, nrm = ctx.mesh(0)[0], ctx.normals(0)
        ctx.set_mesh_options(normals=False)
    out = {"workload": "one 64^3-cell fbm terrain chunk, K=4, plain path", "gpu_ms": round(gpu_ms, 4),
           "_check": {"workload": wl, "pos": pos, "normals_pos": npos, "normals": nrm}}
    if with_cpu:
        from oracle import pyoracle

        spec = pyoracle.SamplerSpec(**wl.sampler)
        c = wl.chunks[0]
        oc = pyoracle.make_chunk(c.pos, c.size, c.dim, c.overlap, c.iters, c.boundary_processing)
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            pyoracle.extract_batch(spec, [oc], 1)
            times.append(time.perf_counter() - t0)
        out["cpu_ms_1_thread"] = round(1e3 * float(np.median(times)), 3)
    log(f"[bench] single chunk: GPU plain {gpu_ms:.3f} ms" +
        (f", CPU 1 thread {out['cpu_ms_1_thread']:.1f} ms" if with_cpu else ""))
    return out


def quads_leg(capi, roofline, workloads, steps, nctx=2):
    """The uniform 512-chunk batch in dual-marching-cubes quad mode: one
    vertex per EdgeTable polygon at the mass point of its edge points, one
    quad per sign-changing edge (Tables.hpp:76-354), then K = 4 smoothing
    steps as MeshProcessor<4> (MeshProcessor.cpp:130-306), batched."""
    wl = workloads.uniform_grid(n_side=8, c=64, iters=4)
    cd = workloads.to_capi(wl.chunks)
    n = len(cd)
    arr = capi.chunk_array(cd)
    def make():
        cx = capi.Context(max_chunks=n, max_samples=n * 65 ** 3, max_vertices=1 << 22, max_indices=1 << 24)
        cx.set_sampler(**wl.sampler)
        cx.set_mesh_options(mesh=capi.MESH_QUADS)
        return cx

    with make() as ctx:
        per, used, same = stream_times(capi, ctx, make, arr, n, steps, nctx)
        wall, wall1 = per[used], per[1]
        rc, res, _, _ = ctx.results(n)
        capi.check(rc, "bmf_result")
        sampled = sum(r.sampled for r in res)
        V = sum(r.vertex_count for r in res)
        Q = sum(r.index_count for r in res) // 4
        meshes = {j: ctx.mesh(j) for j in range(n) if res[j].vertex_count}
        counts = [(r.vertex_count, r.index_count, r.contains_mesh) for r in res]
        stages = ctx.profile_stages(cd)
        capi.check(ctx.extract_raw(arr, n, capi.MODE_PLAIN), "bmf_extract plain")
        ctx.sync()
        t0 = time.perf_counter()
        capi.check(ctx.extract_raw(arr, n, capi.MODE_PLAIN), "bmf_extract plain")
        ctx.sync()
        plain = time.perf_counter() - t0
    flops = sampled * 65 ** 3 * roofline.terrain_flops_per_sample(wl.sampler["octaves"]) + \
        roofline.smoothing_flops(V, 2 * Q, 4)
    nbytes = roofline.mesh_bytes(V, 2 * Q)
    log(f"[bench] quads: {n / wall:.1f} chunks/s ({1e3 * wall:.3f} ms/step over {used} context(s); one context "
        f"{n / wall1:.1f}), plain {n / plain:.1f} chunks/s, "
        f"V={V} quads={Q}, stages {', '.join(f'{k}={v:.3f}' for k, v in stages.items())}")
    return {"workload": "uniform-512x64-quads: 8x8x8 chunks of 64^3 cells, fbm terrain seed 1337, "
                        "DMC quads (mass point), K=4, batched",
            "chunks_per_s": round(n / wall, 2), "ms_per_step": round(1e3 * wall, 4),
            **stream_fields(n, per, used, same),
            "cells_per_s": round(n * 64 ** 3 / wall, 1), "surface_chunks_per_s": round(sampled / wall, 2),
            "plain_chunks_per_s": round(n / plain, 2), "batched_over_plain": round(plain / wall, 2),
            "sampled_chunks": sampled, "vertices": V, "quads": Q,
            "stage_ms": {k: round(v, 4) for k, v in stages.items()},
            "roofline": roofline.frame_fractions(flops, nbytes, wall),
            "_check": {"workload": wl, "meshes": meshes, "counts": counts}}


def diff_leg(capi, roofline, workloads, steps):
    """Differentiable path on one 128^3-cell terrain chunk (129^3 samples),
    8 smoothing iterations, checkpoints every 2, loss = mean squared distance
    to the nearest of 200,000 targets on a sphere (r 40, centre 64^3)."""
    dim, K = 129, 8
    with capi.Context(max_chunks=4, max_samples=dim ** 3, max_vertices=1 << 20, max_indices=1 << 22) as ctx:
        ctx.set_sampler(**workloads.terrain_sampler())
        with capi.Tape(ctx, max_dim=dim, max_vertices=1 << 20, max_indices=1 << 22, checkpoint_every=2) as tape:
            cd = capi.chunk_desc((0.0, 0.0, 0.0), float(dim - 1), dim, 0.0, K, 0)
            targets_np = workloads.sphere_targets()
            targets = capi.DeviceArray.from_numpy(targets_np)
            info = tape.forward(cd)
            # the targets are fixed: their search hierarchy is built once
            t0 = time.perf_counter()
            tape.set_targets(targets)
            t_build = time.perf_counter() - t0
            L, g, near = tape.loss_nearest_set()
            grad = capi.DeviceArray((dim, dim, dim), np.float32)
            capi.check(capi.lib().bmf_backward_diff(tape.h, g.ptr, grad.ptr), "bmf_backward_diff")
            t_f = t_l = t_b = 0.0
            dev = [0.0, 0.0, 0.0]
            for _ in range(steps):
                t0 = time.perf_counter()
                capi.check(capi.lib().bmf_forward_diff(ctx.h, capi.C.byref(cd), None, tape.h), "bmf_forward_diff")
                t1 = time.perf_counter()
                capi.check(capi.lib().bmf_tape_loss_nearest(tape.h, g.ptr, None, None), "bmf_tape_loss_nearest")
                t2 = time.perf_counter()
                capi.check(capi.lib().bmf_backward_diff(tape.h, g.ptr, grad.ptr), "bmf_backward_diff")
                t3 = time.perf_counter()
                t_f, t_l, t_b = t_f + t1 - t0, t_l + t2 - t1, t_b + t3 - t2
                dev = [a + b for a, b in zip(dev, tape.timing())]
            dev = [d / steps for d in dev]
            check = {"chunk": workloads.Chunk((0.0, 0.0, 0.0), float(dim - 1), dim, 0.0, K, 0),
                     "densities": tape.densities(), "targets": targets_np, "nearest": near,
                     "gradient": grad.numpy()}
            V = info.vertex_count
            T = info.index_count // 3
            n3 = dim ** 3
            per = (t_f + t_l + t_b) / steps
            # compulsory bytes of the backward: dL/dp in, the gradient out
            bwd_bytes = 12 * V + 4 * n3
            fwd_flops = n3 * roofline.terrain_flops_per_sample(5) + roofline.smoothing_flops(V, T, K)
            log(f"[bench] fwd+bwd: {1e3 * per:.3f} ms/step (forward {1e3 * t_f / steps:.3f}, loss {1e3 * t_l / steps:.3f},"
                f" backward {1e3 * t_b / steps:.3f}; device {dev[0]:.3f} / {dev[1]:.3f} / {dev[2]:.3f}), V={V}, "
                f"loss {L:.4f}")
            return {"workload": "one 128^3-cell fbm terrain chunk, K=8, 200k sphere targets",
                    "steps": steps, "ms_per_step": round(1e3 * per, 4), "steps_per_s": round(1.0 / per, 2),
                    "forward_ms": round(1e3 * t_f / steps, 4), "loss_ms": round(1e3 * t_l / steps, 4),
                    "backward_ms": round(1e3 * t_b / steps, 4), "vertices": V, "triangles": T, "loss": L,
                    "device_ms": {"forward": round(dev[0], 4), "loss": round(dev[1], 4), "backward": round(dev[2], 4)},
                    "forward_fp32_frac": round(fwd_flops / (dev[0] * 1e-3) / 1e12 / roofline.PEAK_FP32_TFLOPS, 4),
                    "backward_hbm_frac": round(bwd_bytes / (dev[2] * 1e-3) / 1e9 / roofline.PEAK_HBM_GBS, 5),
                    "targets_build_ms": round(1e3 * t_build, 4),
                    "timing": "ms_per_step and *_ms: host wall clock per call (each call ends with a stream "
                              "synchronisation); device_ms and the roofline fractions: HIP events on the library's "
                              "stream around each call's GPU work; the target hierarchy is built once "
                              "(targets_build_ms) and each step's loss queries it",
                    "_check": check}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every box here has one GPU, "
              "and the extraction path shards by chunk with no collective on the hot path)", file=sys.stderr)
        return 2
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)

    from binarymeshfitting_amd import capi, roofline, workloads

    wl = workloads.uniform_grid(n_side=8, c=64, iters=4)
    n = len(wl.chunks)
    dim = wl.chunks[0].dim
    cd = workloads.to_capi(wl.chunks)
    arr = capi.chunk_array(cd)
    def make():
        cx = capi.Context(max_chunks=n, max_samples=n * dim ** 3, max_vertices=1 << 22, max_indices=1 << 24)
        cx.set_sampler(**wl.sampler)
        return cx

    ctx = make()
    # launch shapes: the autotuner's pick on this batch (tile size, workgroups per kernel, row split)
    t0 = time.perf_counter()
    tuned = ctx.autotune(cd, reps=3) if args.autotune else ctx.tuning()
    log(f"[bench] tuning {tuned} ({'autotuned' if args.autotune else 'defaults'} in {time.perf_counter() - t0:.2f} s)")

    # ---- per-stage device times (HIP events on the library stream), one context,
    # before the stream timings -------------------------------------------------
    stages = None
    for _ in range(max(1, args.profile_reps)):
        st = ctx.profile_stages(cd)
        stages = {k: v + (stages or {}).get(k, 0.0) for k, v in st.items()}
    stages = {k: v / max(1, args.profile_reps) for k, v in stages.items()}
    log("[bench] stages (ms): " + ", ".join(f"{k}={v:.4f}" for k, v in stages.items()))
    # ---- batched path: warmup, then K timed steps ----------------------------
    # The stream of batches is timed on one context, then alternating over
    # `contexts` contexts (each its own arenas, stream and captured frame, the
    # tuned shapes of the first), so consecutive batches overlap on the GPU.
    nctx = max(1, args.contexts)
    log(f"[bench] {args.warmup} warmup and {args.steps} timed batched steps of {n} chunks, on one context and "
        f"alternating over 1 .. {nctx} contexts")
    per, used, ctx_identical = stream_times(capi, ctx, make, arr, n, args.steps, nctx, args.warmup)
    streams = stream_fields(n, per, used, ctx_identical)
    log("[bench] per context count: " + ", ".join(f"{k}: {n / w:.1f} chunks/s ({1e3 * w:.3f} ms/step)"
                                                  for k, w in per.items()))
    wall = per[used] * args.steps
    value = n * args.steps / wall
    rc, res, _, _ = ctx.results(n)
    capi.check(rc, "bmf_result")
    work = ctx.frame_stats()
    sampled = sum(r.sampled for r in res)
    V = sum(r.vertex_count for r in res)
    T = sum(r.index_count for r in res) // 3
    log(f"[bench] batched: {value:.1f} chunks/s ({1e3 * wall / args.steps:.3f} ms/step over {used} context(s), "
        f"meshes identical across contexts: {ctx_identical}), {sampled}/{n} chunks sampled, V={V} T={T}")

    octaves = wl.sampler["octaves"]
    samples = sampled * dim ** 3
    flops = samples * roofline.terrain_flops_per_sample(octaves)
    sample_s = stages["sample"] * 1e-3
    achieved_tflops = flops / sample_s / 1e12
    exec_flops = (work["octave_evals"] * (roofline.SIMPLEX_FLOPS + roofline.OCTAVE_FLOPS)
                  + work["samples"] * roofline.TERRAIN_FLOPS)
    traffic = None
    tfile = ROOT / "profiles" / "sample_kernel_uniform512x64_traffic.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text()).get("hbm_bytes_per_launch")
        except (ValueError, OSError):
            traffic = None
    roof = {
        "kernel": "k_sample",
        "bound": "valu",
        "achieved": round(achieved_tflops, 3),
        "peak": roofline.PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved_tflops / roofline.PEAK_FP32_TFLOPS, 4),
        "frac_of_unpacked_peak": round(achieved_tflops / roofline.PEAK_FP32_UNPACKED_TFLOPS, 4),
        "traffic": traffic,
        "algorithmic_flops_per_launch": flops,
        "algorithmic_bytes_per_launch": int(samples * roofline.LABEL_BYTES),
        "avg_launch_ms": round(stages["sample"], 4),
        # what the kernel executed: the adaptive evaluation stops a sample's octaves once its
        # sign is decided (octaves counted by the kernel itself), against the FP32 peak
        # without packed math (the kernels are built without v_pk_* FP32 instructions)
        "executed": {"octave_evals": work["octave_evals"], "samples": work["samples"],
                     "octaves_per_sample": round(work["octave_evals"] / max(1, work["samples"]), 4),
                     "flops_per_launch": exec_flops, "tflops": round(exec_flops / sample_s / 1e12, 3),
                     "frac_of_unpacked_peak": round(exec_flops / sample_s / 1e12
                                                    / roofline.PEAK_FP32_UNPACKED_TFLOPS, 4)},
        "hbm_frac": round((traffic or samples * roofline.LABEL_BYTES) / sample_s / 1e9 / roofline.PEAK_HBM_GBS, 5),
    }
    # whole frame: full-octave sampling + smoothing FLOPs and the compulsory
    # mesh bytes over the step time
    frame_s = wall / args.steps
    path_flops = flops + sum(roofline.smoothing_flops(r.vertex_count, r.index_count // 3, c.iters)
                              for r, c in zip(res, wl.chunks))
    path_bytes = sum(roofline.mesh_bytes(r.vertex_count, r.index_count // 3) for r in res)
    path = {
        "flops_per_step": path_flops,
        "bytes_per_step": path_bytes,
        "tflops": round(path_flops / frame_s / 1e12, 3),
        "fp32_frac": round(path_flops / frame_s / 1e12 / roofline.PEAK_FP32_TFLOPS, 4),
        "hbm_gbs": round(path_bytes / frame_s / 1e9, 2),
        "hbm_frac": round(path_bytes / frame_s / 1e9 / roofline.PEAK_HBM_GBS, 5),
    }

    # ---- plain path ---------------------------------------------------------------
    plain = None
    plain_roof = None
    if not args.no_plain and args.plain_steps > 0:
        capi.check(ctx.extract_raw(arr, n, capi.MODE_PLAIN), "bmf_extract plain")
        ctx.sync()
        log(f"[bench] timing {args.plain_steps} plain steps")
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            capi.check(ctx.extract_raw(arr, n, capi.MODE_PLAIN), "bmf_extract plain")
        ctx.sync()
        plain = n * args.plain_steps / (time.perf_counter() - t0)
        log(f"[bench] plain: {plain:.1f} chunks/s")
        plain_roof = roofline.frame_fractions(path_flops, path_bytes, n / plain)

    # ---- uneven LOD: 1,536 chunks of 32/64/128 cells, autotuned shapes ------------------
    lod = None
    if args.lod_steps > 0:
        lod = lod_leg(capi, roofline, workloads, args.lod_steps, args.lod_class, max(1, args.contexts))

    # ---- the batch prologue at growing batch sizes ---------------------------------------
    scaling = plan_scaling(capi, workloads) if args.lod_steps > 0 else None

    # ---- the reference's default world sampler over its LOD leaves ----------------------
    world = None
    if args.world_steps > 0:
        world = world_leg(capi, roofline, workloads, args.world_steps, max(1, args.contexts))

    # ---- Hermite / QEF: 64 CSG chunks, dual quads placed by FP32 QEF ---------------------
    hermite = None
    if args.hermite_steps > 0:
        hermite = hermite_leg(capi, roofline, workloads, args.hermite_steps, max(1, args.contexts))

    # ---- config 2 as worded: the same batch in dual-marching-cubes quad mode --------------
    quads = None
    if args.quads_steps > 0:
        quads = quads_leg(capi, roofline, workloads, args.quads_steps, max(1, args.contexts))

    # ---- single chunk, plain path: latency against one CPU thread ----------------------
    single = None
    if args.single_steps > 0:
        single = single_leg(capi, workloads, args.single_steps, not args.no_cpu)

    # ---- forward + loss + backward: one 128^3-cell chunk, K=8, 200k targets ---------
    diff = None
    if args.diff_steps > 0:
        diff = diff_leg(capi, roofline, workloads, args.diff_steps)

    # ---- CPU baseline + parity -----------------------------------------------------
    cpu = None
    parity = None
    fp64 = None
    if not args.no_cpu:
        from oracle import pyoracle

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        spec = pyoracle.SamplerSpec(**wl.sampler)
        ochunks = [pyoracle.make_chunk(c.pos, c.size, c.dim, c.overlap, c.iters, c.boundary_processing)
                   for c in wl.chunks]
        runs = max(1, args.cpu_runs)
        log(f"[bench] CPU oracle on all {n} chunks with {threads} threads: 1 warm-up + {runs} timed passes")
        oV, onI, ocm, ock = pyoracle.extract_batch(spec, ochunks, threads)
        times = []
        for _ in range(runs):
            t0 = time.perf_counter()
            pyoracle.extract_batch(spec, ochunks, threads)
            times.append(time.perf_counter() - t0)
        cpu_s = float(np.median(times))
        # like for like: the oracle on the chunks the GPU samples only (the others are classified
        # by the plan's |fbm| <= 1 bound with no sampling, which a CPU loop can apply as cheaply)
        rc0, res0, _, _ = ctx.results(n)
        sub = [oc for oc, r in zip(ochunks, res0) if r.sampled]
        sub_times = []
        for _ in range(runs):
            t0 = time.perf_counter()
            pyoracle.extract_batch(spec, sub, threads)
            sub_times.append(time.perf_counter() - t0)
        sub_s = float(np.median(sub_times))
        cpu = {"value": round(n / cpu_s, 3), "unit": "chunks/s", "cores": threads, "kind": "port",
               "sample": f"all {n} chunks of the workload; median of {runs} passes after a warm-up "
                         f"({cpu_s:.2f} s wall per pass)",
               "culled_value": round(n / sub_s, 3),
               "culled_sample": f"the same batch with the GPU's analytic culling applied: the oracle on the "
                                f"{len(sub)} chunks the GPU samples, the other {n - len(sub)} taken as classified "
                                f"({sub_s:.3f} s wall per pass, median of {runs}); every octave is still evaluated "
                                f"(the GPU stops early where the sign is decided)",
               "host": host_cpu()}
        log(f"[bench] CPU: {n / cpu_s:.2f} chunks/s; with the same culling {n / sub_s:.2f} chunks/s")
        # the batched path's meshes vs the CPU run
        capi.check(ctx.extract_raw(arr, n, capi.MODE_BATCHED), "bmf_extract")
        rc, res, _, _ = ctx.results(n)
        bad = 0
        for j in range(n):
            r = res[j]
            if r.vertex_count != oV[j] or r.index_count != onI[j] or bool(r.contains_mesh) != bool(ocm[j]):
                bad += 1
                continue
            if r.vertex_count:
                pos, idx, _ = ctx.mesh(j)
                if pyoracle.checksum_mesh(pos, idx) != int(ock[j]):
                    bad += 1
        parity = {"chunks": n, "mismatched": bad, "check": "bit-exact positions and triangle lists vs CPU"}
        log(f"[bench] parity: {bad} of {n} chunks differ")
        # ---- FP32-vs-FP64 error of every leg's results (the oracle's FP64 twins) ---------
        from oracle import fp64_errors as fe

        t0 = time.perf_counter()
        js = [j for j in range(n) if res[j].vertex_count]
        gpos = {j: ctx.mesh(j)[0] for j in js}
        e = fe.merge(fe.over_chunks(lambda j: fe.triangle_positions(spec, wl.chunks[j], gpos[j]), js, threads))
        fp64 = {"positions": fe.stats(e, dim - 1)}
        if lod:
            ck = lod["_check"]
            lw = ck["workload"]
            parts = fe.over_chunks(lambda j: (fe.triangle_positions(pyoracle.SamplerSpec(**lw.sampler), lw.chunks[j],
                                                                    ck["meshes"][j]) / (lw.chunks[j].dim - 1)),
                                   sorted(ck["meshes"]), threads)
            lod["fp64_error"] = {"positions_over_extent": fe.stats(fe.merge(parts))}
        if world:
            ck = world["_check"]
            ww = ck["workload"]
            wspec = pyoracle.SamplerSpec(**ww.sampler)

            def world_ref(j):
                c = ww.chunks[j]
                return pyoracle.extract_chunk(wspec, pyoracle.make_chunk(c.pos, c.size, c.dim, c.overlap, c.iters,
                                                                         c.boundary_processing))

            wbad = 0
            for j, m in enumerate(fe.over_chunks(world_ref, range(len(ww.chunks)), threads)):
                V_, nI_, cm_ = ck["counts"][j]
                if V_ != m["V"] or nI_ != m["nI"] or bool(cm_) != bool(m["contains_mesh"]):
                    wbad += 1
                elif V_:
                    wp, wi, wb = ck["meshes"][j]
                    if not (np.array_equal(wp, m["pos"]) and np.array_equal(wi, m["idx"])
                            and np.array_equal(wb, m["boundary"])):
                        wbad += 1
            world["parity"] = {"chunks": len(ww.chunks), "mismatched": wbad,
                               "check": "bit-exact positions, triangle lists and boundary flags vs CPU"}
            parts = fe.over_chunks(lambda j: (fe.triangle_positions(wspec, ww.chunks[j], ck["meshes"][j][0])
                                              / (ww.chunks[j].dim - 1)), sorted(ck["meshes"]), threads)
            world["fp64_error"] = {"positions_over_extent": fe.stats(fe.merge(parts))}
            log(f"[bench] default world parity: {wbad} of {len(ww.chunks)} chunks differ")
        if hermite:
            ck = hermite["_check"]
            hw = ck["workload"]
            hspec = pyoracle.SamplerSpec(**hw.sampler)
            parts = fe.over_chunks(lambda j: fe.quad_positions(hspec, hw.chunks[j], ck["meshes"][j],
                                                               placement=pyoracle.PLACE_QEF, tol=0.1),
                                   sorted(ck["meshes"]), threads)
            hermite["fp64_error"] = {"qef_positions": fe.stats(fe.merge(parts), 64)}
        if quads:
            ck = quads["_check"]
            qw = ck["workload"]
            qspec = pyoracle.SamplerSpec(**qw.sampler)

            def quad_ref(j):
                c = qw.chunks[j]
                oc = pyoracle.make_chunk(c.pos, c.size, c.dim, c.overlap, c.iters, c.boundary_processing)
                delta, opos = pyoracle.geometry(oc)
                m = pyoracle.dmc_mesh(pyoracle.sample_block(qspec, c.dim, opos, delta), qspec, opos, delta)
                if c.iters > 0 and m["V"] and m["nI"]:
                    m["pos"] = pyoracle.smooth_n(m["pos"], m["boundary"], m["idx"], 4, c.iters, c.boundary_processing)
                return m

            qbad = 0
            for j, m in enumerate(fe.over_chunks(quad_ref, range(n), threads)):
                V_, nI_, _ = ck["counts"][j]
                if V_ != m["V"] or nI_ != m["nI"]:
                    qbad += 1
                elif V_:
                    qp, qi, qb = ck["meshes"][j]
                    if not (np.array_equal(qp, m["pos"]) and np.array_equal(qi, m["idx"])
                            and np.array_equal(qb, m["boundary"])):
                        qbad += 1
            quads["parity"] = {"chunks": n, "mismatched": qbad,
                               "check": "bit-exact positions, quad lists and boundary flags vs CPU"}
            parts = fe.over_chunks(lambda j: fe.quad_positions(qspec, qw.chunks[j], ck["meshes"][j][0]),
                                   sorted(ck["meshes"]), threads)
            quads["fp64_error"] = {"positions": fe.stats(fe.merge(parts), 64)}
            log(f"[bench] quads parity: {qbad} of {n} chunks differ")
        if single:
            ck = single["_check"]
            c0 = ck["workload"].chunks[0]
            ep = fe.triangle_positions(spec, c0, ck["pos"])
            epn, en = fe.triangle_normals(spec, c0, ck["normals_pos"], ck["normals"])
            single["fp64_error"] = {"positions": fe.stats(ep, c0.dim - 1), "normals": fe.stats(en),
                                    "positions_with_normals": fe.stats(epn, c0.dim - 1)}
        if diff:
            ck = diff["_check"]
            dc = ck["chunk"]
            dmesh = pyoracle.extract_chunk(spec, pyoracle.make_chunk(dc.pos, dc.size, dc.dim, dc.overlap, dc.iters,
                                                                      dc.boundary_processing))
            diff["fp64_error"] = fe.gradient(ck["densities"], dmesh, dc.iters, dc.boundary_processing,
                                             ck["targets"], ck["nearest"], ck["gradient"], threads=threads)
        log(f"[bench] FP64 error statements in {time.perf_counter() - t0:.1f} s")

    ms_per_step = 1e3 * wall / args.steps
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "chunks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "uniform-512x64: 8x8x8 chunks of 64^3 cells, fbm terrain seed 1337, K=4, batched",
                   "chunks": n, "cells_per_chunk": (dim - 1) ** 3, "smoothing_iters": 4, "octaves": octaves,
                   "sampled_chunks": sampled, "vertices": V, "triangles": T, "tuning": tuned,
                   "pipeline": {"contexts": used, "note": "the timed batches alternate over this many contexts "
                                "(own arenas, stream and captured frame each; the same tuned shapes), so "
                                "consecutive batches overlap on the GPU; ms_per_step is the time per batch "
                                "of the stream; the fastest of 1 .. --contexts contexts",
                                "meshes_identical_across_contexts": ctx_identical}},
        "one_context": streams["one_context"],
        "stream_over_contexts": streams["stream_over_contexts"],
        "cells_per_s": round(value * (dim - 1) ** 3, 1),
        "surface_chunks_per_s": round(sampled * args.steps / wall, 2),
        "plain_chunks_per_s": round(plain, 2) if plain else None,
        "plain_roofline": plain_roof,
        "batched_over_plain": round(value / plain, 2) if plain else None,
        "stage_ms": {k: round(v, 4) for k, v in stages.items()},
        "roofline": roof,
        "path_roofline": path,
        "fp64_error": fp64,
        "uneven_lod": lod,
        "default_world": world,
        "plan_scaling": scaling,
        "hermite_qef": hermite,
        "dmc_quads": quads,
        "single_chunk": single,
        "fwd_bwd": diff,
        "cpu_baseline": cpu,
        "parity": parity,
    }
    for leg in (lod, world, hermite, quads, single, diff):
        if leg:
            leg.pop("_check", None)
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
