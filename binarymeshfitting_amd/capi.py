"""ctypes binding of libbmf.so (include/bmf.h).

This is the Python face of the C ABI used by the tests, the smoke check and
bench.py.  It loads only the in-tree HIP library and raises if it is missing:
there is no CPU fallback.
"""
import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libbmf.so"

OK, EINVAL, ECAPACITY, EHIP, ENODEV, ESTATE = range(6)
SAMPLER_TERRAIN, SAMPLER_SPHERE, SAMPLER_CUBOID, SAMPLER_TORUS_Z, SAMPLER_PLANE_Y, SAMPLER_CSG, SAMPLER_EXTERNAL = range(7)
SAMPLER_NOISE = 7
NOISE_VALUE, NOISE_SIMPLEX = 0, 1
FRACTAL_FBM, FRACTAL_RIGID = 0, 1
PERTURB_NONE, PERTURB_GRADIENT, PERTURB_OFFSET_XZ = 0, 1, 2
MAX_PERTURB_OCTAVES = 24
PRESET_TERRAIN2D, PRESET_TERRAIN2D_PERT, PRESET_TERRAIN3D, PRESET_TERRAIN3D_PERT, PRESET_WINDY3D, PRESET_WORLD = range(6)
NOISE_FIELDS = ("noise_dims", "noise_basis", "noise_fractal", "g_scale", "y_scale", "height", "perturb",
                "perturb_octaves", "perturb_amp", "perturb_freq", "perturb_gain")
MODE_PLAIN, MODE_BATCHED = 0, 1
MAX_DIM = 256

EXPORTS = [
    "bmf_create", "bmf_destroy", "bmf_strerror", "bmf_version", "bmf_set_sampler", "bmf_validate_chunks",
    "bmf_extract", "bmf_result", "bmf_mesh_view_get", "bmf_mesh_to_host", "bmf_sync", "bmf_sample_block",
    "bmf_label_block", "bmf_cell_masks", "bmf_last_timing", "bmf_profile_stages", "bmf_tape_create",
    "bmf_tape_destroy", "bmf_forward_diff", "bmf_tape_get", "bmf_tape_positions", "bmf_loss_nearest",
    "bmf_backward_diff", "bmf_get_tuning", "bmf_set_tuning", "bmf_autotune", "bmf_set_mesh_options",
    "bmf_get_mesh_options", "bmf_qef_error_to_host", "bmf_normals_to_host", "bmf_tape_set_targets",
    "bmf_tape_loss_nearest", "bmf_loss_nearest_noalloc", "bmf_format", "bmf_lod_queue", "bmf_extract_queued", "bmf_stitch",
    "bmf_tape_timing", "bmf_noise_preset", "bmf_get_frame_stats",
]
SPACE_GRID, SPACE_WORLD = 0, 1
COLOR_NONE, COLOR_WHITE, COLOR_NOISE = 0, 1, 2
NOALLOC_MAX_TARGETS = 262144
MESH_TRIANGLES, MESH_QUADS = 0, 1
PLACE_MASS_POINT, PLACE_QEF = 0, 1
NUM_STAGES = 7
STAGES = ["plan", "sample", "count", "scan", "emit", "smooth", "finalize"]


class Limits(C.Structure):
    _fields_ = [("max_chunks", C.c_int32), ("max_samples", C.c_int64), ("max_vertices", C.c_int64),
                ("max_indices", C.c_int64)]


class SamplerDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("world_size", C.c_float), ("amp", C.c_float), ("y0", C.c_float),
        ("freq", C.c_float), ("lacunarity", C.c_float), ("gain", C.c_float), ("octaves", C.c_int32),
        ("seed", C.c_uint32), ("spheres", C.POINTER(C.c_float)), ("n_spheres", C.c_int32),
        ("boxes", C.POINTER(C.c_float)), ("n_boxes", C.c_int32), ("densities", C.c_void_p),
        ("noise_dims", C.c_int32), ("noise_basis", C.c_int32), ("noise_fractal", C.c_int32),
        ("g_scale", C.c_float), ("y_scale", C.c_float), ("height", C.c_float), ("perturb", C.c_int32),
        ("perturb_octaves", C.c_int32), ("perturb_amp", C.c_float), ("perturb_freq", C.c_float),
        ("perturb_gain", C.c_float),
    ]


def noise_preset(which, world_size=256.0, seed=1337):
    """bmf_noise_preset as a dict of set_sampler keyword arguments (the
    reference's block sampler `which` stands for, with the project's noise)."""
    d = SamplerDesc()
    check(lib().bmf_noise_preset(which, world_size, seed & 0xFFFFFFFF, C.byref(d)), "bmf_noise_preset")
    keys = ("kind", "world_size", "freq", "lacunarity", "gain", "octaves", "seed") + NOISE_FIELDS
    return {k: getattr(d, k) for k in keys}


class ChunkDesc(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("size", C.c_float), ("dim", C.c_int32), ("overlap", C.c_float),
                ("iters", C.c_int32), ("boundary_processing", C.c_int32)]


class ChunkResult(C.Structure):
    _fields_ = [("vertex_offset", C.c_int64), ("index_offset", C.c_int64), ("vertex_count", C.c_int32),
                ("index_count", C.c_int32), ("contains_mesh", C.c_int32), ("status", C.c_int32),
                ("sampled", C.c_int32), ("delta", C.c_float), ("overlap_pos", C.c_float * 3)]


class MeshView(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("indices", C.c_void_p), ("boundary", C.c_void_p),
                ("normals", C.c_void_p), ("total_vertices", C.c_int64), ("total_indices", C.c_int64)]


class MeshOptions(C.Structure):
    _fields_ = [("mesh", C.c_int32), ("placement", C.c_int32), ("qef_tolerance", C.c_float),
                ("quads_as_triangles", C.c_int32), ("normals", C.c_int32)]


class Tuning(C.Structure):
    _fields_ = [("sample_tile", C.c_int32), ("sample_blocks", C.c_int32), ("row_blocks", C.c_int32),
                ("smooth_blocks", C.c_int32), ("row_split", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TapeInfo(C.Structure):
    _fields_ = [("dim", C.c_int32), ("iters", C.c_int32), ("checkpoint_every", C.c_int32),
                ("vertex_count", C.c_int32), ("index_count", C.c_int32), ("contains_mesh", C.c_int32),
                ("positions", C.c_void_p), ("indices", C.c_void_p), ("boundary", C.c_void_p),
                ("densities", C.c_void_p)]


class LodParams(C.Structure):
    _fields_ = [("focus", C.c_float * 3), ("split_multiplier", C.c_float), ("group_multiplier", C.c_float),
                ("size_modifier", C.c_float), ("max_level", C.c_int32), ("min_level", C.c_int32),
                ("max_gen", C.c_int32), ("chunk_resolution", C.c_int32), ("process_iters", C.c_int32),
                ("boundary_processing", C.c_int32), ("overlap", C.c_float)]

    @classmethod
    def of(cls, p):
        """From a workloads.LodParams."""
        return cls((C.c_float * 3)(*p.focus), p.split_multiplier, p.group_multiplier, p.size_modifier, p.max_level,
                   p.min_level, p.max_gen, p.chunk_resolution, p.process_iters, p.boundary_processing, p.overlap)


class FormatOptions(C.Structure):
    _fields_ = [("space", C.c_int32), ("colors", C.c_int32), ("global_indices", C.c_int32)]


class Vbo(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("colors", C.c_void_p), ("normals", C.c_void_p), ("indices", C.c_void_p),
                ("cap_vertices", C.c_int64), ("cap_indices", C.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"libbmf.so not built at {LIB_PATH}; run __graft_entry__.build()")
        L = C.CDLL(str(LIB_PATH))
        P = C.POINTER
        ctxp = C.c_void_p
        L.bmf_create.argtypes = [P(ctxp), C.c_int, P(Limits)]
        L.bmf_destroy.argtypes = [ctxp]
        L.bmf_destroy.restype = None
        L.bmf_strerror.argtypes = [C.c_int]
        L.bmf_strerror.restype = C.c_char_p
        L.bmf_version.restype = C.c_char_p
        L.bmf_set_sampler.argtypes = [ctxp, P(SamplerDesc)]
        L.bmf_validate_chunks.argtypes = [P(ChunkDesc), C.c_int, P(Limits)]
        L.bmf_extract.argtypes = [ctxp, P(ChunkDesc), C.c_int, C.c_int]
        L.bmf_result.argtypes = [ctxp, P(ChunkResult), C.c_int, P(C.c_int64), P(C.c_int64)]
        L.bmf_mesh_view_get.argtypes = [ctxp, P(MeshView)]
        L.bmf_mesh_to_host.argtypes = [ctxp, C.c_int, P(C.c_float), P(C.c_uint32), P(C.c_uint8), C.c_int64,
                                       C.c_int64, P(C.c_int64), P(C.c_int64)]
        L.bmf_sync.argtypes = [ctxp]
        L.bmf_sample_block.argtypes = [ctxp, P(ChunkDesc), P(C.c_float)]
        L.bmf_label_block.argtypes = [ctxp, P(ChunkDesc), P(C.c_uint32), P(C.c_int32)]
        L.bmf_cell_masks.argtypes = [ctxp, P(ChunkDesc), P(C.c_uint64)]
        L.bmf_last_timing.argtypes = [ctxp, P(C.c_float)]
        L.bmf_profile_stages.argtypes = [ctxp, P(ChunkDesc), C.c_int, P(C.c_float)]
        L.bmf_set_mesh_options.argtypes = [ctxp, P(MeshOptions)]
        L.bmf_get_mesh_options.argtypes = [ctxp, P(MeshOptions)]
        L.bmf_qef_error_to_host.argtypes = [ctxp, C.c_int, P(C.c_float), C.c_int64]
        L.bmf_normals_to_host.argtypes = [ctxp, C.c_int, P(C.c_float), C.c_int64]
        L.bmf_get_tuning.argtypes = [ctxp, P(Tuning)]
        L.bmf_set_tuning.argtypes = [ctxp, P(Tuning)]
        L.bmf_autotune.argtypes = [ctxp, P(ChunkDesc), C.c_int, C.c_int, P(Tuning)]
        tp = C.c_void_p
        L.bmf_tape_create.argtypes = [P(tp), ctxp, C.c_int32, C.c_int64, C.c_int64, C.c_int32]
        L.bmf_tape_destroy.argtypes = [tp]
        L.bmf_tape_destroy.restype = None
        L.bmf_forward_diff.argtypes = [ctxp, P(ChunkDesc), C.c_void_p, tp]
        L.bmf_tape_get.argtypes = [tp, P(TapeInfo)]
        L.bmf_tape_positions.argtypes = [tp, C.c_int32, C.c_void_p]
        L.bmf_loss_nearest.argtypes = [tp, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, P(C.c_float)]
        L.bmf_backward_diff.argtypes = [tp, C.c_void_p, C.c_void_p]
        L.bmf_tape_timing.argtypes = [tp, P(C.c_float)]
        L.bmf_noise_preset.argtypes = [C.c_int32, C.c_float, C.c_uint32, P(SamplerDesc)]
        L.bmf_get_frame_stats.argtypes = [ctxp, P(C.c_int64 * 3)]
        L.bmf_tape_set_targets.argtypes = [tp, C.c_void_p, C.c_int64]
        L.bmf_tape_loss_nearest.argtypes = [tp, C.c_void_p, C.c_void_p, P(C.c_float)]
        L.bmf_loss_nearest_noalloc.argtypes = [tp, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
        L.bmf_format.argtypes = [ctxp, P(FormatOptions), P(Vbo), P(C.c_int64), P(C.c_int64)]
        L.bmf_lod_queue.argtypes = [ctxp, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, P(LodParams), C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_int32, P(C.c_int32), P(C.c_int32)]
        L.bmf_extract_queued.argtypes = [ctxp, C.c_int]
        L.bmf_stitch.argtypes = [ctxp, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, P(C.c_int32), C.c_void_p,
                                 C.c_int64, P(C.c_int64)]
        _lib = L
    return _lib


class BmfError(RuntimeError):
    def __init__(self, code, what):
        self.code = code
        super().__init__(f"{what}: {lib().bmf_strerror(code).decode()} ({code})")


def check(code, what):
    if code != OK:
        raise BmfError(code, what)


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def chunk_desc(pos, size, dim, overlap=0.0, iters=0, boundary_processing=0):
    return ChunkDesc((C.c_float * 3)(*pos), size, dim, overlap, iters, boundary_processing)


def chunk_array(chunks):
    return (ChunkDesc * len(chunks))(*chunks)


class Context:
    """A bmf_ctx with its arenas; one per host thread."""

    def __init__(self, max_chunks=1024, max_samples=64 * 65 ** 3, max_vertices=1 << 22, max_indices=1 << 24,
                 device=0):
        self.limits = Limits(max_chunks, max_samples, max_vertices, max_indices)
        h = C.c_void_p()
        check(lib().bmf_create(C.byref(h), device, C.byref(self.limits)), "bmf_create")
        self.h = h
        self._keep = []

    def close(self):
        if self.h:
            lib().bmf_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_sampler(self, kind=SAMPLER_TERRAIN, world_size=256.0, amp=24.0, y0=32.0, freq=0.02, lacunarity=2.0,
                    gain=0.5, octaves=5, seed=1337, spheres=None, boxes=None, densities=None, **noise):
        """densities: a DeviceArray of packed dim^3 blocks for SAMPLER_EXTERNAL;
        noise: the NOISE sampler's fields (NOISE_FIELDS, e.g. from noise_preset)."""
        sp = np.ascontiguousarray(spheres if spheres is not None else np.zeros((0, 4)), np.float32)
        bx = np.ascontiguousarray(boxes if boxes is not None else np.zeros((0, 6)), np.float32)
        unknown = set(noise) - set(NOISE_FIELDS)
        if unknown:
            raise TypeError(f"unknown sampler fields {sorted(unknown)}")
        d = SamplerDesc(kind, world_size, amp, y0, freq, lacunarity, gain, octaves, seed & 0xFFFFFFFF,
                        _ptr(sp, C.c_float), len(sp), _ptr(bx, C.c_float), len(bx),
                        densities.ptr if densities is not None else None)
        for k, v in noise.items():
            setattr(d, k, v)
        self._sampler_keep = densities
        check(lib().bmf_set_sampler(self.h, C.byref(d)), "bmf_set_sampler")

    def extract(self, chunks, mode=MODE_BATCHED):
        arr = chunk_array(chunks)
        self._keep = [arr]
        check(lib().bmf_extract(self.h, arr, len(chunks), mode), "bmf_extract")
        return len(chunks)

    def extract_raw(self, arr, n, mode=MODE_BATCHED):
        return lib().bmf_extract(self.h, arr, n, mode)

    def sync(self):
        check(lib().bmf_sync(self.h), "bmf_sync")

    def results(self, n):
        out = (ChunkResult * n)()
        nv, ni = C.c_int64(), C.c_int64()
        rc = lib().bmf_result(self.h, out, n, C.byref(nv), C.byref(ni))
        return rc, list(out), nv.value, ni.value

    def mesh(self, chunk, cap_v=None, cap_i=None):
        nv, ni = C.c_int64(), C.c_int64()
        rc = lib().bmf_mesh_to_host(self.h, chunk, None, None, None, 0, 0, C.byref(nv), C.byref(ni))
        if rc not in (OK, ECAPACITY):
            check(rc, "bmf_mesh_to_host")
        V, I = nv.value, ni.value
        pos = np.zeros((V, 3), np.float32)
        idx = np.zeros(I, np.uint32)
        bnd = np.zeros(V, np.uint8)
        check(lib().bmf_mesh_to_host(self.h, chunk, _ptr(pos, C.c_float), _ptr(idx, C.c_uint32),
                                     _ptr(bnd, C.c_uint8), V, I, C.byref(nv), C.byref(ni)), "bmf_mesh_to_host")
        return pos, idx, bnd

    def sample_block(self, chunk):
        out = np.zeros(chunk.dim ** 3, np.float32)
        check(lib().bmf_sample_block(self.h, C.byref(chunk), _ptr(out, C.c_float)), "bmf_sample_block")
        return out.reshape(chunk.dim, chunk.dim, chunk.dim)

    def label_block(self, chunk):
        dim = chunk.dim
        out = np.zeros(dim * dim * ((dim + 31) // 32), np.uint32)
        cm = C.c_int32()
        check(lib().bmf_label_block(self.h, C.byref(chunk), _ptr(out, C.c_uint32), C.byref(cm)), "bmf_label_block")
        return out.reshape(dim, dim, (dim + 31) // 32), bool(cm.value)

    def cell_masks(self, chunk):
        dim = chunk.dim
        out = np.zeros(dim * dim * ((dim + 7) // 8), np.uint64)
        check(lib().bmf_cell_masks(self.h, C.byref(chunk), _ptr(out, C.c_uint64)), "bmf_cell_masks")
        return out.reshape(dim, dim, (dim + 7) // 8)

    def last_timing(self):
        t = C.c_float()
        check(lib().bmf_last_timing(self.h, C.byref(t)), "bmf_last_timing")
        return t.value

    def set_mesh_options(self, mesh=MESH_TRIANGLES, placement=PLACE_MASS_POINT, qef_tolerance=0.1,
                         quads_as_triangles=False, normals=False):
        o = MeshOptions(mesh, placement, qef_tolerance, int(quads_as_triangles), int(normals))
        check(lib().bmf_set_mesh_options(self.h, C.byref(o)), "bmf_set_mesh_options")

    def mesh_options(self):
        o = MeshOptions()
        check(lib().bmf_get_mesh_options(self.h, C.byref(o)), "bmf_get_mesh_options")
        return {k: getattr(o, k) for k, _ in o._fields_}

    def normals(self, chunk):
        rc, res, _, _ = self.results(chunk + 1)
        V = res[chunk].vertex_count
        out = np.zeros((V, 3), np.float32)
        check(lib().bmf_normals_to_host(self.h, chunk, _ptr(out, C.c_float), V), "bmf_normals_to_host")
        return out

    def format(self, space=SPACE_GRID, colors=COLOR_WHITE, global_indices=True, normals=False, cap_vertices=None,
               cap_indices=None):
        """bmf_format into fresh device arrays; returns (rc, positions, colors, normals, indices) as numpy
        (None for arrays not asked for), cut to the formatted totals."""
        view = MeshView()
        check(lib().bmf_mesh_view_get(self.h, C.byref(view)), "bmf_mesh_view_get")
        cv = view.total_vertices if cap_vertices is None else cap_vertices
        ci = view.total_indices if cap_indices is None else cap_indices
        pos = DeviceArray((max(cv, 1), 3), np.float32)
        col = DeviceArray((max(cv, 1), 3), np.float32) if colors != COLOR_NONE else None
        nrm = DeviceArray((max(cv, 1), 3), np.float32) if normals else None
        idx = DeviceArray((max(ci, 1),), np.uint32)
        o = FormatOptions(space, colors, int(global_indices))
        vb = Vbo(pos.ptr, col.ptr if col else None, nrm.ptr if nrm else None, idx.ptr, cv, ci)
        nv, ni = C.c_int64(), C.c_int64()
        rc = lib().bmf_format(self.h, C.byref(o), C.byref(vb), C.byref(nv), C.byref(ni))
        V, I = nv.value, ni.value
        if rc != OK:
            return rc, V, I, None, None, None, None
        return (rc, V, I, pos.numpy()[:V], col.numpy()[:V] if col else None, nrm.numpy()[:V] if nrm else None,
                idx.numpy()[:I])

    def lod_queue(self, nodes, renderables, params, cap_items=None):
        """bmf_lod_queue on host arrays (workloads.WORLD_NODE nodes, int32 list, workloads.LodParams).
        Returns (rc, flags, dirty, items, n_chunks); the descriptors stay in the context for extract_queued."""
        cap = self.limits.max_chunks if cap_items is None else cap_items
        dn = DeviceArray.from_numpy(np.ascontiguousarray(nodes))
        dr = DeviceArray.from_numpy(np.ascontiguousarray(renderables, np.int32))
        fl = DeviceArray((max(len(nodes), 1),), np.uint32)
        di = DeviceArray((max(len(renderables), 1),), np.int32)
        it = DeviceArray((max(cap, 1),), np.dtype([("node", "<i4"), ("child", "<i4"), ("level", "<i4"),
                                                    ("action", "<i4"), ("morton", "<u8")], align=True))
        nd, nc = C.c_int32(), C.c_int32()
        rc = lib().bmf_lod_queue(self.h, dn.ptr, len(nodes), dr.ptr, len(renderables), C.byref(LodParams.of(params)),
                                 fl.ptr, di.ptr, it.ptr, cap, C.byref(nd), C.byref(nc))
        if rc != OK:
            return rc, None, None, None, nc.value
        return rc, fl.numpy()[: len(nodes)], di.numpy()[: nd.value], it.numpy()[: nc.value], nc.value

    def stitch(self, nodes, leaves, chunk_resolution, cap_triangles=1 << 20):
        """bmf_stitch on host arrays (workloads.WORLD_NODE nodes, int32 leaf list).
        Returns (rc, triangles [T][3][3] numpy or None, T)."""
        dn = DeviceArray.from_numpy(np.ascontiguousarray(nodes))
        dl = DeviceArray.from_numpy(np.ascontiguousarray(leaves, np.int32))
        out = DeviceArray((max(cap_triangles, 1), 3, 3), np.float32)
        nt = C.c_int64()
        res = C.c_int32(chunk_resolution)
        rc = lib().bmf_stitch(self.h, dn.ptr, len(nodes), dl.ptr, len(leaves), C.byref(res), out.ptr, cap_triangles,
                              C.byref(nt))
        if rc != OK:
            return rc, None, nt.value
        return rc, out.numpy()[: nt.value], nt.value

    def extract_queued(self, mode=MODE_BATCHED):
        check(lib().bmf_extract_queued(self.h, mode), "bmf_extract_queued")

    def qef_errors(self, chunk):
        rc, res, _, _ = self.results(chunk + 1)
        V = res[chunk].vertex_count
        out = np.zeros(V, np.float32)
        check(lib().bmf_qef_error_to_host(self.h, chunk, _ptr(out, C.c_float), V), "bmf_qef_error_to_host")
        return out

    def tuning(self):
        t = Tuning()
        check(lib().bmf_get_tuning(self.h, C.byref(t)), "bmf_get_tuning")
        return t.as_dict()

    def set_tuning(self, sample_tile=0, sample_blocks=0, row_blocks=0, smooth_blocks=0, row_split=0):
        t = Tuning(sample_tile, sample_blocks, row_blocks, smooth_blocks, row_split)
        check(lib().bmf_set_tuning(self.h, C.byref(t)), "bmf_set_tuning")

    def autotune(self, chunks, reps=3):
        arr = chunk_array(chunks)
        t = Tuning()
        check(lib().bmf_autotune(self.h, arr, len(chunks), reps, C.byref(t)), "bmf_autotune")
        return t.as_dict()

    def frame_stats(self):
        """bmf_get_frame_stats of the last extraction (after results())."""
        out = (C.c_int64 * 3)()
        check(lib().bmf_get_frame_stats(self.h, C.byref(out)), "bmf_get_frame_stats")
        return {"samples": out[0], "octave_evals": out[1], "refine_evals": out[2]}

    def profile_stages(self, chunks):
        arr = chunk_array(chunks)
        out = (C.c_float * NUM_STAGES)()
        check(lib().bmf_profile_stages(self.h, arr, len(chunks), out), "bmf_profile_stages")
        return dict(zip(STAGES, list(out)))


def validate_chunks(chunks, limits=None):
    arr = chunk_array(chunks) if chunks else None
    return lib().bmf_validate_chunks(arr, len(chunks), C.byref(limits) if limits is not None else None)


# ---- device buffers (HIP runtime through ctypes; the library's own runtime) ----------------
_hip = None
_H2D, _D2H = 1, 2


def hip():
    global _hip
    if _hip is None:
        lib()
        H = C.CDLL("libamdhip64.so.7")
        H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        H.hipFree.argtypes = [C.c_void_p]
        H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        H.hipDeviceSynchronize.argtypes = []
        _hip = H
    return _hip


class DeviceArray:
    """A device allocation holding a numpy array's bytes."""

    def __init__(self, shape, dtype):
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        if hip().hipMalloc(C.byref(p), max(self.nbytes, 1)) != 0:
            raise BmfError(EHIP, "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if a.nbytes and hip().hipMemcpy(d.ptr, a.ctypes.data, a.nbytes, _H2D) != 0:
            raise BmfError(EHIP, "hipMemcpy")
        return d

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes and hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, _D2H) != 0:
            raise BmfError(EHIP, "hipMemcpy")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_to_numpy(ptr, shape, dtype):
    out = np.empty(shape, dtype)
    if out.nbytes and hip().hipMemcpy(out.ctypes.data, ptr, out.nbytes, _D2H) != 0:
        raise BmfError(EHIP, "hipMemcpy")
    return out


class Tape:
    """bmf_tape: forward / loss / backward of one chunk on a context."""

    def __init__(self, ctx, max_dim=129, max_vertices=1 << 20, max_indices=1 << 22, checkpoint_every=2):
        h = C.c_void_p()
        check(lib().bmf_tape_create(C.byref(h), ctx.h, max_dim, max_vertices, max_indices, checkpoint_every),
              "bmf_tape_create")
        self.h = h
        self.ctx = ctx

    def close(self):
        if self.h:
            lib().bmf_tape_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def forward(self, chunk, densities=None):
        """densities: DeviceArray (dim^3 FP32) or None for the context's sampler."""
        check(lib().bmf_forward_diff(self.ctx.h, C.byref(chunk), densities.ptr if densities is not None else None,
                                     self.h), "bmf_forward_diff")
        return self.info()

    def info(self):
        i = TapeInfo()
        check(lib().bmf_tape_get(self.h, C.byref(i)), "bmf_tape_get")
        return i

    def mesh(self):
        i = self.info()
        V, nI = i.vertex_count, i.index_count
        return (device_to_numpy(i.positions, (V, 3), np.float32), device_to_numpy(i.indices, (nI,), np.uint32),
                device_to_numpy(i.boundary, (V,), np.uint8))

    def densities(self):
        i = self.info()
        return device_to_numpy(i.densities, (i.dim,) * 3, np.float32)

    def positions(self, iteration):
        i = self.info()
        out = DeviceArray((max(i.vertex_count, 1), 3), np.float32)
        check(lib().bmf_tape_positions(self.h, iteration, out.ptr), "bmf_tape_positions")
        return out.numpy()[: i.vertex_count]

    def loss_nearest(self, targets):
        """targets: DeviceArray [n][3] FP32.  Returns (loss, dL/dp DeviceArray, nearest numpy)."""
        i = self.info()
        g = DeviceArray((max(i.vertex_count, 1), 3), np.float32)
        near = DeviceArray((max(i.vertex_count, 1),), np.int32)
        L = C.c_float()
        check(lib().bmf_loss_nearest(self.h, targets.ptr, targets.shape[0], g.ptr, near.ptr, C.byref(L)),
              "bmf_loss_nearest")
        return L.value, g, near.numpy()[: i.vertex_count]

    def timing(self):
        """Device ms of the last forward, loss and backward (-1 where not run)."""
        ms = (C.c_float * 3)()
        check(lib().bmf_tape_timing(self.h, ms), "bmf_tape_timing")
        return list(ms)

    def set_targets(self, targets):
        """Builds the search hierarchy of a fixed target set (DeviceArray [n][3] FP32) once."""
        check(lib().bmf_tape_set_targets(self.h, targets.ptr, targets.shape[0]), "bmf_tape_set_targets")

    def loss_nearest_set(self):
        """The loss against the set of set_targets: (loss, dL/dp DeviceArray, nearest numpy)."""
        i = self.info()
        g = DeviceArray((max(i.vertex_count, 1), 3), np.float32)
        near = DeviceArray((max(i.vertex_count, 1),), np.int32)
        L = C.c_float()
        check(lib().bmf_tape_loss_nearest(self.h, g.ptr, near.ptr, C.byref(L)), "bmf_tape_loss_nearest")
        return L.value, g, near.numpy()[: i.vertex_count]

    def backward(self, dldp):
        i = self.info()
        out = DeviceArray((i.dim,) * 3, np.float32)
        check(lib().bmf_backward_diff(self.h, dldp.ptr if dldp is not None else None, out.ptr), "bmf_backward_diff")
        return out.numpy()
