"""Seeded synthetic workloads of the extraction path.

All randomness comes from splitmix64 (a fixed, portable generator) so that
every consumer draws the same chunks, primitives and targets.

Terrain field: d(x,y,z) = 24 * fbm(x,y,z) - (y - 32), fbm = FBM of 3D simplex
noise normalised by the amplitude sum.
"""
from dataclasses import dataclass, field

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK64

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        """float64 in [lo, hi) from the top 53 bits."""
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / (1 << 53)))

    def randint(self, lo, hi):
        """integer in [lo, hi] inclusive."""
        return lo + int(self.next() % (hi - lo + 1))

    def uniforms(self, n):
        """The next n uniform() values at once (same stream, vectorised)."""
        k = np.arange(1, n + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = np.uint64(self.state) + k * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        self.state = (self.state + n * 0x9E3779B97F4A7C15) & MASK64
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


@dataclass
class TerrainParams:
    amp: float = 24.0
    y0: float = 32.0
    freq: float = 0.02
    lacunarity: float = 2.0
    gain: float = 0.5
    octaves: int = 5
    seed: int = 1337


@dataclass
class Chunk:
    pos: tuple
    size: float
    dim: int
    overlap: float = 0.0
    iters: int = 0
    boundary_processing: int = 0


@dataclass
class Workload:
    name: str
    sampler: dict
    chunks: list
    extra: dict = field(default_factory=dict)

    @property
    def cells(self):
        return sum((c.dim - 1) ** 3 for c in self.chunks)


def terrain_sampler(p=None):
    p = p or TerrainParams()
    return dict(kind=0, world_size=256.0, amp=p.amp, y0=p.y0, freq=p.freq, lacunarity=p.lacunarity, gain=p.gain,
                octaves=p.octaves, seed=p.seed)


def single_chunk(iters=4):
    """One 64^3-cell chunk (65^3 samples, unit spacing, origin 0)."""
    return Workload("single-64", terrain_sampler(), [Chunk((0.0, 0.0, 0.0), 64.0, 65, 0.0, iters, 0)])


def uniform_grid(n_side=8, c=64, iters=4):
    """n_side^3 chunks of c^3 cells tiling the terrain, chunk (i,j,k) at c*(i,j,k)."""
    chunks = [Chunk((float(c * i), float(c * j), float(c * k)), float(c), c + 1, 0.0, iters, 0)
              for i in range(n_side) for j in range(n_side) for k in range(n_side)]
    return Workload(f"uniform-{n_side ** 3}x{c}", terrain_sampler(), chunks)


def mixed_lod(n=1536, seed=7, world=4096):
    """Uneven-LOD batch: 60/30/10 % chunks of 32/64/128 cells at spacing
    1/2/4, origins uniform in [0, world - extent]^3 snapped to the spacing,
    K ~ U{0..8}; 7-octave field at base frequency 0.005."""
    rng = SplitMix64(seed)
    n32 = round(0.6 * n)
    n64 = round(0.3 * n)
    n128 = n - n32 - n64
    classes = [(32, 1)] * n32 + [(64, 2)] * n64 + [(128, 4)] * n128
    # Fisher-Yates shuffle so the queue mixes sizes like an octree frontier
    for i in range(len(classes) - 1, 0, -1):
        j = rng.randint(0, i)
        classes[i], classes[j] = classes[j], classes[i]
    chunks = []
    for c, sp in classes:
        ext = c * sp
        steps = (world - ext) // sp
        pos = tuple(float(sp * rng.randint(0, steps)) for _ in range(3))
        chunks.append(Chunk(pos, float(ext), c + 1, 0.0, rng.randint(0, 8), 0))
    p = TerrainParams(octaves=7, freq=0.005)
    return Workload(f"mixed-lod-{n}", terrain_sampler(p), chunks)


def csg_scene(n_side=4, c=64, seed=42, n_spheres=200, n_boxes=100, iters=0):
    """n_side^3 chunks of c^3 cells over [0, n_side*c)^3 with a union of
    spheres (r ~ U[4,12]) and boxes (half extents ~ U[3,10]) at uniform
    centres."""
    rng = SplitMix64(seed)
    ext = n_side * c
    sph = np.array([[rng.uniform(0, ext), rng.uniform(0, ext), rng.uniform(0, ext), rng.uniform(4, 12)]
                    for _ in range(n_spheres)], np.float32)
    box = np.array([[rng.uniform(0, ext), rng.uniform(0, ext), rng.uniform(0, ext), rng.uniform(3, 10),
                     rng.uniform(3, 10), rng.uniform(3, 10)] for _ in range(n_boxes)], np.float32)
    chunks = [Chunk((float(c * i), float(c * j), float(c * k)), float(c), c + 1, 0.0, iters, 0)
              for i in range(n_side) for j in range(n_side) for k in range(n_side)]
    sampler = dict(kind=5, world_size=256.0, spheres=sph, boxes=box)
    return Workload(f"csg-{n_side ** 3}x{c}", sampler, chunks)


def sphere_targets(n=200_000, seed=3, radius=40.0, centre=(64.0, 64.0, 64.0)):
    """Points uniform on a sphere (normalised Gaussian directions from
    Box-Muller over splitmix64 uniforms)."""
    rng = SplitMix64(seed)
    u = rng.uniforms(4 * n).reshape(n, 4)
    u1 = np.maximum(u[:, 0], 1e-300)
    r = np.sqrt(-2.0 * np.log(u1))
    g0 = r * np.cos(2 * np.pi * u[:, 1])
    g1 = r * np.sin(2 * np.pi * u[:, 1])
    u2 = np.maximum(u[:, 2], 1e-300)
    g2 = np.sqrt(-2.0 * np.log(u2)) * np.cos(2 * np.pi * u[:, 3])
    d = np.stack([g0, g1, g2], 1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return (np.asarray(centre) + radius * d).astype(np.float32)


def to_capi(chunks):
    from . import capi

    return [capi.chunk_desc(c.pos, c.size, c.dim, c.overlap, c.iters, c.boundary_processing) for c in chunks]


# ---- world octree for the LOD pass (bmf_lod_queue) ----------------------------------------
# numpy layouts of bmf_world_node / bmf_lod_params / bmf_lod_item (include/bmf.h)
WORLD_NODE = np.dtype([("pos", "<f4", 3), ("size", "<f4"), ("level", "<i4"), ("flags", "<u4"), ("world_leaf", "<i4"),
                       ("parent", "<i4"), ("children", "<i4", 8), ("morton", "<u8")], align=True)
LOD_ITEM = np.dtype([("node", "<i4"), ("child", "<i4"), ("level", "<i4"), ("action", "<i4"), ("morton", "<u8")],
                    align=True)
NODE_DRAW, NODE_DESTROY, NODE_SPLIT, NODE_GROUP, NODE_DIRTY, NODE_DRAW_CHILDREN, NODE_SUPERCEDED = \
    1, 2, 4, 8, 32, 64, 256
_MCD = ((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), (0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1))  # Tables.hpp:7-9


@dataclass
class LodParams:
    """WorldProperties defaults (WorldOctree.cpp:20-33) and the watcher's max_gen (WorldWatcher.cpp:39)."""
    focus: tuple = (0.0, 0.0, 0.0)
    split_multiplier: float = 1.0
    group_multiplier: float = 2.0
    size_modifier: float = 0.0
    max_level: int = 7
    min_level: int = 1
    max_gen: int = 400
    chunk_resolution: int = 33
    process_iters: int = 0
    boundary_processing: int = 0
    overlap: float = 0.035


def _needs_split(pos, size, level, p):
    """WorldOctree::node_needs_split, for building the input tree (float32)."""
    if level >= p.max_level:
        return False
    if level < p.min_level:
        return True
    f = np.float32
    h = f(size) * f(0.5)
    d = [f(p.focus[a]) - (f(pos[a]) + h) for a in range(3)]
    dist = np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
    return dist < (f(size) * f(p.split_multiplier) + f(p.size_modifier)) + f(size) * f(0.5)


def lod_world(world_size=4096.0, build_focus=(0.0, 0.0, 0.0), params=None, seed=0, perturb=0.0):
    """A world octree as WorldOctree::init + split_leaves leave it
    (WorldOctree.cpp:115-173): the root of side 2 * world_size centred on the
    origin, split while node_needs_split(build_focus); leaves are drawn world
    leaves, inner nodes superseded.  Returns (nodes, renderables) with the
    leaves in depth-first order as the list.  perturb > 0 flips that fraction
    of flags at random (DESTROY, SPLIT, DRAW_CHILDREN, GROUP, SUPERCEDED,
    world_leaf) to exercise every branch of the pass."""
    p = params or LodParams(focus=build_focus)
    if params is not None:
        p = LodParams(**{**p.__dict__, "focus": tuple(build_focus)})
    rows = []
    size = np.float32(2.0 * world_size)
    rows.append(dict(pos=(np.float32(-0.5) * size,) * 3, size=size, level=0, parent=-1, morton=1))
    leaves = []
    stack = [0]
    while stack:
        i = stack.pop()
        r = rows[i]
        if not _needs_split(r["pos"], r["size"], r["level"], p):
            leaves.append(i)
            r["children"] = [-1] * 8
            continue
        cs = np.float32(r["size"]) * np.float32(0.5)
        kids = []
        for k, (dx, dy, dz) in enumerate(_MCD):
            cp = tuple(np.float32(r["pos"][a]) + np.float32(d) * cs for a, d in enumerate((dx, dy, dz)))
            kids.append(len(rows))
            rows.append(dict(pos=cp, size=cs, level=r["level"] + 1, parent=i,
                             morton=(r["morton"] << 3) | (dx | (dy << 1) | (dz << 2))))
        r["children"] = kids
        stack.extend(kids)
    nodes = np.zeros(len(rows), WORLD_NODE)
    for i, r in enumerate(rows):
        leaf = r["children"][0] < 0
        nodes[i]["pos"] = r["pos"]
        nodes[i]["size"] = r["size"]
        nodes[i]["level"] = r["level"]
        nodes[i]["flags"] = NODE_DRAW if leaf else NODE_SUPERCEDED
        nodes[i]["world_leaf"] = int(leaf)
        nodes[i]["parent"] = r["parent"]
        nodes[i]["children"] = r["children"]
        nodes[i]["morton"] = r["morton"]
    if perturb > 0:
        rng = SplitMix64(seed)
        u = rng.uniforms(3 * len(nodes)).reshape(len(nodes), 3)
        bits = (NODE_DESTROY, NODE_SPLIT, NODE_DRAW_CHILDREN, NODE_GROUP, NODE_SUPERCEDED, NODE_DRAW)
        for i in range(len(nodes)):
            if u[i, 0] < perturb:
                nodes[i]["flags"] ^= bits[int(u[i, 1] * len(bits)) % len(bits)]
            if u[i, 2] < perturb / 4:
                nodes[i]["world_leaf"] ^= 1
    return nodes, np.array(leaves, np.int32)
