"""Deterministic synthetic scenes for the drawable-processing pass.

Counter-based SplitMix64 (every value is a pure function of seed, stream,
item and component, so any slice can be generated independently); floats
take the top 24 bits.  Seed = 0xCAD00000 + workload index.  The workloads:

  plain100k  100,000 drawables, 64 meshes (indexCount U{36..3000}, firstIndex
             cumulative, radius U[0.5, 2.5)), flat transforms: translation
             U[-500,500]^3, uniform scale exp(U[ln 0.1, ln 2]); 1 StateSet;
             camera at the origin looking -z, vfov 60 deg, 16:9, near 0.1,
             far 2000, 1080-pixel viewport.
  full10m    transform tree 1,000 roots -> 100 children -> 100 leaves
             (10,101,000 nodes, 10,000,000 leaf drawables); local matrices are
             a uniform random rotation (Shoemake quaternion) plus translation
             U[-50,50]^3; 64 meshes with 3 LODs (LOD k index count roughly
             count/4^k), LOD thresholds 64 px and 8 px at 1080 px; StateSet of
             each leaf U{0..63}, drawables grouped by StateSet in tree order.
  batch256   the full10m scene with 256 views: eye U[-600,600]^3, uniform
             random direction, vfov U[40, 90] deg, 16:9.
  uneven64   64 scenes of round(exp(U[ln 1e4, ln 5e6])) drawables, tree depth
             U{1..4} (equal fan-out per level), LOD count U{1..4}, 64
             StateSets, 1..32 views each.
  served128  128 flat scenes of round(exp(U[ln 1e4, ln 2e7])) drawables
             (plain100k structure, 64 StateSets), trimmed to a byte budget.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from ._native import MESH_DTYPE

GAMMA = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
SEED_BASE = 0xCAD00000

WORKLOADS = ("plain100k", "full10m", "batch256", "uneven64", "served128")


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * M1
    z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def rand_u64(seed: int, stream: int, start: int, count: int, comps: int = 1, items=None) -> np.ndarray:
    """SplitMix64 outputs for items start..start+count-1 (or the given item
    indices), comps values each, shape (count, comps).
    Counter = (stream << 40) + item*comps + comp + 1."""
    with np.errstate(over="ignore"):
        if items is None:
            idx = np.arange(start * comps, (start + count) * comps, dtype=np.uint64)
        else:
            items = np.asarray(items, dtype=np.uint64)
            count = items.shape[0]
            idx = (items[:, None] * np.uint64(comps) + np.arange(comps, dtype=np.uint64)[None, :]).reshape(-1)
        ctr = (np.uint64(stream) << np.uint64(40)) + idx + np.uint64(1)
        x = np.uint64(seed) + ctr * GAMMA
        return _mix(x).reshape(count, comps)


def rand_unit(seed: int, stream: int, start: int, count: int, comps: int = 1, items=None) -> np.ndarray:
    """Uniform floats in [0, 1) from the top 24 bits (float32)."""
    u = rand_u64(seed, stream, start, count, comps, items) >> np.uint64(40)
    return (u.astype(np.float32) * np.float32(1.0 / 16777216.0)).astype(np.float32)


def rand_int(seed: int, stream: int, start: int, count: int, lo: int, hi: int) -> np.ndarray:
    """Uniform integers in [lo, hi] (inclusive) from the top 32 bits."""
    out = np.empty(count, dtype=np.int64)
    span = np.uint64(hi - lo + 1)
    for c0 in range(0, count, 1 << 22):
        c1 = min(count, c0 + (1 << 22))
        u = rand_u64(seed, stream, start + c0, c1 - c0, 1)[:, 0] >> np.uint64(32)
        out[c0:c1] = lo + ((u * span) >> np.uint64(32)).astype(np.int64)
    return out


# ------------------------------------------------------------------ meshes

def make_meshes(seed: int, count: int, lod_counts) -> np.ndarray:
    """Mesh table: centre 0, radius U[0.5, 2.5), LOD k index count derived
    from a base count U{36..3000} as max(12, (c0 >> 2k) rounded down to a
    multiple of 3), firstIndex cumulative over (mesh, LOD)."""
    lod_counts = np.broadcast_to(np.asarray(lod_counts, dtype=np.int64), (count,))
    base = rand_int(seed, 1, 0, count, 36, 3000)
    radius = np.float32(0.5) + np.float32(2.0) * rand_unit(seed, 2, 0, count)[:, 0]
    m = np.zeros(count, dtype=MESH_DTYPE)
    m["radius"] = radius
    m["lodCount"] = lod_counts
    first = 0
    for i in range(count):
        for k in range(int(lod_counts[i])):
            c = int(base[i]) >> (2 * k)
            c = max(12, c - c % 3)
            m["indexCount"][i, k] = c
            m["firstIndex"][i, k] = first
            first += c
    return m


# --------------------------------------------------------------- matrices

CHUNK = 1 << 18   # keeps numpy temporaries under glibc's mmap threshold


def rotation_translation(seed: int, stream: int, start: int, count: int, trange: float, items=None) -> np.ndarray:
    """count x 16 float32 column-major matrices: uniform random rotation
    (Shoemake's quaternion from three uniforms) and translation U[-t, t]^3.
    Matrix i is a pure function of its item index (start + i, or items[i])."""
    if items is not None:
        count = len(items)
    out = np.empty((count, 16), dtype=np.float32)
    for c0 in range(0, count, CHUNK):
        c1 = min(count, c0 + CHUNK)
        it = None if items is None else items[c0:c1]
        out[c0:c1] = _rotation_translation_chunk(seed, stream, start + c0, c1 - c0, trange, it)
    return out


def _rotation_translation_chunk(seed: int, stream: int, start: int, count: int, trange: float,
                                items=None) -> np.ndarray:
    u = rand_unit(seed, stream, start, count, 6, items).astype(np.float64)
    u1, u2, u3 = u[:, 0], u[:, 1], u[:, 2]
    a = np.sqrt(1.0 - u1)
    b = np.sqrt(u1)
    x = a * np.sin(2 * math.pi * u2)
    y = a * np.cos(2 * math.pi * u2)
    z = b * np.sin(2 * math.pi * u3)
    w = b * np.cos(2 * math.pi * u3)
    m = np.zeros((count, 16), dtype=np.float32)
    # column 0
    m[:, 0] = 1 - 2 * (y * y + z * z)
    m[:, 1] = 2 * (x * y + z * w)
    m[:, 2] = 2 * (x * z - y * w)
    # column 1
    m[:, 4] = 2 * (x * y - z * w)
    m[:, 5] = 1 - 2 * (x * x + z * z)
    m[:, 6] = 2 * (y * z + x * w)
    # column 2
    m[:, 8] = 2 * (x * z + y * w)
    m[:, 9] = 2 * (y * z - x * w)
    m[:, 10] = 1 - 2 * (x * x + y * y)
    # column 3: translation
    t = u[:, 3:6].astype(np.float32)
    m[:, 12:15] = (np.float32(-trange) + np.float32(2 * trange) * t).astype(np.float32)
    m[:, 15] = 1.0
    return m


def scale_translation(seed: int, stream: int, start: int, count: int) -> np.ndarray:
    """Flat transforms of plain100k: scale exp(U[ln 0.1, ln 2]), translation U[-500,500]^3."""
    u = rand_unit(seed, stream, start, count, 4)
    lo, hi = math.log(0.1), math.log(2.0)
    s = np.exp(lo + u[:, 0].astype(np.float64) * (hi - lo)).astype(np.float32)
    m = np.zeros((count, 16), dtype=np.float32)
    m[:, 0] = s
    m[:, 5] = s
    m[:, 10] = s
    m[:, 12:15] = (np.float32(-500.0) + np.float32(1000.0) * u[:, 1:4]).astype(np.float32)
    m[:, 15] = 1.0
    return m


# ------------------------------------------------------------------ views

@dataclass
class Camera:
    eye: tuple
    dir: tuple
    up: tuple
    fovy: float            # radians
    aspect: float = 16.0 / 9.0
    near: float = 0.1
    far: float = 2000.0
    viewport_height: float = 1080.0
    thresholds: tuple = (64.0, 8.0, 1.0)


DEFAULT_CAMERA = Camera(eye=(0.0, 0.0, 0.0), dir=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0),
                        fovy=float(np.float32(math.radians(60.0))))


def random_cameras(seed: int, stream: int, start: int, count: int, box: float = 600.0) -> list:
    u = rand_unit(seed, stream, start, count, 6).astype(np.float64)
    cams = []
    for i in range(count):
        eye = tuple(float(np.float32(-box + 2 * box * u[i, k])) for k in range(3))
        z = 2.0 * u[i, 3] - 1.0
        phi = 2.0 * math.pi * u[i, 4]
        r = math.sqrt(max(0.0, 1.0 - z * z))
        d = (float(np.float32(r * math.cos(phi))), float(np.float32(r * math.sin(phi))), float(np.float32(z)))
        up = (0.0, 1.0, 0.0) if abs(d[1]) < 0.99 else (1.0, 0.0, 0.0)
        fovy = float(np.float32(math.radians(40.0 + 50.0 * u[i, 5])))
        cams.append(Camera(eye=eye, dir=d, up=up, fovy=fovy))
    return cams


# ------------------------------------------------------------------ scenes

@dataclass
class Scene:
    """A scene in the pass's input layout.

    levels[d] = (local float32[n_d,16], parent uint32[n_d] or None); the last
    level holds the drawables (leaves) in drawable order.  mesh_index and
    stateset_counts describe the drawables; meshes is the mesh table.
    """
    name: str
    levels: list
    mesh_index: np.ndarray
    meshes: np.ndarray
    stateset_counts: np.ndarray
    cameras: list = field(default_factory=list)

    @property
    def num_drawables(self) -> int:
        return int(self.mesh_index.shape[0])


def _group_by_stateset(seed, stream, leaf_count, num_statesets, start=0):
    ss = rand_int(seed, stream, start, leaf_count, 0, num_statesets - 1)
    order = np.argsort(ss, kind="stable")
    counts = np.bincount(ss, minlength=num_statesets).astype(np.uint32)
    return order, counts


def plain_scene(n: int = 100_000, seed: int = SEED_BASE + 0, num_statesets: int = 1,
                num_meshes: int = 64) -> Scene:
    meshes = make_meshes(seed, num_meshes, 1)
    mesh_index = rand_int(seed, 3, 0, n, 0, num_meshes - 1).astype(np.uint32)
    world = scale_translation(seed, 4, 0, n)
    if num_statesets > 1:
        order, counts = _group_by_stateset(seed, 5, n, num_statesets)
        world = world[order]
        mesh_index = mesh_index[order]
    else:
        counts = np.array([n], dtype=np.uint32)
    return Scene("plain", [(np.ascontiguousarray(world), None)], np.ascontiguousarray(mesh_index), meshes,
                 counts, [DEFAULT_CAMERA])


class TreeSpec:
    """A transform tree whose level sizes are fanouts[0], fanouts[0]*fanouts[1],
    ...; the leaf level (the drawables) is ordered by StateSet, tree order
    inside a StateSet.  Levels can be produced chunk by chunk (each matrix is
    a function of its tree index), so large scenes stream to the device
    without whole-level host arrays."""

    def __init__(self, fanouts=(1000, 100, 100), seed: int = SEED_BASE + 1, num_statesets: int = 64,
                 lod_count: int = 3, num_meshes: int = 64, trange: float = 50.0, name: str = "tree"):
        self.fanouts, self.seed, self.trange, self.name = tuple(fanouts), seed, trange, name
        self.meshes = make_meshes(seed, num_meshes, lod_count)
        self.sizes = [int(x) for x in np.cumprod(fanouts)]
        leaves = self.sizes[-1]
        self.order, self.stateset_counts = _group_by_stateset(seed, 5, leaves, num_statesets)
        self.mesh_index = np.ascontiguousarray(
            rand_int(seed, 3, 0, leaves, 0, num_meshes - 1).astype(np.uint32)[self.order])
        self.cameras = [DEFAULT_CAMERA]

    @property
    def num_drawables(self) -> int:
        return self.sizes[-1]

    def level_chunk(self, d: int, c0: int, c1: int):
        last = d == len(self.fanouts) - 1
        items = self.order[c0:c1] if last else None
        local = rotation_translation(self.seed, 10 + d, c0, c1 - c0, self.trange, items)
        if d == 0:
            parent = None
        else:
            tree_index = items.astype(np.uint64) if last else np.arange(c0, c1, dtype=np.uint64)
            parent = (tree_index // np.uint64(self.fanouts[d])).astype(np.uint32)
        return local, parent

    def to_scene(self) -> Scene:
        levels = []
        for d, n in enumerate(self.sizes):
            local, parent = self.level_chunk(d, 0, n)
            levels.append((local, parent))
        return Scene(self.name, levels, self.mesh_index, self.meshes, self.stateset_counts, list(self.cameras))


def tree_scene(fanouts=(1000, 100, 100), seed: int = SEED_BASE + 1, num_statesets: int = 64,
               lod_count: int = 3, num_meshes: int = 64, trange: float = 50.0) -> Scene:
    """Materialised TreeSpec (host arrays for every level)."""
    return TreeSpec(fanouts, seed, num_statesets, lod_count, num_meshes, trange).to_scene()


def full10m_spec(scale: float = 1.0) -> TreeSpec:
    """The full10m workload; scale < 1 shrinks the root count."""
    roots = max(1, int(round(1000 * scale)))
    return TreeSpec((roots, 100, 100), seed=SEED_BASE + 1, name="full10m")


def full10m_scene(scale: float = 1.0) -> Scene:
    return full10m_spec(scale).to_scene()


def batch_cameras(count: int = 256) -> list:
    return random_cameras(SEED_BASE + 2, 20, 0, count)


def uneven_specs(num_scenes: int = 64, lo: float = 1e4, hi: float = 5e6, max_views: int = 32,
                 seed: int = SEED_BASE + 3):
    """(drawables, depth, lod_count, views) per scene of the uneven64 workload."""
    u = rand_unit(seed, 30, 0, num_scenes, 1)[:, 0].astype(np.float64)
    sizes = np.rint(np.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))).astype(np.int64)
    depth = rand_int(seed, 31, 0, num_scenes, 1, 4)
    lods = rand_int(seed, 32, 0, num_scenes, 1, 4)
    views = rand_int(seed, 33, 0, num_scenes, 1, max_views)
    return [(int(sizes[i]), int(depth[i]), int(lods[i]), int(views[i])) for i in range(num_scenes)]


def uneven_scene(index: int, drawables: int, depth: int, lod_count: int, num_views: int,
                 seed: int = SEED_BASE + 3, num_statesets: int = 64) -> Scene:
    """Scene `index` of uneven64: `depth` levels with equal fan-out
    f = ceil(drawables^(1/depth)); level d has ceil(n_{d+1}/f) nodes."""
    sseed = seed + 0x1000 * (index + 1)
    f = max(1, int(math.ceil(drawables ** (1.0 / depth) - 1e-9)))
    sizes = [drawables]
    for _ in range(depth - 1):
        sizes.append((sizes[-1] + f - 1) // f)
    sizes = sizes[::-1]
    meshes = make_meshes(sseed, 64, lod_count)
    levels = []
    for d, n in enumerate(sizes):
        local = rotation_translation(sseed, 10 + d, 0, n, 50.0)
        parent = None if d == 0 else (np.arange(n, dtype=np.uint64) // f).astype(np.uint32)
        levels.append([local, parent])
    mesh_index = rand_int(sseed, 3, 0, drawables, 0, 63).astype(np.uint32)
    order, counts = _group_by_stateset(sseed, 5, drawables, num_statesets)
    local, parent = levels[-1]
    levels[-1] = [np.ascontiguousarray(local[order]), None if parent is None else np.ascontiguousarray(parent[order])]
    cams = random_cameras(sseed, 20, 0, num_views)
    return Scene(f"uneven{index}", [tuple(l) for l in levels], np.ascontiguousarray(mesh_index[order]), meshes,
                 counts, cams)


def balanced_counts(n: int, k: int) -> np.ndarray:
    """StateSet sizes floor((j+1)n/k) - floor(jn/k): positions split into k near-equal ranges."""
    edges = (np.arange(k + 1, dtype=np.int64) * n) // k
    return np.diff(edges).astype(np.uint32)


class SynthTree:
    """A transform tree generated in device memory by cadrSynthesizeLevel.

    Level sizes as uneven_scene: `depth` levels with equal fan-out
    f = ceil(drawables^(1/depth)).  Inner node i is tree node i with parent
    i // f; the drawable at leaf position p is tree node perm(p) (the
    generator's Feistel permutation, key = seed), so the near-equal StateSet
    ranges of positions hold uniformly random sets of leaves.  Matrices are
    rotation_translation(seed, 10 + d, node), mesh indexes rand_int(seed, 3,
    node) over the 64 meshes."""
    synth = True

    def __init__(self, name: str, drawables: int, depth: int, lod_count: int, seed: int, num_statesets: int = 64,
                 cameras=None, trange: float = 50.0, num_meshes: int = 64):
        self.name, self.seed, self.trange = name, seed, trange
        f = max(1, int(math.ceil(drawables ** (1.0 / depth) - 1e-9)))
        sizes = [drawables]
        for _ in range(depth - 1):
            sizes.append((sizes[-1] + f - 1) // f)
        self.sizes = sizes[::-1]
        self.fanout = f
        self.meshes = make_meshes(seed, num_meshes, lod_count)
        self.stateset_counts = balanced_counts(drawables, num_statesets)
        self.cameras = list(cameras or [])
        self.flat = False

    @property
    def num_drawables(self) -> int:
        return self.sizes[-1]

    def level_desc(self, d: int) -> np.ndarray:
        """cadrSynthLevel of level d without output addresses."""
        from ._native import SYNTH_LEVEL_DTYPE, SYNTH_ROTATION_TRANSLATION
        L = np.zeros(1, dtype=SYNTH_LEVEL_DTYPE)
        L["seed"], L["count"], L["stream"] = self.seed, self.sizes[d], 10 + d
        L["transform"], L["translationRange"] = SYNTH_ROTATION_TRANSLATION, self.trange
        L["fanout"] = self.fanout if d > 0 else 0
        if d == len(self.sizes) - 1:
            L["permuteKey"] = self.seed
            L["meshStream"], L["numMeshes"] = 3, self.meshes.shape[0]
        return L


class SynthFlat:
    """A flat scene generated in device memory: the drawables' world spheres
    directly (plain100k transforms, scale_translation(seed, 4, i) with
    translation range `trange`), mesh indexes rand_int(seed, 3, i); StateSets
    are near-equal ranges of the i.i.d. drawables."""
    synth = True

    def __init__(self, name: str, drawables: int, seed: int, lod_count: int = 3, num_statesets: int = 64,
                 trange: float = 500.0, num_meshes: int = 64):
        self.name, self.seed, self.trange = name, seed, trange
        self.sizes = [drawables]
        self.meshes = make_meshes(seed, num_meshes, lod_count)
        self.stateset_counts = balanced_counts(drawables, num_statesets)
        self.cameras = []
        self.flat = True

    @property
    def num_drawables(self) -> int:
        return self.sizes[0]

    def level_desc(self) -> np.ndarray:
        from ._native import SYNTH_LEVEL_DTYPE, SYNTH_SCALE_TRANSLATION
        L = np.zeros(1, dtype=SYNTH_LEVEL_DTYPE)
        L["seed"], L["count"], L["stream"] = self.seed, self.sizes[0], 4
        L["transform"], L["translationRange"] = SYNTH_SCALE_TRANSLATION, self.trange
        L["meshStream"], L["numMeshes"] = 3, self.meshes.shape[0]
        return L


def uneven_synth(index: int, drawables: int, depth: int, lod_count: int, num_views: int,
                 seed: int = SEED_BASE + 3) -> SynthTree:
    """Scene `index` of uneven64, generated in device memory."""
    sseed = seed + 0x1000 * (index + 1)
    return SynthTree(f"uneven{index}", drawables, depth, lod_count, sseed,
                     cameras=random_cameras(sseed, 20, 0, num_views))


def served_synth(index: int, drawables: int, seed: int = SEED_BASE + 4) -> SynthFlat:
    return SynthFlat(f"served{index}", drawables, seed + 0x1000 * (index + 1))


def served_specs(num_scenes: int = 128, lo: float = 1e4, hi: float = 2e7, budget_drawables: int | None = None,
                 seed: int = SEED_BASE + 4):
    u = rand_unit(seed, 40, 0, num_scenes, 1)[:, 0].astype(np.float64)
    sizes = np.rint(np.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))).astype(np.int64)
    if budget_drawables is not None and sizes.sum() > budget_drawables:
        sizes = np.maximum(1, (sizes * (budget_drawables / sizes.sum())).astype(np.int64))
    return [int(s) for s in sizes]


# ------------------------------------------------------------ gather input

class GatherScene:
    """Input of the reference's gather pass (processDrawables.comp) in one
    device image, built with numpy in bulk: handle tables of the reference's
    layout (2048 uint64 entries per table, HandleTable.h:33-35), `geometries`
    Geometry allocations (vertex, index, 4 PrimitiveSet records {count,
    first}), one MatrixList {numMatrices = 1, capacity = 1, pad} + mat4 per
    drawable (128 B, MatrixList.h:54-59), one shared drawable-data
    allocation, and the 48-B DrawableGpuData list (Drawable.h:32-43).

    Handles are numbered consecutively from `first_handle` (geometries
    first, then MatrixLists, then the drawable-data handle), as the
    reference's HandleTable numbers allocations; first_handle >= 2^22 puts
    the scene at handle level 3 whatever its size."""

    TABLE = 2048

    def __init__(self, n: int, seed: int = SEED_BASE + 6, geometries: int = 64, first_handle: int = 1):
        self.n, self.G = n, geometries
        self.h0 = first_handle
        self.hmax = first_handle + 3 * geometries + n  # the drawable-data handle
        self.level = 1 if self.hmax < 2048 else (2 if self.hmax < (1 << 22) else 3)
        u = rand_u64(seed, 50, 0, n, 2)
        self.geom = ((u[:, 0] >> np.uint64(32)) * np.uint64(geometries) >> np.uint64(32)).astype(np.int64)
        self.ps = ((u[:, 1] >> np.uint64(32)) * np.uint64(4) >> np.uint64(32)).astype(np.int64)
        pv = rand_u64(seed, 51, 0, geometries, 8)
        self.ps_count = (pv[:, 0:4] >> np.uint64(47)).astype(np.uint32) + np.uint32(1)
        self.ps_first = (pv[:, 4:8] >> np.uint64(33)).astype(np.uint32)

    def layout(self, base: int):
        """Byte offsets of every region of the image at device address base."""
        T = self.TABLE * 8
        lo, hi = self.h0 >> 11, self.hmax >> 11
        L = {"root": 0}
        off = T
        if self.level == 3:
            mlo, mhi = self.h0 >> 22, self.hmax >> 22
            L["mid"], L["mid_range"] = off, (mlo, mhi)
            off += (mhi - mlo + 1) * T
        if self.level >= 2:
            L["leaf"], L["leaf_range"] = off, (lo, hi)
            off += (hi - lo + 1) * T
        L["geom"] = off
        off += self.G * 192
        L["dd"] = off
        off += 64
        L["ml"] = off
        off += self.n * 128
        L["list"] = off
        off += self.n * 48
        L["size"] = off
        self.base = base
        return L

    def addresses(self, base: int) -> np.ndarray:
        """Device address of every handle h0..hmax (index h - h0)."""
        L = self.layout(base)
        a = np.zeros(self.hmax - self.h0 + 1, np.uint64)
        g = np.arange(self.G, dtype=np.uint64)
        gbase = np.uint64(base + L["geom"]) + g * np.uint64(192)
        a[0:3 * self.G:3] = gbase                    # vertex data
        a[1:3 * self.G:3] = gbase + np.uint64(64)    # index data
        a[2:3 * self.G:3] = gbase + np.uint64(128)   # PrimitiveSet records
        a[3 * self.G:3 * self.G + self.n] = np.uint64(base + L["ml"]) + np.arange(self.n, dtype=np.uint64) * np.uint64(128)
        a[-1] = base + L["dd"]
        return a

    def image(self, base: int):
        """(uint8 image, push constants (table, list, 0, 0), level)."""
        L = self.layout(base)
        img = np.zeros(L["size"], np.uint8)
        addr = self.addresses(base)
        h = np.arange(self.h0, self.hmax + 1, dtype=np.int64)
        T = self.TABLE
        if self.level == 1:
            img[:T * 8].view(np.uint64)[h] = addr
        else:
            lo, hi = L["leaf_range"]
            leaf = img[L["leaf"]:L["leaf"] + (hi - lo + 1) * T * 8].view(np.uint64)
            leaf[h - lo * T] = addr
            leaf_addr = np.uint64(base + L["leaf"]) + np.arange(hi - lo + 1, dtype=np.uint64) * np.uint64(T * 8)
            if self.level == 2:
                img[:T * 8].view(np.uint64)[lo:hi + 1] = leaf_addr
            else:
                mlo, mhi = L["mid_range"]
                mid = img[L["mid"]:L["mid"] + (mhi - mlo + 1) * T * 8].view(np.uint64)
                leaves = np.arange(lo, hi + 1, dtype=np.int64)
                mid[leaves - mlo * T] = leaf_addr           # mid table (L >> 11), entry L & 2047
                img[:T * 8].view(np.uint64)[mlo:mhi + 1] = (np.uint64(base + L["mid"]) +
                                                            np.arange(mhi - mlo + 1, dtype=np.uint64) * np.uint64(T * 8))
        geo = img[L["geom"]:L["geom"] + self.G * 192].reshape(self.G, 192)
        ps = np.zeros((self.G, 4, 2), np.uint32)
        ps[:, :, 0], ps[:, :, 1] = self.ps_count, self.ps_first
        geo[:, 128:160] = ps.reshape(self.G, 8).view(np.uint8)
        ml = img[L["ml"]:L["ml"] + self.n * 128].view(np.uint32).reshape(self.n, 32)
        ml[:, 0] = 1   # numMatrices
        ml[:, 1] = 1   # capacity
        ml[:, 16] = 0x3F800000  # the matrix: 1.0 on the diagonal
        ml[:, 21] = 0x3F800000
        ml[:, 26] = 0x3F800000
        ml[:, 31] = 0x3F800000
        from ._native import DRAWABLE_DTYPE
        d = img[L["list"]:L["list"] + self.n * 48].view(DRAWABLE_DTYPE)
        gh = np.uint64(self.h0) + np.uint64(3) * self.geom.astype(np.uint64)
        d["vertexDataHandle"] = gh
        d["indexDataHandle"] = gh + np.uint64(1)
        d["primitiveSetHandle"] = gh + np.uint64(2)
        d["matrixListHandle"] = np.uint64(self.h0 + 3 * self.G) + np.arange(self.n, dtype=np.uint64)
        d["drawableDataHandle"] = np.uint64(self.hmax)
        d["primitiveSetOffset"] = (8 * self.ps).astype(np.uint32)
        return img, (base, base + L["list"], 0, 0), self.level

    def expected(self, base: int):
        """Closed-form outputs: indirect uint32[n,4], pointers uint64[n,4]."""
        L = self.layout(base)
        ind = np.zeros((self.n, 4), np.uint32)
        ind[:, 0] = self.ps_count[self.geom, self.ps]
        ind[:, 1] = 1
        ind[:, 2] = self.ps_first[self.geom, self.ps]
        ptr = np.zeros((self.n, 4), np.uint64)
        gbase = np.uint64(base + L["geom"]) + self.geom.astype(np.uint64) * np.uint64(192)
        ptr[:, 0] = gbase
        ptr[:, 1] = gbase + np.uint64(64)
        ptr[:, 2] = np.uint64(base + L["ml"]) + np.arange(self.n, dtype=np.uint64) * np.uint64(128)
        ptr[:, 3] = base + L["dd"]
        return ind, ptr
