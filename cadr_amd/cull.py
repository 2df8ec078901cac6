"""Python driver of the cull pass over the C ABI: views, tiles, jobs, device
scenes and batched launches.  Used by tests and the benchmark; C++ callers use
include/cadr_hip.h directly (or the CadR facade)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N
from .device import DeviceBuffer
from .scenes import Camera, Scene

DEFAULT_CONFIG = (256, 4)


def make_config(block: int = 256, items: int = 4, per_cu: int = 0, flags: int = 0) -> N.CullConfig:
    return N.CullConfig(block, items, per_cu, flags)


def tile_size(config: N.CullConfig) -> int:
    return int(N.load().cadrCullTileSize(ctypes.byref(config)))


def make_view(cam: Camera) -> np.ndarray:
    """cadrMakeView for one camera -> 1-element VIEW_DTYPE array."""
    out = np.zeros(1, dtype=N.VIEW_DTYPE)
    eye = np.asarray(cam.eye, dtype=np.float32)
    d = np.asarray(cam.dir, dtype=np.float32)
    up = np.asarray(cam.up, dtype=np.float32)
    th = np.asarray(cam.thresholds, dtype=np.float32)
    N.check(N.load().cadrMakeView(N.ptr(eye), N.ptr(d), N.ptr(up), float(np.float32(cam.fovy)),
                                  float(np.float32(cam.aspect)), float(np.float32(cam.near)),
                                  float(np.float32(cam.far)), float(np.float32(cam.viewport_height)),
                                  N.ptr(th), N.ptr(out)))
    return out


def make_views(cams) -> np.ndarray:
    return np.concatenate([make_view(c) for c in cams]) if cams else np.zeros(0, dtype=N.VIEW_DTYPE)


def build_tiles(stateset_counts, tile: int) -> np.ndarray:
    counts = np.ascontiguousarray(np.asarray(stateset_counts, dtype=np.uint32))
    n = ctypes.c_uint32(0)
    lib = N.load()
    N.check(lib.cadrBuildCullTiles(N.ptr(counts) if counts.size else None, counts.size, tile, None, 0,
                                   ctypes.byref(n)))
    tiles = np.zeros(max(n.value, 1), dtype=N.TILE_DTYPE)
    N.check(lib.cadrBuildCullTiles(N.ptr(counts) if counts.size else None, counts.size, tile, N.ptr(tiles),
                                   tiles.size, ctypes.byref(n)))
    return tiles[: n.value]


class DeviceScene:
    """A scene resident in HBM: transform levels, world spheres, mesh table."""

    def __init__(self, scene, keep_levels: bool = True, chunk: int = 1 << 18):
        """scene: a Scene (host arrays), a TreeSpec (streamed level by level in
        chunks) or a SynthTree / SynthFlat (generated in device memory)."""
        self.scene = scene
        self.n = scene.num_drawables
        self.K = int(len(scene.stateset_counts))
        self.stateset_counts = np.asarray(scene.stateset_counts, dtype=np.uint32)
        self.num_meshes = int(scene.meshes.shape[0])
        self.meshes = DeviceBuffer.from_array(scene.meshes)
        self.spheres = DeviceBuffer(max(16 * self.n, 16))
        self.levels = []
        synth = getattr(scene, "synth", False)
        if synth:
            self.mesh_index = DeviceBuffer(max(4 * self.n, 16))
            self._synthesize(scene)
        else:
            self.mesh_index = DeviceBuffer.from_array(scene.mesh_index)
        if synth:
            pass
        elif keep_levels and isinstance(scene, Scene):
            for local, parent in scene.levels:
                cnt = local.shape[0]
                lb = DeviceBuffer.from_array(local)
                wb = DeviceBuffer(max(64 * cnt, 16))
                pb = DeviceBuffer.from_array(parent) if parent is not None else None
                self.levels.append((cnt, lb, pb, wb))
        elif keep_levels:
            for d, cnt in enumerate(scene.sizes):
                lb = DeviceBuffer(max(64 * cnt, 16))
                wb = DeviceBuffer(max(64 * cnt, 16))
                pb = DeviceBuffer(max(4 * cnt, 16)) if d > 0 else None
                for c0 in range(0, cnt, chunk):
                    c1 = min(cnt, c0 + chunk)
                    local, parent = scene.level_chunk(d, c0, c1)
                    lb.upload(local, offset=64 * c0)
                    if pb is not None:
                        pb.upload(parent, offset=4 * c0)
                self.levels.append((cnt, lb, pb, wb))
        self._tiles = {}

    def _synthesize(self, spec) -> None:
        """Generates the scene's levels (or, for a flat scene, its world
        spheres) in device memory with cadrSynthesizeLevel."""
        lib = N.load()
        if spec.flat:
            L = spec.level_desc()
            L["meshIndex"], L["meshes"], L["spheres"] = self.mesh_index.address, self.meshes.address, self.spheres.address
            N.check(lib.cadrSynthesizeLevel(N.ptr(L), None))
            return
        for d, cnt in enumerate(spec.sizes):
            lb = DeviceBuffer(max(64 * cnt, 16))
            wb = DeviceBuffer(max(64 * cnt, 16))
            pb = DeviceBuffer(max(4 * cnt, 16)) if d > 0 else None
            L = spec.level_desc(d)
            L["localMatrices"] = lb.address
            if pb is not None:
                L["parentIndex"] = pb.address
            if d == len(spec.sizes) - 1:
                L["meshIndex"] = self.mesh_index.address
            N.check(lib.cadrSynthesizeLevel(N.ptr(L), None))
            self.levels.append((cnt, lb, pb, wb))

    def tiles(self, tile: int):
        if tile not in self._tiles:
            t = build_tiles(self.stateset_counts, tile)
            self._tiles[tile] = (DeviceBuffer.from_array(t), int(t.shape[0]))
        return self._tiles[tile]

    def propagate(self, stream=None, spheres: bool = True) -> None:
        """Runs every tree level in order; the last level writes world spheres."""
        lib = N.load()
        prev = None
        for d, (cnt, lb, pb, wb) in enumerate(self.levels):
            lv = np.zeros(1, dtype=N.LEVEL_DTYPE)
            lv["localMatrices"] = lb.address
            lv["worldMatrices"] = wb.address
            lv["count"] = cnt
            if prev is not None and pb is not None:
                lv["parentIndex"] = pb.address
                lv["parentWorld"] = prev[1].address
                lv["parentCount"] = prev[0]
            if spheres and d == len(self.levels) - 1:
                lv["meshIndex"] = self.mesh_index.address
                lv["meshes"] = self.meshes.address
                lv["numMeshes"] = self.num_meshes
                lv["spheres"] = self.spheres.address
            N.check(lib.cadrPropagateTransforms(N.ptr(lv), stream))
            prev = (cnt, wb)

    def level_descs(self, spheres: bool = True) -> list:
        """cadrTransformLevel of every level, root first; the last one writes
        the world spheres when `spheres`."""
        out, prev = [], None
        for d, (cnt, lb, pb, wb) in enumerate(self.levels):
            lv = np.zeros(1, dtype=N.LEVEL_DTYPE)
            lv["localMatrices"], lv["worldMatrices"], lv["count"] = lb.address, wb.address, cnt
            if prev is not None and pb is not None:
                lv["parentIndex"], lv["parentWorld"], lv["parentCount"] = pb.address, prev[1].address, prev[0]
            if spheres and d == len(self.levels) - 1:
                lv["meshIndex"], lv["meshes"], lv["numMeshes"] = (self.mesh_index.address, self.meshes.address,
                                                                  self.num_meshes)
                lv["spheres"] = self.spheres.address
            out.append(lv)
            prev = (cnt, wb)
        return out

    def world_matrices(self, level: int = -1) -> np.ndarray:
        cnt, _, _, wb = self.levels[level]
        return wb.download(np.float32, cnt * 16).reshape(cnt, 16)

    def download_spheres(self) -> np.ndarray:
        return self.spheres.download(np.float32, 4 * self.n).reshape(self.n, 4)

    def upload_spheres(self, spheres: np.ndarray) -> None:
        self.spheres.upload(np.ascontiguousarray(spheres, dtype=np.float32))

    def close(self):
        for _, lb, pb, wb in self.levels:
            lb.close(); wb.close()
            if pb is not None:
                pb.close()
        self.meshes.close(); self.mesh_index.close(); self.spheres.close()
        for b, _ in self._tiles.values():
            b.close()


class CullOutputs:
    def __init__(self, num_views: int, stride: int, K: int):
        self.num_views, self.stride, self.K = num_views, stride, K
        self.commands = DeviceBuffer(max(20 * num_views * stride, 16))
        self.counts = DeviceBuffer(max(4 * num_views * K, 16))

    def download(self):
        cmds = self.commands.download(N.CMD_DTYPE, self.num_views * self.stride).reshape(self.num_views, self.stride)
        return cmds, self.download_counts()

    def download_counts(self):
        """The per-view StateSet counts alone (the commands of many views can
        be tens of GB)."""
        return self.counts.download(np.uint32, self.num_views * self.K).reshape(self.num_views, self.K)

    def close(self):
        self.commands.close(); self.counts.close()


class Workspace:
    def __init__(self, status_words: int, max_jobs: int = 1024):
        w = ctypes.c_void_p(0)
        N.check(N.load().cadrCullWorkspaceCreate(int(status_words), int(max_jobs), ctypes.byref(w)))
        self.handle = w.value
        self.status_words = status_words

    def set_wait_limit(self, microseconds: int) -> None:
        N.check(N.load().cadrCullWorkspaceSetWaitLimit(self.handle, int(microseconds)))

    def check(self) -> None:
        """Waits for the device; raises CadrError (CADR_ERROR_TIMEOUT) if a
        look-back wait of a launch on this workspace gave up."""
        N.check(N.load().cadrCullWorkspaceCheck(self.handle))

    def close(self):
        if self.handle:
            N.check(N.load().cadrCullWorkspaceDestroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CullBatch:
    """A launch: jobs (scene, views, outputs) in one persistent kernel."""

    def __init__(self, config: N.CullConfig, entries, workspace: Workspace | None = None):
        """entries: list of (DeviceScene, views VIEW_DTYPE array or DeviceBuffer+count, CullOutputs)."""
        self.config = config
        self.tile = tile_size(config)
        if self.tile == 0:
            raise N.CadrError(N.ERROR_INVALID_VALUE, "unsupported cull config")
        jobs = np.zeros(len(entries), dtype=N.JOB_DTYPE)
        self._keep = []
        status = 0
        for j, (ds, views, out) in enumerate(entries):
            if isinstance(views, np.ndarray):
                vb = DeviceBuffer.from_array(views)
                nv = int(views.shape[0])
                self._keep.append(vb)
            else:
                vb, nv = views
            tb, nt = ds.tiles(self.tile)
            jobs[j] = (ds.spheres.address, ds.mesh_index.address, ds.meshes.address, tb.address, vb.address,
                       out.commands.address, out.counts.address, ds.n, nt, ds.K, nv, out.stride, ds.num_meshes)
            status += nv * nt
        self.host_jobs = jobs
        self.device_jobs = DeviceBuffer.from_array(jobs)
        self.status_words = status
        self.workspace = workspace or Workspace(max(status, 1), max(len(entries), 1))
        self._own_ws = workspace is None

    def run(self, stream=None) -> None:
        N.check(N.load().cadrCullBatch(ctypes.byref(self.config), self.device_jobs.address, N.ptr(self.host_jobs),
                                       self.host_jobs.shape[0], self.workspace.handle, stream))

    def run_view(self, stream=None) -> None:
        """The single-view pass (cadrCullView) for a batch of one job with one view."""
        assert self.host_jobs.shape[0] == 1 and self.host_jobs["numViews"][0] == 1
        N.check(N.load().cadrCullView(ctypes.byref(self.config), self.device_jobs.address, N.ptr(self.host_jobs),
                                      self.workspace.handle, stream))

    def run_fused(self, leaf: np.ndarray, stream=None) -> None:
        """cadrPropagateCullView: the leaf level (LEVEL_DTYPE, 1 element) fused
        with the one-view cull of this batch's single job."""
        assert self.host_jobs.shape[0] == 1 and self.host_jobs["numViews"][0] == 1
        N.check(N.load().cadrPropagateCullView(N.ptr(leaf), ctypes.byref(self.config), self.device_jobs.address,
                                               N.ptr(self.host_jobs), self.workspace.handle, stream))

    def set_leaves(self, leaves: np.ndarray) -> None:
        """Leaf levels (LEVEL_DTYPE, one per job) for run_fused_batch; kept
        in device memory until close()."""
        leaves = np.ascontiguousarray(leaves, dtype=N.LEVEL_DTYPE)
        assert leaves.shape == self.host_jobs.shape
        if getattr(self, "_leaves_dev", None) is not None:
            self._leaves_dev.close()
        self.host_leaves = leaves
        self._leaves_dev = DeviceBuffer.from_array(leaves)

    def run_fused_batch(self, stream=None) -> None:
        """cadrPropagateCullBatch: each job's leaf level (set_leaves) fused
        with the batched cull of all jobs, one persistent launch."""
        N.check(N.load().cadrPropagateCullBatch(self._leaves_dev.address, N.ptr(self.host_leaves),
                                                ctypes.byref(self.config), self.device_jobs.address,
                                                N.ptr(self.host_jobs), self.host_jobs.shape[0],
                                                self.workspace.handle, stream))

    def close(self):
        self.device_jobs.close()
        if getattr(self, "_leaves_dev", None) is not None:
            self._leaves_dev.close()
            self._leaves_dev = None
        for b in self._keep:
            b.close()
        if self._own_ws:
            self.workspace.close()


def valid_commands(cmds: np.ndarray, counts: np.ndarray, stateset_counts) -> list:
    """Per view, the list of per-StateSet compacted command arrays."""
    base = np.concatenate([[0], np.cumsum(np.asarray(stateset_counts, dtype=np.int64))])
    out = []
    for v in range(cmds.shape[0]):
        out.append([cmds[v, base[k]: base[k] + counts[v, k]] for k in range(len(stateset_counts))])
    return out
