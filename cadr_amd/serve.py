"""Python driver of the serving loop (cadrServer* in include/cadr_hip.h)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N


class Server:
    def __init__(self, block=256, items=4, max_batch_views=256, inflight=2, budget_bytes=8 << 30,
                 max_queued=1 << 20, flags=0):
        cfg = N.ServerConfig(N.CullConfig(block, items, 0, flags), max_batch_views, inflight, budget_bytes,
                             max_queued, 0)
        h = ctypes.c_void_p(0)
        N.check(N.load().cadrServerCreate(ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h.value
        self._keep = []
        self.scenes = []

    def add_scene(self, device_scene) -> int:
        """device_scene: cadr_amd.cull.DeviceScene whose world spheres are current."""
        counts = np.ascontiguousarray(device_scene.stateset_counts, dtype=np.uint32)
        self._keep.append(counts)
        sc = N.ServerScene(device_scene.spheres.address, device_scene.mesh_index.address, device_scene.meshes.address,
                           counts.ctypes.data, counts.size, device_scene.num_meshes, device_scene.n, 0)
        sid = ctypes.c_uint32(0)
        N.check(N.load().cadrServerAddScene(self.handle, ctypes.byref(sc), ctypes.byref(sid)))
        self.scenes.append(device_scene)
        return sid.value

    def start(self):
        N.check(N.load().cadrServerStart(self.handle))

    def submit(self, scene: int, view: np.ndarray) -> int:
        v = np.ascontiguousarray(view, dtype=N.VIEW_DTYPE).reshape(1)
        rid = ctypes.c_uint64(0)
        N.check(N.load().cadrServerSubmit(self.handle, scene, v.ctypes.data, ctypes.byref(rid)))
        return rid.value

    def poll(self, capacity=1024):
        res = (N.ServerResult * capacity)()
        n = ctypes.c_uint32(0)
        N.check(N.load().cadrServerPoll(self.handle, res, capacity, ctypes.byref(n)))
        return [res[i] for i in range(n.value)]

    def counts(self, request_id: int, K: int) -> np.ndarray:
        out = np.zeros(K, np.uint32)
        N.check(N.load().cadrServerCopyCounts(self.handle, request_id, out.ctypes.data, K))
        return out

    def release(self, request_id: int):
        N.check(N.load().cadrServerRelease(self.handle, request_id))

    def drain(self, timeout_ms=30000):
        N.check(N.load().cadrServerDrain(self.handle, timeout_ms))

    def run_load(self, views: np.ndarray, rate: float, duration: float, seed=1, hot=16, hot_fraction=0.8,
                 sample_every: int = 0, sample_capacity: int = 0):
        """Open-loop Poisson load (cadrServerRunLoad).  With sample_every > 0,
        every sample_every-th completion (up to sample_capacity) is kept with
        its outputs allocated and returned under "samples" (ServerResult,
        tag = view index); the caller releases each one."""
        v = np.ascontiguousarray(views, dtype=N.VIEW_DTYPE)
        samples = (N.LoadSample * max(sample_capacity, 1))()
        spec = N.LoadSpec(rate, duration, seed, hot, hot_fraction, v.ctypes.data, v.shape[0], sample_every,
                          ctypes.cast(samples, ctypes.c_void_p), sample_capacity, 0)
        st = N.LoadStats()
        N.check(N.load().cadrServerRunLoad(self.handle, ctypes.byref(spec), ctypes.byref(st)))
        out = {f: getattr(st, f) for f, _ in N.LoadStats._fields_}
        out["samples"] = [samples[i].result for i in range(st.sampled)]
        return out

    def close(self):
        if self.handle:
            N.check(N.load().cadrServerDestroy(self.handle))
            self.handle = None
