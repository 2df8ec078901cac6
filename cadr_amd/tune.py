"""Autotuning of the cull launch shape (cadrCullAutotune in include/cadr_hip.h).

    best, table = autotune([(device_scene, views, outputs), ...])
    best, table = autotune([(device_scene, one_view, outputs)], single_view=True)

The tuner runs real launches on the given jobs (their outputs are written)
and returns the fastest cadrCullConfig together with every candidate's mean
time per launch and whether its counts agreed with the first candidate's.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N
from .device import DeviceBuffer


def autotune(entries, single_view: bool = False, reps: int = 5, candidates=None, stream=None):
    """entries: (DeviceScene, views VIEW_DTYPE array, CullOutputs) per job.
    candidates: optional list of (block, items, workgroupsPerCU, flags)."""
    jobs = np.zeros(len(entries), dtype=N.TUNE_JOB_DTYPE)
    keep = []
    for j, (ds, views, out) in enumerate(entries):
        vb = DeviceBuffer.from_array(np.ascontiguousarray(views, dtype=N.VIEW_DTYPE))
        counts = np.ascontiguousarray(ds.stateset_counts, dtype=np.uint32)
        keep += [vb, counts]
        jobs[j]["job"] = (ds.spheres.address, ds.mesh_index.address, ds.meshes.address, 0, vb.address,
                          out.commands.address, out.counts.address, ds.n, 0, ds.K, views.shape[0], out.stride,
                          ds.num_meshes)
        jobs[j]["stateSetCounts"] = counts.ctypes.data
    cand = None
    ncand = 0
    if candidates:
        cand = (N.CullConfig * len(candidates))(*[N.CullConfig(*c) for c in candidates])
        ncand = len(candidates)
    cap = 64
    results = np.zeros(cap, dtype=N.TUNE_RESULT_DTYPE)
    nres = ctypes.c_uint32(0)
    best = N.CullConfig()
    try:
        N.check(N.load().cadrCullAutotune(N.ptr(jobs), len(entries), N.TUNE_SINGLE_VIEW if single_view else 0, reps,
                                          cand, ncand, N.ptr(results), cap, ctypes.byref(nres), ctypes.byref(best),
                                          stream))
    finally:
        for k in keep:
            if isinstance(k, DeviceBuffer):
                k.close()
    table = [{"block": int(r["blockSize"]), "items": int(r["itemsPerThread"]), "per_cu": int(r["workgroupsPerCU"]),
              "flags": int(r["flags"]), "ms": float(r["milliseconds"]), "agrees": bool(r["agrees"])}
             for r in results[: nres.value]]
    return best, table


def describe(config: N.CullConfig) -> str:
    sph = "" if config.flags & N.CULL_NO_TILE_SPHERES else "+tile spheres"
    return f"{config.blockSize}x{config.itemsPerThread}{sph}"
