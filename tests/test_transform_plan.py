"""cadrTransformPlan: many independent levels sharing one parent array,
their 64-node units ordered by parent.  World matrices and spheres bit for
bit against the oracle and against cadrPropagateTransformsBatch, with more
levels than one batch launch takes, sorted and unsorted parent indices,
empty and ragged levels, a root level and out-of-range parents."""
import ctypes

import numpy as np
import pytest

import oracle
from cadr_amd import _native as N
from cadr_amd import scenes
from cadr_amd.device import DeviceBuffer, synchronize

pytestmark = pytest.mark.gpu


def _matrices(rng, n):
    m = rng.standard_normal((n, 16)).astype(np.float32)
    m[:, 3] = m[:, 7] = m[:, 11] = 0.0
    m[:, 15] = 1.0
    return m


def test_planned_propagation_matches_oracle_and_batch(gpu):
    rng = np.random.default_rng(scenes.SEED_BASE + 77)
    P = 5000
    parent_world = _matrices(rng, P)
    d_parent = DeviceBuffer.from_array(parent_world)
    keep = [d_parent]
    levels, expect = [], []
    sizes = rng.integers(0, 300, size=1500)
    sizes[:5] = [0, 1, 63, 64, 65]
    for i, n in enumerate(sizes):
        n = int(n)
        local = _matrices(rng, max(n, 1))[:n]
        if i % 3 == 0:
            parent = np.sort(rng.integers(0, P, size=n)).astype(np.uint32)   # ordered by parent
        else:
            parent = rng.integers(0, P, size=n).astype(np.uint32)
        if i % 97 == 5 and n > 0:
            parent[0] = P + 3  # out of range: world = local
        lv = np.zeros(1, dtype=N.LEVEL_DTYPE)
        if i == 7:  # a root level among them
            world_exp, _ = oracle.propagate_level(local, None, None)
        else:
            world_exp, _ = oracle.propagate_level(local, parent, parent_world)
        dl = DeviceBuffer.from_array(local if n else np.zeros((1, 16), np.float32))
        dp = DeviceBuffer.from_array(parent if n else np.zeros(1, np.uint32))
        dw = DeviceBuffer(max(n, 1) * 64)
        keep += [dl, dp, dw]
        lv["localMatrices"], lv["worldMatrices"], lv["count"] = dl.address, dw.address, n
        if i != 7:
            lv["parentIndex"], lv["parentWorld"], lv["parentCount"] = dp.address, d_parent.address, P
        levels.append(lv)
        expect.append((dw, n, world_exp))
    host = np.concatenate(levels)
    plan = ctypes.c_void_p(0)
    lib = N.load()
    N.check(lib.cadrTransformPlanCreate(N.ptr(host), host.shape[0], None, ctypes.byref(plan)))
    try:
        for rep in range(2):
            for dw, n, _ in expect:
                dw.fill_u32(0xFFFFFFFF)
            N.check(lib.cadrPropagateTransformsPlanned(plan, None))
            synchronize()
            for i, (dw, n, world_exp) in enumerate(expect):
                if n:
                    got = dw.download(np.float32, 16 * n).reshape(n, 16)
                    assert got.tobytes() == world_exp.tobytes(), f"planned, level {i} (n={n}), rep {rep}"
        # the same levels through batch launches of at most 1024 levels
        for dw, n, _ in expect:
            dw.fill_u32(0xFFFFFFFF)
        for b in range(0, host.shape[0], 1024):
            part = np.ascontiguousarray(host[b:b + 1024])
            dev = DeviceBuffer.from_array(part)
            keep.append(dev)
            N.check(lib.cadrPropagateTransformsBatch(dev.address, N.ptr(part), part.shape[0], None))
        synchronize()
        for i, (dw, n, world_exp) in enumerate(expect):
            if n:
                assert dw.download(np.float32, 16 * n).reshape(n, 16).tobytes() == world_exp.tobytes(), f"batch {i}"
    finally:
        lib.cadrTransformPlanDestroy(plan)
        for b in keep:
            b.close()


def test_plan_rejects_bad_levels(gpu):
    lib = N.load()
    plan = ctypes.c_void_p(0)
    lv = np.zeros(1, dtype=N.LEVEL_DTYPE)
    lv["count"] = 10  # no matrix arrays
    assert lib.cadrTransformPlanCreate(N.ptr(lv), 1, None, ctypes.byref(plan)) == N.ERROR_INVALID_VALUE
    assert plan.value is None
    assert lib.cadrTransformPlanCreate(None, 1, None, ctypes.byref(plan)) == N.ERROR_INVALID_VALUE
    assert lib.cadrPropagateTransformsPlanned(None, None) == N.ERROR_INVALID_VALUE
    # an empty plan runs and does nothing
    N.check(lib.cadrTransformPlanCreate(None, 0, None, ctypes.byref(plan)))
    N.check(lib.cadrPropagateTransformsPlanned(plan, None))
    lib.cadrTransformPlanDestroy(plan)
