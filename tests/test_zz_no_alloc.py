"""cadrCullBatchNoAlloc (the batched cull with its workspace in a fixed device
array instead of allocated memory) against cadrCullBatch on the same jobs:
counts and compacted commands bit-identical, on two calls in a row, the
second one smaller, through the C ABI."""
import ctypes

import numpy as np
import pytest

import oracle
from cadr_amd import _native as N
from cadr_amd import cull, scenes
from cadr_amd.device import synchronize

pytestmark = pytest.mark.gpu


def _outputs(entries_spec):
    return [cull.CullOutputs(views.shape[0], ds.n, ds.K) for ds, views in entries_spec]


def _same(a_out, b_out, ds, what):
    ca, na = a_out.download()
    cb, nb = b_out.download()
    assert np.array_equal(na, nb), f"{what}: counts differ"
    ga = cull.valid_commands(ca, na, ds.stateset_counts)
    gb = cull.valid_commands(cb, nb, ds.stateset_counts)
    for v in range(len(ga)):
        for k in range(ds.K):
            assert ga[v][k].tobytes() == gb[v][k].tobytes(), f"{what}: view {v} StateSet {k} differs"
    return na


@pytest.fixture(scope="module")
def two_scenes(gpu):
    a = scenes.tree_scene((2000, 100), seed=scenes.SEED_BASE + 71, num_statesets=64, lod_count=3, trange=200.0)
    b = scenes.plain_scene(100_000)
    da, db = cull.DeviceScene(a), cull.DeviceScene(b)
    da.propagate()
    db.propagate()
    yield (a, da), (b, db)
    da.close()
    db.close()


def test_no_alloc_matches_cull_batch_on_two_calls(two_scenes):
    (a, da), (b, db) = two_scenes
    lib = N.load()
    config = cull.make_config(256, 8)
    views_a = cull.make_views(scenes.random_cameras(scenes.SEED_BASE + 71, 20, 0, 70, box=400.0))
    views_b = cull.make_views([scenes.DEFAULT_CAMERA] + scenes.random_cameras(scenes.SEED_BASE + 72, 20, 0, 2))
    calls = [
        ("call 1: 200k x 70 views + 100k x 3 views", [(da, views_a), (db, views_b)]),
        ("call 2: 200k x 5 views", [(da, views_a[:5])]),
    ]
    for what, spec in calls:
        print(f"no-alloc {what}", flush=True)
        ref_out, got_out = _outputs(spec), _outputs(spec)
        ref = cull.CullBatch(config, [(ds, v, o) for (ds, v), o in zip(spec, ref_out)])
        got = cull.CullBatch(config, [(ds, v, o) for (ds, v), o in zip(spec, got_out)])
        ref.run()
        N.check(lib.cadrCullBatchNoAlloc(ctypes.byref(config), got.device_jobs.address, N.ptr(got.host_jobs),
                                         got.host_jobs.shape[0], None))
        synchronize()
        for (ds, v), ro, go in zip(spec, ref_out, got_out):
            counts = _same(ro, go, ds, what)
            assert counts.sum() > 0
        # and both agree with the oracle on the first view of the first job
        ds, v = spec[0]
        scene = a if ds is da else b
        _, spheres = oracle.propagate_scene(scene)
        _, exp_counts = oracle.cull(spheres, scene.mesh_index, scene.meshes, scene.stateset_counts, v[:1],
                                    cmd_dtype=N.CMD_DTYPE)
        assert np.array_equal(got_out[0].download()[1][:1], exp_counts)
        for x in (ref, got):
            x.close()
        for o in ref_out + got_out:
            o.close()


def test_no_alloc_refuses_calls_past_its_scratch(gpu):
    lib = N.load()
    config = cull.make_config(256, 4)
    fake = 1 << 40  # never dereferenced: the size checks come first
    jobs = np.zeros(1, dtype=N.JOB_DTYPE)
    jobs["numTiles"] = 4097
    jobs["numViews"] = 512  # 2,097,664 status words > 2^21
    assert lib.cadrCullBatchNoAlloc(ctypes.byref(config), fake, N.ptr(jobs), 1, None) == N.ERROR_INVALID_VALUE
    assert "status words" in lib.cadrGetLastErrorMessage().decode()
    many = np.zeros(4097, dtype=N.JOB_DTYPE)
    assert lib.cadrCullBatchNoAlloc(ctypes.byref(config), fake, N.ptr(many), 4097, None) == N.ERROR_INVALID_VALUE
    assert "jobs" in lib.cadrGetLastErrorMessage().decode()
    assert lib.cadrCullBatchNoAlloc(ctypes.byref(cull.make_config(100, 3)), fake, N.ptr(jobs), 1, None) \
        == N.ERROR_INVALID_VALUE
