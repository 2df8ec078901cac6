"""cadrTimeNextLaunch: events recorded by a kernel's own dispatch.  The timed
launch computes exactly what an untimed one does, start / stop bracket the
kernel, a stop-only event closes the interval from the launch before it, an
entry point with nothing to launch records both events, and a cleared
request leaves the next launch untimed."""
import numpy as np
import pytest

import oracle
from cadr_amd import _native as N
from cadr_amd import cull, scenes
from cadr_amd.device import DeviceBuffer, Event, Stream, time_next_launch

pytestmark = pytest.mark.gpu


def _matrices(rng, n):
    m = rng.standard_normal((n, 16)).astype(np.float32)
    m[:, 3] = m[:, 7] = m[:, 11] = 0.0
    m[:, 15] = 1.0
    return m


def test_timed_propagation_matches_untimed_and_oracle(gpu):
    lib = N.load()
    rng = np.random.default_rng(scenes.SEED_BASE + 91)
    n, P = 300_000, 4000
    local, parent_world = _matrices(rng, n), _matrices(rng, P)
    parent = rng.integers(0, P, size=n).astype(np.uint32)
    dl, dp, dpw = DeviceBuffer.from_array(local), DeviceBuffer.from_array(parent), DeviceBuffer.from_array(parent_world)
    worlds = [DeviceBuffer(n * 64) for _ in range(2)]
    stream = Stream()
    s = stream.handle
    levels = []
    for w in worlds:
        lv = np.zeros(1, dtype=N.LEVEL_DTYPE)
        lv["localMatrices"], lv["parentIndex"], lv["parentWorld"] = dl.address, dp.address, dpw.address
        lv["worldMatrices"], lv["count"], lv["parentCount"] = w.address, n, P
        levels.append(lv)
    before, start, stop, after = Event(), Event(), Event(), Event()
    N.check(lib.cadrPropagateTransforms(N.ptr(levels[0]), s))  # untimed
    before.record(s)
    time_next_launch(start, stop)
    N.check(lib.cadrPropagateTransforms(N.ptr(levels[1]), s))  # timed
    after.record(s)
    stream.synchronize()
    kernel_ms = start.elapsed_ms(stop)
    assert 0.0 < kernel_ms < 100.0
    # the dispatch's own interval lies inside the marker events around it
    assert before.elapsed_ms(start) >= 0.0 and stop.elapsed_ms(after) >= 0.0
    assert kernel_ms <= before.elapsed_ms(after)
    expect, _ = oracle.propagate_level(local, parent, parent_world)
    got0 = worlds[0].download(np.float32, n * 16).reshape(n, 16)
    got1 = worlds[1].download(np.float32, n * 16).reshape(n, 16)
    assert np.array_equal(got0.view(np.uint32), expect.view(np.uint32))
    assert np.array_equal(got1.view(np.uint32), got0.view(np.uint32))

    # stop only: the interval from the previous launch's end to this one's
    mid, end = Event(), Event()
    time_next_launch(None, mid)
    N.check(lib.cadrPropagateTransforms(N.ptr(levels[0]), s))
    time_next_launch(None, end)
    N.check(lib.cadrPropagateTransforms(N.ptr(levels[1]), s))
    stream.synchronize()
    assert 0.0 < mid.elapsed_ms(end) < 100.0

    # nothing to launch: both events are recorded back to back
    empty = levels[0].copy()
    empty["count"] = 0
    e0, e1 = Event(), Event()
    time_next_launch(e0, e1)
    N.check(lib.cadrPropagateTransforms(N.ptr(empty), s))
    stream.synchronize()
    assert 0.0 <= e0.elapsed_ms(e1) < 1.0

    # a cleared request leaves the next launch untimed: its events stay
    # unrecorded, so the later launch (timed with fresh events) is unaffected
    stale0, stale1, f0, f1 = Event(), Event(), Event(), Event()
    time_next_launch(stale0, stale1)
    time_next_launch(None, None)
    N.check(lib.cadrPropagateTransforms(N.ptr(levels[0]), s))
    time_next_launch(f0, f1)
    N.check(lib.cadrPropagateTransforms(N.ptr(levels[1]), s))
    stream.synchronize()
    assert 0.0 < f0.elapsed_ms(f1) < 100.0
    for e in (before, start, stop, after, mid, end, e0, e1, stale0, stale1, f0, f1):
        e.close()
    stream.close()


def test_timed_cull_counts_match_untimed(gpu):
    spec = scenes.TreeSpec((20, 50, 100), seed=scenes.SEED_BASE + 92, name="timed")
    stream = Stream()
    s = stream.handle
    ds = cull.DeviceScene(spec)
    ds.propagate(s)
    views = cull.make_views([scenes.DEFAULT_CAMERA])
    outs = [cull.CullOutputs(1, ds.n, ds.K) for _ in range(2)]
    runs = [cull.CullBatch(cull.make_config(256, 8), [(ds, views, o)]) for o in outs]
    runs[0].run_view(s)
    start, stop = Event(), Event()
    time_next_launch(start, stop)
    runs[1].run_view(s)
    stream.synchronize()
    assert 0.0 < start.elapsed_ms(stop) < 100.0
    c0, c1 = outs[0].download_counts(), outs[1].download_counts()
    assert c0.sum() > 0 and np.array_equal(c0, c1)
    for r in runs:
        r.close()
    start.close()
    stop.close()
    stream.close()


def test_timed_batched_and_fused_launches_match_untimed(gpu):
    """cadrCullBatch, cadrPropagateCullView and cadrPropagateCullBatch with
    dispatch-bound events: each timed launch gives the counts of the same
    launch untimed, and its interval is positive."""
    spec = scenes.TreeSpec((20, 50, 100), seed=scenes.SEED_BASE + 93, name="timed3")
    stream = Stream()
    s = stream.handle
    ds = cull.DeviceScene(spec)
    ds.propagate(s)
    views = cull.make_views(scenes.batch_cameras(8))
    one = cull.make_views([scenes.DEFAULT_CAMERA])
    leaf = ds.level_descs(spheres=True)[-1].copy()
    leaf["spheres"] = 0
    cases = [("batch", cull.make_config(256, 8), views, lambda b: b.run(s)),
             ("fused view", cull.make_config(512, 4), one, lambda b: b.run_fused(leaf, s))]
    for name, cfg, v, run in cases:
        outs = [cull.CullOutputs(v.shape[0], ds.n, ds.K) for _ in range(2)]
        runs = [cull.CullBatch(cfg, [(ds, v, o)]) for o in outs]
        run(runs[0])
        start, stop = Event(), Event()
        time_next_launch(start, stop)
        run(runs[1])
        stream.synchronize()
        assert 0.0 < start.elapsed_ms(stop) < 100.0, name
        c0, c1 = outs[0].download_counts(), outs[1].download_counts()
        assert c0.sum() > 0 and np.array_equal(c0, c1), name
        for r in runs:
            r.close()
        for o in outs:
            o.close()
        start.close()
        stop.close()
    # the fused batched form (leaf levels in device memory)
    outs = [cull.CullOutputs(views.shape[0], ds.n, ds.K) for _ in range(2)]
    runs = [cull.CullBatch(cull.make_config(256, 8), [(ds, views, o)]) for o in outs]
    for r in runs:
        r.set_leaves(leaf)
    runs[0].run_fused_batch(s)
    start, stop = Event(), Event()
    time_next_launch(start, stop)
    runs[1].run_fused_batch(s)
    stream.synchronize()
    assert 0.0 < start.elapsed_ms(stop) < 100.0
    c0, c1 = outs[0].download_counts(), outs[1].download_counts()
    assert c0.sum() > 0 and np.array_equal(c0, c1)
    for r in runs:
        r.close()
    for o in outs:
        o.close()
    start.close()
    stop.close()
    stream.close()
