"""Serving loop: each request's draw list is the same alone and batched with
others, equals the CPU oracle, and memory admission holds requests until
outputs are released."""
import time

import numpy as np
import pytest

import oracle
from cadr_amd import _native as N
from cadr_amd import cull, device, scenes
from cadr_amd.serve import Server

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def served(gpu):
    specs = [(30_000, 1), (120_000, 2), (5_000, 3), (64_000, 2)]
    ds_list, spheres = [], []
    for i, (n, depth) in enumerate(specs):
        sc = scenes.uneven_scene(100 + i, n, depth, 3, 1)
        ds = cull.DeviceScene(sc)
        ds.propagate()
        _, sph = oracle.propagate_scene(sc)
        assert ds.download_spheres().tobytes() == sph.tobytes()
        ds_list.append(ds)
        spheres.append(sph)
    views = cull.make_views(scenes.random_cameras(scenes.SEED_BASE + 77, 20, 0, 40, box=200.0))
    yield ds_list, spheres, views
    for ds in ds_list:
        ds.close()


def _result_arrays(ds, r):
    cmds = np.empty(ds.n, N.CMD_DTYPE)
    counts = np.empty(ds.K, np.uint32)
    N.check(N.load().cadrMemcpyDtoH(cmds.ctypes.data, r.commands, cmds.nbytes, None))
    N.check(N.load().cadrMemcpyDtoH(counts.ctypes.data, r.counts, counts.nbytes, None))
    N.check(N.load().cadrStreamSynchronize(None))
    return cmds, counts


def _wait(server, want, timeout=30.0):
    got = {}
    t0 = time.time()
    while len(got) < want:
        for r in server.poll():
            got[r.requestId] = r
        assert time.time() - t0 < timeout, "requests did not complete"
        time.sleep(0.001)
    return got


def _valid(cmds, counts, stateset_counts):
    return cull.valid_commands(cmds[None], counts[None], stateset_counts)[0]


def test_serve_alone_equals_mixed_equals_oracle(served):
    ds_list, spheres, views = served
    srv = Server(max_batch_views=32, inflight=3, budget_bytes=256 << 20)
    for ds in ds_list:
        srv.add_scene(ds)
    srv.start()
    probe_scene, probe_view = 1, 7
    # alone
    rid = srv.submit(probe_scene, views[probe_view])
    alone = _wait(srv, 1)[rid]
    a_cmds, a_counts = _result_arrays(ds_list[probe_scene], alone)
    assert np.array_equal(srv.counts(rid, ds_list[probe_scene].K), a_counts)
    srv.release(rid)
    # in a mix of 120 requests over all scenes, submitted in bursts
    rng = np.random.default_rng(4)
    ids = {}
    for i in range(120):
        s = probe_scene if i == 61 else int(rng.integers(0, len(ds_list)))
        v = probe_view if i == 61 else int(rng.integers(0, views.shape[0]))
        ids[srv.submit(s, views[v])] = (s, v)
        if i % 17 == 0:
            time.sleep(0.0005)
    got = _wait(srv, len(ids))
    for r_id, (s, v) in ids.items():
        r = got[r_id]
        cmds, counts = _result_arrays(ds_list[s], r)
        assert r.visible == counts.sum()
        exp_cmds, exp_counts = oracle.cull(spheres[s], ds_list[s].scene.mesh_index, ds_list[s].scene.meshes,
                                           ds_list[s].stateset_counts, views[v:v + 1], cmd_dtype=N.CMD_DTYPE)
        assert np.array_equal(counts, exp_counts[0]), f"request {r_id} counts"
        for a, b in zip(_valid(cmds, counts, ds_list[s].stateset_counts),
                        _valid(exp_cmds[0], exp_counts[0], ds_list[s].stateset_counts)):
            assert a.tobytes() == b.tobytes(), f"request {r_id} commands"
        if (s, v) == (probe_scene, probe_view):
            assert np.array_equal(counts, a_counts)
            for a, b in zip(_valid(cmds, counts, ds_list[s].stateset_counts),
                            _valid(a_cmds, a_counts, ds_list[s].stateset_counts)):
                assert a.tobytes() == b.tobytes(), "probe differs between alone and mixed"
        srv.release(r_id)
    srv.drain()
    srv.close()


def test_serve_memory_admission(served):
    """A budget that holds one output at a time: the queue advances only as
    results are released."""
    ds_list, spheres, views = served
    ds = ds_list[1]
    need = ((ds.n * 20 + 255) // 256 + (ds.K * 4 + 255) // 256) * 256
    srv = Server(max_batch_views=8, inflight=2, budget_bytes=need)
    srv.add_scene(ds)
    srv.start()
    ids = [srv.submit(0, views[i]) for i in range(5)]
    done = []
    t0 = time.time()
    while len(done) < 5:
        rs = srv.poll()
        assert len(rs) <= 1, "two outputs were admitted into a one-output budget"
        for r in rs:
            done.append(r.requestId)
            time.sleep(0.002)  # hold the output a little: the next request must wait
            srv.release(r.requestId)
        assert time.time() - t0 < 30
        time.sleep(0.0005)
    assert sorted(done) == sorted(ids)
    srv.drain()
    srv.close()
    # a request bigger than the whole budget is refused at submission
    small = Server(max_batch_views=8, inflight=1, budget_bytes=4096)
    small.add_scene(ds)
    small.start()
    with pytest.raises(N.CadrError) as e:
        small.submit(0, views[0])
    assert e.value.code == N.ERROR_OUT_OF_RESOURCES
    small.close()


def test_serve_load_runner(served):
    ds_list, spheres, views = served
    srv = Server(max_batch_views=64, inflight=4, budget_bytes=1 << 30)
    for ds in ds_list:
        srv.add_scene(ds)
    srv.start()
    st = srv.run_load(views, rate=3000.0, duration=0.3, seed=5, hot=2, hot_fraction=0.8, sample_every=10,
                      sample_capacity=40)
    samples = st.pop("samples")
    print(st)
    assert st["refused"] == 0 and st["completed"] == st["submitted"] > 500
    assert st["failed"] == 0 and st["sampled"] == len(samples) == 40
    assert 0 < st["p50Ms"] <= st["p99Ms"] <= st["p999Ms"] <= st["maxMs"]
    assert st["launches"] > 0 and st["meanViewsPerLaunch"] >= 1.0
    # the kept results are the oracle's for their scene and view, whatever they were batched with
    for r in samples:
        s, v = r.sceneId, r.tag
        assert r.status == 0 and v < views.shape[0]
        cmds, counts = _result_arrays(ds_list[s], r)
        exp_cmds, exp_counts = oracle.cull(spheres[s], ds_list[s].scene.mesh_index, ds_list[s].scene.meshes,
                                           ds_list[s].stateset_counts, views[v:v + 1], cmd_dtype=N.CMD_DTYPE)
        assert np.array_equal(counts, exp_counts[0]) and r.visible == counts.sum()
        for a, b in zip(_valid(cmds, counts, ds_list[s].stateset_counts),
                        _valid(exp_cmds[0], exp_counts[0], ds_list[s].stateset_counts)):
            assert a.tobytes() == b.tobytes()
        srv.release(r.requestId)
    srv.drain()
    srv.close()
