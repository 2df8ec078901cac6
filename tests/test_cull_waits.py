"""Bounded look-back waits: a caller-built tile list that names a predecessor
no tile publishes makes every launch end anyway, and the workspace reports
CADR_ERROR_TIMEOUT.  The condition is forced from the host: a tile list
without the first tile of its StateSet and a 1-ms wait limit, so each kernel
exits after its waits give up.  Well-formed launches on the same workspace
stay bit-exact against the oracle."""
import numpy as np
import pytest

import oracle
from cadr_amd import _native as N
from cadr_amd import cull, scenes
from cadr_amd.device import DeviceBuffer, synchronize

pytestmark = pytest.mark.gpu

N_DRAWABLES = 3000


@pytest.fixture(scope="module")
def small(gpu):
    scene = scenes.plain_scene(N_DRAWABLES)
    ds = cull.DeviceScene(scene)
    ds.propagate()
    _, spheres = oracle.propagate_scene(scene)
    yield scene, ds, spheres
    ds.close()


def _job(ds, tiles_buf, num_tiles, views_buf, num_views, out):
    jobs = np.zeros(1, dtype=N.JOB_DTYPE)
    jobs[0] = (ds.spheres.address, ds.mesh_index.address, ds.meshes.address, tiles_buf.address, views_buf.address,
               out.commands.address, out.counts.address, ds.n, num_tiles, ds.K, num_views, out.stride,
               ds.num_meshes)
    return jobs


@pytest.mark.parametrize("path", ["view", "batch"])
def test_missing_predecessor_times_out_and_workspace_recovers(small, path):
    import ctypes
    scene, ds, spheres = small
    config = cull.make_config(256, 4)
    tile = cull.tile_size(config)
    good = cull.build_tiles(scene.stateset_counts, tile)
    assert good.shape[0] == 3 and int(good[0]["stateSetAndFlags"]) & N.TILE_FIRST
    bad = good[1:].copy()  # tiles 1 and 2: both wait (through chain 0) on a tile that is not in the list
    print(f"missing-predecessor tile list, {path} kernel, 1 ms wait limit", flush=True)
    views = cull.make_views([scenes.DEFAULT_CAMERA])
    vb = DeviceBuffer.from_array(views)
    good_b, bad_b = DeviceBuffer.from_array(good), DeviceBuffer.from_array(bad)
    out = cull.CullOutputs(1, ds.n, ds.K)
    ws = cull.Workspace(16, 1)
    ws.set_wait_limit(1000)
    lib = N.load()

    def launch(tb, nt):
        jobs = _job(ds, tb, nt, vb, 1, out)
        jb = DeviceBuffer.from_array(jobs)
        if path == "view":
            rc = lib.cadrCullView(ctypes.byref(config), jb.address, N.ptr(jobs), ws.handle, None)
        else:
            rc = lib.cadrCullBatch(ctypes.byref(config), jb.address, N.ptr(jobs), 1, ws.handle, None)
        N.check(rc)
        synchronize()
        jb.close()

    exp_cmds, exp_counts = oracle.cull(spheres, scene.mesh_index, scene.meshes, scene.stateset_counts, views,
                                       cmd_dtype=N.CMD_DTYPE)
    # the malformed list ends and is reported once
    launch(bad_b, bad.shape[0])
    with pytest.raises(N.CadrError) as e:
        ws.check()
    assert e.value.code == N.ERROR_TIMEOUT
    ws.check()  # the record was cleared
    # a well-formed launch after it is exact and reports nothing
    launch(good_b, good.shape[0])
    ws.check()
    cmds, counts = out.download()
    assert np.array_equal(counts, exp_counts)
    n = int(exp_counts[0, 0])
    assert cmds[0, :n].tobytes() == exp_cmds[0, :n].tobytes()
    # the record is sticky across later good launches until checked
    launch(bad_b, bad.shape[0])
    launch(good_b, good.shape[0])
    cmds, counts = out.download()
    assert np.array_equal(counts, exp_counts) and cmds[0, :n].tobytes() == exp_cmds[0, :n].tobytes()
    with pytest.raises(N.CadrError) as e:
        ws.check()
    assert e.value.code == N.ERROR_TIMEOUT
    for b in (vb, good_b, bad_b):
        b.close()
    out.close()
    ws.close()


def test_missing_predecessor_in_the_persistent_view_kernel(gpu):
    """The same through the persistent one-view kernel, which runs when a
    view has more tiles than resident workgroups (800k drawables in tiles
    of 1024; two 1024-thread workgroups fit per CU): a list without its first
    tile ends after the 1-ms limit and is reported; the full list then gives
    the oracle's draw list on the same workspace."""
    import ctypes
    scene = scenes.plain_scene(800_000, seed=scenes.SEED_BASE + 23)
    ds = cull.DeviceScene(scene)
    ds.propagate()
    spheres = ds.download_spheres()
    config = cull.make_config(1024, 1)
    good = cull.build_tiles(scene.stateset_counts, cull.tile_size(config))
    assert good.shape[0] > 2 * 256
    bad = good[1:].copy()
    print(f"missing-predecessor tile list, persistent view kernel, {bad.shape[0]} tiles", flush=True)
    views = cull.make_views([scenes.DEFAULT_CAMERA])
    vb = DeviceBuffer.from_array(views)
    good_b, bad_b = DeviceBuffer.from_array(good), DeviceBuffer.from_array(bad)
    out = cull.CullOutputs(1, ds.n, ds.K)
    ws = cull.Workspace(good.shape[0], 1)
    ws.set_wait_limit(1000)
    lib = N.load()

    def launch(tb, nt):
        jobs = _job(ds, tb, nt, vb, 1, out)
        jb = DeviceBuffer.from_array(jobs)
        N.check(lib.cadrCullView(ctypes.byref(config), jb.address, N.ptr(jobs), ws.handle, None))
        synchronize()
        jb.close()

    launch(bad_b, bad.shape[0])
    with pytest.raises(N.CadrError) as e:
        ws.check()
    assert e.value.code == N.ERROR_TIMEOUT
    launch(good_b, good.shape[0])
    ws.check()
    cmds, counts = out.download()
    exp_cmds, exp_counts = oracle.cull(spheres, scene.mesh_index, scene.meshes, scene.stateset_counts, views,
                                       cmd_dtype=N.CMD_DTYPE)
    assert np.array_equal(counts, exp_counts)
    n = int(exp_counts[0, 0])
    assert n > 0 and cmds[0, :n].tobytes() == exp_cmds[0, :n].tobytes()
    for b in (vb, good_b, bad_b):
        b.close()
    out.close()
    ws.close()
    ds.close()
