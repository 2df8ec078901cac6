"""Autotuner (cadrCullAutotune): every candidate agrees, the best one is the
fastest agreeing candidate, and the outputs it leaves equal the oracle's."""
import numpy as np
import pytest

import oracle
from cadr_amd import _native as N
from cadr_amd import cull, scenes
from cadr_amd.tune import autotune

pytestmark = pytest.mark.gpu


def _compare(out, sc, sph, views):
    exp_cmds, exp_counts = oracle.cull(sph, sc.mesh_index, sc.meshes, sc.stateset_counts, views,
                                       cmd_dtype=N.CMD_DTYPE, nthreads=8)
    cmds, counts = out.download()
    assert np.array_equal(counts, exp_counts)
    for a, b in zip(cull.valid_commands(cmds, counts, sc.stateset_counts),
                    cull.valid_commands(exp_cmds, exp_counts, sc.stateset_counts)):
        for x, y in zip(a, b):
            assert x.tobytes() == y.tobytes()


def test_autotune_batched_and_single_view(gpu):
    sc = scenes.tree_scene((300, 200), seed=scenes.SEED_BASE + 21, num_statesets=64, lod_count=3, trange=200.0)
    ds = cull.DeviceScene(sc)
    ds.propagate()
    _, sph = oracle.propagate_scene(sc)
    views = cull.make_views(scenes.random_cameras(scenes.SEED_BASE + 21, 20, 0, 40, box=300.0))
    out = cull.CullOutputs(views.shape[0], ds.n, ds.K)
    best, table = autotune([(ds, views, out)], reps=2)
    print(table)
    assert len(table) == 16 and all(t["agrees"] for t in table)
    fastest = min(table, key=lambda t: t["ms"])
    assert (best.blockSize, best.itemsPerThread, best.flags) == (fastest["block"], fastest["items"], fastest["flags"])
    _compare(out, sc, sph, views)

    one = views[3:4]
    out1 = cull.CullOutputs(1, ds.n, ds.K)
    best1, table1 = autotune([(ds, one, out1)], single_view=True, reps=2)
    # the one-view kernel at 256, 512 and 1024 threads
    assert len(table1) == 13 and all(t["agrees"] for t in table1)
    assert {(t["block"], t["items"]) for t in table1} == {(b, i) for b in (256, 512) for i in (2, 4, 8, 16)} | {
        (1024, i) for i in (1, 2, 4, 8, 16)}
    fastest1 = min(table1, key=lambda t: t["ms"])
    assert (best1.blockSize, best1.itemsPerThread) == (fastest1["block"], fastest1["items"])
    _compare(out1, sc, sph, one)

    # explicit candidates, and an unsupported one is refused
    best2, table2 = autotune([(ds, views, out)], candidates=[(256, 4, 0, 0), (512, 4, 2, 0)], reps=1)
    assert len(table2) == 2
    with pytest.raises(N.CadrError):
        autotune([(ds, views, out)], candidates=[(256, 3, 0, 0)])
    out.close(); out1.close(); ds.close()
