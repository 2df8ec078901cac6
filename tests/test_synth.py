"""Device scene generator (include/cadr_synth.h): the permutation is a
bijection, and generated levels equal the host generator's arrays."""
import numpy as np
import pytest

import oracle
from cadr_amd import _native as N
from cadr_amd import cull, scenes


def _perm(n, key):
    lib = N.load()
    return np.array([lib.cadrSynthPermute(i, n, key) for i in range(n)], dtype=np.uint64)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 64, 1000, 4097, 65536])
def test_permutation_is_a_bijection(n):
    p = _perm(n, 0xCAD01234)
    assert np.array_equal(np.sort(p), np.arange(n, dtype=np.uint64))
    if n >= 1000:
        # not the identity, and not a shift: well mixed
        assert np.mean(p == np.arange(n)) < 0.01
        assert abs(np.corrcoef(p.astype(np.float64), np.arange(n))[0, 1]) < 0.1


def test_permutation_key_zero_is_identity():
    lib = N.load()
    assert [lib.cadrSynthPermute(i, 100, 0) for i in range(100)] == list(range(100))


def test_balanced_counts():
    c = scenes.balanced_counts(1003, 64)
    assert c.sum() == 1003 and c.max() - c.min() <= 1
    assert scenes.balanced_counts(5, 64).sum() == 5


def _ulps(a, b):
    ai = a.view(np.int32).astype(np.int64)
    bi = b.view(np.int32).astype(np.int64)
    ai = np.where(ai < 0, np.int64(-(2 ** 31)) - ai, ai)
    bi = np.where(bi < 0, np.int64(-(2 ** 31)) - bi, bi)
    return np.abs(ai - bi)


@pytest.mark.gpu
def test_synth_tree_matches_host_generator(gpu):
    spec = scenes.uneven_synth(3, 40_000, 3, 2, 2)
    ds = cull.DeviceScene(spec)
    perm = _perm(spec.num_drawables, spec.seed)
    for d, (cnt, lb, pb, wb) in enumerate(ds.levels):
        last = d == len(ds.levels) - 1
        items = perm if last else np.arange(cnt, dtype=np.uint64)
        host = scenes.rotation_translation(spec.seed, 10 + d, 0, cnt, spec.trange, items=items)
        dev = lb.download(np.float32, cnt * 16).reshape(cnt, 16)
        u = _ulps(dev, host)
        assert u.max() <= 1 and np.mean(u == 0) > 0.999, f"level {d}: {u.max()} ulp"
        if d > 0:
            parent = pb.download(np.uint32, cnt)
            assert np.array_equal(parent, (items // np.uint64(spec.fanout)).astype(np.uint32))
    mesh = ds.mesh_index.download(np.uint32, ds.n)
    assert np.array_equal(mesh, scenes.rand_int(spec.seed, 3, 0, ds.n, 0, 63)[perm.astype(np.int64)].astype(np.uint32))
    # the generated tree propagates like any other
    ds.propagate()
    levels = []
    for d, (cnt, lb, pb, wb) in enumerate(ds.levels):
        levels.append((lb.download(np.float32, cnt * 16).reshape(cnt, 16), None if pb is None else pb.download(np.uint32, cnt)))
    host_scene = scenes.Scene("t", levels, mesh, spec.meshes, spec.stateset_counts)
    _, sph = oracle.propagate_scene(host_scene)
    assert ds.download_spheres().tobytes() == sph.tobytes()
    ds.close()


@pytest.mark.gpu
def test_synth_flat_matches_host_generator(gpu):
    spec = scenes.served_synth(7, 30_000)
    ds = cull.DeviceScene(spec)
    n = ds.n
    mesh = ds.mesh_index.download(np.uint32, n)
    assert np.array_equal(mesh, scenes.rand_int(spec.seed, 3, 0, n, 0, 63).astype(np.uint32))
    host_world = scenes.scale_translation(spec.seed, 4, 0, n)
    dev = ds.download_spheres()
    _, exp = oracle.propagate_level(host_world, None, None, mesh, spec.meshes)
    u = _ulps(dev, exp)
    assert u.max() <= 1 and np.mean(u == 0) > 0.999
    ds.close()
