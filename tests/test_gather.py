"""Gather pass (processDrawables.comp:92-113): oracle pinning on the CPU and
kernel parity on the GPU, handle levels 1-3, table boundaries and dispatch
sizes around the reference's 32768-workgroup split (Renderer.cpp:684-692)."""
import ctypes

import numpy as np
import pytest

import oracle
from cadr_amd import _native as N
from gather_fixtures import build_gather_scene


def _oracle_run(img, push, n, level):
    rc = oracle.process_drawables([(img.base, img.buf)], push, level, n)
    ind = img.read(push[2], np.uint32, 4 * n).reshape(n, 4)
    ptr = img.read(push[3], np.uint64, 4 * n).reshape(n, 4)
    return rc, ind, ptr


@pytest.mark.parametrize("level,n", [(1, 1), (1, 500), (2, 0), (2, 1), (2, 3000), (3, 2000)])
def test_oracle_gather_matches_fixture_truth(level, n):
    img, push, exp_ind, exp_ptr = build_gather_scene(1 << 40, level, n, np.random.default_rng(level * 100 + n))
    rc, ind, ptr = _oracle_run(img, push, n, level)
    assert rc == 0
    assert np.array_equal(ind, exp_ind)
    assert np.array_equal(ptr, exp_ptr)


def test_oracle_lookup_levels_closed_form():
    """Handle h at level 2 is table[h>>11][h&0x7ff]; level 3 table[h>>22][(h>>11)&0x7ff][h&0x7ff]."""
    from gather_fixtures import HandleTables, Image
    for level, h in [(1, 2047), (2, 2048), (2, 4194303), (3, 4194304), (3, (7 << 22) + (2047 << 11) + 1)]:
        img = Image(0x10000, 1 << 20)
        t = HandleTables(img, level)
        t.set(h, 0xABCDEF00)
        out = ctypes.c_uint64(0)
        reg = (oracle.Region * 1)(oracle.Region(img.base, img.buf.nbytes, img.buf.ctypes.data))
        assert oracle.lib().oracle_lookup_handle(reg, 1, t.root, level, h, ctypes.byref(out)) == 0
        assert out.value == 0xABCDEF00
        # a neighbouring handle in an absent table resolves to 0
        assert oracle.lib().oracle_lookup_handle(reg, 1, t.root, level, h ^ (1 << 11) if level > 1 else 3,
                                                 ctypes.byref(out)) == 0
        assert out.value == 0


def _gpu_gather(level, n, seed, **kw):
    from cadr_amd.device import DeviceBuffer, synchronize
    probe, *_ = build_gather_scene(0, level, n, np.random.default_rng(seed), **kw)
    buf = DeviceBuffer(probe.buf.size)
    img, push, exp_ind, exp_ptr = build_gather_scene(buf.address, level, n, np.random.default_rng(seed),
                                                     image_size=probe.buf.size, **kw)
    buf.upload(img.buf)
    pc = N.PushConstants(*push)
    N.check(N.load().cadrProcessDrawables(ctypes.byref(pc), level, n, None))
    synchronize()
    got = buf.download(np.uint8)
    base = buf.address
    ind = got[push[2] - base: push[2] - base + 16 * n].view(np.uint32).reshape(n, 4)
    ptr = got[push[3] - base: push[3] - base + 32 * n].view(np.uint64).reshape(n, 4)
    rc, o_ind, o_ptr = _oracle_run(img, push, n, level)
    buf.close()
    return rc, ind, ptr, o_ind, o_ptr, exp_ind, exp_ptr


@pytest.mark.gpu
@pytest.mark.parametrize("level,n", [(1, 1), (1, 600), (2, 255), (2, 256), (2, 257), (2, 32768), (2, 32769),
                                     (3, 65537), (3, 100_000)])
def test_gpu_gather_parity(gpu, level, n):
    print(f"gather level={level} n={n}", flush=True)
    rc, ind, ptr, o_ind, o_ptr, exp_ind, exp_ptr = _gpu_gather(level, n, 1000 + level * 7 + n,
                                                               with_drawable_data=(level != 1 or n < 600))
    assert rc == 0
    assert np.array_equal(o_ind, exp_ind) and np.array_equal(o_ptr, exp_ptr)
    assert np.array_equal(ind, o_ind), f"indirect mismatch at {np.argwhere(ind != o_ind)[:5]}"
    assert np.array_equal(ptr, o_ptr), f"pointers mismatch at {np.argwhere(ptr != o_ptr)[:5]}"


@pytest.mark.gpu
def test_gpu_gather_zero_and_errors(gpu):
    pc = N.PushConstants(0, 0, 0, 0)
    assert N.load().cadrProcessDrawables(ctypes.byref(pc), 2, 0, None) == N.SUCCESS  # no-op
    assert N.load().cadrProcessDrawables(ctypes.byref(pc), 2, 5, None) == N.ERROR_INVALID_VALUE
    pc = N.PushConstants(64, 64, 64, 64)
    assert N.load().cadrProcessDrawables(ctypes.byref(pc), 4, 5, None) == N.ERROR_INVALID_VALUE
    pc = N.PushConstants(64, 72, 64, 64)  # misaligned drawable list
    assert N.load().cadrProcessDrawables(ctypes.byref(pc), 2, 5, None) == N.ERROR_INVALID_VALUE


@pytest.mark.gpu
def test_gpu_scatter_drawables(gpu):
    """cadrScatterDrawables: list[slots[i]] = records[i]; slots past the list
    are skipped; untouched slots keep their bytes; count 0 is a no-op."""
    from cadr_amd.device import DeviceBuffer, synchronize
    lib = N.load()
    rng = np.random.default_rng(21)
    slots_total = 70_000
    before = rng.integers(0, 2**32, (slots_total, 12), dtype=np.uint64).astype(np.uint32)
    lst = DeviceBuffer.from_array(before)
    count = 20_000
    slots = rng.choice(slots_total, count - 3, replace=False).astype(np.uint32)
    slots = np.concatenate([slots, np.array([slots_total, slots_total + 5, 0xFFFFFFFF], np.uint32)])
    records = rng.integers(0, 2**32, (count, 12), dtype=np.uint64).astype(np.uint32)
    dr, ds_ = DeviceBuffer.from_array(records), DeviceBuffer.from_array(slots)
    assert lib.cadrScatterDrawables(lst.address, slots_total, dr.address, ds_.address, 0, None) == N.SUCCESS
    N.check(lib.cadrScatterDrawables(lst.address, slots_total, dr.address, ds_.address, count, None))
    synchronize()
    got = lst.download(np.uint32).reshape(slots_total, 12)
    exp = before.copy()
    exp[slots[:-3]] = records[:-3]
    assert np.array_equal(got, exp)
    assert lib.cadrScatterDrawables(0, slots_total, dr.address, ds_.address, 5, None) == N.ERROR_INVALID_VALUE
    assert lib.cadrScatterDrawables(lst.address + 8, slots_total, dr.address, ds_.address, 5, None) == \
        N.ERROR_INVALID_VALUE
    for b in (lst, dr, ds_):
        b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n,first_handle", [(300_000, 1), (2_000_000, 1 << 22), (1_000_000, (1 << 22) - 5000)])
def test_gpu_gather_bulk_scene(gpu, n, first_handle):
    """Large consecutive-handle scenes (the reference's numbering) at levels 2 and
    3, including a level-2/3 table boundary inside the scene: every output
    record against the closed form."""
    from cadr_amd import scenes
    from cadr_amd.device import DeviceBuffer, synchronize
    g = scenes.GatherScene(n, first_handle=first_handle)
    buf = DeviceBuffer(g.layout(0)["size"])
    img, push, level = g.image(buf.address)
    buf.upload(img)
    ind, ptr = DeviceBuffer(16 * n), DeviceBuffer(32 * n)
    pc = N.PushConstants(push[0], push[1], ind.address, ptr.address)
    N.check(N.load().cadrProcessDrawables(ctypes.byref(pc), level, n, None))
    synchronize()
    ei, ep = g.expected(buf.address)
    assert np.array_equal(ind.download(np.uint32, 4 * n).reshape(n, 4), ei)
    assert np.array_equal(ptr.download(np.uint64, 4 * n).reshape(n, 4), ep)
    for b in (buf, ind, ptr):
        b.close()


def test_gather_scene_generator_matches_oracle():
    """The bulk generator's closed form agrees with the oracle at every level."""
    from cadr_amd import scenes
    for n, h0 in ((1000, 1), (50_000, 1), (30_000, 1 << 22), (20_000, (1 << 22) - 100)):
        g = scenes.GatherScene(n, first_handle=h0)
        base, ib, pb = 1 << 44, 1 << 45, 3 << 44
        img, push, level = g.image(base)
        ind, ptr = np.zeros((n, 4), np.uint32), np.zeros((n, 4), np.uint64)
        rc = oracle.process_drawables([(base, img), (ib, ind.view(np.uint8).reshape(-1)),
                                       (pb, ptr.view(np.uint8).reshape(-1))], (push[0], push[1], ib, pb), level, n)
        assert rc == 0
        ei, ep = g.expected(base)
        assert np.array_equal(ind, ei) and np.array_equal(ptr, ep), f"n={n} h0={h0}"


@pytest.mark.gpu
def test_gpu_compacted_drawable_pointers(gpu):
    """The reference-layout companion of the cull: the gather writes every
    drawable's DrawablePointers, the cull compacts the visible ones per
    (view, StateSet), and cadrCompactDrawablePointers lays the records out
    beside the commands, so DrawablePointers[gl_DrawID] of a StateSet's draw
    is its own drawable's record.  Checked against the gather's closed form
    and the oracle's command lists; slots past each count stay untouched."""
    from cadr_amd import cull, scenes
    from cadr_amd.device import DeviceBuffer, synchronize
    n, K = 60_000, 9
    g = scenes.GatherScene(n, first_handle=1)
    buf = DeviceBuffer(g.layout(0)["size"])
    img, push, level = g.image(buf.address)
    buf.upload(img)
    ind, ptr = DeviceBuffer(16 * n), DeviceBuffer(32 * n)
    pc = N.PushConstants(push[0], push[1], ind.address, ptr.address)
    N.check(N.load().cadrProcessDrawables(ctypes.byref(pc), level, n, None))
    _, records = g.expected(buf.address)  # uint64 [n, 4], the gather's closed form

    sc = scenes.plain_scene(n, seed=scenes.SEED_BASE + 61, num_statesets=K)
    sc.mesh_index[::97] = 1000  # invalid meshes: never visible
    cams = scenes.random_cameras(scenes.SEED_BASE + 62, 20, 0, 5, box=300.0) + [scenes.DEFAULT_CAMERA]
    views = cull.make_views(cams)
    _, sph = oracle.propagate_scene(sc)
    exp_cmds, exp_counts = oracle.cull(sph, sc.mesh_index, sc.meshes, sc.stateset_counts, views, cmd_dtype=N.CMD_DTYPE)
    assert exp_counts.sum() > 1000
    ds = cull.DeviceScene(sc)
    ds.propagate()
    V, stride = views.shape[0], n + 3
    out = cull.CullOutputs(V, stride, K)
    b = cull.CullBatch(cull.make_config(256, 8), [(ds, views, out)])
    b.run()
    base = np.concatenate([[0], np.cumsum(sc.stateset_counts)[:-1]]).astype(np.uint32)
    bb = DeviceBuffer.from_array(base)
    po = DeviceBuffer(32 * V * stride)
    po.fill_u32(0xA5A5A5A5)
    N.check(N.load().cadrCompactDrawablePointers(out.commands.address, out.counts.address, bb.address, V, K, stride,
                                                 ptr.address, n, po.address, None))
    synchronize()
    cmds, counts = out.download()
    assert np.array_equal(counts, exp_counts)
    got = po.download(np.uint64, 4 * V * stride).reshape(V, stride, 4)
    sentinel = np.uint64(0xA5A5A5A5A5A5A5A5)
    for w in range(V):
        for k in range(K):
            b0, c = int(base[k]), int(counts[w, k])
            seg = exp_cmds[w, b0:b0 + c]
            assert np.array_equal(cmds[w, b0:b0 + c], seg)
            # the records of the visible drawables, in the order of their commands
            assert np.array_equal(got[w, b0:b0 + c], records[seg["firstInstance"].astype(np.int64)]), (w, k)
            end = b0 + int(sc.stateset_counts[k])
            assert np.all(got[w, b0 + c:end] == sentinel), (w, k)
    assert np.all(got[:, n:] == sentinel)
    # an index past the drawable count gives a zero record
    bad = np.zeros(1, N.CMD_DTYPE)
    bad["firstInstance"] = n + 5
    c1 = DeviceBuffer.from_array(np.concatenate([bad, np.zeros(stride - 1, N.CMD_DTYPE)]))
    cnt = DeviceBuffer.from_array(np.array([1] + [0] * (K - 1), np.uint32))
    N.check(N.load().cadrCompactDrawablePointers(c1.address, cnt.address, bb.address, 1, K, stride, ptr.address, n,
                                                 po.address, None))
    synchronize()
    assert np.all(po.download(np.uint64, 4) == 0)
    for x in (buf, ind, ptr, bb, po, c1, cnt, out, ds):
        x.close()
    b.close()
