"""Parity at the workloads' stated sizes: the GPU pass against the CPU oracle
at 10M drawables (one view, fused and unfused), 10M drawables x 256 views
(every view's counts and draw lists, including the last view, whose output
slice starts about 51 GB into the command buffer), the 64-scene uneven batch
(52.8M drawables, 1,106 views, unfused and fused frames, scene by scene) and
a sample of 500 requests taken from the 128-scene served set under a
50,000 req/s load.  Everything is bit-exact: counts, compacted commands,
world matrices and world spheres.

Scenes generated in device memory (uneven, served) are read back and given
to the oracle as its input, so the oracle checks the pass, not the
generator."""
import ctypes
import os
import time

import numpy as np
import pytest

import oracle
from cadr_amd import _native as N
from cadr_amd import cull, scenes
from cadr_amd.device import DeviceBuffer, synchronize

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

THREADS = max(1, min(16, os.cpu_count() or 1))


def _log(msg):
    print(msg, flush=True)


def _bases(stateset_counts):
    return np.concatenate([[0], np.cumsum(np.asarray(stateset_counts, np.int64))])[:-1]


def _check_view(got_cmds, got_counts, exp_cmds, exp_counts, bases, what):
    """Counts equal, and the compacted commands of every StateSet equal byte
    for byte (slots past a StateSet's count are unspecified)."""
    if not np.array_equal(got_counts, exp_counts):
        bad = np.argwhere(got_counts != exp_counts)[:5, 0]
        raise AssertionError(f"{what}: counts differ at StateSets {bad}: {got_counts[bad]} vs {exp_counts[bad]}")
    g = np.concatenate([got_cmds[b:b + c] for b, c in zip(bases, got_counts)])
    e = np.concatenate([exp_cmds[b:b + c] for b, c in zip(bases, exp_counts)])
    if g.tobytes() != e.tobytes():
        i = int(np.argwhere(g != e)[0, 0])
        raise AssertionError(f"{what}: command {i} of the concatenated lists differs: {g[i]} vs {e[i]}")


def _download_view(out, v, n):
    return out.commands.download(N.CMD_DTYPE, n, offset=20 * v * out.stride)


@pytest.fixture(scope="module")
def full10m(gpu):
    oracle.set_threads(THREADS)
    t0 = time.time()
    scene = scenes.full10m_scene()
    assert scene.mesh_index.shape[0] == 10_000_000
    ds = cull.DeviceScene(scene)
    worlds, spheres = oracle.propagate_scene(scene)
    _log(f"full10m: 1000 x 100 x 100 tree generated, uploaded and propagated by the oracle in {time.time() - t0:.1f}s")
    yield scene, ds, worlds, spheres
    ds.close()


def test_full10m_unfused_and_fused_frames(full10m):
    scene, ds, worlds, spheres = full10m
    lib = N.load()
    views = cull.make_views([scenes.DEFAULT_CAMERA])
    exp_cmds, exp_counts = oracle.cull(spheres, scene.mesh_index, scene.meshes, scene.stateset_counts, views,
                                       cmd_dtype=N.CMD_DTYPE, nthreads=THREADS)
    bases = _bases(scene.stateset_counts)
    assert exp_counts.sum() > 1_000_000

    # unfused: every level on its own (leaf level writes world spheres), then the one-view cull
    _log("full10m unfused: propagation by level + cullViewKernel<256,16> and <1024,8>")
    ds.propagate()
    out = cull.CullOutputs(1, ds.n, ds.K)
    synchronize()
    assert ds.download_spheres().tobytes() == spheres.tobytes(), "world spheres differ"
    for d in range(3):
        assert ds.world_matrices(d).tobytes() == worlds[d].tobytes(), f"level {d} world matrices differ"
    for shape in ((256, 16), (1024, 8)):
        out.commands.fill_u32(0)
        b = cull.CullBatch(cull.make_config(*shape), [(ds, views, out)])
        b.run_view()
        synchronize()
        b.workspace.check()
        cmds, counts = out.download()
        _check_view(cmds[0], counts[0], exp_cmds[0], exp_counts[0], bases, f"full10m unfused {shape}")
        b.close()

    # fused: root and inner level, then the leaf level fused with the cull (the headline frame's launches)
    for shape in ((512, 4), (256, 8)):
        _log(f"full10m fused: cullFusedKernel<{shape[0]},{shape[1]}>")
        ds.levels[-1][3].fill_u32(0xFFFFFFFF)  # leaf world matrices: the fused kernel must write them all
        out.commands.fill_u32(0)
        descs = ds.level_descs(spheres=False)
        for lv in descs[:-1]:
            N.check(lib.cadrPropagateTransforms(N.ptr(lv), None))
        fb = cull.CullBatch(cull.make_config(*shape), [(ds, views, out)])
        fb.run_fused(descs[-1])
        synchronize()
        fb.workspace.check()
        assert ds.world_matrices(-1).tobytes() == worlds[-1].tobytes(), "fused leaf world matrices differ"
        cmds, counts = out.download()
        _check_view(cmds[0], counts[0], exp_cmds[0], exp_counts[0], bases, f"full10m fused {shape}")
        fb.close()
    out.close()


def test_batch256_every_view(full10m):
    """10M drawables x 256 views in one persistent launch (and with the leaf
    level fused into it): all 256 views' counts and draw lists against the
    oracle, in groups of 16 views so host memory stays bounded."""
    scene, ds, worlds, spheres = full10m
    ds.propagate()
    views = cull.make_views(scenes.batch_cameras(256))
    bases = _bases(scene.stateset_counts)
    out = cull.CullOutputs(256, ds.n, ds.K)  # 51.2 GB of commands
    assert out.commands.nbytes == 256 * 10_000_000 * 20
    batch = cull.CullBatch(cull.make_config(256, 8), [(ds, views, out)])
    _log("batch256: cullKernel<256,8> over 256 views")
    batch.run()
    synchronize()
    batch.workspace.check()
    counts_all = out.counts.download(np.uint32, 256 * ds.K).reshape(256, ds.K)
    group = 16
    for g0 in range(0, 256, group):
        exp_cmds, exp_counts = oracle.cull(spheres, scene.mesh_index, scene.meshes, scene.stateset_counts,
                                           views[g0:g0 + group], cmd_dtype=N.CMD_DTYPE, nthreads=THREADS)
        for i in range(group):
            v = g0 + i
            _check_view(_download_view(out, v, ds.n), counts_all[v], exp_cmds[i], exp_counts[i], bases,
                        f"batch256 view {v}")
        if g0 == 0:
            first_group = (exp_cmds[:2].copy(), exp_counts.copy())
        last_group = (exp_cmds[-1].copy(), exp_counts.copy())
        del exp_cmds
    _log(f"batch256: 256 views x 10M bit-exact, {int(counts_all.sum())} visible in total")

    # the leaf level fused into the batched launch: world matrices, all counts, views 0, 1 and 255
    ds.levels[-1][3].fill_u32(0xFFFFFFFF)
    out.counts.fill_u32(0)
    lib = N.load()
    descs = ds.level_descs(spheres=False)
    for lv in descs[:-1]:
        N.check(lib.cadrPropagateTransforms(N.ptr(lv), None))
    batch.set_leaves(descs[-1])
    _log("batch256: cullKernel<256,8,fused>")
    batch.run_fused_batch()
    synchronize()
    batch.workspace.check()
    assert ds.world_matrices(-1).tobytes() == worlds[-1].tobytes(), "fused batched world matrices differ"
    counts_f = out.counts.download(np.uint32, 256 * ds.K).reshape(256, ds.K)
    assert np.array_equal(counts_f, counts_all), "fused batched counts differ"
    for i, v in enumerate((0, 1)):
        _check_view(_download_view(out, v, ds.n), counts_f[v], first_group[0][i], first_group[1][i], bases,
                    f"batch256 fused view {v}")
    _check_view(_download_view(out, 255, ds.n), counts_f[255], last_group[0], last_group[1][-1], bases,
                "batch256 fused view 255")
    batch.close()
    out.close()


def _host_levels(ds):
    """Local matrices and parent indices of a device-generated tree, read back."""
    levels = []
    for d, (cnt, lb, pb, wb) in enumerate(ds.levels):
        local = lb.download(np.float32, 16 * cnt).reshape(cnt, 16)
        parent = None if pb is None else pb.download(np.uint32, cnt)
        levels.append((local, parent))
    return levels


def test_uneven64_scene_by_scene(gpu):
    """The 64-scene uneven batch: every scene propagated, all 1,106 views of
    all scenes culled in one persistent launch (unfused), and again with the
    leaf levels fused into that launch; each scene's world spheres, counts and
    draw lists against the oracle run on that scene's read-back inputs."""
    lib = N.load()
    oracle.set_threads(THREADS)
    t0 = time.time()
    specs = scenes.uneven_specs()
    dss, entries_a, entries_b, views_l = [], [], [], []
    for i, (n, depth, lods, nv) in enumerate(specs):
        ds = cull.DeviceScene(scenes.uneven_synth(i, n, depth, lods, nv))
        views = cull.make_views(ds.scene.cameras)
        dss.append(ds)
        views_l.append(views)
        entries_a.append((ds, views, cull.CullOutputs(nv, n, ds.K)))
        entries_b.append((ds, views, cull.CullOutputs(nv, n, ds.K)))
    total = sum(ds.n for ds in dss)
    nviews = sum(v.shape[0] for v in views_l)
    assert len(dss) == 64 and nviews > 1000
    _log(f"uneven64: {total} drawables, {nviews} views generated in {time.time() - t0:.1f}s")

    # unfused frame
    levels = [ds.level_descs(spheres=True) for ds in dss]
    for lvs in levels:
        for lv in lvs:
            N.check(lib.cadrPropagateTransforms(N.ptr(lv), None))
    batch_a = cull.CullBatch(cull.make_config(256, 8), entries_a)
    batch_a.run()
    synchronize()
    batch_a.workspace.check()
    # fused frame: inner levels, then every job's leaf level fused into the batched launch
    for ds in dss:
        ds.levels[-1][3].fill_u32(0xFFFFFFFF)
    for lvs in levels:
        for lv in lvs[:-1]:
            N.check(lib.cadrPropagateTransforms(N.ptr(lv), None))
    batch_b = cull.CullBatch(cull.make_config(256, 8), entries_b)
    batch_b.set_leaves(np.concatenate([ds.level_descs(spheres=False)[-1] for ds in dss]))
    batch_b.run_fused_batch()
    synchronize()
    batch_b.workspace.check()
    _log("uneven64: unfused and fused frames done on the GPU")

    checked = 0
    for j, ds in enumerate(dss):
        sc = ds.scene
        host = _host_levels(ds)
        mesh_index = ds.mesh_index.download(np.uint32, ds.n)
        prev = None
        for d, (local, parent) in enumerate(host):
            last = d == len(host) - 1
            w, sph = oracle.propagate_level(local, parent, prev, mesh_index if last else None,
                                            sc.meshes if last else None)
            prev = w
        assert ds.download_spheres().tobytes() == sph.tobytes(), f"scene {j}: world spheres differ"
        assert ds.world_matrices(-1).tobytes() == w.tobytes(), f"scene {j}: fused leaf world matrices differ"
        views = views_l[j]
        exp_cmds, exp_counts = oracle.cull(sph, mesh_index, sc.meshes, ds.stateset_counts, views,
                                           cmd_dtype=N.CMD_DTYPE, nthreads=THREADS)
        bases = _bases(ds.stateset_counts)
        for which, (_, _, out) in (("unfused", entries_a[j]), ("fused", entries_b[j])):
            cmds, counts = out.download()
            for v in range(views.shape[0]):
                _check_view(cmds[v], counts[v], exp_cmds[v], exp_counts[v], bases,
                            f"uneven64 scene {j} ({ds.n} drawables, depth {len(host)}) view {v} {which}")
        checked += views.shape[0]
        del exp_cmds
    assert checked == nviews
    batch_a.close()
    batch_b.close()
    for ds, _, out in entries_a:
        out.close()
    for ds, _, out in entries_b:
        out.close()
        ds.close()


def test_served128_sample_under_load(gpu):
    """128 resident scenes (3.4e8 drawables) served under 50,000 req/s Poisson
    arrivals (80% to the 16 hottest scenes); 500 requests kept from the run,
    each checked against the oracle for its scene and view."""
    from cadr_amd.serve import Server
    t0 = time.time()
    sizes = scenes.served_specs()
    dss = [cull.DeviceScene(scenes.served_synth(i, n)) for i, n in enumerate(sizes)]
    synchronize()
    views = cull.make_views(scenes.batch_cameras(256))
    _log(f"served128: {sum(sizes)} drawables generated in {time.time() - t0:.1f}s")
    srv = Server(block=256, items=8, max_batch_views=256, inflight=4, budget_bytes=160 << 30)
    for ds in dss:
        srv.add_scene(ds)
    srv.start()
    st = srv.run_load(views, rate=50_000.0, duration=1.0, seed=50_000, hot=16, hot_fraction=0.8, sample_every=90,
                      sample_capacity=500)
    samples = st.pop("samples")
    _log(f"served128 @ 50k req/s: {st['completed']} completed, p99 {st['p99Ms']:.3f} ms, "
         f"{st['meanViewsPerLaunch']:.1f} views/launch, {len(samples)} kept")
    assert st["failed"] == 0 and st["refused"] == 0 and st["completed"] == st["submitted"] > 40_000
    assert len(samples) == 500
    host_spheres = {}
    scenes_seen = set()
    for r in samples:
        s, v = r.sceneId, r.tag
        ds = dss[s]
        assert r.status == 0
        if s not in host_spheres:
            host_spheres[s] = (ds.download_spheres(), ds.mesh_index.download(np.uint32, ds.n))
        sph, mi = host_spheres[s]
        exp_cmds, exp_counts = oracle.cull(sph, mi, ds.scene.meshes, ds.stateset_counts, views[v:v + 1],
                                           cmd_dtype=N.CMD_DTYPE, nthreads=THREADS)
        cmds = np.empty(ds.n, N.CMD_DTYPE)
        counts = np.empty(ds.K, np.uint32)
        N.check(N.load().cadrMemcpyDtoH(cmds.ctypes.data, r.commands, cmds.nbytes, None))
        N.check(N.load().cadrMemcpyDtoH(counts.ctypes.data, r.counts, counts.nbytes, None))
        N.check(N.load().cadrStreamSynchronize(None))
        assert r.visible == counts.sum()
        _check_view(cmds, counts, exp_cmds[0], exp_counts[0], _bases(ds.stateset_counts),
                    f"served128 request {r.requestId} (scene {s}, {ds.n} drawables, view {v})")
        srv.release(r.requestId)
        scenes_seen.add(s)
    _log(f"served128: 500 sampled requests bit-exact over {len(scenes_seen)} scenes")
    assert len(scenes_seen) >= 16
    srv.drain()
    srv.close()
    for ds in dss:
        ds.close()
