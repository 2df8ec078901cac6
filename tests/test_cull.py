"""Transform propagation, world spheres and the cull / LOD / compaction pass:
GPU results against the CPU oracle, bit for bit."""
import numpy as np
import pytest

import oracle
from cadr_amd import _native as N
from cadr_amd import cull, scenes

pytestmark = pytest.mark.gpu


def _compare(cmds, counts, exp_cmds, exp_counts, stateset_counts, what=""):
    assert np.array_equal(counts, exp_counts), f"{what}: counts differ, first at {np.argwhere(counts != exp_counts)[:5]}"
    got = cull.valid_commands(cmds, counts, stateset_counts)
    exp = cull.valid_commands(exp_cmds, exp_counts, stateset_counts)
    for v in range(len(got)):
        for k in range(len(stateset_counts)):
            if got[v][k].tobytes() != exp[v][k].tobytes():
                a, b = got[v][k], exp[v][k]
                bad = np.argwhere(a != b)[:5]
                raise AssertionError(f"{what}: view {v} StateSet {k} commands differ at {bad}: {a[bad]} vs {b[bad]}")


def _run(ds, views, config, stride=None, single=False):
    out = cull.CullOutputs(views.shape[0], stride or ds.n, ds.K)
    b = cull.CullBatch(config, [(ds, views, out)])
    if single:
        b.run_view()
    else:
        b.run()
    from cadr_amd.device import synchronize
    synchronize()
    cmds, counts = out.download()
    b.close(); out.close()
    return cmds, counts


def test_propagation_and_spheres_bit_exact(gpu):
    scene = scenes.tree_scene((10, 100, 100), seed=scenes.SEED_BASE + 11)
    print("propagation 3-level 100k leaves", flush=True)
    ds = cull.DeviceScene(scene)
    ds.propagate()
    worlds, spheres = oracle.propagate_scene(scene)
    for d in range(3):
        assert ds.world_matrices(d).tobytes() == worlds[d].tobytes(), f"level {d} world matrices differ"
    got = ds.download_spheres()
    bad = np.argwhere(got.view(np.uint32) != spheres.view(np.uint32))
    assert bad.size == 0, f"spheres differ at {bad[:5]}: {got[bad[:5, 0]]} vs {spheres[bad[:5, 0]]}"
    ds.close()


def test_world_spheres_flat_bit_exact(gpu):
    from cadr_amd.device import DeviceBuffer, synchronize
    import ctypes
    scene = scenes.plain_scene(100_000)
    world = scene.levels[0][0]
    # include unusual matrices: shear, mirror, huge and tiny scales, an invalid mesh index
    rng = np.random.default_rng(5)
    world = world.copy()
    world[:1000] = (rng.standard_normal((1000, 16)) * 10).astype(np.float32)
    mi = scene.mesh_index.copy()
    mi[7] = 10_000
    wb, mb, meb = DeviceBuffer.from_array(world), DeviceBuffer.from_array(mi), DeviceBuffer.from_array(scene.meshes)
    sb = DeviceBuffer(16 * world.shape[0])
    N.check(N.load().cadrComputeWorldSpheres(wb.address, mb.address, meb.address, scene.meshes.shape[0],
                                             world.shape[0], sb.address, None))
    synchronize()
    got = sb.download(np.float32, 4 * world.shape[0]).reshape(-1, 4)
    _, exp = oracle.propagate_level(world, None, None, mi, scene.meshes)
    assert got.tobytes() == exp.tobytes()
    assert got[7, 3] == -np.inf


@pytest.fixture(scope="module")
def plain(gpu):
    scene = scenes.plain_scene(100_000)
    ds = cull.DeviceScene(scene)
    ds.propagate()
    _, spheres = oracle.propagate_scene(scene)
    assert ds.download_spheres().tobytes() == spheres.tobytes()
    yield scene, ds, spheres
    ds.close()


@pytest.mark.parametrize("block,items", [(256, 2), (256, 4), (256, 8), (256, 16), (512, 2), (512, 4), (128, 8),
                                         (128, 16)])
@pytest.mark.parametrize("flags", [0, N.CULL_NO_TILE_SPHERES])
@pytest.mark.parametrize("path", ["view", "batch"])
def test_cull_plain100k_single_view(plain, block, items, flags, path):
    scene, ds, spheres = plain
    print(f"plain100k {path} block={block} items={items} flags={flags}", flush=True)
    views = cull.make_views([scenes.DEFAULT_CAMERA])
    cmds, counts = _run(ds, views, cull.make_config(block, items, flags=flags), single=(path == "view"))
    exp_cmds, exp_counts = oracle.cull(spheres, scene.mesh_index, scene.meshes, scene.stateset_counts, views,
                                       cmd_dtype=N.CMD_DTYPE)
    assert exp_counts.sum() > 1000
    _compare(cmds, counts, exp_cmds, exp_counts, scene.stateset_counts, "plain100k")


@pytest.mark.parametrize("block,items", [(512, 8), (512, 16), (1024, 1), (1024, 2), (1024, 4), (1024, 8),
                                         (1024, 16)])
def test_cull_one_view_large_workgroups(plain, block, items):
    """The one-view kernel at 512 and 1024 threads (tiles up to 16384
    drawables), against the oracle; shapes with no batched kernel are
    refused by the batched entry point."""
    scene, ds, spheres = plain
    print(f"plain100k view block={block} items={items}", flush=True)
    views = cull.make_views([scenes.DEFAULT_CAMERA])
    cmds, counts = _run(ds, views, cull.make_config(block, items), single=True)
    exp_cmds, exp_counts = oracle.cull(spheres, scene.mesh_index, scene.meshes, scene.stateset_counts, views,
                                       cmd_dtype=N.CMD_DTYPE)
    _compare(cmds, counts, exp_cmds, exp_counts, scene.stateset_counts, f"plain100k {block}x{items}")
    with pytest.raises(N.CadrError):
        _run(ds, views, cull.make_config(block, items))


def test_cull_one_view_persistent_loop(gpu):
    """A view with more tiles than workgroups fit on the GPU runs the
    persistent one-view kernel (next tile's loads in flight during the
    current tile's look-back): 2M drawables in 8 StateSets, shapes whose
    tile count exceeds their resident workgroups, against the oracle, for
    the default camera and a random one."""
    scene = scenes.plain_scene(2_000_000, seed=scenes.SEED_BASE + 21, num_statesets=8)
    ds = cull.DeviceScene(scene)
    ds.propagate()
    spheres = ds.download_spheres()
    cams = [scenes.DEFAULT_CAMERA] + scenes.random_cameras(scenes.SEED_BASE + 21, 20, 0, 1, box=400.0)
    for c in cams:
        views = cull.make_views([c])
        exp_cmds, exp_counts = oracle.cull(spheres, scene.mesh_index, scene.meshes, scene.stateset_counts, views,
                                           cmd_dtype=N.CMD_DTYPE, nthreads=8)
        for block, items in [(256, 2), (256, 8), (512, 4), (1024, 1), (1024, 4)]:
            print(f"plain 2M view block={block} items={items} visible={int(exp_counts.sum())}", flush=True)
            cmds, counts = _run(ds, views, cull.make_config(block, items), single=True)
            _compare(cmds, counts, exp_cmds, exp_counts, scene.stateset_counts, f"plain 2M {block}x{items}")
    ds.close()


def test_cull_views_match_oracle_views(plain):
    cams = scenes.batch_cameras(64)
    for c in cams:
        a = cull.make_view(c)
        b = oracle.make_view(c, N.VIEW_DTYPE)
        assert a.tobytes() == b.tobytes()


def test_cull_many_views_and_statesets(gpu):
    """300 views (crosses the 64-view look-back groups) on a 2-level 200k scene with 64 StateSets."""
    scene = scenes.tree_scene((2000, 100), seed=scenes.SEED_BASE + 12, num_statesets=64, lod_count=3,
                              trange=200.0)
    ds = cull.DeviceScene(scene)
    ds.propagate()
    _, spheres = oracle.propagate_scene(scene)
    assert ds.download_spheres().tobytes() == spheres.tobytes()
    cams = scenes.random_cameras(scenes.SEED_BASE + 12, 20, 0, 300, box=400.0)
    views = cull.make_views(cams)
    exp_cmds, exp_counts = oracle.cull(spheres, scene.mesh_index, scene.meshes, scene.stateset_counts, views,
                                       cmd_dtype=N.CMD_DTYPE, nthreads=8)
    assert exp_counts.sum() > 10000
    # a job's last (persistent grid) tiles run as one work item per 64-view
    # group: every tile at the default grid; 256 x 2 with one workgroup per CU
    # splits only the last 256 of its ~450 tiles; CULL_WHOLE_TILES none
    for block, items, per_cu, flags in [(256, 4, 0, 0), (256, 8, 0, 0), (512, 2, 0, N.CULL_NO_TILE_SPHERES),
                                        (256, 2, 1, 0), (256, 2, 1, N.CULL_WHOLE_TILES),
                                        (256, 8, 0, N.CULL_WHOLE_TILES)]:
        print(f"300 views block={block} items={items} per_cu={per_cu} flags={flags}", flush=True)
        cmds, counts = _run(ds, views, cull.make_config(block, items, per_cu, flags=flags))
        _compare(cmds, counts, exp_cmds, exp_counts, scene.stateset_counts, f"{block}x{items} flags {flags}")
    ds.close()


def test_cull_uneven_batch(gpu):
    """Several scenes of different sizes, depths and LOD counts in one launch."""
    specs = [(3000, 1, 1, 3), (50_000, 2, 4, 17), (120_000, 3, 2, 1), (1, 1, 1, 5), (70_000, 4, 3, 65)]
    entries, exp = [], []
    for i, (n, depth, lods, nv) in enumerate(specs):
        sc = scenes.uneven_scene(i, n, depth, lods, nv)
        ds = cull.DeviceScene(sc)
        ds.propagate()
        _, sph = oracle.propagate_scene(sc)
        assert ds.download_spheres().tobytes() == sph.tobytes(), f"scene {i} spheres"
        views = cull.make_views(sc.cameras)
        out = cull.CullOutputs(nv, ds.n, ds.K)
        entries.append((ds, views, out))
        exp.append(oracle.cull(sph, sc.mesh_index, sc.meshes, sc.stateset_counts, views, cmd_dtype=N.CMD_DTYPE,
                               nthreads=8) + (sc,))
    from cadr_amd.device import synchronize
    for rep in range(2):  # a second launch reuses status words under a new epoch
        b = cull.CullBatch(cull.make_config(256, 4), entries)
        b.run()
        synchronize()
        for (ds, views, out), (ec, en, sc) in zip(entries, exp):
            cmds, counts = out.download()
            _compare(cmds, counts, ec, en, sc.stateset_counts, f"{sc.name} rep {rep}")
        b.close()
    for ds, _, out in entries:
        out.close(); ds.close()


def test_cull_edge_cases(gpu):
    from cadr_amd.device import synchronize
    rng = np.random.default_rng(3)
    meshes = scenes.make_meshes(77, 4, [1, 2, 3, 4])
    cam = scenes.Camera(eye=(0.0, 0.0, 0.0), dir=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fovy=1.2)
    views = cull.make_views([cam, scenes.Camera(eye=(0.0, 0.0, 0.0), dir=(0.0, 0.0, 1.0), up=(0.0, 1.0, 0.0),
                                                fovy=1.2)])
    cases = {
        "empty statesets": np.array([0, 5, 0, 0, 3000, 0], np.uint32),
        "all empty": np.array([0, 0, 0], np.uint32),
        "single": np.array([1], np.uint32),
        "tile boundaries": np.array([1023, 1024, 1025, 2047, 2048, 2049, 4097], np.uint32),
        "many small": rng.integers(0, 9, 700).astype(np.uint32),
    }
    for name, counts_in in cases.items():
        n = int(counts_in.sum())
        c = rng.uniform(-60, 60, (n, 3)).astype(np.float32)
        c[: n // 3, 2] = -np.abs(c[: n // 3, 2]) - 5  # a third in front of the first camera
        r = rng.uniform(0.1, 3, (n, 1)).astype(np.float32)
        s = np.concatenate([c, r], axis=1).astype(np.float32)
        mi = rng.integers(0, 5, n).astype(np.uint32)  # 4 = invalid mesh index
        if n:
            s[0] = (0.0, 0.0, 0.0, 1.0)  # contains the eye
        sc = scenes.Scene(name, [(np.zeros((max(n, 1), 16), np.float32), None)], mi, meshes, counts_in, [cam])
        ds = cull.DeviceScene(sc, keep_levels=False)
        if n:
            ds.upload_spheres(s)
        exp_cmds, exp_counts = oracle.cull(s, mi, meshes, counts_in, views, cmd_dtype=N.CMD_DTYPE)
        for block, items in [(256, 4), (128, 8)]:
            print(f"edge {name} {block}x{items}", flush=True)
            out = cull.CullOutputs(2, max(n, 1), len(counts_in))
            out.counts.fill_u32(0xDEADBEEF)
            synchronize()
            b = cull.CullBatch(cull.make_config(block, items), [(ds, views, out)])
            b.run()
            synchronize()
            cmds, counts = out.download()
            _compare(cmds[:, :max(n, 1)], counts, exp_cmds if n else np.zeros((2, 1), N.CMD_DTYPE), exp_counts,
                     counts_in, name)
            b.close(); out.close()
        ds.close()


def _boundary_spheres(views, reps, rng):
    """Spheres that touch one frustum plane of one view to within a few ulps:
    for drawable i, view k = i % V and plane p = (i // V) % 4 (near, far, x
    side, y side); the radius is the one that puts the glm-order comparison
    of that plane exactly on its threshold, moved by -4..4 ulps.  Centres lie
    inside the other planes where the plane allows it."""
    f32 = np.float32
    V = views.shape[0]
    n = V * 4 * 9 * reps
    out = np.zeros((n, 4), f32)
    for i in range(n):
        k, p, step = i % V, (i // V) % 4, (i // (4 * V)) % 9 - 4
        v = views[k]
        rows = v["viewRow"].astype(np.float64)
        near, far = float(v["nearZ"]), float(v["farZ"])
        tx = float(v["sideX"][1]) / float(v["sideX"][0])   # tan of the half angles
        ty = float(v["sideY"][1]) / float(v["sideY"][0])
        if p == 0:
            qz = near * rng.uniform(0.2, 0.95)
        elif p == 1:
            qz = far * rng.uniform(1.01, 1.2)
        else:
            qz = rng.uniform(2 * near, 0.5 * far)
        qx = rng.uniform(-0.9, 0.9) * tx * qz
        qy = rng.uniform(-0.9, 0.9) * ty * qz
        if p == 2:
            qx = rng.choice([-1, 1]) * rng.uniform(1.05, 1.5) * tx * qz
        if p == 3:
            qy = rng.choice([-1, 1]) * rng.uniform(1.05, 1.5) * ty * qz
        c = (rows[:, :3].T @ (np.array([qx, qy, qz]) - rows[:, 3])).astype(f32)
        r = v["viewRow"]
        vc = [(r[a][0] * c[0] + r[a][1] * c[1]) + (r[a][2] * c[2] + r[a][3]) for a in range(3)]
        if p == 0:
            w = f32(v["nearZ"] - vc[2])
        elif p == 1:
            w = f32(vc[2] - v["farZ"])
        elif p == 2:
            w = f32(abs(vc[0]) * v["sideX"][0]) - f32(vc[2] * v["sideX"][1])
        else:
            w = f32(abs(vc[1]) * v["sideY"][0]) - f32(vc[2] * v["sideY"][1])
        w = f32(w)
        for _ in range(abs(step)):
            w = np.nextafter(w, f32(np.inf) if step > 0 else f32(0))
        out[i, :3], out[i, 3] = c, max(w, f32(0))
    return out


def test_cull_boundary_spheres(gpu):
    """Spheres within a few ulps of a plane: the batched kernel's FMA estimate
    must defer to the glm-order test there, so results stay bit-exact."""
    from cadr_amd.device import synchronize
    rng = np.random.default_rng(11)
    cams = scenes.random_cameras(scenes.SEED_BASE + 31, 20, 0, 8, box=100.0)
    views = cull.make_views(cams)
    s = _boundary_spheres(views, 40, rng)
    n = s.shape[0]
    meshes = scenes.make_meshes(91, 16, 3)
    mi = rng.integers(0, 16, n).astype(np.uint32)
    counts_in = scenes.balanced_counts(n, 8)
    sc = scenes.Scene("boundary", [(np.zeros((n, 16), np.float32), None)], mi, meshes, counts_in, cams)
    ds = cull.DeviceScene(sc, keep_levels=False)
    ds.upload_spheres(s)
    exp_cmds, exp_counts = oracle.cull(s, mi, meshes, counts_in, views, cmd_dtype=N.CMD_DTYPE, nthreads=8)
    assert 0 < exp_counts.sum() < n * views.shape[0]
    for block, items in [(256, 2), (256, 4), (256, 8), (512, 4), (128, 16)]:
        print(f"boundary {block}x{items}", flush=True)
        cmds, counts = _run(ds, views, cull.make_config(block, items))
        _compare(cmds, counts, exp_cmds, exp_counts, counts_in, f"boundary {block}x{items}")
    synchronize()
    ds.close()


def test_cull_view_after_batch_shares_workspace(plain):
    """Single-view launches, a batched launch and single-view launches again on one
    workspace: each launch leaves the ticket reset and the epoch advanced."""
    from cadr_amd.device import synchronize
    scene, ds, spheres = plain
    cams = scenes.batch_cameras(5)
    views = cull.make_views(cams)
    exp_cmds, exp_counts = oracle.cull(spheres, scene.mesh_index, scene.meshes, scene.stateset_counts, views,
                                       cmd_dtype=N.CMD_DTYPE)
    cfg = cull.make_config(256, 4)
    ws = cull.Workspace(5 * 200, 8)
    outs = [cull.CullOutputs(1, ds.n, ds.K) for _ in range(5)]
    singles = [cull.CullBatch(cfg, [(ds, views[i:i + 1], outs[i])], workspace=ws) for i in range(5)]
    multi_out = cull.CullOutputs(5, ds.n, ds.K)
    multi = cull.CullBatch(cfg, [(ds, views, multi_out)], workspace=ws)
    for rep in range(3):
        for b in singles[:2]:
            b.run_view()
        multi.run()
        for b in singles[2:]:
            b.run_view()
        synchronize()
        for i in range(5):
            c, k = outs[i].download()
            _compare(c, k, exp_cmds[i:i + 1], exp_counts[i:i + 1], scene.stateset_counts, f"single {i} rep {rep}")
        c, k = multi_out.download()
        _compare(c, k, exp_cmds, exp_counts, scene.stateset_counts, f"multi rep {rep}")
    for b in singles + [multi]:
        b.close()
    for o in outs + [multi_out]:
        o.close()
    ws.close()


@pytest.mark.parametrize("block,items", [(256, 2), (256, 4), (256, 8), (256, 16), (512, 2), (512, 4), (1024, 1),
                                         (1024, 2)])
def test_fused_propagation_and_cull(gpu, block, items):
    """cadrPropagateCullView against cadrPropagateTransforms + cadrCullView and the
    oracle: world matrices, spheres, commands and counts bit for bit; tree and
    flat (root) leaf levels, a mesh table too large for LDS, spheres on and off."""
    from cadr_amd.device import DeviceBuffer, synchronize
    cases = [
        ("tree", scenes.tree_scene((40, 50, 60), seed=scenes.SEED_BASE + 31, num_statesets=64, lod_count=3,
                                   trange=60.0)),
        ("flat", scenes.plain_scene(70_001, seed=scenes.SEED_BASE + 32, num_statesets=7)),
        ("100 meshes", scenes.tree_scene((300, 130), seed=scenes.SEED_BASE + 33, num_statesets=5, lod_count=2,
                                         num_meshes=100, trange=200.0)),
    ]
    cam = scenes.Camera(eye=(10.0, -20.0, 150.0), dir=(-0.1, 0.1, -1.0), up=(0.0, 1.0, 0.0), fovy=1.1)
    views = cull.make_views([cam])
    config = cull.make_config(block, items)
    for name, sc in cases:
        print(f"fused {name} {block}x{items}", flush=True)
        worlds, sph = oracle.propagate_scene(sc)
        exp_cmds, exp_counts = oracle.cull(sph, sc.mesh_index, sc.meshes, sc.stateset_counts, views,
                                           cmd_dtype=N.CMD_DTYPE)
        assert exp_counts.sum() > 100
        ds = cull.DeviceScene(sc)
        descs = ds.level_descs(spheres=True)
        for lv in descs[:-1]:
            N.check(N.load().cadrPropagateTransforms(N.ptr(lv), None))
        for with_spheres in (True, False):
            out = cull.CullOutputs(1, ds.n, ds.K)
            ds.spheres.upload(np.full((ds.n, 4), np.nan, np.float32))
            ds.levels[-1][3].upload(np.full((ds.n, 16), np.nan, np.float32))
            leaf = descs[-1].copy()
            if not with_spheres:
                leaf["spheres"] = 0
            b = cull.CullBatch(config, [(ds, views, out)])
            b.run_fused(leaf)
            synchronize()
            cmds, counts = out.download()
            _compare(cmds, counts, exp_cmds, exp_counts, sc.stateset_counts, f"fused {name}")
            assert ds.world_matrices(-1).tobytes() == worlds[-1].tobytes(), f"{name}: world matrices"
            got_sph = ds.download_spheres()
            if with_spheres:
                assert got_sph.tobytes() == sph.tobytes(), f"{name}: spheres"
            else:
                assert np.isnan(got_sph).all(), f"{name}: spheres written although not requested"
            b.close(); out.close()
        ds.close()


def test_fused_rejects_bad_arguments(gpu):
    sc = scenes.plain_scene(1000, seed=scenes.SEED_BASE + 34)
    ds = cull.DeviceScene(sc)
    views = cull.make_views([scenes.DEFAULT_CAMERA])
    out = cull.CullOutputs(1, ds.n, ds.K)
    leaf = ds.level_descs()[-1].copy()
    for cfg in (cull.make_config(128, 8), cull.make_config(256, 3)):
        b = cull.CullBatch(cull.make_config(256, 4), [(ds, views, out)])
        b.config = cfg
        with pytest.raises(N.CadrError):
            b.run_fused(leaf)
        b.close()
    b = cull.CullBatch(cull.make_config(256, 4), [(ds, views, out)])
    bad = leaf.copy(); bad["count"] = 999
    with pytest.raises(N.CadrError):
        b.run_fused(bad)
    bad = leaf.copy(); bad["worldMatrices"] = 0
    with pytest.raises(N.CadrError):
        b.run_fused(bad)
    b.close(); out.close(); ds.close()


def test_batched_propagation_matches_per_level(gpu):
    """cadrPropagateTransformsBatch (one launch per tree depth over several
    trees) against the oracle: world matrices and spheres bit for bit."""
    from cadr_amd.device import DeviceBuffer, synchronize
    specs = [(3000, 1, 1), (50_000, 2, 4), (120_000, 3, 2), (1, 1, 1), (70_000, 4, 3), (6_500, 3, 3)]
    sc_list = [scenes.uneven_scene(40 + i, n, depth, lods, 1) for i, (n, depth, lods) in enumerate(specs)]
    dss = [cull.DeviceScene(sc) for sc in sc_list]
    by_depth = {}
    for ds in dss:
        for d, lv in enumerate(ds.level_descs(spheres=True)):
            by_depth.setdefault(d, []).append(lv)
    bufs = []
    for d in sorted(by_depth):
        host = np.concatenate(by_depth[d])
        dev = DeviceBuffer.from_array(host)
        bufs.append(dev)
        N.check(N.load().cadrPropagateTransformsBatch(dev.address, N.ptr(host), host.shape[0], None))
    synchronize()
    for sc, ds in zip(sc_list, dss):
        worlds, sph = oracle.propagate_scene(sc)
        for d in range(len(worlds)):
            assert ds.world_matrices(d).tobytes() == worlds[d].tobytes(), f"{sc.name} level {d}"
        assert ds.download_spheres().tobytes() == sph.tobytes(), f"{sc.name} spheres"
        ds.close()
    for b in bufs:
        b.close()
    with pytest.raises(N.CadrError):
        N.check(N.load().cadrPropagateTransformsBatch(0, None, 1, None))


def test_root_and_child_level_in_one_launch(gpu):
    """A root level's world matrices equal its local ones, so the level below
    can name the root's local matrices as parentWorld and share the root's
    launch: every level bit for bit as the oracle's level-by-level result."""
    from cadr_amd.device import DeviceBuffer, synchronize
    sc = scenes.tree_scene((300, 40, 30), seed=scenes.SEED_BASE + 17, num_statesets=8)
    ds = cull.DeviceScene(sc)
    lv = ds.level_descs(spheres=True)
    lv[1]["parentWorld"] = lv[0]["localMatrices"]
    both = np.concatenate([lv[0], lv[1]])
    dev = DeviceBuffer.from_array(both)
    N.check(N.load().cadrPropagateTransformsBatch(dev.address, N.ptr(both), 2, None))
    N.check(N.load().cadrPropagateTransforms(N.ptr(lv[2]), None))
    synchronize()
    worlds, sph = oracle.propagate_scene(sc)
    for d in range(3):
        assert ds.world_matrices(d).tobytes() == worlds[d].tobytes(), f"level {d}"
    assert ds.download_spheres().tobytes() == sph.tobytes()
    dev.close()
    ds.close()


def _fused_batch_cases():
    """Scenes for the fused batched launch: tree depths 1 (flat: the leaf is
    a root level) to 4, LOD counts 1 to 4, 1 to 65 views (two view groups),
    a mesh table too large for LDS, a scene with invalid mesh indices and
    parent indices past the parent level (those nodes act as roots), and two
    scenes without views."""
    out = []
    for i, (n, depth, lods, nv) in enumerate([(3000, 1, 1, 3), (50_000, 2, 4, 17), (90_000, 3, 2, 1), (1, 1, 1, 5),
                                              (40_000, 4, 3, 65)]):
        out.append(scenes.uneven_scene(60 + i, n, depth, lods, nv))
    sc = scenes.tree_scene((300, 130), seed=scenes.SEED_BASE + 33, num_statesets=5, lod_count=2, num_meshes=100,
                           trange=200.0)
    sc.cameras = scenes.random_cameras(scenes.SEED_BASE + 35, 20, 0, 9, box=300.0)
    out.append(sc)
    sc = scenes.uneven_scene(70, 20_000, 2, 3, 7)
    rng = np.random.default_rng(8)
    mi = sc.mesh_index.copy()
    mi[rng.integers(0, mi.shape[0], 500)] = 64 + rng.integers(0, 1000, 500).astype(np.uint32)
    local, parent = sc.levels[1]
    parent = parent.copy()
    parent[rng.integers(0, parent.shape[0], 700)] = sc.levels[0][0].shape[0] + rng.integers(0, 5, 700).astype(np.uint32)
    sc.mesh_index = mi
    sc.levels[1] = (local, parent)
    out.append(sc)
    # jobs without views still propagate their leaf level (spheres off, on)
    out.append(scenes.uneven_scene(71, 5000, 2, 2, 0))
    out.append(scenes.uneven_scene(72, 7000, 3, 1, 0))
    return out


def test_fused_batch_propagation_and_cull(gpu):
    """cadrPropagateCullBatch (every job's leaf level fused with the batched
    cull) against the oracle: leaf world matrices, spheres where requested,
    commands and counts bit for bit, for each supported shape, twice per
    shape (the second launch reuses the status words under a new epoch)."""
    from cadr_amd.device import synchronize
    cases = _fused_batch_cases()
    entries, exp, leaves = [], [], []
    for j, sc in enumerate(cases):
        worlds, sph = oracle.propagate_scene(sc)
        views = cull.make_views(sc.cameras)
        ec, en = oracle.cull(sph, sc.mesh_index, sc.meshes, sc.stateset_counts, views, cmd_dtype=N.CMD_DTYPE,
                             nthreads=8)
        ds = cull.DeviceScene(sc)
        descs = ds.level_descs(spheres=j % 2 == 0)  # spheres requested for every other job
        for lv in descs[:-1]:
            N.check(N.load().cadrPropagateTransforms(N.ptr(lv), None))
        leaves.append(descs[-1])
        entries.append((ds, views, cull.CullOutputs(views.shape[0], ds.n, ds.K)))
        exp.append((worlds[-1], sph, ec, en))
    assert sum(int(e[3].sum()) for e in exp) > 1000
    leaves = np.concatenate(leaves)
    for block, items in ((256, 4), (256, 8), (512, 4)):
        b = cull.CullBatch(cull.make_config(block, items), entries)
        b.set_leaves(leaves)
        for rep in range(2):
            print(f"fused batch {block}x{items} rep {rep}", flush=True)
            for ds, _, _ in entries:
                ds.spheres.upload(np.full((ds.n, 4), np.nan, np.float32))
                ds.levels[-1][3].upload(np.full((ds.n, 16), np.nan, np.float32))
            b.run_fused_batch()
            synchronize()
            for j, ((ds, views, out), (world, sph, ec, en)) in enumerate(zip(entries, exp)):
                what = f"{ds.scene.name} {block}x{items} rep {rep}"
                cmds, counts = out.download()
                _compare(cmds, counts, ec, en, ds.stateset_counts, what)
                assert ds.world_matrices(-1).tobytes() == world.tobytes(), f"{what}: world matrices"
                got = ds.download_spheres()
                if j % 2 == 0:
                    assert got.tobytes() == sph.tobytes(), f"{what}: spheres"
                else:
                    assert np.isnan(got).all(), f"{what}: spheres written although not requested"
        b.close()
    # unsupported shapes and a leaf that does not match its job are refused
    b = cull.CullBatch(cull.make_config(256, 4), entries)
    b.set_leaves(leaves)
    for cfg in (cull.make_config(256, 2), cull.make_config(128, 8)):
        b.config = cfg
        with pytest.raises(N.CadrError):
            b.run_fused_batch()
    b.config = cull.make_config(256, 4)
    bad = leaves.copy()
    bad["count"][1] += 1
    b.set_leaves(bad)
    with pytest.raises(N.CadrError):
        b.run_fused_batch()
    bad = leaves.copy()
    bad["parentIndex"][2] = 0
    b.set_leaves(bad)
    with pytest.raises(N.CadrError):
        b.run_fused_batch()
    b.close()
    for ds, _, out in entries:
        out.close(); ds.close()


@pytest.mark.parametrize("seed", range(6))
def test_cull_batched_random_scenes(gpu, seed):
    """Random scenes against the oracle: 1-80 StateSets (some empty), 1-130
    views (up to three view groups), mesh tables of 1-100 meshes with 1-4
    LODs and invalid mesh indices, and irregular spheres (NaN components,
    infinite, negative and empty radii) that switch the tile sphere off for
    their tiles; two batched shapes per seed, command stride past the
    drawable count."""
    from cadr_amd.device import synchronize
    rng = np.random.default_rng(200 + seed)
    K = int(rng.integers(1, 81))
    counts_in = rng.integers(0, 3000, K).astype(np.uint32)
    counts_in[rng.random(K) < 0.2] = 0
    n = int(counts_in.sum())
    c = rng.uniform(-200, 200, (n, 3)).astype(np.float32)
    r = rng.uniform(0, 20, (n, 1)).astype(np.float32)
    s = np.concatenate([c, r], axis=1).astype(np.float32)
    for frac, val, col in ((0.002, np.nan, 0), (0.002, np.inf, 3), (0.002, -1.0, 3), (0.002, -np.inf, 3)):
        s[rng.random(n) < frac, col] = val
    nm = int(rng.integers(1, 101))
    meshes = scenes.make_meshes(300 + seed, nm, rng.integers(1, 5, nm))
    mi = rng.integers(0, nm + 2, n).astype(np.uint32)  # a few invalid mesh indices
    V = int(rng.integers(1, 131))
    cams = scenes.random_cameras(scenes.SEED_BASE + 50 + seed, 20, 0, V, box=250.0)
    views = cull.make_views(cams)
    exp_cmds, exp_counts = oracle.cull(s, mi, meshes, counts_in, views, cmd_dtype=N.CMD_DTYPE, nthreads=8)
    sc = scenes.Scene(f"random{seed}", [(np.zeros((max(n, 1), 16), np.float32), None)], mi, meshes, counts_in, cams)
    ds = cull.DeviceScene(sc, keep_levels=False)
    if n:
        ds.upload_spheres(s)
    stride = max(n, 1) + int(rng.integers(0, 5))
    shapes = [(256, 2), (256, 4), (256, 8), (256, 16), (512, 2), (512, 4), (128, 8), (128, 16)]
    for i in rng.choice(len(shapes), 2, replace=False):
        block, items = shapes[int(i)]
        print(f"random seed {seed}: {n} drawables, {K} StateSets, {V} views, {nm} meshes, {block}x{items}", flush=True)
        out = cull.CullOutputs(V, stride, K)
        b = cull.CullBatch(cull.make_config(block, items), [(ds, views, out)])
        b.run()
        synchronize()
        cmds, counts = out.download()
        exp_full = np.zeros((V, stride), N.CMD_DTYPE)
        exp_full[:, :n] = exp_cmds[:, :n]
        _compare(cmds, counts, exp_full, exp_counts, counts_in, f"seed {seed} {block}x{items}")
        b.close(); out.close()
    ds.close()


@pytest.mark.parametrize("seed", range(4))
def test_fused_one_view_random_scenes(gpu, seed):
    """Random trees through cadrPropagateCullView against the oracle: depth
    1-4, 1-70 StateSets (some empty), parent indices past the parent level,
    invalid mesh indices, a few non-finite local matrices (their tiles lose
    the tile sphere), every fused shape; world matrices, commands and
    counts bit for bit."""
    from cadr_amd.device import synchronize
    rng = np.random.default_rng(400 + seed)
    depth = int(rng.integers(1, 5))
    n = int(rng.integers(1, 120_000))
    sc = scenes.uneven_scene(80 + seed, n, depth, int(rng.integers(1, 5)), 1, num_statesets=int(rng.integers(1, 71)))
    local, parent = sc.levels[-1]
    local = local.copy()
    local[rng.random(n) < 0.001, int(rng.integers(0, 16))] = np.inf
    if parent is not None:
        parent = parent.copy()
        parent[rng.random(n) < 0.01] = sc.levels[-2][0].shape[0] + 3
    sc.levels[-1] = (local, parent)
    mi = sc.mesh_index.copy()
    mi[rng.random(n) < 0.01] = 64 + 7
    sc.mesh_index = mi
    e = rng.normal(size=3)
    e = e / np.linalg.norm(e) * rng.uniform(120, 250)  # outside the scene, looking at its centre
    d = -e / np.linalg.norm(e)
    cam = scenes.Camera(eye=tuple(float(np.float32(x)) for x in e), dir=tuple(float(np.float32(x)) for x in d),
                        up=(0.0, 1.0, 0.0) if abs(d[1]) < 0.99 else (1.0, 0.0, 0.0), fovy=float(rng.uniform(0.5, 1.5)))
    views = cull.make_views([cam])
    worlds, sph = oracle.propagate_scene(sc)
    exp_cmds, exp_counts = oracle.cull(sph, sc.mesh_index, sc.meshes, sc.stateset_counts, views, cmd_dtype=N.CMD_DTYPE)
    ds = cull.DeviceScene(sc)
    descs = ds.level_descs(spheres=False)
    for lv in descs[:-1]:
        N.check(N.load().cadrPropagateTransforms(N.ptr(lv), None))
    for block, items in ((256, 2), (256, 4), (256, 8), (256, 16), (512, 2), (512, 4), (1024, 1), (1024, 2)):
        print(f"fused random seed {seed}: {n} drawables, depth {depth}, {ds.K} StateSets, {block}x{items}", flush=True)
        ds.levels[-1][3].upload(np.full((n, 16), np.nan, np.float32))
        out = cull.CullOutputs(1, n, ds.K)
        b = cull.CullBatch(cull.make_config(block, items), [(ds, views, out)])
        b.run_fused(descs[-1])
        synchronize()
        cmds, counts = out.download()
        _compare(cmds, counts, exp_cmds, exp_counts, sc.stateset_counts, f"seed {seed} {block}x{items}")
        # bit for bit except the payload of generated NaNs (inf * 0), which the
        # host and the GPU encode differently
        got, exp = ds.world_matrices(-1), worlds[-1]
        nan = np.isnan(exp)
        assert np.array_equal(np.isnan(got), nan)
        assert np.array_equal(got.view(np.uint32)[~nan], exp.view(np.uint32)[~nan])
        b.close(); out.close()
    ds.close()


@pytest.mark.parametrize("seed", range(3))
def test_fused_batch_random_jobs(gpu, seed):
    """Random multi-scene launches of cadrPropagateCullBatch against the
    oracle: 2-6 scenes of random depth, size and StateSet count, 0-70 views
    each, out-of-range parents and invalid meshes; each fused shape."""
    from cadr_amd.device import synchronize
    rng = np.random.default_rng(500 + seed)
    entries, exp, leaves = [], [], []
    for j in range(int(rng.integers(2, 7))):
        n = int(rng.integers(1, 60_000))
        nv = int(rng.integers(0, 71))
        sc = scenes.uneven_scene(100 + 10 * seed + j, n, int(rng.integers(1, 5)), int(rng.integers(1, 5)), nv,
                                 num_statesets=int(rng.integers(1, 65)))
        local, parent = sc.levels[-1]
        if parent is not None:
            parent = parent.copy()
            parent[rng.random(n) < 0.01] = sc.levels[-2][0].shape[0] + 1
            sc.levels[-1] = (local, parent)
        mi = sc.mesh_index.copy()
        mi[rng.random(n) < 0.01] = 99
        sc.mesh_index = mi
        sc.cameras = scenes.random_cameras(scenes.SEED_BASE + 120 + seed, 20 + j, 0, nv, box=250.0)
        worlds, sph = oracle.propagate_scene(sc)
        views = cull.make_views(sc.cameras)
        ec, en = oracle.cull(sph, sc.mesh_index, sc.meshes, sc.stateset_counts, views, cmd_dtype=N.CMD_DTYPE,
                             nthreads=8)
        ds = cull.DeviceScene(sc)
        descs = ds.level_descs(spheres=False)
        for lv in descs[:-1]:
            N.check(N.load().cadrPropagateTransforms(N.ptr(lv), None))
        leaves.append(descs[-1])
        entries.append((ds, views, cull.CullOutputs(nv, ds.n, ds.K)))
        exp.append((worlds[-1], ec, en))
    leaves = np.concatenate(leaves)
    for block, items, flags in ((256, 4, 0), (256, 8, 0), (512, 4, 0), (256, 8, N.CULL_NO_TILE_SPHERES)):
        print(f"fused batch random seed {seed}: {len(entries)} scenes, {block}x{items} flags {flags}", flush=True)
        for ds, _, _ in entries:
            ds.levels[-1][3].upload(np.full((ds.n, 16), np.nan, np.float32))
        b = cull.CullBatch(cull.make_config(block, items, flags=flags), entries)
        b.set_leaves(leaves)
        b.run_fused_batch()
        synchronize()
        for (ds, views, out), (world, ec, en) in zip(entries, exp):
            cmds, counts = out.download()
            _compare(cmds, counts, ec, en, ds.stateset_counts, f"{ds.scene.name} {block}x{items}")
            assert ds.world_matrices(-1).tobytes() == world.tobytes(), f"{ds.scene.name}: world matrices"
        b.close()
    for ds, _, out in entries:
        out.close(); ds.close()
