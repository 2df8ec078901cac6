"""The C ABI library loads, exports every entry point include/*.h declares, and
its host-only functions (views, tiles, argument checks) behave without a GPU."""
import ctypes
import glob
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from cadr_amd import _native as N
from cadr_amd import cull, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(cadr[A-Z]\w*)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = N.load()
    names = declared_functions()
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", N.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in names if n not in exported]
    assert not missing, f"declared but not exported: {missing}"
    for n in names:
        getattr(lib, n)
    # every binding in _native matches a declaration
    assert set(N.SIGNATURES) <= set(names)


def test_make_view_matches_oracle_bit_exact():
    cams = scenes.batch_cameras(256) + [scenes.DEFAULT_CAMERA]
    for c in cams:
        assert cull.make_view(c).tobytes() == oracle.make_view(c, N.VIEW_DTYPE).tobytes()


def test_make_view_rejects_bad_cameras():
    bad = [scenes.Camera(eye=(0.0, 0.0, 0.0), dir=(0.0, 1.0, 0.0), up=(0.0, 1.0, 0.0), fovy=1.0),  # up // dir
           scenes.Camera(eye=(0.0, 0.0, 0.0), dir=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fovy=0.0),
           scenes.Camera(eye=(0.0, 0.0, 0.0), dir=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fovy=1.0),
           scenes.Camera(eye=(0.0, 0.0, 0.0), dir=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fovy=1.0, near=5.0, far=1.0)]
    for c in bad:
        with pytest.raises(N.CadrError) as e:
            cull.make_view(c)
        assert e.value.code == N.ERROR_INVALID_VALUE


def test_build_tiles_layout():
    counts = np.array([0, 1, 1024, 1025, 3000, 0], np.uint32)
    t = cull.build_tiles(counts, 1024)
    base = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    # every drawable covered once, tiles within their StateSet, flags at the ends
    covered = np.zeros(int(counts.sum()), np.int64)
    per_k = {}
    for tile in t:
        k = int(tile["stateSetAndFlags"]) & N.TILE_STATESET_MASK
        per_k.setdefault(k, []).append(tile)
        assert tile["stateSetBase"] == base[k]
        assert base[k] <= tile["start"] and tile["start"] + tile["count"] <= base[k + 1]
        assert tile["count"] <= 1024
        covered[tile["start"]: tile["start"] + tile["count"]] += 1
    assert np.all(covered == 1)
    chain_base = 0
    for k in range(len(counts)):
        ts = per_k[k]
        assert int(ts[0]["stateSetAndFlags"]) & N.TILE_FIRST and int(ts[-1]["stateSetAndFlags"]) & N.TILE_LAST
        assert len(ts) == max(1, -(-int(counts[k]) // 1024))
        if counts[k] == 0:
            assert ts[0]["count"] == 0
        # chain indices number the tiles StateSet by StateSet, in drawable order
        assert [int(x["chainIndex"]) for x in ts] == list(range(chain_base, chain_base + len(ts)))
        assert all(ts[i]["start"] < ts[i + 1]["start"] for i in range(len(ts) - 1))
        chain_base += len(ts)
    # round-robin issue order: the q-th tile of every StateSet before any (q+1)-th
    q_of = [(int(x["start"]) - int(x["stateSetBase"])) // 1024 for x in t]
    assert q_of == sorted(q_of)
    assert sorted(int(x["chainIndex"]) for x in t) == list(range(len(t)))
    # counting mode agrees; a short buffer is an error
    n = ctypes.c_uint32(0)
    assert N.load().cadrBuildCullTiles(N.ptr(counts), counts.size, 1024, None, 0, ctypes.byref(n)) == 0
    assert n.value == len(t)
    small = np.zeros(2, N.TILE_DTYPE)
    assert N.load().cadrBuildCullTiles(N.ptr(counts), counts.size, 1024, N.ptr(small), 2, ctypes.byref(n)) \
        == N.ERROR_INVALID_VALUE


def test_cull_config_tile_sizes():
    for (b, i) in [(256, 2), (256, 4), (256, 8), (512, 2), (512, 4), (128, 8), (128, 16)]:
        assert cull.tile_size(cull.make_config(b, i)) == b * i
    assert cull.tile_size(cull.make_config(100, 3)) == 0


def test_cull_config_paths():
    """Which entry points have a kernel of each shape: every one-view shape
    (256/512/1024 threads), the batched shapes, the fused ones."""
    lib = N.load()

    def paths(b, i):
        return lib.cadrCullConfigPaths(ctypes.byref(cull.make_config(b, i)))

    V, B, FV, FB = N.CULL_PATH_VIEW, N.CULL_PATH_BATCH, N.CULL_PATH_FUSED_VIEW, N.CULL_PATH_FUSED_BATCH
    assert paths(256, 8) == V | B | FV | FB
    assert paths(512, 4) == V | B | FV | FB
    assert paths(256, 16) == V | B | FV
    assert paths(1024, 8) == V  # one view only
    assert paths(1024, 2) == V | FV
    assert paths(128, 8) == V | B  # the one-view entry runs the batched kernel
    assert paths(100, 3) == 0
    for b, i in [(256, 2), (256, 4), (512, 8), (512, 16), (1024, 1), (1024, 4), (1024, 16)]:
        assert paths(b, i) & V and cull.tile_size(cull.make_config(b, i)) == b * i


def test_result_strings():
    lib = N.load()
    assert lib.cadrGetResultString(0) == b"CADR_SUCCESS"
    assert lib.cadrGetResultString(2) == b"CADR_ERROR_OUT_OF_RESOURCES"


def test_fused_batch_argument_checks_before_any_device_call():
    """cadrPropagateCullBatch refuses unsupported shapes, missing arrays and
    leaf levels that do not match their jobs on the host, before it touches
    the device (the addresses below are never dereferenced)."""
    lib = N.load()
    fake = 1 << 40  # a 16-byte aligned address the checks only inspect
    jobs = np.zeros(2, dtype=N.JOB_DTYPE)
    jobs["numDrawables"] = (100, 7)
    leaves = np.zeros(2, dtype=N.LEVEL_DTYPE)
    leaves["count"] = (100, 7)
    leaves["localMatrices"] = leaves["worldMatrices"] = fake

    def call(cfg, dl=fake, hl=leaves, hj=jobs, n=2, ws=None):
        return lib.cadrPropagateCullBatch(dl, None if hl is None else N.ptr(hl), ctypes.byref(cfg), fake,
                                          None if hj is None else N.ptr(hj), n, ws, None)

    ok = cull.make_config(256, 8)
    for block, items in ((256, 2), (256, 16), (512, 2), (128, 8), (1024, 1), (256, 3)):
        assert call(cull.make_config(block, items)) == N.ERROR_INVALID_VALUE, (block, items)
    assert call(ok, n=0) == 0  # nothing to do
    assert call(ok, dl=0) == N.ERROR_INVALID_VALUE
    assert call(ok, hl=None) == N.ERROR_INVALID_VALUE
    assert call(ok, hj=None) == N.ERROR_INVALID_VALUE
    assert call(ok, dl=fake + 4) == N.ERROR_INVALID_VALUE  # misaligned descriptors
    bad = leaves.copy(); bad["count"][1] = 8
    assert call(ok, hl=bad) == N.ERROR_INVALID_VALUE
    assert "leaf level must hold" in lib.cadrGetLastErrorMessage().decode()
    bad = leaves.copy(); bad["worldMatrices"][0] = 0
    assert call(ok, hl=bad) == N.ERROR_INVALID_VALUE
    bad = leaves.copy(); bad["localMatrices"][1] = fake + 8
    assert call(ok, hl=bad) == N.ERROR_INVALID_VALUE
    bad = leaves.copy(); bad["parentWorld"][0] = fake  # a non-root level needs parent indices
    assert call(ok, hl=bad) == N.ERROR_INVALID_VALUE
    bad["parentIndex"][0] = fake + 2
    assert call(ok, hl=bad) == N.ERROR_INVALID_VALUE
    # well-formed leaves reach the workspace check
    assert call(ok) == N.ERROR_INVALID_VALUE
    assert "workspace is NULL" in lib.cadrGetLastErrorMessage().decode()


def test_build_tiles_refuses_statesets_past_30_bits():
    """Look-back status words carry 30-bit counts: a StateSet of 2^30 drawables
    is refused on the host (counting mode, no device call)."""
    n = ctypes.c_uint32(0)
    lib = N.load()
    ok = np.array([N.TILE_STATESET_MASK, 5], np.uint32)
    assert lib.cadrBuildCullTiles(N.ptr(ok), 2, 1 << 20, None, 0, ctypes.byref(n)) == 0
    assert n.value == 1024 + 1
    big = np.array([5, N.TILE_STATESET_MASK + 1], np.uint32)
    assert lib.cadrBuildCullTiles(N.ptr(big), 2, 1 << 20, None, 0, ctypes.byref(n)) == N.ERROR_INVALID_VALUE
    assert "2^30" in lib.cadrGetLastErrorMessage().decode()


def test_workspace_calls_check_arguments():
    lib = N.load()
    assert lib.cadrCullWorkspaceCheck(None) == N.ERROR_INVALID_VALUE
    assert lib.cadrCullWorkspaceSetWaitLimit(None, 10) == N.ERROR_INVALID_VALUE
