"""Shared buffers for a Vulkan renderer (cadrSharedBuffer*): the one-view
cull writes its commands and counts into exported device memory, and the
same bytes are read back through an import of the exported descriptor, as a
renderer's vkAllocateMemory + VkImportMemoryFdInfoKHR would map them.  The
Vulkan side itself is not run (no Vulkan driver here); the descriptor
round trip is HIP's own external-memory import."""
import ctypes
import os

import numpy as np
import pytest

import oracle
from cadr_amd import _native as N
from cadr_amd import cull, scenes
from cadr_amd.device import synchronize

pytestmark = pytest.mark.gpu


def _shared(lib, size):
    h = ctypes.c_void_p(0)
    N.check(lib.cadrSharedBufferCreate(size, ctypes.byref(h)))
    return h


def _imported(lib, created):
    fd = ctypes.c_int(-1)
    N.check(lib.cadrSharedBufferExportFd(created, ctypes.byref(fd)))
    assert fd.value >= 0
    h = ctypes.c_void_p(0)
    rc = lib.cadrSharedBufferImport(fd.value, lib.cadrSharedBufferSize(created), ctypes.byref(h))
    if rc != N.SUCCESS:
        os.close(fd.value)
        N.check(rc)
    return h


def _download(lib, address, dtype, count):
    out = np.zeros(count, dtype=dtype)
    N.check(lib.cadrMemcpyDtoH(out.ctypes.data, address, out.nbytes, None))
    synchronize()
    return out


def test_cull_outputs_in_an_exported_buffer(gpu):
    lib = N.load()
    scene = scenes.plain_scene(100_000)
    ds = cull.DeviceScene(scene)
    ds.propagate()
    _, spheres = oracle.propagate_scene(scene)
    views = cull.make_views([scenes.DEFAULT_CAMERA])
    exp_cmds, exp_counts = oracle.cull(spheres, scene.mesh_index, scene.meshes, scene.stateset_counts, views,
                                       cmd_dtype=N.CMD_DTYPE)
    n, K = ds.n, ds.K
    cmd_buf = _shared(lib, 20 * n)
    cnt_buf = _shared(lib, 4 * K)
    assert lib.cadrSharedBufferSize(cmd_buf) >= 20 * n and lib.cadrSharedBufferAddress(cmd_buf) % 4096 == 0
    try:
        cmd_view = _imported(lib, cmd_buf)
        cnt_view = _imported(lib, cnt_buf)
        try:
            out = cull.CullOutputs(1, n, K)
            batch = cull.CullBatch(cull.make_config(1024, 8), [(ds, views, out)])
            # point the job at the shared buffers
            batch.host_jobs["commands"] = lib.cadrSharedBufferAddress(cmd_buf)
            batch.host_jobs["counts"] = lib.cadrSharedBufferAddress(cnt_buf)
            batch.device_jobs.upload(batch.host_jobs)
            N.check(lib.cadrFillU32(lib.cadrSharedBufferAddress(cmd_buf), 0xFFFFFFFF, 20 * n, None))
            batch.run_view()
            synchronize()
            batch.workspace.check()
            counts = _download(lib, lib.cadrSharedBufferAddress(cnt_view), np.uint32, K).reshape(1, K)
            cmds = _download(lib, lib.cadrSharedBufferAddress(cmd_view), N.CMD_DTYPE, n).reshape(1, n)
            assert np.array_equal(counts, exp_counts)
            got = cull.valid_commands(cmds, counts, scene.stateset_counts)
            exp = cull.valid_commands(exp_cmds, exp_counts, scene.stateset_counts)
            for k in range(K):
                assert got[0][k].tobytes() == exp[0][k].tobytes(), f"StateSet {k}"
            # and a write through the import is seen through the exporting mapping
            N.check(lib.cadrFillU32(lib.cadrSharedBufferAddress(cnt_view), 0x5A5A5A5A, 4 * K, None))
            synchronize()
            assert (_download(lib, lib.cadrSharedBufferAddress(cnt_buf), np.uint32, K) == 0x5A5A5A5A).all()
            batch.close(); out.close()
        finally:
            lib.cadrSharedBufferDestroy(cmd_view)
            lib.cadrSharedBufferDestroy(cnt_view)
    finally:
        lib.cadrSharedBufferDestroy(cmd_buf)
        lib.cadrSharedBufferDestroy(cnt_buf)
        ds.close()


def test_shared_buffer_argument_checks(gpu):
    lib = N.load()
    h = ctypes.c_void_p(0)
    assert lib.cadrSharedBufferCreate(0, ctypes.byref(h)) == N.ERROR_INVALID_VALUE
    assert lib.cadrSharedBufferImport(-1, 4096, ctypes.byref(h)) == N.ERROR_INVALID_VALUE
    assert lib.cadrSharedBufferAddress(None) == 0 and lib.cadrSharedBufferSize(None) == 0
    lib.cadrSharedBufferDestroy(None)
    b = _shared(lib, 1000)
    assert lib.cadrSharedBufferSize(b) >= 1000
    v = _imported(lib, b)
    fd = ctypes.c_int(-1)
    assert lib.cadrSharedBufferExportFd(v, ctypes.byref(fd)) == N.ERROR_INVALID_VALUE  # imported: not ours to export
    lib.cadrSharedBufferDestroy(v)
    lib.cadrSharedBufferDestroy(b)
