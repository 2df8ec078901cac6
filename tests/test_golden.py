"""The oracle and the HIP kernels against golden vectors of the reference's own
FP32 functions (tests/golden/ref_fp32.npz, written by tests/golden/make_golden.py
from GLM 0.9.9 and src/CadR/BoundingSphere.h compiled from the reference tree).
Bit-exact throughout."""
import os

import numpy as np
import pytest

import oracle
from cadr_amd import _native as N
from cadr_amd import cull
from cadr_amd.device import DeviceBuffer
from cadr_amd.scenes import Camera

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_fp32.npz"))


def _tree_levels():
    sizes = [int(x) for x in G["tree_sizes"]]
    starts = np.concatenate([[0], np.cumsum(sizes)])
    levels = []
    for d, n in enumerate(sizes):
        sl = slice(starts[d], starts[d + 1])
        parent = None if d == 0 else (G["tree_parent"][sl] - starts[d - 1]).astype(np.uint32)
        levels.append((G["tree_local"][sl], parent, G["tree_world"][sl]))
    return levels


def test_oracle_mat4_mul_golden():
    for a, b, ab in zip(G["mat_a"], G["mat_b"], G["mat_ab"]):
        assert oracle.mat4_mul(a, b).tobytes() == ab.tobytes()


def test_oracle_sphere_transform_golden():
    for m, s, o in zip(G["sph_m"], G["sph_s"], G["sph_out"]):
        assert oracle.sphere_transform(m, s).tobytes() == o.tobytes()


def test_oracle_tree_propagation_golden():
    prev = None
    for local, parent, world in _tree_levels():
        w, _ = oracle.propagate_level(local, parent, prev)
        assert w.tobytes() == world.tobytes()
        prev = w


def _golden_cams():
    for i in range(G["cam_eye"].shape[0]):
        yield i, Camera(eye=tuple(map(float, G["cam_eye"][i])), dir=tuple(map(float, G["cam_dir"][i])),
                        up=tuple(map(float, G["cam_up"][i])), fovy=float(G["cam_fovy"][i]),
                        aspect=float(G["cam_aspect"][i]), near=float(G["cam_near"][i]), far=float(G["cam_far"][i]))


def _check_view(v, i):
    L = G["cam_look"][i].reshape(4, 4)  # [column][row]
    for r in range(3):
        assert v["viewRow"][r].tobytes() == np.array([L[0, r], L[1, r], L[2, r], L[3, r]], np.float32).tobytes()
    p11 = G["cam_proj"][i][5]
    assert v["lodScale"] == np.float32(np.float32(540.0) * p11)
    ny = np.float32(np.sqrt(np.float32(p11 * p11 + np.float32(1))))
    assert v["sideY"][0] == np.float32(p11 / ny) and v["sideY"][1] == np.float32(np.float32(1) / ny)
    p00 = G["cam_proj"][i][0]
    nx = np.float32(np.sqrt(np.float32(p00 * p00 + np.float32(1))))
    assert v["sideX"][0] == np.float32(p00 / nx) and v["sideX"][1] == np.float32(np.float32(1) / nx)


def test_oracle_views_golden():
    for i, cam in _golden_cams():
        _check_view(oracle.make_view(cam, N.VIEW_DTYPE)[0], i)


def test_cadr_make_view_golden():
    """cadrMakeView (host code of the C ABI) against glm::lookAtLH / perspectiveLH_ZO."""
    for i, cam in _golden_cams():
        _check_view(cull.make_view(cam)[0], i)


@pytest.mark.gpu
def test_propagation_kernel_golden(gpu):
    lib = N.load()
    prev = None
    bufs = []
    for local, parent, world in _tree_levels():
        n = local.shape[0]
        lb = DeviceBuffer.from_array(np.ascontiguousarray(local))
        wb = DeviceBuffer(64 * n)
        lv = np.zeros(1, dtype=N.LEVEL_DTYPE)
        lv["localMatrices"], lv["worldMatrices"], lv["count"] = lb.address, wb.address, n
        if prev is not None:
            pb = DeviceBuffer.from_array(parent)
            bufs.append(pb)
            lv["parentIndex"], lv["parentWorld"], lv["parentCount"] = pb.address, prev[1].address, prev[0]
        N.check(lib.cadrPropagateTransforms(N.ptr(lv), None))
        got = wb.download(np.float32, 16 * n).reshape(n, 16)
        assert got.tobytes() == world.tobytes()
        bufs += [lb, wb]
        prev = (n, wb)
    for b in bufs:
        b.close()


@pytest.mark.gpu
def test_world_sphere_kernel_golden(gpu):
    n = G["sph_m"].shape[0]
    meshes = np.zeros(n, dtype=N.MESH_DTYPE)
    meshes["center"] = G["sph_s"][:, :3]
    meshes["radius"] = G["sph_s"][:, 3]
    meshes["lodCount"] = 1
    mb, wb = DeviceBuffer.from_array(meshes), DeviceBuffer.from_array(np.ascontiguousarray(G["sph_m"]))
    ib = DeviceBuffer.from_array(np.arange(n, dtype=np.uint32))
    sb = DeviceBuffer(16 * n)
    N.check(N.load().cadrComputeWorldSpheres(wb.address, ib.address, mb.address, n, n, sb.address, None))
    got = sb.download(np.float32, 4 * n).reshape(n, 4)
    assert got.tobytes() == G["sph_out"].tobytes()
    for b in (mb, wb, ib, sb):
        b.close()
