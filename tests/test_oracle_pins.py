"""Pins the CPU oracle's FP32 math before it is trusted as the parity checker.

Checks, per oracle function:
  - bit-exact agreement with the reference's own code (oracle/_ref: GLM 0.9.9 and
    src/CadR/BoundingSphere.h compiled from the reference tree) on random inputs;
  - closed forms that a dropped term, swapped operand or transposed index breaks
    (identity, permutations, power-of-two scales, exact translations);
  - a float64 restatement within 2 ULP.
"""
import math

import numpy as np
import pytest

import oracle
from cadr_amd import _native as N
from cadr_amd.scenes import Camera

rng = np.random.default_rng(1234)
needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no reference tree)")


def rand_mats(n, scale=10.0):
    return (rng.standard_normal((n, 16)) * scale).astype(np.float32)


def ulp_diff(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7FFFFFFF), a)
    b = np.where(b < 0, -(b & 0x7FFFFFFF), b)
    return np.abs(a - b)


# ------------------------------------------------------------ mat4 * mat4

@needs_ref
def test_mat4_mul_matches_glm_bit_exact():
    R = oracle.ref()
    for _ in range(2000):
        A, B = rand_mats(2)
        exp = np.zeros(16, np.float32)
        R.ref_mat4_mul(A.ctypes.data, B.ctypes.data, exp.ctypes.data)
        got = oracle.mat4_mul(A, B)
        assert got.tobytes() == exp.tobytes()


def test_mat4_mul_closed_forms():
    I = np.eye(4, dtype=np.float32).reshape(16)
    A = rand_mats(1)[0]
    # identity on the right is exact (x*1 + y*0 ... with finite y)
    assert np.array_equal(oracle.mat4_mul(A, I), A)
    # column-major: B = translation(t) puts A*(t,1) into column 3
    t = np.array([3.0, -5.0, 7.0], np.float32)
    T = np.eye(4, dtype=np.float32)
    T[3, :3] = t  # column 3 (column-major storage: m[col][row])
    got = oracle.mat4_mul(A, T.reshape(16)).reshape(4, 4)
    Am = A.reshape(4, 4)  # Am[col][row]
    col3 = ((Am[0] * t[0] + Am[1] * t[1]) + Am[2] * t[2]) + Am[3]
    assert np.array_equal(got[3], col3.astype(np.float32))
    # a permutation on the left permutes rows exactly
    P = np.zeros((4, 4), np.float32)
    P[0, 1] = P[1, 0] = P[2, 2] = P[3, 3] = 1  # swaps x and y
    got = oracle.mat4_mul(P.reshape(16), A).reshape(4, 4)
    assert np.array_equal(got[:, 0], Am[:, 1]) and np.array_equal(got[:, 1], Am[:, 0])
    # products of power-of-two diagonal matrices are exact
    D1 = np.diag([2.0, 0.5, 4.0, 1.0]).astype(np.float32).reshape(16)
    D2 = np.diag([8.0, 0.25, 0.5, 1.0]).astype(np.float32).reshape(16)
    assert np.array_equal(oracle.mat4_mul(D1, D2), np.diag([16.0, 0.125, 2.0, 1.0]).astype(np.float32).reshape(16))


def test_mat4_mul_within_2ulp_of_float64():
    for _ in range(500):
        A, B = rand_mats(2, 1.0)
        got = oracle.mat4_mul(A, B)
        exact = (A.reshape(4, 4).astype(np.float64).T @ B.reshape(4, 4).astype(np.float64).T).T.reshape(16)
        # 2 ULP of the largest term magnitude (cancellation makes per-element ULP meaningless)
        mag = (np.abs(A.reshape(4, 4).T.astype(np.float64)) @ np.abs(B.reshape(4, 4).T.astype(np.float64))).T.reshape(16)
        assert np.all(np.abs(got - exact) <= 2 * mag * 2.0 ** -23 + 1e-30)


# -------------------------------------------------------- sphere transform

@needs_ref
def test_sphere_transform_matches_boundingsphere_bit_exact():
    R = oracle.ref()
    for _ in range(2000):
        M = rand_mats(1)[0]
        s = np.array([*rng.standard_normal(3) * 50, abs(rng.standard_normal()) * 3], np.float32)
        exp = np.zeros(4, np.float32)
        R.ref_sphere_transform(M.ctypes.data, s.ctypes.data, exp.ctypes.data)
        assert oracle.sphere_transform(M, s).tobytes() == exp.tobytes()


def test_sphere_transform_closed_forms():
    s = np.array([1.5, -2.0, 3.25, 0.75], np.float32)
    I = np.eye(4, dtype=np.float32).reshape(16)
    assert np.array_equal(oracle.sphere_transform(I, s), s)
    # uniform power-of-two scale and exact translation
    M = np.diag([4.0, 4.0, 4.0, 1.0]).astype(np.float32)
    M[3, :3] = [10.0, 20.0, -30.0]
    got = oracle.sphere_transform(M.reshape(16), s)
    assert np.array_equal(got, np.array([16.0, 12.0, -17.0, 3.0], np.float32))
    # the radius uses the largest axis scale, not the first
    M = np.diag([1.0, 8.0, 2.0, 1.0]).astype(np.float32).reshape(16)
    assert oracle.sphere_transform(M, np.array([0, 0, 0, 1], np.float32))[3] == 8.0
    # rows of mat3(M) multiply the centre (a transposed index would swap these)
    M = np.eye(4, dtype=np.float32)
    M[0, 1] = 2.0  # column 0, row 1: x contributes 2x to y
    got = oracle.sphere_transform(M.reshape(16), np.array([1.0, 0.0, 0.0, 0.0], np.float32))
    assert np.array_equal(got[:3], np.array([1.0, 2.0, 0.0], np.float32))


# ----------------------------------------------------------- propagation

@needs_ref
def test_tree_propagation_matches_reference_recursion():
    R = oracle.ref()
    # 3-level tree: 4 roots, 3 children each, 5 leaves each
    sizes = [4, 12, 60]
    parents = [None, np.arange(12) // 3, np.arange(60) // 5]
    locals_ = [rand_mats(n, 3.0) for n in sizes]
    prev = None
    worlds = []
    for d in range(3):
        w, _ = oracle.propagate_level(locals_[d], None if d == 0 else parents[d].astype(np.uint32), prev)
        worlds.append(w)
        prev = w
    # flatten into topological order for the reference-style recursion
    flat_local = np.concatenate(locals_)
    flat_parent = np.concatenate([np.full(4, -1), parents[1], 4 + parents[2]]).astype(np.int64)
    exp = np.zeros_like(flat_local)
    R.ref_propagate_tree(flat_local.ctypes.data, flat_parent.ctypes.data, flat_local.shape[0], exp.ctypes.data)
    assert np.concatenate(worlds).tobytes() == exp.tobytes()


def test_propagation_roots_copy_bits_and_out_of_range_parent():
    L = rand_mats(3)
    L[0, 0] = -0.0
    w, _ = oracle.propagate_level(L, None, None)
    assert w.tobytes() == L.tobytes()  # -0.0 survives: copy, not identity product
    P = rand_mats(2)
    w, _ = oracle.propagate_level(L, np.array([1, 7, 0], np.uint32), P)
    assert w[1].tobytes() == L[1].tobytes()  # parent 7 >= 2 acts as a root
    assert np.array_equal(w[0], oracle.mat4_mul(P[1], L[0]))


# ----------------------------------------------------------------- views

def _cam(**kw):
    base = dict(eye=(1.0, 2.0, 3.0), dir=(0.3, -0.2, -1.0), up=(0.0, 1.0, 0.0), fovy=float(np.float32(1.0)))
    base.update(kw)
    return Camera(**base)


@needs_ref
def test_view_matrix_matches_glm_lookatlh_and_perspective():
    R = oracle.ref()
    for i in range(200):
        eye = (rng.uniform(-600, 600, 3)).astype(np.float32)
        d = rng.standard_normal(3).astype(np.float32)
        cam = _cam(eye=tuple(map(float, eye)), dir=tuple(map(float, d)), fovy=float(np.float32(rng.uniform(0.5, 1.6))))
        v = oracle.make_view(cam, N.VIEW_DTYPE)[0]
        center = (eye + d).astype(np.float32)
        L = np.zeros(16, np.float32)
        R.ref_look_at_lh(eye.ctypes.data, center.ctypes.data, np.array(cam.up, np.float32).ctypes.data, L.ctypes.data)
        Lm = L.reshape(4, 4)  # [col][row]
        for r in range(3):
            assert v["viewRow"][r].tobytes() == np.array([Lm[0, r], Lm[1, r], Lm[2, r], Lm[3, r]], np.float32).tobytes()
        P = np.zeros(16, np.float32)
        R.ref_perspective_lh_zo(np.float32(cam.fovy), np.float32(cam.aspect), np.float32(cam.near),
                                np.float32(cam.far), P.ctypes.data)
        p00, p11 = P[0], P[5]
        ny = np.float32(np.sqrt(np.float32(p11 * p11 + np.float32(1))))
        assert v["sideY"][0] == np.float32(p11 / ny)
        assert v["lodScale"] == np.float32(np.float32(540.0) * p11)
        nx = np.float32(np.sqrt(np.float32(p00 * p00 + np.float32(1))))
        assert v["sideX"][1] == np.float32(np.float32(1) / nx)


@needs_ref
def test_view_coordinates_follow_glm_mat4_vec4():
    """The cull's v = V*(c,1) uses glm's pairwise (m0x+m1y)+(m2z+m3w) order."""
    R = oracle.ref()
    cam = _cam()
    vv = oracle.make_view(cam, N.VIEW_DTYPE)
    v = vv[0]
    M = np.eye(4, dtype=np.float32)
    for r in range(3):
        M[:, r] = v["viewRow"][r]  # M[col][row]
    for _ in range(1000):
        c = (rng.standard_normal(3) * 500).astype(np.float32)
        out = np.zeros(4, np.float32)
        R.ref_mat4_vec4(M.reshape(16).ctypes.data, np.array([*c, 1.0], np.float32).ctypes.data, out.ctypes.data)
        # a sphere exactly at a plane: r = 0 reduces each inequality to one coordinate
        m = np.zeros(1, N.MESH_DTYPE); m["lodCount"] = 1
        s = np.array([*c, 0.0], np.float32)
        vis = oracle.lib().oracle_test_drawable(s.ctypes.data, vv.ctypes.data, m.ctypes.data)
        near_ok = out[2] >= v["nearZ"]
        far_ok = out[2] <= v["farZ"]
        x_ok = np.float32(abs(out[0]) * v["sideX"][0]) - np.float32(out[2] * v["sideX"][1]) <= 0
        y_ok = np.float32(abs(out[1]) * v["sideY"][0]) - np.float32(out[2] * v["sideY"][1]) <= 0
        assert (vis >= 0) == bool(near_ok and far_ok and x_ok and y_ok)


# ------------------------------------------------------------------ cull

def _one_mesh(lods=3):
    m = np.zeros(1, N.MESH_DTYPE)
    m["radius"] = 1.0
    m["lodCount"] = lods
    m["indexCount"][0] = [300, 75, 18, 6]
    m["firstIndex"][0] = [0, 300, 375, 393]
    return m


def test_cull_geometry_against_float64_frustum():
    """Spheres classified in float64 by distance to the six frustum planes with a
    margin: clearly inside -> visible, clearly outside -> culled."""
    cam = Camera(eye=(0.0, 0.0, 0.0), dir=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fovy=float(np.float32(math.radians(60))))
    v = oracle.make_view(cam, N.VIEW_DTYPE)
    mesh = _one_mesh()
    n = 20000
    c = rng.uniform(-2500, 2500, (n, 3)).astype(np.float32)
    r = rng.uniform(0, 50, n).astype(np.float32)
    s = np.concatenate([c, r[:, None]], axis=1).astype(np.float32)
    t = math.tan(float(np.float32(cam.fovy)) / 2)
    ty, tx = t, t * 16 / 9
    # camera looks down -z: depth d = -z, right = -x (lookAtLH with up +y)
    d = -c[:, 2].astype(np.float64)
    x = c[:, 0].astype(np.float64)
    y = c[:, 1].astype(np.float64)
    dist = np.stack([d - 0.1, 2000 - d,
                     (tx * d - np.abs(x)) / math.sqrt(1 + tx * tx),
                     (ty * d - np.abs(y)) / math.sqrt(1 + ty * ty)], axis=1)
    inside = np.all(dist >= -r[:, None] + 1e-2, axis=1)
    outside = np.any(dist < -r[:, None] - 1e-2, axis=1)
    res = np.array([oracle.lib().oracle_test_drawable(s[i].ctypes.data, v.ctypes.data, mesh.ctypes.data) for i in range(n)])
    assert np.all(res[inside] >= 0)
    assert np.all(res[outside] < 0)
    assert inside.sum() > 100 and outside.sum() > 1000


def test_cull_lod_by_projected_radius():
    cam = Camera(eye=(0.0, 0.0, 0.0), dir=(0.0, 0.0, -1.0), up=(0.0, 1.0, 0.0), fovy=float(np.float32(math.radians(60))))
    v = oracle.make_view(cam, N.VIEW_DTYPE)
    mesh = _one_mesh(3)
    t = math.tan(float(np.float32(cam.fovy)) / 2)
    lods = []
    for depth in np.geomspace(0.5, 1900, 400):
        s = np.array([0, 0, -depth, 1.0], np.float32)
        px = 1.0 * 540 / (depth * t)  # projected radius in pixels (float64)
        lod = oracle.lib().oracle_test_drawable(s.ctypes.data, v.ctypes.data, mesh.ctypes.data)
        exp = 0 if px >= 64 * 1.001 else (1 if 8 * 1.001 <= px < 64 * 0.999 else (2 if px < 8 * 0.999 else None))
        if exp is not None:
            assert lod == exp, (depth, px, lod)
        lods.append(lod)
    assert lods == sorted(lods) and set(lods) == {0, 1, 2}  # monotone in distance
    # an eye-containing sphere is always visible, LOD 0
    s = np.array([0.5, 0.5, 0.5, 3.0], np.float32)
    assert oracle.lib().oracle_test_drawable(s.ctypes.data, v.ctypes.data, mesh.ctypes.data) == 0


def test_cull_compaction_invariants():
    n, K = 5000, 7
    counts = rng.multinomial(n, np.ones(K) / K).astype(np.uint32)
    counts[3] = 0  # an empty StateSet
    n = int(counts.sum())
    s = np.concatenate([rng.uniform(-300, 300, (n, 3)), rng.uniform(0.5, 5, (n, 1))], axis=1).astype(np.float32)
    meshes = np.zeros(4, N.MESH_DTYPE)
    meshes["lodCount"] = [1, 2, 3, 4]
    meshes["indexCount"] = np.arange(16).reshape(4, 4) + 100
    meshes["firstIndex"] = np.arange(16).reshape(4, 4) * 1000
    meshes["vertexOffset"] = [0, -5, 7, 9]
    mi = rng.integers(0, 5, n).astype(np.uint32)  # 4 is out of range: never visible
    cams = [Camera(eye=(0.0, 0.0, 0.0), dir=tuple(map(float, rng.standard_normal(3))), up=(0.0, 1.0, 0.0),
                   fovy=float(np.float32(1.2))) for _ in range(3)]
    views = np.concatenate([oracle.make_view(c, N.VIEW_DTYPE) for c in cams])
    cmds, cnt = oracle.cull(s, mi, meshes, counts, views, cmd_dtype=N.CMD_DTYPE)
    base = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    for vi in range(3):
        total = 0
        for k in range(K):
            sel = cmds[vi, base[k]: base[k] + cnt[vi, k]]
            ids = sel["firstInstance"].astype(np.int64)
            assert np.all(np.diff(ids) > 0)  # stable order
            assert np.all((ids >= base[k]) & (ids < base[k + 1]))  # survivors of this StateSet only
            # brute force: exactly the drawables of the range that test visible
            exp_ids = [i for i in range(base[k], base[k + 1]) if mi[i] < 4 and
                       oracle.lib().oracle_test_drawable(s[i].ctypes.data, views[vi:vi + 1].ctypes.data,
                                                         meshes[mi[i]:mi[i] + 1].ctypes.data) >= 0]
            assert list(ids) == exp_ids
            assert np.all(sel["instanceCount"] == 1)
            assert np.all(sel["vertexOffset"] == meshes["vertexOffset"][mi[ids]])
            total += cnt[vi, k]
        assert cnt[vi, 3] == 0
        assert total > 0
