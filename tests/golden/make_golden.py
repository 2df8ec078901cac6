"""Writes tests/golden/ref_fp32.npz: input/output vectors of the reference's own
FP32 functions, computed by oracle/_ref/libcadr_ref.so (GLM 0.9.9 and
src/CadR/BoundingSphere.h compiled from the reference tree by
`make -C oracle ref`).

The fixtures let the parity tests check the oracle and the HIP kernels
against the reference's arithmetic where the reference tree (and so
oracle/_ref) is absent.

  mat_a, mat_b -> mat_ab        glm operator*(mat4, mat4)       type_mat4x4.inl:630-648
  sph_m, sph_s -> sph_out       CadR operator*(mat4, BoundingSphere)  BoundingSphere.h:70-87
  tree_local, tree_parent -> tree_world   node recursion of gltfReader/main.cpp:1538-1576
  cam_eye, cam_dir (center = eye + dir), cam_up -> cam_look      glm::lookAtLH   matrix_transform.inl:122-142
  cam_fovy, cam_aspect, cam_near, cam_far -> cam_proj   glm::perspectiveLH_ZO  matrix_clip_space.inl:255-268

Run: python tests/golden/make_golden.py   (needs oracle/_ref built)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libcadr_ref.so is not built: make -C oracle ref")
    R = oracle.ref()
    P = oracle._p
    rng = np.random.default_rng(0xCAD0601D)
    out = {}

    # mat4 * mat4: general, ill-conditioned magnitudes, and cancellation-prone inputs
    n = 512
    a = (rng.standard_normal((n, 16)) * np.exp(rng.uniform(-6, 6, (n, 1)))).astype(np.float32)
    b = (rng.standard_normal((n, 16)) * np.exp(rng.uniform(-6, 6, (n, 1)))).astype(np.float32)
    b[: n // 4] = np.round(b[: n // 4] * 64) / 64  # exactly representable, sums cancel
    ab = np.zeros_like(a)
    for i in range(n):
        R.ref_mat4_mul(P(a[i]), P(b[i]), P(ab[i]))
    out.update(mat_a=a, mat_b=b, mat_ab=ab)

    # world bounding spheres: rotations with non-uniform scales, translations, odd radii
    sm = (rng.standard_normal((n, 16)) * 3).astype(np.float32)
    sm[:, 3] = sm[:, 7] = sm[:, 11] = 0.0
    sm[:, 15] = 1.0
    ss = np.concatenate([rng.standard_normal((n, 3)) * 5, rng.uniform(0.01, 40, (n, 1))], axis=1).astype(np.float32)
    so = np.zeros_like(ss)
    for i in range(n):
        R.ref_sphere_transform(P(sm[i]), P(ss[i]), P(so[i]))
    out.update(sph_m=sm, sph_s=ss, sph_out=so)

    # a 3-level tree (8 roots -> 4 children -> 6 leaves), parents before children
    sizes = [8, 32, 192]
    local = (rng.standard_normal((sum(sizes), 16)) * 2).astype(np.float32)
    local[:, 3] = local[:, 7] = local[:, 11] = 0.0
    local[:, 15] = 1.0
    parent = np.full(sum(sizes), -1, dtype=np.int64)
    parent[8:40] = np.arange(32) // 4
    parent[40:] = 8 + np.arange(192) // 6
    world = np.zeros_like(local)
    R.ref_propagate_tree(P(local), P(parent), sum(sizes), P(world))
    out.update(tree_local=local, tree_parent=parent, tree_sizes=np.array(sizes, np.int64), tree_world=world)

    # cameras
    c = 64
    eye = (rng.uniform(-600, 600, (c, 3))).astype(np.float32)
    d = rng.standard_normal((c, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(1, 100, (c, 1))).astype(np.float32)
    center = (eye + d).astype(np.float32)  # as cadrMakeView forms it: center = eye + dir in float
    up = np.tile(np.array([0, 1, 0], np.float32), (c, 1))
    look = np.zeros((c, 16), np.float32)
    fovy = np.radians(rng.uniform(30, 100, c)).astype(np.float32)
    aspect = rng.uniform(1.0, 2.5, c).astype(np.float32)
    near = rng.uniform(0.05, 1.0, c).astype(np.float32)
    far = rng.uniform(500, 5000, c).astype(np.float32)
    proj = np.zeros((c, 16), np.float32)
    for i in range(c):
        R.ref_look_at_lh(P(eye[i]), P(center[i]), P(up[i]), P(look[i]))
        R.ref_perspective_lh_zo(float(fovy[i]), float(aspect[i]), float(near[i]), float(far[i]), P(proj[i]))
    out.update(cam_eye=eye, cam_dir=d, cam_center=center, cam_up=up, cam_look=look, cam_fovy=fovy, cam_aspect=aspect,
               cam_near=near, cam_far=far, cam_proj=proj)

    path = os.path.join(HERE, "ref_fp32.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
