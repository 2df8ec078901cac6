"""glTF node hierarchy as the transform-propagation input.

Follows examples/gltfReader/main.cpp:1422-1576 of the reference:

- node matrix (:1437-1500): "matrix" (16 numbers, column-major), which must
  not come with "translation", "rotation" or "scale"; otherwise
  scale (default 1,1,1), then with a "rotation" quaternion {x,y,z,w} the
  columns of glm::mat3(q) (mat3_cast) multiplied by scale.x, .y, .z and the
  row 0,0,0,1, without one diag(scale, 1); a "translation" replaces column 3's
  x, y, z.  All in float32, each operation rounded as glm does it;
- the default scene ("scene", else scene 0) is walked depth first from its
  root nodes (:1518-1576); a root node's matrix has its y and z rows negated
  (glTF to Vulkan axes, :1547-1553) and every other node's world matrix is
  parentWorld * node matrix (glm mat4 product);
- a node with a "mesh" appends its world matrix to that mesh's instance list,
  in depth-first order (:1529-1531, :1556-1558).

Instead of the recursion, the walk is flattened into levels by depth: level d
holds every node instance at depth d in depth-first order, with the index of
its parent instance in level d-1, which is what cadrPropagateTransforms takes
(one launch per level).  A node reached along several paths gets one instance
per path, as the recursion visits it once per path.  The world matrices are
computed on the GPU; instance lists per mesh are gathered from them.
"""
from __future__ import annotations

import json
import os
import struct
from dataclasses import dataclass, field

import numpy as np

from . import _native as N


class GltfError(ValueError):
    """A malformed glTF node (the reference's GltfError)."""


F = np.float32


def _floats(node, key, n, what):
    a = node[key]
    if not isinstance(a, list) or len(a) != n:
        raise GltfError(f"Node.{key} is not vector of {what} components.")
    return np.array([F(x) for x in a], dtype=F)


def node_matrix(node: dict) -> np.ndarray:
    """Column-major float32 mat4 of one glTF node (main.cpp:1437-1500)."""
    if "matrix" in node:
        if "translation" in node or "rotation" in node or "scale" in node:
            raise GltfError("If matrix is provided for the node, translation, rotation and scale must not be present.")
        return _floats(node, "matrix", 16, "16")
    s = _floats(node, "scale", 3, "three") if "scale" in node else np.ones(3, dtype=F)
    M = np.zeros(16, dtype=F)
    if "rotation" in node:
        x, y, z, w = _floats(node, "rotation", 4, "four")
        qxx, qyy, qzz = x * x, y * y, z * z
        qxz, qxy, qyz = x * z, x * y, y * z
        qwx, qwy, qwz = w * x, w * y, w * z
        one, two = F(1), F(2)
        cols = ((one - two * (qyy + qzz), two * (qxy + qwz), two * (qxz - qwy)),
                (two * (qxy - qwz), one - two * (qxx + qzz), two * (qyz + qwx)),
                (two * (qxz + qwy), two * (qyz - qwx), one - two * (qxx + qyy)))
        for c in range(3):
            for k in range(3):
                M[4 * c + k] = F(cols[c][k] * s[c])
    else:
        M[0], M[5], M[10] = s[0], s[1], s[2]
    M[15] = F(1)
    if "translation" in node:
        M[12:15] = _floats(node, "translation", 3, "three")
    return M


def root_flip(M: np.ndarray) -> np.ndarray:
    """y and z rows negated in every column (main.cpp:1547-1553)."""
    out = np.array(M, dtype=F).reshape(4, 4).copy()
    out[:, 1] = -out[:, 1]
    out[:, 2] = -out[:, 2]
    return out.reshape(16)


def read_document(path: str) -> dict:
    """The JSON document of a .gltf file or of a .glb container's JSON chunk."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:4] == b"glTF":
        if len(data) < 20:
            raise GltfError("truncated glb header")
        length, ctype = struct.unpack_from("<II", data, 12)
        if ctype != 0x4E4F534A:  # "JSON"
            raise GltfError("the first glb chunk is not JSON")
        return json.loads(data[20:20 + length].decode("utf-8"))
    return json.loads(data.decode("utf-8"))


@dataclass
class Hierarchy:
    """Node instances of the default scene by depth.  levels[d] = (local
    float32 [n_d,16], parent uint32 [n_d] or None for d = 0), the layout of
    scenes.Scene levels; node[d] / mesh[d] give each instance's glTF node and
    mesh index (-1: none); order lists (depth, index) in depth-first order."""
    levels: list = field(default_factory=list)
    node: list = field(default_factory=list)
    mesh: list = field(default_factory=list)
    order: list = field(default_factory=list)
    num_meshes: int = 0

    @property
    def num_instances(self) -> int:
        return sum(l[0].shape[0] for l in self.levels)

    def mesh_instances(self, world_levels) -> list:
        """Per mesh, the world matrices of its instances in depth-first order
        (the reference's meshMatrixList)."""
        out = [[] for _ in range(self.num_meshes)]
        for d, i in self.order:
            m = int(self.mesh[d][i])
            if m >= 0:
                out[m].append(world_levels[d][i])
        return [np.array(x, dtype=F).reshape(len(x), 16) for x in out]


def load(doc) -> Hierarchy:
    """Hierarchy of a glTF document (dict, or a path to .gltf / .glb)."""
    if isinstance(doc, (str, os.PathLike)):
        doc = read_document(os.fspath(doc))
    nodes = doc.get("nodes", [])
    mats = [node_matrix(n) for n in nodes]
    meshes = doc.get("meshes", [])
    h = Hierarchy(num_meshes=len(meshes))
    scenes = doc.get("scenes", [])
    if not scenes:
        return h
    si = doc.get("scene", 0)
    if not isinstance(si, int) or not 0 <= si < len(scenes):
        raise GltfError(f"scene index {si} out of range")
    scene = scenes[si]
    local, parent, node_of, mesh_of = [], [], [], []

    def level(d):
        while len(local) <= d:
            local.append([]); parent.append([]); node_of.append([]); mesh_of.append([])

    # iterative depth-first walk; a path longer than the node count has a cycle
    for root in scene.get("nodes", []):
        stack = [(root, 0, -1)]
        while stack:
            n, d, p = stack.pop()
            if not isinstance(n, int) or n < 0 or n >= len(nodes):
                raise GltfError(f"node index {n} out of range")
            if d > len(nodes):
                raise GltfError("the node hierarchy has a cycle")
            level(d)
            i = len(local[d])
            local[d].append(root_flip(mats[n]) if d == 0 else mats[n])
            parent[d].append(p)
            node_of[d].append(n)
            m = nodes[n].get("mesh", -1)
            if m != -1 and not (isinstance(m, int) and 0 <= m < len(meshes)):
                raise GltfError(f"mesh index {m} out of range")
            mesh_of[d].append(int(m))
            h.order.append((d, i))
            for c in reversed(nodes[n].get("children", [])):
                stack.append((c, d + 1, i))
    for d in range(len(local)):
        L = np.array(local[d], dtype=F).reshape(len(local[d]), 16)
        P = None if d == 0 else np.array(parent[d], dtype=np.uint32)
        h.levels.append((L, P))
        h.node.append(np.array(node_of[d], dtype=np.int64))
        h.mesh.append(np.array(mesh_of[d], dtype=np.int64))
    return h


def propagate(h: Hierarchy, stream=None) -> list:
    """World matrices of every instance, level by level on the GPU
    (cadrPropagateTransforms, world = parentWorld * local, roots copied)."""
    from .device import DeviceBuffer, synchronize
    lib = N.load()
    bufs, worlds, prev = [], [], None
    try:
        for L, P in h.levels:
            n = L.shape[0]
            lb = DeviceBuffer.from_array(L)
            wb = DeviceBuffer(max(64 * n, 16))
            bufs += [lb, wb]
            worlds.append((n, wb))
            lv = np.zeros(1, dtype=N.LEVEL_DTYPE)
            lv["localMatrices"], lv["worldMatrices"], lv["count"] = lb.address, wb.address, n
            if prev is not None:
                pb = DeviceBuffer.from_array(P)
                bufs.append(pb)
                lv["parentIndex"], lv["parentWorld"], lv["parentCount"] = pb.address, prev[1].address, prev[0]
            N.check(lib.cadrPropagateTransforms(N.ptr(lv), stream))
            prev = (n, wb)
        synchronize()
        return [wb.download(np.float32, 16 * n).reshape(n, 16) for n, wb in worlds]
    finally:
        for b in bufs:
            b.close()
