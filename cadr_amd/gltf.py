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

Mesh data (main.cpp:1392-1418, 2700-3770), for the cull's mesh table:

- buffers: a "uri" relative to the file's folder, percent-decoded, as the
  reference reads them; also (beyond the reference) a base64 data URI and
  the BIN chunk of a .glb for a buffer without "uri";
- accessors: "bufferView" required, "sparse" refused; stride = the view's
  "byteStride" or the element size; the accessor's range
  (count-1)*stride + elementSize from its "byteOffset" must lie inside the
  view, and the view's inside the buffer (main.cpp:2827-2869);
- primitives: attributes POSITION (VEC3 float, "min"/"max" required),
  NORMAL, COLOR_0, TEXCOORD_n and TANGENT are checked (type, component type,
  normalized flag, one count for all); any other attribute is refused.
  Indices (SCALAR, unsigned byte / short / int) or, without them, the
  vertices in order; mode 4/1/0 copy them, 5 (strip), 6 (fan), 3 (line
  strip) and 2 (line loop) are converted to triangle / line lists exactly
  as main.cpp:3259-3585 does (strip triangles alternate v1 v2 v3 / v2 v1 v3;
  fans keep the first vertex first);
- bounds: a primitive's box is its POSITION accessor's min/max (float32 of
  the JSON numbers); its sphere is centred on the box centre (min+max)/2
  with radius sqrt of the largest squared distance of a position
  (BoundingSphere::extendRadiusByPointUsingRadius2, BoundingSphere.h:65-69,
  glm::distance2 = ((dx*dx + dy*dy) + dz*dz)); a mesh's sphere is centred on
  the union of its primitives' boxes and grown by each primitive sphere
  (BoundingSphere::extendRadiusBy, BoundingSphere.h:50-58).

Each primitive becomes one entry of the cull's mesh table (its sphere, one
LOD, index count, and offsets into one shared index / position buffer), and
each (mesh instance, primitive) pair one drawable.  The reference draws a
primitive once with all instances of its mesh (instanceCount = number of
matrices); the cull needs a sphere per instance, so here every instance is
its own drawable.  StateSets group the drawables by material (the default
material first), the drawables of a StateSet in (mesh, primitive, instance)
order; the reference's StateSets also split by pipeline, which follows from
shader state outside this path.

Instead of the recursion, the walk is flattened into levels by depth: level d
holds every node instance at depth d in depth-first order, with the index of
its parent instance in level d-1, which is what cadrPropagateTransforms takes
(one launch per level).  A node reached along several paths gets one instance
per path, as the recursion visits it once per path.  The world matrices are
computed on the GPU; instance lists per mesh are gathered from them.
"""
from __future__ import annotations

import base64
import json
import os
import struct
import urllib.parse
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


def read_gltf(path: str):
    """(JSON document, BIN chunk or None) of a .gltf file or .glb container."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:4] == b"glTF":
        if len(data) < 20:
            raise GltfError("truncated glb header")
        length, ctype = struct.unpack_from("<II", data, 12)
        if ctype != 0x4E4F534A:  # "JSON"
            raise GltfError("the first glb chunk is not JSON")
        doc = json.loads(data[20:20 + length].decode("utf-8"))
        pos = 20 + length
        bin_chunk = None
        if pos + 8 <= len(data):
            blen, btype = struct.unpack_from("<II", data, pos)
            if btype == 0x004E4942:  # "BIN"
                if pos + 8 + blen > len(data):
                    raise GltfError("truncated glb BIN chunk")
                bin_chunk = data[pos + 8:pos + 8 + blen]
        return doc, bin_chunk
    return json.loads(data.decode("utf-8")), None


def read_document(path: str) -> dict:
    """The JSON document of a .gltf file or of a .glb container's JSON chunk."""
    return read_gltf(path)[0]


def load_buffers(doc: dict, base_dir: str = ".", bin_chunk: bytes | None = None) -> list:
    """Contents of doc["buffers"] (main.cpp:1392-1418): a relative "uri" is
    percent-decoded and read from base_dir; a base64 data URI and, for a
    buffer without "uri", the .glb BIN chunk are accepted too."""
    out = []
    for b in doc.get("buffers", []):
        uri = b.get("uri")
        if uri is None:
            if bin_chunk is None:
                raise GltfError("Unsupported functionality: Undefined buffer.uri.")
            out.append(bytes(bin_chunk))
        elif uri.startswith("data:"):
            head, _, payload = uri.partition(",")
            if not head.endswith(";base64"):
                raise GltfError("Unsupported functionality: data URI without base64 encoding.")
            out.append(base64.b64decode(payload))
        else:
            p = urllib.parse.unquote(uri)
            if not os.path.isabs(p):
                p = os.path.join(base_dir, p)
            try:
                with open(p, "rb") as f:
                    out.append(f.read())
            except OSError as e:
                raise GltfError(f"Error opening file {p}.") from e
    return out


_COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_TYPE_SIZE = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}


def _data_range(doc: dict, buffers: list, accessor: dict, count: int, elem_size: int):
    """(buffer bytes, byte offset, stride) of an accessor's elements, with
    the reference's checks (main.cpp:2827-2869)."""
    if "sparse" in accessor:
        raise GltfError("Unsupported functionality: Property sparse.")
    if "bufferView" not in accessor:
        raise GltfError("Unsupported functionality: Omitted bufferView.")
    views = doc.get("bufferViews", [])
    vi = accessor["bufferView"]
    if not isinstance(vi, int) or not 0 <= vi < len(views):
        raise GltfError(f"bufferView index {vi} out of range")
    view = views[vi]
    stride = int(view.get("byteStride", elem_size))
    size = (count - 1) * stride + elem_size
    offset = int(accessor.get("byteOffset", 0))
    if offset + size > int(view["byteLength"]):
        raise GltfError("Accessor range is not completely inside its BufferView.")
    offset += int(view.get("byteOffset", 0))
    bi = view["buffer"]
    if not isinstance(bi, int) or not 0 <= bi < len(doc.get("buffers", [])):
        raise GltfError(f"buffer index {bi} out of range")
    if offset + size > int(doc["buffers"][bi]["byteLength"]):
        raise GltfError("BufferView range is not completely inside its Buffer.")
    if offset + size > len(buffers[bi]):
        raise GltfError("BufferView range is not completely inside data range.")
    return buffers[bi], offset, stride


def _elements(data: bytes, offset: int, stride: int, count: int, comps: int, dtype) -> np.ndarray:
    """count x comps elements of dtype starting at offset, stride bytes apart."""
    item = np.dtype(dtype).itemsize
    a = np.ndarray((count, comps), dtype=np.dtype(dtype).newbyteorder("<"), buffer=data, offset=offset,
                   strides=(stride, item))
    return np.array(a, dtype=dtype)


def _accessor(doc: dict, index) -> dict:
    acc = doc.get("accessors", [])
    if not isinstance(index, int) or not 0 <= index < len(acc):
        raise GltfError(f"accessor index {index} out of range")
    return acc[index]


def _vec3_minmax(accessor: dict, key: str) -> np.ndarray:
    if key not in accessor:
        raise GltfError(f"Accessor.{key} must be defined for POSITION accessor.")
    a = accessor[key]
    if not isinstance(a, list) or len(a) != 3:
        raise GltfError(f"POSITION's Accessor.{key} is not vector of three components.")
    return np.array([np.float32(float(x)) for x in a], dtype=F)


def _count(accessor: dict, num_vertices: int) -> int:
    """main.cpp updateNumVertices: every attribute has the same count >= 1."""
    count = int(accessor["count"])
    if num_vertices != count:
        if num_vertices == 0:
            if count == 0:
                raise GltfError("Accessor's count member must be greater than zero.")
            return count
        raise GltfError("Number of elements is not the same for all primitive attributes.")
    return num_vertices


def _check_attribute(name: str, accessor: dict):
    """Type rules of the non-position attributes; returns the element size."""
    t, ct = accessor["type"], int(accessor["componentType"])
    normalized = bool(accessor.get("normalized", False))
    if name in ("NORMAL", "TANGENT"):
        want = "VEC3" if name == "NORMAL" else "VEC4"
        if t != want:
            raise GltfError(f"{name.capitalize()} attribute is not of {want} type.")
        if ct != 5126:
            raise GltfError(f"{name.capitalize()} attribute componentType is not float.")
        if normalized:
            raise GltfError(f"{name.capitalize()} attribute normalized flag is true.")
        return 12 if name == "NORMAL" else 16
    if name == "COLOR_0":
        if t not in ("VEC3", "VEC4"):
            raise GltfError("Color attribute is not of VEC3 or VEC4 type.")
        if ct not in (5126, 5121, 5123):
            raise GltfError("Color attribute componentType is not float, unsigned byte, or unsigned short.")
        if normalized and ct == 5126:
            raise GltfError("Color attribute component type is set to float while normalized flag is true.")
        if not normalized and ct != 5126:
            raise GltfError("Color attribute component type is set to unsigned byte or unsigned short while "
                            "normalized flag is not true.")
        return _TYPE_SIZE[t] * np.dtype(_COMPONENT[ct]).itemsize
    # TEXCOORD_n
    if t != "VEC2":
        raise GltfError("TexCoord attribute is not of VEC2 type.")
    if ct not in (5126, 5121, 5123):
        raise GltfError("TexCoord attribute componentType is not float, unsigned byte, or unsigned short.")
    if normalized and ct == 5126:
        raise GltfError("TexCoord attribute component type is set to float while normalized flag is true.")
    if not normalized and ct != 5126:
        raise GltfError("TexCoord attribute component type is set to unsigned byte or unsigned short while "
                        "normalized flag is not true.")
    return {5126: 8, 5121: 2, 5123: 4}[ct]


def convert_indices(mode: int, src, num_vertices: int) -> np.ndarray:
    """uint32 index list of a primitive (main.cpp:3259-3585): src holds the
    primitive's indices or is None (the vertices in order); strips, fans and
    line strips / loops become triangle / line lists."""
    n = num_vertices if src is None else int(src.shape[0])
    s = np.arange(n, dtype=np.uint32) if src is None else np.asarray(src, dtype=np.uint32)
    need = {4: 3, 1: 2, 0: 1, 5: 3, 6: 3, 3: 2, 2: 2}
    names = {4: "TRIANGLES", 1: "LINES", 0: "POINTS", 5: "TRIANGLE_STRIP", 6: "TRIANGLE_FAN", 3: "LINE_STRIP",
             2: "LINE_LOOP"}
    if mode not in need:
        raise GltfError("Invalid value for mesh.primitive.mode.")
    if n < need[mode]:
        raise GltfError(f"Invalid number of indices for {names[mode]}.")
    out_n = {4: n, 1: n, 0: n, 5: (n - 2) * 3, 6: (n - 2) * 3, 3: (n - 1) * 2, 2: n * 2}[mode]
    if src is None and out_n >= 0xFFFFFFFF - 1:
        raise GltfError("Too large primitive. Index out of 32-bit integer range.")
    if mode in (4, 1, 0):
        return s.copy()
    if mode == 5:  # triangle k: (k, k+1, k+2), the first two swapped for odd k
        k = np.arange(n - 2)
        odd = (k & 1).astype(bool)
        a, b = np.where(odd, k + 1, k), np.where(odd, k, k + 1)
        return np.stack([s[a], s[b], s[k + 2]], axis=1).reshape(-1)
    if mode == 6:  # triangle k: (0, k+1, k+2)
        k = np.arange(n - 2)
        return np.stack([np.full(n - 2, s[0], np.uint32), s[k + 1], s[k + 2]], axis=1).reshape(-1)
    k = np.arange(n - 1)
    lines = np.stack([s[k], s[k + 1]], axis=1).reshape(-1)
    if mode == 3:
        return lines
    return np.concatenate([lines, np.array([s[n - 1], s[0]], np.uint32)])


def primitive_sphere(positions: np.ndarray, box_min: np.ndarray, box_max: np.ndarray):
    """(centre, radius) float32: centre (min+max)/2 of the POSITION box,
    radius sqrt of the largest squared distance of a position (float32 as
    glm computes it: d = centre - p, (d.x*d.x + d.y*d.y) + d.z*d.z)."""
    c = ((box_min + box_max) / F(2)).astype(F)
    p = np.asarray(positions, dtype=F)
    r2 = F(0)
    if p.shape[0]:
        d = (c[None, :] - p).astype(F)
        d2 = ((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]).astype(F)
        d2 = d2[~np.isnan(d2)]  # `d2 > r2` is false for NaN
        if d2.size:
            r2 = max(r2, F(d2.max()))
    return c, F(np.sqrt(r2))


def mesh_sphere(boxes, spheres):
    """(centre, radius) of a mesh from its primitives' boxes and spheres:
    centre of the union box, radius grown by each primitive sphere in order
    (BoundingSphere::extendRadiusBy)."""
    mn = np.full(3, np.inf, F)
    mx = np.full(3, -np.inf, F)
    for bmin, bmax in boxes:
        for a in range(3):
            if bmin[a] < mn[a]:
                mn[a] = bmin[a]
            if bmax[a] > mx[a]:
                mx[a] = bmax[a]
    c = ((mn + mx) / F(2)).astype(F)
    r = F(0)
    for sc, sr in spheres:
        d = (np.asarray(sc, F) - c).astype(F)
        d2 = F(F(F(d[0] * d[0]) + F(d[1] * d[1])) + F(d[2] * d[2]))
        dif = F(r - F(sr))
        if dif < 0 or d2 > F(dif * dif):
            # the unqualified sqrt(float) of BoundingSphere.h:57, inside
            # namespace CadR, is ::sqrt(double): one rounding to float
            r = F(np.sqrt(np.float64(d2)) + np.float64(sr))
    return c, r


@dataclass
class Primitive:
    mesh: int
    index: int            # within the mesh
    mode: int
    indices: np.ndarray   # uint32, converted to a list
    positions: np.ndarray  # float32 [n, 3]
    box: tuple            # (min, max) float32
    sphere: tuple         # (centre float32[3], radius float32)
    material: int         # -1: the default material


@dataclass
class MeshData:
    """Primitives of every mesh, their shared index / position buffers and
    the cull's mesh table (one entry per primitive)."""
    primitives: list = field(default_factory=list)
    mesh_primitives: list = field(default_factory=list)  # per mesh: indices into primitives
    mesh_spheres: list = field(default_factory=list)     # per mesh: (centre, radius) or None (no instances)

    def buffers(self):
        """(uint32 index buffer, float32 [n, 4] position buffer with w = 1,
        per-primitive first index and vertex offset)."""
        first, vofs, idx, pos = [], [], [], []
        fi = vo = 0
        for p in self.primitives:
            first.append(fi); vofs.append(vo)
            idx.append(p.indices); pos.append(p.positions)
            fi += p.indices.shape[0]; vo += p.positions.shape[0]
        ib = np.concatenate(idx) if idx else np.zeros(0, np.uint32)
        pb = np.concatenate(pos) if pos else np.zeros((0, 3), F)
        pb4 = np.concatenate([pb, np.ones((pb.shape[0], 1), F)], axis=1)
        return ib, pb4, np.array(first, np.int64), np.array(vofs, np.int64)

    def mesh_table(self) -> np.ndarray:
        """cadrMesh per primitive: its sphere, one LOD, its index count and
        offsets into buffers()."""
        from ._native import MESH_DTYPE
        _, _, first, vofs = self.buffers()
        t = np.zeros(len(self.primitives), dtype=MESH_DTYPE)
        for i, p in enumerate(self.primitives):
            t["center"][i] = p.sphere[0]
            t["radius"][i] = p.sphere[1]
            t["lodCount"][i] = 1
            t["vertexOffset"][i] = vofs[i]
            t["indexCount"][i, 0] = p.indices.shape[0]
            t["firstIndex"][i, 0] = first[i]
        return t


def load_meshes(doc: dict, buffers: list, h: "Hierarchy | None" = None) -> MeshData:
    """Primitives and bounds of every mesh (main.cpp:2700-3770).  With a
    Hierarchy, meshes without instances get no sphere (the reference skips
    them) and a document without meshes is refused as the reference does."""
    meshes = doc.get("meshes", [])
    if h is not None and not meshes:
        raise GltfError("No meshes in the model.")
    instanced = None
    if h is not None:
        instanced = set(int(h.mesh[d][i]) for d, i in h.order if int(h.mesh[d][i]) >= 0)
    md = MeshData()
    for mi, mesh in enumerate(meshes):
        prims = mesh.get("primitives")
        if prims is None or (not prims and (instanced is None or mi in instanced)):
            raise GltfError("No primitives in the mesh.")
        ids, boxes, spheres = [], [], []
        for pi, prim in enumerate(prims):
            attrs = prim["attributes"]
            num_vertices, positions, box = 0, None, None
            for name, ai in attrs.items():
                acc = _accessor(doc, ai)
                if name == "POSITION":
                    if acc["type"] != "VEC3":
                        raise GltfError("Position attribute is not of VEC3 type.")
                    if int(acc["componentType"]) != 5126:
                        raise GltfError("Position attribute componentType is not float.")
                    if acc.get("normalized", False):
                        raise GltfError("Position attribute normalized flag is true.")
                    num_vertices = _count(acc, num_vertices)
                    data, off, stride = _data_range(doc, buffers, acc, num_vertices, 12)
                    positions = _elements(data, off, stride, num_vertices, 3, F)
                    box = (_vec3_minmax(acc, "min"), _vec3_minmax(acc, "max"))
                elif name in ("NORMAL", "COLOR_0", "TANGENT") or (
                        name.startswith("TEXCOORD_") and name[9:].isdigit()):
                    size = _check_attribute(name if not name.startswith("TEXCOORD_") else "TEXCOORD", acc)
                    num_vertices = _count(acc, num_vertices)
                    _data_range(doc, buffers, acc, num_vertices, size)
                else:
                    raise GltfError(f"Unsupported functionality: {name} attribute.")
            src = None
            if "indices" in prim:
                acc = _accessor(doc, prim["indices"])
                if acc["type"] != "SCALAR":
                    raise GltfError("Indices are not of SCALAR type.")
                ct = int(acc["componentType"])
                if ct not in (5125, 5123, 5121):
                    raise GltfError("Index componentType is not unsigned int, unsigned short or unsigned byte.")
                if acc.get("normalized", False):
                    raise GltfError("Indices cannot have normalized flag set to true.")
                n = int(acc["count"])
                if n == 0:
                    raise GltfError("Accessor's count member must be greater than zero.")
                dt = _COMPONENT[ct]
                data, off, _ = _data_range(doc, buffers, acc, n, np.dtype(dt).itemsize)
                src = _elements(data, off, np.dtype(dt).itemsize, n, 1, dt)[:, 0]
            if src is None and num_vertices == 0:
                continue  # an empty primitive (main.cpp:3183-3185)
            if positions is None:
                # the reference leaves the bounds undefined here; refuse instead
                raise GltfError("A primitive without POSITION attribute.")
            mode = int(prim.get("mode", 4))
            indices = convert_indices(mode, src, num_vertices)
            sphere = primitive_sphere(positions, *box)
            md.primitives.append(Primitive(mi, pi, mode, indices, positions, box, sphere,
                                           int(prim.get("material", -1))))
            ids.append(len(md.primitives) - 1)
            boxes.append(box)
            spheres.append(sphere)
        md.mesh_primitives.append(ids)
        if (instanced is None or mi in instanced) and ids:
            md.mesh_spheres.append(mesh_sphere(boxes, spheres))
        else:
            md.mesh_spheres.append(None)
    return md


@dataclass
class Drawables:
    """Cull drawables of a glTF scene: instance[i] = (depth, index) of the
    node instance whose world matrix drawable i uses, mesh_index[i] its
    primitive's entry in the mesh table, stateset_counts per StateSet
    (materials, the default one first)."""
    instance: np.ndarray
    mesh_index: np.ndarray
    stateset_counts: np.ndarray
    materials: list

    @property
    def count(self) -> int:
        return int(self.mesh_index.shape[0])

    def world_matrices(self, world_levels) -> np.ndarray:
        """float32 [n, 16]: each drawable's node-instance world matrix."""
        out = np.zeros((self.count, 16), dtype=F)
        for i, (d, k) in enumerate(self.instance):
            out[i] = world_levels[int(d)][int(k)]
        return out


def drawables(h: "Hierarchy", md: MeshData) -> Drawables:
    """One drawable per (mesh instance, primitive), grouped by material."""
    inst = [[] for _ in range(h.num_meshes)]
    for d, i in h.order:
        m = int(h.mesh[d][i])
        if m >= 0:
            inst[m].append((d, i))
    rows = []
    for m, ids in enumerate(md.mesh_primitives):
        for pid in ids:
            for di in inst[m]:
                rows.append((md.primitives[pid].material, di, pid))
    mats = sorted(set(r[0] for r in rows))
    key = {m: k for k, m in enumerate(mats)}
    order = sorted(range(len(rows)), key=lambda r: key[rows[r][0]])  # stable: creation order inside a material
    instance = np.array([rows[r][1] for r in order], dtype=np.int64).reshape(-1, 2)
    mesh_index = np.array([rows[r][2] for r in order], dtype=np.uint32)
    counts = np.bincount([key[rows[r][0]] for r in order], minlength=len(mats)).astype(np.uint32)
    return Drawables(instance, mesh_index, counts, mats)


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


def cull_scene(h: Hierarchy, md: MeshData, dr: Drawables, world_levels, cameras=()):
    """The cull's input for a glTF scene as a flat scenes.Scene: one root
    level whose matrices are the drawables' world matrices (a root level's
    world matrices are its local ones, so propagating it copies them and
    writes the world spheres), the mesh table of md, the drawables' mesh
    indices and StateSet counts."""
    from .scenes import Scene
    W = dr.world_matrices(world_levels)
    return Scene("gltf", [(W, None)], dr.mesh_index.copy(), md.mesh_table(), dr.stateset_counts.copy(),
                 list(cameras))
