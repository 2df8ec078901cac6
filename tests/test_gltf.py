"""glTF ingestion (examples/gltfReader/main.cpp:1392-1576, 2700-3770): node
matrices, the root axis flip, the depth-first walk flattened into levels, GPU
propagation against the reference's recursion restated on the oracle; mesh
data: buffers, accessors, index lists per primitive mode, primitive and mesh
bounding spheres against the reference's BoundingSphere code, and a document
culled end to end on the GPU against the oracle."""
import json
import struct

import numpy as np
import pytest

import oracle
from cadr_amd import gltf

F = np.float32


def random_document(seed, n_nodes=300, n_meshes=7, shared=0.05, depth_limit=40):
    """Nodes whose children have larger indices (no cycles), a few reached from
    two parents, random matrix / TRS content, two scenes with scene 1 default."""
    rng = np.random.default_rng(seed)
    nodes = [dict() for _ in range(n_nodes)]
    depth = [0] * n_nodes
    for i in range(1, n_nodes):
        if rng.random() < 0.15:
            continue  # a root candidate
        p = int(rng.integers(max(0, i - 60), i))
        if depth[p] >= depth_limit:
            continue
        nodes[p].setdefault("children", []).append(i)
        depth[i] = depth[p] + 1
        if rng.random() < shared and p > 0:
            q = int(rng.integers(0, p))
            if depth[q] < depth_limit and i not in nodes[q].get("children", []):
                nodes[q].setdefault("children", []).append(i)
    for n in nodes:
        r = rng.random()
        if r < 0.3:
            n["matrix"] = [float(x) for x in rng.uniform(-2, 2, 16)]
        else:
            if rng.random() < 0.7:
                n["translation"] = [float(x) for x in rng.uniform(-50, 50, 3)]
            if rng.random() < 0.7:
                q = rng.normal(size=4)
                n["rotation"] = [float(x) for x in q / np.linalg.norm(q)]
            if rng.random() < 0.5:
                n["scale"] = [float(x) for x in rng.uniform(0.2, 3, 3)]
        if rng.random() < 0.6:
            n["mesh"] = int(rng.integers(0, n_meshes))
    children = {c for n in nodes for c in n.get("children", [])}
    roots = [i for i in range(n_nodes) if i not in children]
    return {"scene": 1, "scenes": [{"nodes": roots[:1]}, {"nodes": roots}], "nodes": nodes,
            "meshes": [{} for _ in range(n_meshes)]}


def oracle_node_matrix(node):
    if "matrix" in node:
        return np.array(node["matrix"], dtype=F)
    return oracle.gltf_trs_matrix(node.get("scale", [1, 1, 1]), node.get("rotation"), node.get("translation"))


def oracle_recursion(doc):
    """main.cpp:1518-1576 on the oracle: per mesh, world matrices in
    depth-first order."""
    nodes = doc["nodes"]
    lists = [[] for _ in doc.get("meshes", [])]

    def process(n, parent):
        m = oracle.mat4_mul(parent, oracle_node_matrix(nodes[n]))
        if "mesh" in nodes[n]:
            lists[nodes[n]["mesh"]].append(m)
        for c in nodes[n].get("children", []):
            process(c, m)

    for r in doc["scenes"][doc.get("scene", 0)].get("nodes", []):
        m = oracle.gltf_root_flip(oracle_node_matrix(nodes[r]))
        if "mesh" in nodes[r]:
            lists[nodes[r]["mesh"]].append(m)
        for c in nodes[r].get("children", []):
            process(c, m)
    return [np.array(x, dtype=F).reshape(len(x), 16) for x in lists]


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (needs the reference tree)")
def test_oracle_node_matrix_matches_reference_glm():
    rng = np.random.default_rng(5)
    for _ in range(5000):
        s = rng.uniform(-3, 3, 3).astype(F)
        q = rng.uniform(-1.5, 1.5, 4).astype(F)  # normalised or not, as the file gives it
        t = rng.uniform(-1e4, 1e4, 3).astype(F)
        for qq, tt in ((q, t), (None, t), (q, None), (None, None)):
            a = oracle.gltf_trs_matrix(s, qq, tt)
            b = oracle.gltf_trs_matrix(s, qq, tt, ref=True)
            assert a.tobytes() == b.tobytes()
        M = rng.uniform(-5, 5, 16).astype(F)
        assert oracle.gltf_root_flip(M).tobytes() == oracle.gltf_root_flip(M, ref=True).tobytes()


def test_oracle_node_matrix_closed_forms():
    # no rotation: diag(scale, 1) with the translation in column 3
    M = oracle.gltf_trs_matrix([2, 3, 4], None, [5, 6, 7]).reshape(4, 4)
    assert np.array_equal(M, np.array([[2, 0, 0, 0], [0, 3, 0, 0], [0, 0, 4, 0], [5, 6, 7, 1]], F))
    # half turns about z and x are exact: diag(-1,-1,1) and diag(1,-1,-1), columns scaled
    M = oracle.gltf_trs_matrix([2, 3, 4], [0, 0, 1, 0], None).reshape(4, 4)
    assert np.array_equal(M, np.diag(np.array([-2, -3, 4, 1], F)))
    M = oracle.gltf_trs_matrix([1, 1, 1], [1, 0, 0, 0], None).reshape(4, 4)
    assert np.array_equal(M, np.diag(np.array([1, -1, -1, 1], F)))
    # quarter turn about y from a non-normalised quaternion (0, 2, 0, 2): glm
    # does not normalise, so the matrix is 1 - 2*4 = -7 on x and z
    M = oracle.gltf_trs_matrix([1, 1, 1], [0, 2, 0, 2], None).reshape(4, 4)
    assert M[0, 0] == -7 and M[2, 2] == -7 and M[0, 2] == -8 and M[2, 0] == 8 and M[1, 1] == 1
    # root flip negates rows 1 and 2 of every column
    A = np.arange(16, dtype=F)
    B = oracle.gltf_root_flip(A).reshape(4, 4)
    assert np.array_equal(B[:, 0], A.reshape(4, 4)[:, 0]) and np.array_equal(B[:, 1], -A.reshape(4, 4)[:, 1])


def test_node_matrix_matches_oracle():
    doc = random_document(1, n_nodes=2000)
    for n in doc["nodes"]:
        assert gltf.node_matrix(n).tobytes() == oracle_node_matrix(n).tobytes()
    M = np.arange(16, dtype=F) - 3
    assert gltf.root_flip(M).tobytes() == oracle.gltf_root_flip(M).tobytes()


def test_levels_follow_the_depth_first_walk():
    doc = random_document(2)
    h = gltf.load(doc)
    nodes = doc["nodes"]
    # the depth-first walk of the default scene, one instance per path
    walk = []

    def visit(n, d):
        walk.append((n, d))
        for c in nodes[n].get("children", []):
            visit(c, d + 1)

    for r in doc["scenes"][1]["nodes"]:
        visit(r, 0)
    assert [(int(h.node[d][i]), d) for d, i in h.order] == walk
    assert h.num_instances == len(walk) > len(nodes)  # shared nodes appear once per path
    for d in range(1, len(h.levels)):
        parent = h.levels[d][1]
        assert parent.max() < h.levels[d - 1][0].shape[0]
        for i, p in enumerate(parent):
            assert int(h.node[d][i]) in nodes[int(h.node[d - 1][p])]["children"]
    # level 0 holds the flipped root matrices, deeper levels the node matrices
    for i, n in enumerate(h.node[0]):
        assert h.levels[0][0][i].tobytes() == oracle.gltf_root_flip(oracle_node_matrix(nodes[n])).tobytes()
    for i, n in enumerate(h.node[1]):
        assert h.levels[1][0][i].tobytes() == oracle_node_matrix(nodes[n]).tobytes()


def test_read_gltf_and_glb(tmp_path):
    doc = random_document(3, n_nodes=50)
    p = tmp_path / "scene.gltf"
    p.write_text(json.dumps(doc))
    js = json.dumps(doc).encode()
    js += b" " * (-len(js) % 4)
    glb = struct.pack("<III", 0x46546C67, 2, 12 + 8 + len(js)) + struct.pack("<II", len(js), 0x4E4F534A) + js
    q = tmp_path / "scene.glb"
    q.write_bytes(glb)
    a, b, c = gltf.load(str(p)), gltf.load(str(q)), gltf.load(doc)
    assert a.order == b.order == c.order
    for x, y in zip(a.levels, b.levels):
        assert x[0].tobytes() == y[0].tobytes()


def test_malformed_nodes_raise():
    with pytest.raises(gltf.GltfError):
        gltf.node_matrix({"matrix": [0.0] * 16, "scale": [1, 1, 1]})
    with pytest.raises(gltf.GltfError):
        gltf.node_matrix({"matrix": [0.0] * 15})
    with pytest.raises(gltf.GltfError):
        gltf.node_matrix({"rotation": [0, 0, 1]})
    with pytest.raises(gltf.GltfError):
        gltf.node_matrix({"translation": [0, 0]})
    cyc = {"scenes": [{"nodes": [0]}], "nodes": [{"children": [1]}, {"children": [0]}]}
    with pytest.raises(gltf.GltfError):
        gltf.load(cyc)
    with pytest.raises(gltf.GltfError):
        gltf.load({"scenes": [{"nodes": [3]}], "nodes": [{}]})
    with pytest.raises(gltf.GltfError):
        gltf.load({"scene": 2, "scenes": [{"nodes": [0]}], "nodes": [{}]})
    with pytest.raises(gltf.GltfError):
        gltf.load({"scenes": [{"nodes": [0]}], "nodes": [{"mesh": 1}], "meshes": [{}]})
    assert gltf.load({"nodes": [{}]}).num_instances == 0  # no scene: nothing to draw


@pytest.mark.gpu
def test_gltf_propagation_matches_reference_recursion(gpu):
    """World matrices from cadrPropagateTransforms level by level equal the
    reference's depth-first recursion (parentMatrix * node.matrix) bit for bit,
    and so do the per-mesh instance lists."""
    doc = random_document(4, n_nodes=3000)
    h = gltf.load(doc)
    print(f"gltf: {h.num_instances} instances in {len(h.levels)} levels", flush=True)
    worlds = gltf.propagate(h)
    got = h.mesh_instances(worlds)
    exp = oracle_recursion(doc)
    assert len(got) == len(exp)
    for m, (a, b) in enumerate(zip(got, exp)):
        assert a.shape == b.shape, f"mesh {m}"
        assert a.tobytes() == b.tobytes(), f"mesh {m}: first difference at {np.argwhere(a != b)[:3]}"


# ------------------------------------------------------------- mesh data

def _random_bounds(rng, t):
    n = int(rng.integers(1, 50))
    scale = 10.0 ** rng.uniform(-3, 4)
    P = (rng.normal(size=(n, 3)) * scale + rng.normal(size=3) * scale * 3).astype(F)
    mn, mx = P.min(0), P.max(0)
    if t % 7 == 0:  # a box that does not contain every position
        mn = (mn + F(scale * 0.3)).astype(F)
    k = int(rng.integers(1, 6))
    boxes = np.sort(rng.normal(size=(k, 2, 3)) * scale, axis=1).reshape(k, 6).astype(F)
    sph = np.concatenate([rng.normal(size=(k, 3)) * scale, np.abs(rng.normal(size=(k, 1))) * scale], 1).astype(F)
    return P, mn, mx, boxes, sph


def _sphere4(c, r):
    return np.concatenate([np.asarray(c, F), [F(r)]]).astype(F)


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (needs the reference tree)")
def test_bounds_match_reference_bounding_sphere_code():
    """Primitive and mesh spheres: the module and the oracle bit for bit as
    the reference's BoundingBox / BoundingSphere code computes them; the
    cases tell the double-precision sqrt of extendRadiusBy from a float one."""
    rng = np.random.default_rng(11)
    float_sqrt_differs = 0
    for t in range(3000):
        P, mn, mx, boxes, sph = _random_bounds(rng, t)
        ref = oracle.gltf_primitive_sphere(P, mn, mx, ref=True)
        assert oracle.gltf_primitive_sphere(P, mn, mx).tobytes() == ref.tobytes()
        assert _sphere4(*gltf.primitive_sphere(P, mn, mx)).tobytes() == ref.tobytes()
        ref = oracle.gltf_mesh_sphere(boxes, sph, ref=True)
        assert oracle.gltf_mesh_sphere(boxes, sph).tobytes() == ref.tobytes()
        got = gltf.mesh_sphere([(b[:3], b[3:]) for b in boxes], [(s[:3], s[3]) for s in sph])
        assert _sphere4(*got).tobytes() == ref.tobytes()
        # the same loop with a float sqrt and a float sum
        c = got[0]
        r = F(0)
        for s in sph:
            d = (s[:3] - c).astype(F)
            d2 = F(F(F(d[0] * d[0]) + F(d[1] * d[1])) + F(d[2] * d[2]))
            dif = F(r - s[3])
            if dif < 0 or d2 > F(dif * dif):
                r = F(F(np.sqrt(d2)) + s[3])
        float_sqrt_differs += r != ref[3]
    assert float_sqrt_differs > 0


def test_bounds_closed_forms():
    # positions at distance 5 from the box centre (3-4-5 triangles)
    P = np.array([[3, 4, 0], [-3, -4, 0], [0, 0, 5], [0, 0, -5]], F) + F(10)
    c, r = gltf.primitive_sphere(P, P.min(0), P.max(0))
    assert np.array_equal(c, np.full(3, 10, F)) and r == 5
    assert _sphere4(c, r).tobytes() == oracle.gltf_primitive_sphere(P, P.min(0), P.max(0)).tobytes()
    # NaN positions are skipped (d2 > r2 is false for NaN)
    Q = np.concatenate([P, np.array([[np.nan, 0, 0]], F)])
    assert gltf.primitive_sphere(Q, P.min(0), P.max(0))[1] == 5
    # two unit spheres at x = -10 and 10: union box centred at 0, radius 10 + 1
    boxes = [(np.array([-11, -1, -1], F), np.array([-9, 1, 1], F)), (np.array([9, -1, -1], F), np.array([11, 1, 1], F))]
    spheres = [(np.array([-10, 0, 0], F), F(1)), (np.array([10, 0, 0], F), F(1))]
    c, r = gltf.mesh_sphere(boxes, spheres)
    assert np.array_equal(c, np.zeros(3, F)) and r == 11
    # a sphere inside the current one leaves the radius alone
    c, r = gltf.mesh_sphere(boxes, spheres + [(np.array([0, 0, 0], F), F(2))])
    assert r == 11


def test_convert_indices_follow_the_gltf_definitions():
    """Triangle / line lists from the glTF 2.0 definitions of each mode
    (strips: triangle i is (i, i+1, i+2), the first two swapped for odd i;
    fans: (i+1, i+2, 0), which the reference rotates to (0, i+1, i+2);
    line strips and loops: consecutive pairs, the loop closed)."""
    rng = np.random.default_rng(2)
    for n in (1, 2, 3, 4, 5, 8, 33):
        s = rng.integers(0, 1000, n).astype(np.uint32)
        for src in (s, None):
            v = np.arange(n, dtype=np.uint32) if src is None else s
            for mode in (0, 1, 4):
                need = {0: 1, 1: 2, 4: 3}[mode]
                if n < need:
                    with pytest.raises(gltf.GltfError):
                        gltf.convert_indices(mode, src, n)
                else:
                    assert np.array_equal(gltf.convert_indices(mode, src, n), v)
            if n >= 3:
                strip = [x for i in range(n - 2) for x in ((v[i], v[i + 1], v[i + 2]) if i % 2 == 0
                                                          else (v[i + 1], v[i], v[i + 2]))]
                assert gltf.convert_indices(5, src, n).tolist() == [int(x) for x in strip]
                fan = [x for i in range(n - 2) for x in (v[0], v[i + 1], v[i + 2])]
                assert gltf.convert_indices(6, src, n).tolist() == [int(x) for x in fan]
            else:
                for mode in (5, 6):
                    with pytest.raises(gltf.GltfError):
                        gltf.convert_indices(mode, src, n)
            if n >= 2:
                line = [x for i in range(n - 1) for x in (v[i], v[i + 1])]
                assert gltf.convert_indices(3, src, n).tolist() == [int(x) for x in line]
                assert gltf.convert_indices(2, src, n).tolist() == [int(x) for x in line] + [int(v[-1]), int(v[0])]
    with pytest.raises(gltf.GltfError):
        gltf.convert_indices(7, None, 9)


def _mesh_document(tmp_path, rng, n_meshes=5, prims=(1, 3)):
    """A document with random primitives: interleaved POSITION / NORMAL
    (stride 24), indices as u8 / u16 / u32 or none, every mode; buffer 0 an
    external file with a percent-encoded name, buffer 1 a base64 data URI."""
    import base64
    b0, b1 = bytearray(), bytearray()
    views, accessors, meshes, expect = [], [], [], []
    for m in range(n_meshes):
        plist = []
        for p in range(int(rng.integers(prims[0], prims[1] + 1))):
            nv = int(rng.integers(3, 40))
            P = (rng.normal(size=(nv, 3)) * 10 + rng.normal(size=3) * 30).astype(F)
            Nrm = rng.normal(size=(nv, 3)).astype(F)
            inter = np.concatenate([P, Nrm], axis=1).astype(F)
            while len(b0) % 4:
                b0.append(0)
            views.append({"buffer": 0, "byteOffset": len(b0), "byteLength": inter.nbytes, "byteStride": 24})
            b0 += inter.tobytes()
            pa = len(accessors)
            accessors.append({"bufferView": len(views) - 1, "componentType": 5126, "count": nv, "type": "VEC3",
                              "min": [float(x) for x in P.min(0)], "max": [float(x) for x in P.max(0)]})
            accessors.append({"bufferView": len(views) - 1, "byteOffset": 12, "componentType": 5126, "count": nv,
                              "type": "VEC3"})
            prim = {"attributes": {"POSITION": pa, "NORMAL": pa + 1}, "mode": int(rng.choice([0, 1, 2, 3, 4, 5, 6]))}
            kind = int(rng.integers(0, 4))
            src = None
            if kind < 3:
                ni = int(rng.integers(3, 60))
                dt = (np.uint8, np.uint16, np.uint32)[kind]
                src = rng.integers(0, nv, ni).astype(dt)
                while len(b1) % 4:
                    b1.append(0)
                views.append({"buffer": 1, "byteOffset": len(b1), "byteLength": src.nbytes})
                b1 += src.tobytes()
                accessors.append({"bufferView": len(views) - 1, "componentType": (5121, 5123, 5125)[kind],
                                  "count": ni, "type": "SCALAR"})
                prim["indices"] = len(accessors) - 1
            if rng.random() < 0.5:
                prim["material"] = int(rng.integers(0, 3))
            plist.append(prim)
            expect.append((m, P, src, prim["mode"], prim.get("material", -1)))
        meshes.append({"primitives": plist})
    name = "mesh data %d.bin" % rng.integers(0, 100)
    (tmp_path / name).write_bytes(bytes(b0))
    buffers = [{"byteLength": len(b0), "uri": name.replace(" ", "%20")},
               {"byteLength": len(b1),
                "uri": "data:application/octet-stream;base64," + base64.b64encode(bytes(b1)).decode()}]
    return {"bufferViews": views, "accessors": accessors, "meshes": meshes, "buffers": buffers}, expect


def test_mesh_data_decodes_accessors(tmp_path):
    """Positions, indices (converted per mode) and bounds of every primitive
    of a written .gltf with external and data-URI buffers; the shared index
    and position buffers and the mesh table; mesh spheres against the oracle;
    drawables per (instance, primitive) grouped by material."""
    doc, expect = _mesh_document(tmp_path, np.random.default_rng(6), n_meshes=12)
    doc.update({k: v for k, v in random_document(7, n_nodes=200, n_meshes=12).items() if k != "meshes"})
    p = tmp_path / "scene.gltf"
    p.write_text(json.dumps(doc))
    d, bin_chunk = gltf.read_gltf(str(p))
    buffers = gltf.load_buffers(d, str(tmp_path), bin_chunk)
    h = gltf.load(d)
    md = gltf.load_meshes(d, buffers, h)
    assert len(md.primitives) == len(expect)
    for prim, (m, P, src, mode, mat) in zip(md.primitives, expect):
        assert prim.mesh == m and prim.mode == mode and prim.material == mat
        assert prim.positions.tobytes() == P.tobytes()
        assert prim.indices.tobytes() == gltf.convert_indices(mode, src, P.shape[0]).tobytes()
        assert _sphere4(*prim.sphere).tobytes() == oracle.gltf_primitive_sphere(P, P.min(0), P.max(0)).tobytes()
    ib, pb, first, vofs = md.buffers()
    table = md.mesh_table()
    for i, prim in enumerate(md.primitives):
        f, c = int(table["firstIndex"][i, 0]), int(table["indexCount"][i, 0])
        assert ib[f:f + c].tobytes() == prim.indices.tobytes()
        vo = int(table["vertexOffset"][i])
        assert pb[vo:vo + prim.positions.shape[0], :3].tobytes() == prim.positions.tobytes()
        assert np.all(pb[:, 3] == 1)
    for mi, ids in enumerate(md.mesh_primitives):
        if md.mesh_spheres[mi] is None:
            continue
        boxes = np.array([np.concatenate(md.primitives[i].box) for i in ids], F)
        sph = np.array([_sphere4(*md.primitives[i].sphere) for i in ids], F)
        assert _sphere4(*md.mesh_spheres[mi]).tobytes() == oracle.gltf_mesh_sphere(boxes, sph).tobytes()
    # drawables: (instance, primitive) pairs grouped by material, creation order inside
    dr = gltf.drawables(h, md)
    pairs = []
    per_mesh = [[] for _ in range(h.num_meshes)]
    for dd, ii in h.order:
        if int(h.mesh[dd][ii]) >= 0:
            per_mesh[int(h.mesh[dd][ii])].append((dd, ii))
    for mi, ids in enumerate(md.mesh_primitives):
        for pid in ids:
            for di in per_mesh[mi]:
                pairs.append((md.primitives[pid].material, di, pid))
    pairs.sort(key=lambda x: x[0])
    assert [tuple(x) for x in dr.instance.tolist()] == [x[1] for x in pairs]
    assert dr.mesh_index.tolist() == [x[2] for x in pairs]
    assert dr.stateset_counts.sum() == dr.count and dr.materials == sorted(set(x[0] for x in pairs))


def test_mesh_data_from_glb_bin_chunk(tmp_path):
    pos = np.array([[0, 0, 0], [2, 0, 0], [0, 2, 0]], F)
    blob = pos.tobytes()
    doc = {"scenes": [{"nodes": [0]}], "nodes": [{"mesh": 0}],
           "meshes": [{"primitives": [{"attributes": {"POSITION": 0}}]}],
           "buffers": [{"byteLength": len(blob)}], "bufferViews": [{"buffer": 0, "byteLength": len(blob)}],
           "accessors": [{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3",
                          "min": [0, 0, 0], "max": [2, 2, 0]}]}
    js = json.dumps(doc).encode()
    js += b" " * (-len(js) % 4)
    glb = (struct.pack("<III", 0x46546C67, 2, 12 + 8 + len(js) + 8 + len(blob)) + struct.pack("<II", len(js), 0x4E4F534A)
           + js + struct.pack("<II", len(blob), 0x004E4942) + blob)
    p = tmp_path / "m.glb"
    p.write_bytes(glb)
    d, bin_chunk = gltf.read_gltf(str(p))
    md = gltf.load_meshes(d, gltf.load_buffers(d, str(tmp_path), bin_chunk), gltf.load(d))
    assert md.primitives[0].positions.tobytes() == pos.tobytes()
    assert md.primitives[0].indices.tolist() == [0, 1, 2]
    assert md.primitives[0].sphere[1] == np.sqrt(F(2))
    with pytest.raises(gltf.GltfError):  # no BIN chunk for a buffer without uri
        gltf.load_buffers(d, str(tmp_path), None)


def test_mesh_data_errors():
    pos = np.zeros((4, 3), F)
    blob = pos.tobytes()

    def doc(**acc):
        a = {"bufferView": 0, "componentType": 5126, "count": 4, "type": "VEC3", "min": [0, 0, 0], "max": [0, 0, 0]}
        a.update(acc)
        a = {k: v for k, v in a.items() if v is not None}
        return {"buffers": [{"byteLength": len(blob)}], "bufferViews": [{"buffer": 0, "byteLength": len(blob)}],
                "accessors": [a], "meshes": [{"primitives": [{"attributes": {"POSITION": 0}}]}]}

    def load(d, attrs=None):
        if attrs is not None:
            d["meshes"][0]["primitives"][0]["attributes"] = attrs
        return gltf.load_meshes(d, [blob])

    load(doc())
    for bad in (doc(count=5), doc(byteOffset=4), doc(type="VEC4"), doc(componentType=5123), doc(normalized=True),
                doc(min=None), doc(max=[0, 0]), doc(bufferView=None), doc(sparse={"count": 1})):
        with pytest.raises(gltf.GltfError):
            load(bad)
    with pytest.raises(gltf.GltfError):
        load(doc(), {"POSITION": 0, "JOINTS_0": 0})
    d = doc()
    d["accessors"].append({"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3"})
    with pytest.raises(gltf.GltfError):  # NORMAL count differs from POSITION's
        load(d, {"POSITION": 0, "NORMAL": 1})
    d = doc()
    d["meshes"][0]["primitives"][0]["mode"] = 4
    d["meshes"][0]["primitives"][0]["indices"] = 0  # a VEC3 accessor as indices
    with pytest.raises(gltf.GltfError):
        load(d)
    d = doc()
    d["buffers"][0]["byteLength"] = 8
    with pytest.raises(gltf.GltfError):
        load(d)


@pytest.mark.gpu
def test_gltf_scene_cull_matches_oracle(gpu, tmp_path):
    """A glTF document end to end: hierarchy propagated on the GPU, mesh data
    decoded into the mesh table, one drawable per (instance, primitive),
    world spheres and the batched cull of 8 views, every stage bit for bit
    against the oracle (world matrices level by level, spheres, commands and
    counts)."""
    from cadr_amd import _native as N
    from cadr_amd import cull, scenes
    from cadr_amd.device import synchronize
    doc, _ = _mesh_document(tmp_path, np.random.default_rng(9), n_meshes=10, prims=(1, 4))
    doc.update({k: v for k, v in random_document(10, n_nodes=1500, n_meshes=10).items() if k != "meshes"})
    h = gltf.load(doc)
    md = gltf.load_meshes(doc, gltf.load_buffers(doc, str(tmp_path)), h)
    dr = gltf.drawables(h, md)
    print(f"gltf scene: {h.num_instances} node instances, {len(md.primitives)} primitives, {dr.count} drawables, "
          f"{len(dr.materials)} StateSets", flush=True)
    worlds = gltf.propagate(h)
    worlds_o, prev = [], None
    for L, P in h.levels:
        w, _ = oracle.propagate_level(L, P, prev)
        worlds_o.append(w)
        prev = w
    W = dr.world_matrices(worlds)
    assert W.tobytes() == dr.world_matrices(worlds_o).tobytes()
    cams = scenes.random_cameras(scenes.SEED_BASE + 40, 20, 0, 8, box=300.0)
    sc = gltf.cull_scene(h, md, dr, worlds, cams)
    _, sph = oracle.propagate_scene(sc)
    views = cull.make_views(cams)
    exp_cmds, exp_counts = oracle.cull(sph, sc.mesh_index, sc.meshes, sc.stateset_counts, views, cmd_dtype=N.CMD_DTYPE)
    assert exp_counts.sum() > 50
    ds = cull.DeviceScene(sc)
    ds.propagate()
    synchronize()
    assert ds.download_spheres().tobytes() == sph.tobytes()
    out = cull.CullOutputs(views.shape[0], ds.n, ds.K)
    b = cull.CullBatch(cull.make_config(256, 4), [(ds, views, out)])
    b.run()
    synchronize()
    cmds, counts = out.download()
    assert np.array_equal(counts, exp_counts)
    got = cull.valid_commands(cmds, counts, sc.stateset_counts)
    exp = cull.valid_commands(exp_cmds, exp_counts, sc.stateset_counts)
    for v in range(len(got)):
        for k in range(len(sc.stateset_counts)):
            assert got[v][k].tobytes() == exp[v][k].tobytes(), f"view {v} StateSet {k}"
    b.close(); out.close(); ds.close()
