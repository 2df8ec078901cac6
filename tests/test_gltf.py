"""glTF hierarchy ingestion (examples/gltfReader/main.cpp:1422-1576): node
matrices, the root axis flip, the depth-first walk flattened into levels, and
GPU propagation against the reference's recursion restated on the oracle."""
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
