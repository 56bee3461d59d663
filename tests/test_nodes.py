"""Node transforms on the CPU: the oracle's GetGlobalMatrix
(NodeDataCompEntity.cpp:63-72) and its glTF node import
(GameImporter.cpp:503-552), pinned independently, and the library's node
import (host code behind cd_gltf_nodes) against the oracle.

Pins:
  - the quaternion-to-matrix convention against scipy's Rotation (scalar
    last), over random unit quaternions;
  - whole hierarchies against float64 compositions gT @ parent @ T @ R @ S
    built with scipy's rotations, for random trees of depth up to 6;
  - closed forms: a child's translation turns with its parent, a global
    translation does not;
  - the import: each node's local matrix T R S equals diag(1,-1,-1,1) M
    diag(1,-1,-1,1) for M the node's glTF matrix (or its T R S composed by the
    glTF 2.0 rule in float64), over every node of the shipped scenes and of
    synthetic documents that take decompose's other branches (trace <= 0,
    a mirrored basis).
glm's decompose and quaternion normalize come from glm, an empty submodule of
the reference; their exact float operation order is therefore pinned only by
the two restatements here agreeing bit for bit (oracle/gltf.py and the
library's gltf.hpp), the round trips and invariants above.
"""
import json
import os

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from conftest import GOLDEN
from oracle import gltf as og

FIXTURES = [os.path.join(GOLDEN, "gltf", p) for p in ("apple/scene.gltf", "table_room.gltf", "environment.gltf")]
F = np.diag([1.0, -1.0, -1.0, 1.0])


def _mat(g16):
    return np.asarray(g16, np.float64).reshape(4, 4).T  # column-major -> row-major


def _trs13(scale, quat_xyzw, t, gt=(0, 0, 0)):
    return np.concatenate([scale, quat_xyzw, t, gt]).astype(np.float32)


def test_quaternion_convention_matches_scipy(orc):
    rng = np.random.default_rng(1)
    q = Rotation.random(500, random_state=2).as_quat().astype(np.float32)  # (x, y, z, w)
    trs = np.stack([_trs13([1, 1, 1], qi, [0, 0, 0]) for qi in q])
    G = orc.node_globals(trs, np.full(len(q), 0xFFFFFFFF, np.uint32))
    for g, qi in zip(G, q):
        np.testing.assert_allclose(_mat(g)[:3, :3], Rotation.from_quat(qi.astype(np.float64)).as_matrix(), atol=2e-6)
    assert rng is not None


def _random_forest(rng, n, depth_max=6):
    parent = np.full(n, 0xFFFFFFFF, np.uint32)
    depth = np.zeros(n, int)
    for i in range(1, n):
        if rng.random() < 0.85:
            p = int(rng.integers(0, i))
            if depth[p] < depth_max:
                parent[i] = p
                depth[i] = depth[p] + 1
    perm = rng.permutation(n)  # children may come before their parents
    inv = np.empty(n, int)
    inv[perm] = np.arange(n)
    par = np.array([0xFFFFFFFF if parent[perm[k]] == 0xFFFFFFFF else inv[parent[perm[k]]] for k in range(n)], np.uint32)
    scale = rng.uniform(0.5, 2.0, (n, 3)) * rng.choice([-1, 1], (n, 3), p=[0.1, 0.9])
    q = Rotation.random(n, random_state=int(rng.integers(1 << 30))).as_quat()
    t = rng.uniform(-3, 3, (n, 3))
    gt = np.where(rng.random((n, 1)) < 0.3, rng.uniform(-2, 2, (n, 3)), 0.0)
    trs = np.concatenate([scale, q, t, gt], axis=1).astype(np.float32)
    return trs, par


def _float64_globals(trs, par):
    n = len(trs)
    out = [None] * n

    def g(i):
        if out[i] is not None:
            return out[i]
        s, q, t, gt = trs[i, 0:3], trs[i, 3:7], trs[i, 7:10], trs[i, 10:13]
        P = np.eye(4) if par[i] == 0xFFFFFFFF else g(int(par[i]))
        T, GT, R, S = np.eye(4), np.eye(4), np.eye(4), np.diag(np.append(s.astype(np.float64), 1.0))
        T[:3, 3], GT[:3, 3] = t, gt
        R[:3, :3] = Rotation.from_quat(q.astype(np.float64)).as_matrix()
        out[i] = GT @ P @ T @ R @ S
        return out[i]

    return np.stack([g(i) for i in range(n)])


def test_hierarchy_matches_float64(orc):
    rng = np.random.default_rng(3)
    trs, par = _random_forest(rng, 400)
    G = orc.node_globals(trs, par)
    want = _float64_globals(trs, par)
    for g, w in zip(G, want):
        scale = max(1.0, np.abs(w).max())
        np.testing.assert_allclose(_mat(g), w, atol=5e-5 * scale)


def test_translation_closed_forms(orc):
    q90z = [0, 0, np.sin(np.pi / 4), np.cos(np.pi / 4)]
    trs = np.stack([_trs13([1, 1, 1], q90z, [10, 0, 0]),                   # root: turned 90 degrees about z
                    _trs13([1, 1, 1], [0, 0, 0, 1], [1, 0, 0], [5, 0, 0])])  # child: 1 along its parent's x, then +5 x
    G = orc.node_globals(trs, np.array([0xFFFFFFFF, 0], np.uint32))
    np.testing.assert_allclose(_mat(G[0])[:3, 3], [10, 0, 0])
    np.testing.assert_allclose(_mat(G[1])[:3, 3], [15, 1, 0], atol=1e-6)  # parent x is world y; gT is not turned
    with pytest.raises(ValueError):
        orc.node_globals(trs, np.array([1, 0], np.uint32))


def _spec_local(nd):
    """The glTF 2.0 local matrix of a node, in float64."""
    if "matrix" in nd:
        return np.asarray(nd["matrix"], np.float64).reshape(4, 4).T
    T, R, S = np.eye(4), np.eye(4), np.eye(4)
    T[:3, 3] = nd.get("translation", [0, 0, 0])
    R[:3, :3] = Rotation.from_quat(nd.get("rotation", [0, 0, 0, 1])).as_matrix()
    S[:3, :3] = np.diag(nd.get("scale", [1, 1, 1]))
    return T @ R @ S


def _check_import(orc, path):
    doc = json.load(open(path)) if path.endswith(".gltf") else None
    trs, par, _, _ = og.nodes(path)
    roots = np.full(len(trs), 0xFFFFFFFF, np.uint32)
    local = orc.node_globals(trs, roots)  # each node alone: its T R S
    for i, nd in enumerate(doc["nodes"]):
        want = F @ _spec_local(nd) @ F
        scale = max(1.0, np.abs(want).max())
        np.testing.assert_allclose(_mat(local[i]), want, atol=2e-5 * scale, err_msg=f"node {i}")
    return trs, par


@pytest.mark.parametrize("path", FIXTURES)
def test_import_matches_gltf_rule(orc, path):
    _check_import(orc, path)


def _synthetic_nodes_doc(tmp_path):
    rng = np.random.default_rng(9)
    nodes = []
    for k in range(40):
        R = Rotation.random(random_state=k).as_matrix()
        if k % 4 == 1:  # about 180 degrees: trace <= 0
            R = Rotation.from_rotvec(rng.normal(size=3) / 1e3 + [np.pi, 0, 0]).as_matrix()
        s = rng.uniform(0.2, 3.0, 3)
        if k % 4 == 3:
            s[rng.integers(3)] *= -1  # a mirrored basis (a matrix node)
        M = np.eye(4)
        M[:3, :3] = R @ np.diag(s)
        M[:3, 3] = rng.uniform(-5, 5, 3)
        if k % 2:
            nodes.append({"matrix": [float(v) for v in M.T.reshape(-1)]})
        else:
            q = Rotation.from_matrix(R).as_quat()
            nodes.append({"translation": [float(v) for v in M[:3, 3]], "rotation": [float(v) for v in q],
                          "scale": [float(v) for v in np.abs(s)]})
    nodes[0]["children"] = [1, 2]
    nodes[2]["children"] = [3]
    doc = {"asset": {"version": "2.0"}, "nodes": nodes}
    path = tmp_path / "nodes.gltf"
    path.write_text(json.dumps(doc))
    return str(path)


def test_import_synthetic_decompose_branches(orc, tmp_path):
    path = _synthetic_nodes_doc(tmp_path)
    trs, par = _check_import(orc, path)
    assert list(par[:4]) == [0xFFFFFFFF, 0, 0, 2]
    # mirrored matrices come back with one negative scale on x, y and z together
    # (glm flips all three and the rotation), still composing to the same matrix
    assert (trs[3::4, 0:3] < 0).all()


@pytest.mark.parametrize("path", FIXTURES + ["synthetic"])
def test_library_import_matches_oracle(tmp_path, path):
    from inmyroom_vulkan_amd import GltfDocument

    if path == "synthetic":
        path = _synthetic_nodes_doc(tmp_path)
    trs, par, mesh, coll = og.nodes(path)
    with GltfDocument(path) as doc:
        n, m, c = doc.nodes()
    got = np.concatenate([n["local_scale"], n["local_rotation"], n["local_translation"], n["global_translation"]], 1)
    assert got.tobytes() == trs.tobytes()
    assert np.array_equal(n["parent"], par) and np.array_equal(m, mesh) and np.array_equal(c, coll)
