"""Node transforms on the GPU (cd_node_global_matrices) against the oracle's
GetGlobalMatrix (NodeDataCompEntity.cpp:63-72), bit for bit, and the shipped
Table_room scene end to end: glTF meshes and nodes in, node globals on the
GPU, one collision frame over every node with a collision component
(GameImporter.cpp:320-326), against the oracle's frame on its own decode.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, case
from inmyroom_vulkan_amd import ENTITY_DTYPE, MODE_BATCHED, MODE_PLAIN, NODE_DTYPE, CollideError, GltfDocument
from oracle import gltf as og
from test_nodes import _random_forest

pytestmark = pytest.mark.gpu

FIXTURES = [os.path.join(GOLDEN, "gltf", p) for p in ("apple/scene.gltf", "table_room.gltf", "environment.gltf")]


def _records(trs, par):
    nd = np.zeros(len(trs), NODE_DTYPE)
    nd["local_scale"], nd["local_rotation"] = trs[:, 0:3], trs[:, 3:7]
    nd["local_translation"], nd["global_translation"] = trs[:, 7:10], trs[:, 10:13]
    nd["parent"] = par
    return nd


@pytest.mark.parametrize("path", FIXTURES)
def test_node_globals_scenes(gpu, orc, path):
    case(f"node globals, {os.path.basename(os.path.dirname(path))}/{os.path.basename(path)}")
    with GltfDocument(path) as doc:
        nodes, _, _ = doc.nodes()
    trs, par, _, _ = og.nodes(path)
    assert gpu.node_globals(nodes).tobytes() == orc.node_globals(trs, par).tobytes()


def test_node_globals_random_forest(gpu, orc):
    case("node globals, 20,000 random nodes, depth up to 12, children before parents")
    trs, par = _random_forest(np.random.default_rng(11), 20000, depth_max=12)
    got = gpu.node_globals(_records(trs, par))
    assert got.tobytes() == orc.node_globals(trs, par).tobytes()


def test_node_globals_rejects(gpu):
    trs, par = _random_forest(np.random.default_rng(12), 10)
    case("node globals, bad hierarchies")
    bad = _records(trs, par)
    bad["parent"][3] = 50
    with pytest.raises(CollideError, match="out of range"):
        gpu.node_globals(bad)
    cyc = _records(trs, par)
    cyc["parent"][0], cyc["parent"][1] = 1, 0
    with pytest.raises(CollideError, match="cyclic"):
        gpu.node_globals(cyc)
    nan = _records(trs, par)
    nan["local_translation"][2, 1] = np.nan
    with pytest.raises(CollideError, match="non-finite"):
        gpu.node_globals(nan)
    assert len(gpu.node_globals(np.zeros(0, NODE_DTYPE))) == 0


def test_table_room_frame_end_to_end(gpu, orc):
    path = FIXTURES[1]
    case("Table_room: glTF meshes and nodes -> node globals -> collision frame")
    with GltfDocument(path) as doc:
        mesh_ids = gpu.register_gltf(doc)
        nodes, mesh, collides = doc.nodes()
    G = gpu.node_globals(nodes)
    ents_idx = np.nonzero(collides)[0]
    ents = np.zeros(len(ents_idx), ENTITY_DTYPE)
    ents["current"] = ents["previous"] = G[ents_idx]
    ents["mesh_id"] = mesh_ids[mesh[ents_idx]]
    ents["entity"] = np.arange(len(ents_idx))
    ents["should_callback"] = 1
    # the oracle, on its own decode of the same file
    trs, par, omesh, ocoll = og.nodes(path)
    assert np.array_equal(ocoll, collides) and np.array_equal(omesh, mesh)
    OG = orc.node_globals(trs, par)
    assert OG.tobytes() == G.tobytes()
    trees, local = [], {}
    for m, prims in enumerate(og.load(path)):
        soup = og.mesh_triangles(prims)[0]
        if len(soup):
            local[m] = len(trees)
            trees.append(orc.tree(soup, 4))
    mesh_of = np.array([local[int(m)] for m in omesh[ents_idx]], np.uint32)
    for mode in (MODE_PLAIN, MODE_BATCHED):
        pairs, cons = gpu.frame(ents, mode)
        res = orc.frame(trees, mesh_of, OG[ents_idx], np.ones(len(ents_idx), np.uint8), mode=mode)
        want = {}
        for c in res.contacts:
            a, b = res.pairs[c["query"]]
            want[(int(a), int(b), int(c["tri_a"]), int(c["tri_b"]))] = c
        got = {}
        for p in pairs:
            for c in cons[p["contact_begin"]: p["contact_begin"] + p["contact_count"]]:
                got[(int(p["first"]), int(p["second"]), int(c["tri_a"]), int(c["tri_b"]))] = c
        assert set(got) == set(want), mode
        if mode == MODE_PLAIN:
            for k, c in want.items():
                g = got[k]
                assert g["s"].tobytes() == c["s"].tobytes() and g["t"].tobytes() == c["t"].tobytes()
        assert len(want) > 1000
