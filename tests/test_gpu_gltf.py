"""glTF mesh ingest on the GPU (cd_register_gltf) against the oracle: the
shipped scenes (Apple, Table_room, Sponza environmentTest) and synthetic
documents over every draw mode.

Trees and tree-order triangles must be bit-identical to the oracle's tree
built from the oracle's decoded triangle list (MeshesOfNodes.cpp:29-64 ->
PrimitivesOfMeshes.cpp:840-863 -> Triangle.cpp:259-280 -> OBBtree).
Contacts between the real meshes: bit-identical in the plain mode, the
exact oracle's pair set in the batched mode.  Uncollide rays on a real mesh
with its vertex normals: the oracle's ray sets.
"""
import os

import numpy as np
import pytest

import gltfgen as gg
from conftest import GOLDEN, case
from inmyroom_vulkan_amd import CD_NO_MESH, MODE_BATCHED, MODE_PLAIN, QUERY_DTYPE, CollideError, Collider, GltfDocument
from inmyroom_vulkan_amd import scenes
from oracle import gltf as og

pytestmark = pytest.mark.gpu

FIXTURES = {
    "apple": os.path.join(GOLDEN, "gltf", "apple", "scene.gltf"),
    "table_room": os.path.join(GOLDEN, "gltf", "table_room.gltf"),
    "environment": os.path.join(GOLDEN, "gltf", "environment.gltf"),
}


def _check_trees(ctx, orc, path, ids):
    meshes = og.load(path)
    assert len(ids) == len(meshes)
    out = []
    for m, prims in enumerate(meshes):
        soup, nrm, raw, vid = og.mesh_triangles(prims)
        if len(soup) == 0:
            assert ids[m] == CD_NO_MESH
            out.append(None)
            continue
        want = orc.tree(soup, ctx.leaf_max).export()
        got = ctx.export_tree(int(ids[m]))
        assert got["root"].tobytes() == want["root"].tobytes(), m
        assert got["obbs"].tobytes() == want["obbs"].tobytes(), m
        assert np.array_equal(got["links"], want["links"]), m
        assert np.array_equal(got["perm"], want["perm"]), m
        assert ctx.export_triangles(int(ids[m])).tobytes() == soup[want["perm"]].tobytes(), m
        out.append((soup, nrm, raw, vid))
    return out


@pytest.mark.parametrize("name", list(FIXTURES))
def test_gltf_trees_match_oracle(gpu, orc, name):
    case(f"glTF ingest, {name}")
    with GltfDocument(FIXTURES[name]) as doc:
        ids = gpu.register_gltf(doc)
    got = _check_trees(gpu, orc, FIXTURES[name], ids)
    if name == "table_room":
        assert sum(g is None for g in got) == 4  # the skinned meshes


@pytest.mark.parametrize("container", ["gltf", "glb"])
def test_gltf_synthetic_modes(gpu, orc, tmp_path, container):
    case(f"glTF ingest, synthetic draw modes ({container})")
    path = gg.write(str(tmp_path / "doc"), gg.mixed_meshes(np.random.default_rng(7)), container)
    with GltfDocument(path) as doc:
        ids = gpu.register_gltf(doc)
    got = _check_trees(gpu, orc, path, ids)
    assert [g is None for g in got] == [False, False, False, False, False, True, True]


def test_gltf_mixed_normals_rejected(tmp_path):
    P = np.eye(3, dtype=np.float32)
    path = gg.write(str(tmp_path / "mixed"), [[gg.prim(P, [0, 1, 2], P), gg.prim(P, [0, 2, 1])]])
    case("glTF ingest, a mesh mixing primitives with and without NORMAL")
    with Collider(leaf_max=4) as c, GltfDocument(path) as doc:
        with pytest.raises(CollideError, match="NORMAL"):
            c.register_gltf(doc)
        assert c.register_soups([scenes.box_soup()]) == [0]  # the context is still usable


def _queries_near(rng, ids, soups, pairs, n_per_pair, scale=1.0):
    """Poses that put mesh a's centre near a random triangle of mesh b: the
    relative matrix is b's pose in a's frame."""
    qs = []
    for a, b in pairs:
        ca = soups[a].reshape(-1, 3).mean(axis=0)
        tb = soups[b].reshape(-1, 3, 3)
        for _ in range(n_per_pair):
            target = tb[rng.integers(len(tb))].mean(axis=0)
            q = scenes.random_quaternions(rng, 1)
            # place b so that its chosen triangle lands on a's centre, turned by q
            R = scenes.quat_to_mat3(q)[0]
            t = ca - R @ target + rng.normal(scale=0.05 * scale, size=3)
            qs.append((ids[a], ids[b], scenes.pose_matrices(t[None], q)[0]))
    out = np.zeros(len(qs), QUERY_DTYPE)
    for i, (a, b, m) in enumerate(qs):
        out[i] = (a, b, m)
    return out


def test_gltf_contacts_real_meshes(gpu, orc):
    case("glTF ingest, contacts between Apple, Table_room and environment meshes")
    soups, trees, ids = [], [], []
    for name in ("apple", "table_room", "environment"):
        with GltfDocument(FIXTURES[name]) as doc:
            gid = gpu.register_gltf(doc)
        for m, prims in enumerate(og.load(FIXTURES[name])):
            soup = og.mesh_triangles(prims)[0]
            if len(soup) == 0:
                continue
            soups.append(soup)
            trees.append(orc.tree(soup, 4))
            ids.append(int(gid[m]))
    ids = np.array(ids, np.uint32)
    rng = np.random.default_rng(21)
    # the apple against itself and against the larger meshes, and a few meshes among themselves
    order = np.argsort([-len(s) for s in soups])
    big = [int(i) for i in order[:8]]
    pairs = [(0, 0)] + [(0, b) for b in big] + [(big[i], big[i + 1]) for i in range(0, 6, 2)]
    q = _queries_near(rng, ids, soups, pairs, 6)
    local = {int(g): k for k, g in enumerate(ids)}
    qa = np.array([local[int(x)] for x in q["mesh_a"]], np.uint32)
    qb = np.array([local[int(x)] for x in q["mesh_b"]], np.uint32)
    got = gpu.query_batch(q, MODE_PLAIN)
    res = orc.batch(trees, qa, qb, q["rel"], mode=0)
    key = lambda c: sorted((int(x["query"]), int(x["tri_a"]), int(x["tri_b"])) for x in c)
    assert key(got) == key(res.contacts)
    gp = {(int(x["query"]), int(x["tri_a"]), int(x["tri_b"])): x for x in got}
    for c in res.contacts:
        g = gp[(int(c["query"]), int(c["tri_a"]), int(c["tri_b"]))]
        assert g["s"].tobytes() == c["s"].tobytes() and g["t"].tobytes() == c["t"].tobytes()
    assert len(got) > 100
    exact = orc.batch(trees, qa, qb, q["rel"], mode=1).contacts
    assert key(gpu.query_batch(q, MODE_BATCHED)) == key(exact)


def test_gltf_rays_with_vertex_normals(gpu, orc):
    """The apple's rays use its NORMALs and its index-list vertex ids."""
    case("glTF ingest, uncollide rays on the apple")
    with GltfDocument(FIXTURES["apple"]) as doc:
        mid = int(gpu.register_gltf(doc)[0])
    soup, nrm, raw, vid = og.mesh_triangles(og.load(FIXTURES["apple"])[0])
    assert np.array_equal(raw, vid)  # one primitive: the reference's ids unchanged
    t = orc.tree(soup, 4)
    t.set_attrs(nrm, raw)
    rng = np.random.default_rng(5)
    q = _queries_near(rng, np.array([mid], np.uint32), [soup], [(0, 0)] * 40, 1)
    gpu.query_batch(q, MODE_PLAIN)
    pairs, rays = gpu.uncollide_rays()
    z = np.zeros(len(q), np.uint32)
    ref = orc.rays_queries([t], z, z, q["rel"])
    want = {i for i in range(len(ref.pairs)) if ref.pairs[i, 2] + ref.pairs[i, 3] > 0}
    got = {int(p["query"]): p for p in pairs}
    assert set(got) == want and len(want) > 10
    rows = lambda r: sorted(map(tuple, np.concatenate([r["origin"], r["direction"]], axis=1).tolist()))
    for i in want:
        p = got[i]
        fa, fb = ref.split(i)
        assert (p["n_first"], p["n_second"]) == (len(fa), len(fb))
        assert rows(rays[p["begin_first"]:p["begin_first"] + p["n_first"]]) == sorted(map(tuple, fa.reshape(-1, 6).tolist()))
        assert rows(rays[p["begin_second"]:p["begin_second"] + p["n_second"]]) == sorted(map(tuple, fb.reshape(-1, 6).tolist()))
