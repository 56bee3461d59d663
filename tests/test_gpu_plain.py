"""Plain path on the GPU against the oracle, bit for bit (reference semantics)."""
import numpy as np
import pytest

from conftest import case, golden
from inmyroom_vulkan_amd import MODE_PLAIN, scenes

pytestmark = pytest.mark.gpu


def _contact_keys(c):
    return {(int(x["query"]), int(x["tri_a"]), int(x["tri_b"])): x for x in c}


def _assert_same_contacts(got, want):
    g, w = _contact_keys(got), _contact_keys(want)
    assert set(g) == set(w), (len(set(g) - set(w)), len(set(w) - set(g)))
    for k in w:
        for f in ("flags",):
            assert g[k][f] == w[k][f], k
        for f in ("s", "t", "na", "nb"):
            assert g[k][f].tobytes() == w[k][f].tobytes(), (k, f, g[k][f], w[k][f])


def test_tritri_plain_matches_oracle(gpu, orc):
    g = golden("moller_ref.npz")
    a1, b1, _ = scenes.config4(200000, seed=41)
    a2, b2, _ = scenes.config4(50000, seed=42, span=1.0)
    a9 = np.concatenate([g["a9"], a1, a2])
    b9 = np.concatenate([g["b9"], b1, b2])
    case(f"tritri plain n={len(a9)}")
    f, s = gpu.tritri(a9, b9, MODE_PLAIN)
    fo, so = orc.tritri(a9, b9)
    assert np.array_equal(f, fo)
    keep = (fo & 3) == 1
    assert np.array_equal(s[keep].view(np.uint32), so[keep].view(np.uint32))


def _meshes():
    rng = np.random.default_rng(77)
    return [
        ("ico1280", scenes.unit_icosphere_soup(8)),
        ("ico20", scenes.unit_icosphere_soup(1)),
        ("box", scenes.box_soup()),
        ("torus", scenes.torus_soup(rng, 32, 16)),
        ("disp_ico", scenes.displaced_icosphere(13, 0.1, rng)),
        ("random", rng.normal(size=(333, 9)).astype(np.float32)),
        ("one_tri", scenes.box_soup()[:1]),
        ("dup_points", np.repeat(np.ones((1, 9), np.float32), 9, axis=0)),
        ("flat", np.concatenate([scenes.box_soup()[:2]] * 5)),
    ]


@pytest.mark.parametrize("L", [4, 1])
def test_tree_build_matches_oracle(gpu, gpu_l1, orc, L):
    ctx = gpu if L == 4 else gpu_l1
    meshes = _meshes()
    case(f"tree build L={L} meshes={[n for n, _ in meshes]}")
    ids = ctx.register_soups([m for _, m in meshes])
    for (name, m), mid in zip(meshes, ids):
        want = orc.tree(m, L).export()
        got = ctx.export_tree(mid)
        assert got["root"].tobytes() == want["root"].tobytes(), name
        assert got["obbs"].shape == want["obbs"].shape, name
        assert got["obbs"].tobytes() == want["obbs"].tobytes(), name
        assert np.array_equal(got["links"], want["links"]), name
        assert np.array_equal(got["perm"], want["perm"]), name


def test_tree_build_hand_derived_splits(gpu_l1, orc):
    """The split cases derived by hand (tie to the left, an axis that leaves a
    side empty, index halves, the larger-surface swap): the GPU build gives
    the hand-derived permutation and links, and the oracle's bits."""
    from inmyroom_vulkan_amd import Collider
    from test_oracle_build_pins import CASES, _mesh

    with Collider(leaf_max=2) as gpu_l2:
        ctxs = {1: gpu_l1, 2: gpu_l2}
        for name, (rows, leaf, perm, links) in sorted(CASES.items()):
            case(f"hand-derived split case {name} L={leaf}")
            tri = _mesh(rows)
            got = ctxs[leaf].export_tree(ctxs[leaf].register_soups([tri])[0])
            assert got["perm"].tolist() == perm, name
            assert got["links"].tolist() == links, name
            want = orc.tree(tri, leaf).export()
            assert got["root"].tobytes() == want["root"].tobytes(), name
            assert got["obbs"].tobytes() == want["obbs"].tobytes(), name


def test_tree_build_levels_wider_than_one_launch(gpu_l1, orc):
    """A mesh whose deepest levels hold more nodes than one level launch has
    workgroups (65,536), so the launch strides over them, and deeper than one
    batch of level launches: bit-identical to the oracle's tree."""
    rng = np.random.default_rng(5)
    tri = scenes.displaced_icosphere(128, 0.1, rng)
    case(f"tree build, one mesh of {len(tri)} triangles, leaf size 1")
    got = gpu_l1.export_tree(gpu_l1.register_soups([tri])[0])
    want = orc.tree(tri, 1).export()
    assert len(got["links"]) > 2 * 65536
    assert got["root"].tobytes() == want["root"].tobytes()
    assert got["obbs"].tobytes() == want["obbs"].tobytes()
    assert np.array_equal(got["links"], want["links"])
    assert np.array_equal(got["perm"], want["perm"])


@pytest.mark.parametrize("L", [4, 1])
def test_plain_queries_config1(gpu, gpu_l1, orc, L):
    ctx = gpu if L == 4 else gpu_l1
    mesh, rel = scenes.config1(300, seed=1)
    case(f"plain queries config1 L={L} n=300")
    mid = ctx.register_soups([mesh])[0]
    q = ctx.make_queries(mid, mid, rel)
    got = ctx.query_batch(q, MODE_PLAIN)
    cnt = ctx.counters()
    tree = orc.tree(mesh, L)
    res = orc.batch([tree], np.zeros(len(rel), np.uint32), np.zeros(len(rel), np.uint32), rel, mode=0)
    _assert_same_contacts(got, res.contacts)
    assert cnt["sat_tests"] == res.total_sat
    assert cnt["leaf_pairs"] == res.total_leafpairs
    assert cnt["tri_tests"] == res.total_tritests
    assert len(got) > 100


def test_plain_leafpairs_match_oracle(gpu, orc):
    mesh, rel = scenes.config1(100, seed=2)
    case("plain leaf pairs config1 n=100")
    mid = gpu.register_soups([mesh])[0]
    lp = gpu.query_leafpairs(gpu.make_queries(mid, mid, rel))
    res = orc.batch([orc.tree(mesh, 4)], np.zeros(len(rel), np.uint32), np.zeros(len(rel), np.uint32), rel, mode=0,
                    keep_leafpairs=True)
    got = sorted(map(tuple, lp.tolist()))
    want = sorted((int(q), int(a), int(b), int(ca) | (int(cb) << 16))
                  for q, (a, ca, b, cb) in zip(res.lp_query, res.leafpairs))
    assert got == want


def test_plain_queries_uneven_subset(gpu, orc):
    soups, qa, qb, rel, _ = scenes.config3(96, seed=3, freq_cap=14)
    case(f"plain queries config3 subset nq={len(qa)}")
    ids = np.array(gpu.register_soups(soups), np.uint32)
    q = np.zeros(len(qa), dtype=gpu.make_queries(0, 0, rel[:1]).dtype)
    q["mesh_a"], q["mesh_b"], q["rel"] = ids[qa], ids[qb], rel
    got = gpu.query_batch(q, MODE_PLAIN)
    trees = [orc.tree(s, 4) for s in soups]
    res = orc.batch(trees, qa, qb, rel, mode=0)
    _assert_same_contacts(got, res.contacts)
    assert len(got) > 100


def test_plain_edge_cases(gpu, orc):
    one = scenes.box_soup()[:1]
    ico = scenes.unit_icosphere_soup(2)
    case("plain edge cases: root-leaf mesh, separated poses, empty batch")
    ida, idb = gpu.register_soups([one, ico])
    rel = scenes.pose_matrices(np.array([[0.0, 0, 0], [10.0, 0, 0], [0.3, 0.2, 0.1]]),
                               scenes.random_quaternions(np.random.default_rng(1), 3))
    q = np.concatenate([gpu.make_queries(ida, idb, rel), gpu.make_queries(idb, ida, rel), gpu.make_queries(ida, ida, rel)])
    got = gpu.query_batch(q, MODE_PLAIN)
    ta, tb = orc.tree(one, 4), orc.tree(ico, 4)
    res = orc.batch([ta, tb], np.array([0] * 3 + [1] * 3 + [0] * 3, np.uint32),
                    np.array([1] * 3 + [0] * 3 + [0] * 3, np.uint32), np.concatenate([rel] * 3), mode=0)
    _assert_same_contacts(got, res.contacts)
    assert len(gpu.query_batch(q[:0], MODE_PLAIN)) == 0
