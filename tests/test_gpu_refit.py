"""Skinning and OBB-tree refit on the GPU (deformable batch).

The reference has no refit, so there is no reference bit pattern to match:
the refit is pinned by the skinning formula (float64 reference), by box
invariants (containment, orthonormal axes, agreement with the build when
nothing moves) and by the exact pair set against the terrain, which the
oracle computes on the skinned triangles with its own tree.
"""
import numpy as np
import pytest

from conftest import case
from inmyroom_vulkan_amd import MODE_BATCHED, QUERY_DTYPE, scenes

pytestmark = pytest.mark.gpu


def _batch(n=4, nu=24, nv=16, seed=5):
    return scenes.DeformableBatch(n_meshes=n, nu=nu, nv=nv, n_bones=16, seed=seed, terrain_n=64, terrain_size=30.0)


def _register(gpu, db):
    ids = [gpu.register_skinned(db.P, db.F, db.joints, db.weights, db.n_bones) for _ in range(db.n_meshes)]
    return np.array(ids, np.uint32)


def test_skinning_matches_formula(gpu):
    db = _batch()
    case("skinning vs float64 formula")
    ids = _register(gpu, db)
    bones = db.bone_matrices()
    gpu.skin_refit(ids, bones)
    for k, mid in enumerate(ids):
        tri = gpu.export_triangles(int(mid))
        perm = gpu.export_tree(int(mid))["perm"]
        want = scenes.DeformableBatch.skin_reference(db.P, db.joints, db.weights, bones[k])[db.F[perm]].reshape(-1, 9)
        assert np.allclose(tri, want, rtol=0, atol=2e-5 * (1 + np.abs(want).max()))


def test_refit_identity_matches_build(gpu):
    db = _batch(n=2)
    case("refit with identity bones")
    ids = _register(gpu, db)
    before = [gpu.export_tree(int(m)) for m in ids]
    eye = np.tile(scenes.identity16(), (len(ids) * db.n_bones, 1))
    gpu.skin_refit(ids, eye)
    for b, m in zip(before, ids):
        a = gpu.export_tree(int(m))
        assert np.array_equal(a["links"], b["links"]) and np.array_equal(a["perm"], b["perm"])
        assert np.allclose(a["obbs"], b["obbs"], rtol=1e-4, atol=1e-5) or \
            np.allclose(np.abs(a["obbs"]), np.abs(b["obbs"]), rtol=1e-4, atol=1e-5)


def _contains(box, pts, tol):
    """Every point within distance tol of the box (float boxes at |x| ~ 20 round their centre by ~1e-6,
    the same order as the 2 FLT_EPSILON padding, so containment is a distance, as the slack is)."""
    c, S = box[:3].astype(np.float64), box[3:].reshape(3, 3).astype(np.float64)
    coef = np.linalg.solve(S.T, (pts.astype(np.float64) - c).T).T
    excess = (np.maximum(0.0, np.abs(coef) - 1.0) * np.linalg.norm(S, axis=1)).sum(axis=1)
    return np.all(excess <= tol)


def test_refit_boxes_contain_their_triangles(gpu):
    db = _batch(n=2, seed=9)
    case("refit containment")
    ids = _register(gpu, db)
    gpu.skin_refit(ids, db.bone_matrices())
    for m in ids:
        t = gpu.export_tree(int(m))
        tri = gpu.export_triangles(int(m))
        info = gpu.mesh_info(int(m))
        assert info["slack"] < 1e-5
        assert _contains(t["root"], tri.reshape(-1, 3), 1e-5)
        links, obbs = t["links"], t["obbs"]
        for i in range(len(links)):
            for s in range(2):
                box = obbs[i, 12 * s:12 * s + 12]
                S = box[3:].reshape(3, 3).astype(np.float64)
                Sn = S / np.linalg.norm(S, axis=1, keepdims=True)
                assert np.allclose(Sn @ Sn.T, np.eye(3), atol=1e-4)
                if links[i, 2 + s]:
                    o, c = int(links[i, s]), int(links[i, 2 + s])
                    assert _contains(box, tri[o:o + c].reshape(-1, 3), 1e-5)


def test_deformable_vs_terrain_exact_pairs(gpu, orc):
    db = _batch(n=6, nu=32, nv=20, seed=11)
    case("deformable batch vs terrain, exact pair set")
    ids = _register(gpu, db)
    terrain = gpu.register_mesh(db.tP.astype(np.float32), db.tF)
    bones = db.bone_matrices()
    gpu.skin_refit(ids, bones)
    q = np.zeros(len(ids), QUERY_DTYPE)
    q["mesh_a"], q["mesh_b"], q["rel"] = ids, terrain, scenes.identity16()
    got = gpu.query_batch(q, MODE_BATCHED)
    tsoup = scenes.soup(db.tP, db.tF)
    tperm = gpu.export_tree(terrain)["perm"]
    ttree = orc.tree(tsoup, 4)
    tperm_o = ttree.export()["perm"]
    total = 0
    for k, m in enumerate(ids):
        skinned = scenes.DeformableBatch.skin_reference(db.P, db.joints, db.weights, bones[k])
        # the GPU's skinned triangles in input order (what the oracle must see)
        tri_tree = gpu.export_triangles(int(m))
        perm = gpu.export_tree(int(m))["perm"]
        soup_in = np.empty_like(tri_tree)
        soup_in[perm] = tri_tree
        assert np.allclose(soup_in, skinned[db.F].reshape(-1, 9), atol=1e-4)
        t = orc.tree(soup_in, 4)
        res = orc.batch([t, ttree], np.array([0], np.uint32), np.array([1], np.uint32),
                        scenes.identity16()[None], mode=1)
        po = t.export()["perm"]
        want = {(int(po[c["tri_a"]]), int(tperm_o[c["tri_b"]])) for c in res.contacts}
        mine = {(int(perm[c["tri_a"]]), int(tperm[c["tri_b"]])) for c in got if c["query"] == k}
        assert mine == want, k
        total += len(want)
    assert total > 0


def test_rays_after_refit_see_the_new_boxes():
    """Ray shooting keeps per-box terms of its box test (computed once per
    mesh); a refit must renew them.  One context refits twice and shoots
    after each; a second context refits once, to the second pose: their
    responses after that pose agree bit for bit."""
    from inmyroom_vulkan_amd import ENTITY_DTYPE, MODE_REFERENCE, Collider
    db = _batch(n=1, nu=48, nv=24)
    case("rays after a refit use the refit boxes")
    bones_a, bones_b = db.bone_matrices(), db.bone_matrices()
    shift = np.eye(4, dtype=np.float32)
    shift[3, :3] = (0.35, 0.1, 0.0)
    moved = np.eye(4, dtype=np.float32)
    moved[3, :3] = (0.3, 0.1, 0.0)
    e = np.zeros(2, ENTITY_DTYPE)
    e["current"] = [scenes.identity16(), shift.reshape(-1)]
    e["previous"] = [scenes.identity16(), moved.reshape(-1)]
    e["should_callback"] = 1
    out = []
    for bones_seq in ((bones_a, bones_b), (bones_b,)):
        gpu = Collider(leaf_max=4)
        try:
            ids = _register(gpu, db)
            e["mesh_id"] = ids[0]
            for bones in bones_seq:
                gpu.skin_refit(ids, bones)
                got = gpu.step(e, MODE_REFERENCE)
            out.append(got.copy())
        finally:
            gpu.close()
    assert len(out[0]) > 0 and int(out[0]["shot"].sum()) > 0
    assert out[0].tobytes() == out[1].tobytes()
