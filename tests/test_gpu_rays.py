"""Uncollide rays on the GPU against the oracle (CreateUncollideRays.cpp:60-200).

The per-triangle sums are formed in the reference's merge order, so every
ray is bit-identical to the oracle's for meshes whose rays do not depend on
the reference's unordered_map walk (vertex normals, or soups).  The rays'
order and the float sum behind each average point follow that walk in the
reference, so rays are compared as sets and averages within 1e-5 relative.
"""
import numpy as np
import pytest

from conftest import case
from inmyroom_vulkan_amd import ENTITY_DTYPE, MODE_BATCHED, MODE_PLAIN, QUERY_DTYPE, scenes

pytestmark = pytest.mark.gpu


def _rows(r):
    return sorted(map(tuple, np.concatenate([r["origin"], r["direction"]], axis=1).tolist()))


def _check(gpu_pairs, gpu_rays, ref, key_of):
    """Pairs with rays, their ray sets (exact) and averages (1e-5 relative)."""
    want = {key_of(i): i for i in range(len(ref.pairs)) if ref.pairs[i, 2] + ref.pairs[i, 3] > 0}
    got = {(int(p["first"]), int(p["second"]), int(p["query"])): p for p in gpu_pairs}
    assert set(want) == set(got)
    for k, i in want.items():
        p = got[k]
        fa, fb = ref.split(i)
        assert (p["n_first"], p["n_second"]) == (len(fa), len(fb)), k
        ga = gpu_rays[p["begin_first"]:p["begin_first"] + p["n_first"]]
        gb = gpu_rays[p["begin_second"]:p["begin_second"] + p["n_second"]]
        ra = sorted(map(tuple, fa.reshape(-1, 6).tolist()))
        rb = sorted(map(tuple, fb.reshape(-1, 6).tolist()))
        assert _rows(ga) == ra, k
        assert _rows(gb) == rb, k
        for s, name in ((0, "avg_first"), (1, "avg_second")):
            np.testing.assert_allclose(p[name], ref.avg[i, s], rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("normals", [True, False])
def test_rays_queries_config1(gpu, orc, normals):
    case(f"rays, config 1 poses, vertex normals={normals}")
    P, F = scenes.icosphere(8)
    P = P.astype(np.float32)
    t = orc.tree(P[F].reshape(-1, 9), 4)
    if normals:
        mid = gpu.register_mesh(P, F)
        vn = (P / np.linalg.norm(P, axis=1, keepdims=True)).astype(np.float32)
        gpu.set_mesh_normals(mid, vn)
        t.set_attrs(vn[F].reshape(-1, 9), F)
    else:  # face normals on a soup: no vertex is shared, so nothing depends on the walk order
        mid = gpu.register_soups([P[F].reshape(-1, 9)])[0]
    _, mats = scenes.config1(300, seed=1)
    q = np.zeros(len(mats), QUERY_DTYPE)
    q["mesh_a"], q["mesh_b"], q["rel"] = mid, mid, mats
    gpu.query_batch(q, MODE_PLAIN)
    pairs, rays = gpu.uncollide_rays()
    ref = orc.rays_queries([t], np.zeros(len(mats), np.uint32), np.zeros(len(mats), np.uint32), mats)
    assert len(pairs) > 40
    _check(pairs, rays, ref, lambda i: (mid, mid, i))


def test_rays_frame(gpu, orc):
    case("rays, config 2 frames")
    sc = scenes.Scene(n_entities=96, seed=2, cube=12.0)
    ids = gpu.register_soups(sc.soups)
    trees = [orc.tree(s, 4) for s in sc.soups]
    checked = 0
    for _ in range(3):
        ents = sc.entities(ids, ENTITY_DTYPE)
        gpu.frame(ents, MODE_PLAIN)
        pairs, rays = gpu.uncollide_rays()
        ref = orc.rays_frame(trees, sc.mesh_of, ents["current"], ents["should_callback"])
        # frame pairs are keyed by (first, second); the query index is the pair's slot in each list
        got = {(int(p["first"]), int(p["second"])): p for p in pairs}
        want = {(int(ref.pairs[i, 0]), int(ref.pairs[i, 1])): i for i in range(len(ref.pairs))}
        assert set(got) == set(want)
        for k, i in want.items():
            p = got[k]
            fa, fb = ref.split(i)
            assert (p["n_first"], p["n_second"]) == (len(fa), len(fb))
            ga = rays[p["begin_first"]:p["begin_first"] + p["n_first"]]
            gb = rays[p["begin_second"]:p["begin_second"] + p["n_second"]]
            assert _rows(ga) == sorted(map(tuple, fa.reshape(-1, 6).tolist()))
            assert _rows(gb) == sorted(map(tuple, fb.reshape(-1, 6).tolist()))
            np.testing.assert_allclose(p["avg_first"], ref.avg[i, 0], rtol=1e-5, atol=1e-5)
            np.testing.assert_allclose(p["avg_second"], ref.avg[i, 1], rtol=1e-5, atol=1e-5)
        checked += len(want)
        sc.step()
    assert checked > 0


def test_rays_batched_mode_invariants(gpu):
    """Exact-predicate contacts give rays by the same rules: unit directions,
    vertex rays at mesh vertices, at most one ray per vertex and side."""
    case("rays after a batched run")
    P, F = scenes.icosphere(8)
    P = P.astype(np.float32)
    mid = gpu.register_mesh(P, F)
    vn = (P / np.linalg.norm(P, axis=1, keepdims=True)).astype(np.float32)
    gpu.set_mesh_normals(mid, vn)
    _, mats = scenes.config1(200, seed=4)
    q = np.zeros(len(mats), QUERY_DTYPE)
    q["mesh_a"], q["mesh_b"], q["rel"] = mid, mid, mats
    gpu.query_batch(q, MODE_BATCHED)
    pairs, rays = gpu.uncollide_rays()
    assert len(pairs) > 20
    np.testing.assert_allclose(np.linalg.norm(rays["direction"], axis=1), 1.0, atol=1e-6)
    verts = {tuple(v) for v in P.tolist()}
    for p in pairs:
        ga = rays[p["begin_first"]:p["begin_first"] + p["n_first"]]
        on = [tuple(o) in verts for o in ga["origin"].tolist()]
        origins = [tuple(o) for o, v in zip(ga["origin"].tolist(), on) if v]
        assert len(origins) == len(set(origins))


def test_rays_need_a_run(gpu):
    from inmyroom_vulkan_amd import CollideError, Collider
    c = Collider(leaf_max=4)
    try:
        with pytest.raises(CollideError):
            c.uncollide_rays()
    finally:
        c.close()
