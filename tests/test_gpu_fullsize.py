"""The batching and deformable workloads at their full stated sizes, on the
GPU, against the oracle.

* Batching (4,096 mesh-pair queries, 8,192 displaced icospheres of 20 to
  64,980 triangles, 62.2M triangles): the plain path bit for bit against the
  oracle's reference semantics on every query (contacts and the traversal's
  counts), the batched path's pair set equal to the exact oracle's on every
  query, with every contact's flags, normals and (conditioned) segment
  checked.  The oracle builds its own 8,192 trees (threads over meshes).
* Deformable batch (64 skinned tori of 20,000 triangles, 64 bones, 4 weights
  per vertex, against the 316 x 316-vertex terrain): skinning against the
  float64 formula, the exact pair set of all 64 meshes against the oracle's
  (its trees built from the GPU's skinned triangles), and every contact's
  flags, normals and segment.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from conftest import case
from contacts_check import (TreePositions, assert_plain_bit_exact, assert_same_keys, check_batched_geometry,
                            corners_from_trees, sort_contacts)
from inmyroom_vulkan_amd import MODE_BATCHED, MODE_PLAIN, QUERY_DTYPE, scenes

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def _trees(orc, soups, leaf=4):
    with ThreadPoolExecutor(THREADS) as ex:  # the oracle's build releases the GIL
        return list(ex.map(lambda s: orc.tree(s, leaf), soups))


class _C3:
    pass


@pytest.fixture(scope="module")
def c3(orc):
    from inmyroom_vulkan_amd import Collider

    w = _C3()
    case("full batch: generating 4,096 queries")
    w.soups, w.qa, w.qb, w.rel, _ = scenes.config3(4096, seed=3)
    w.qa, w.qb = w.qa.astype(np.uint32), w.qb.astype(np.uint32)
    w.gpu = Collider(leaf_max=4)
    case(f"full batch: GPU trees of {len(w.soups)} meshes, {sum(len(s) for s in w.soups):,} triangles")
    ids = np.array(w.gpu.register_soups(w.soups), np.uint32)
    w.q = np.zeros(len(w.qa), QUERY_DTYPE)
    w.q["mesh_a"], w.q["mesh_b"], w.q["rel"] = ids[w.qa], ids[w.qb], w.rel
    case(f"full batch: oracle trees ({THREADS} threads)")
    w.trees = _trees(orc, w.soups)
    yield w
    w.gpu.close()


def test_full_batch_plain_bit_exact(c3, orc):
    case("full batch, plain path vs the oracle (reference semantics), 4,096 queries")
    got = sort_contacts(c3.gpu.query_batch(c3.q, MODE_PLAIN))
    cnt = c3.gpu.counters()
    res = orc.batch(c3.trees, c3.qa, c3.qb, c3.rel, mode=0, nthreads=THREADS)
    want = sort_contacts(res.contacts)
    assert len(want) > 150_000
    assert_plain_bit_exact(got, want)
    assert cnt["sat_tests"] == res.total_sat
    assert cnt["leaf_pairs"] == res.total_leafpairs
    assert cnt["tri_tests"] == res.total_tritests
    assert max(len(s) for s in c3.soups) > 60_000  # the largest meshes take part


def test_full_batch_reference_mode_bit_exact(c3, orc):
    """Reference semantics on the batched machinery, all 4,096 queries: the
    oracle's reference-semantics contacts and counts, bit for bit."""
    from inmyroom_vulkan_amd import MODE_REFERENCE

    case("full batch, reference mode vs the oracle (reference semantics), 4,096 queries")
    got = sort_contacts(c3.gpu.query_batch(c3.q, MODE_REFERENCE))
    cnt = c3.gpu.counters()
    res = orc.batch(c3.trees, c3.qa, c3.qb, c3.rel, mode=0, nthreads=THREADS)
    assert_plain_bit_exact(got, sort_contacts(res.contacts))
    assert (cnt["sat_tests"], cnt["leaf_pairs"], cnt["tri_tests"]) == (res.total_sat, res.total_leafpairs,
                                                                        res.total_tritests)


def test_full_batch_batched_vs_exact_oracle(c3, orc):
    case("full batch, batched path vs the exact oracle, 4,096 queries")
    got = sort_contacts(c3.gpu.query_batch(c3.q, MODE_BATCHED))
    res = orc.batch(c3.trees, c3.qa, c3.qb, c3.rel, mode=1, nthreads=THREADS)
    want = sort_contacts(res.contacts)
    assert len(want) > 150_000
    assert_same_keys(got, want)  # zero pair-set mismatches
    A, B = corners_from_trees(want, TreePositions(c3.soups, c3.trees), c3.qa, c3.qb, c3.rel)
    checked, worst = check_batched_geometry(got, want, A, B)
    print(f"[full batch] {len(want)} contacts, {checked} segments checked, worst relative error {worst:.3g}")


def test_full_deformable_batch_vs_exact_oracle(orc):
    from inmyroom_vulkan_amd import Collider

    case("deformable batch: 64 skinned tori x 20,000 triangles vs the 316 x 316 terrain")
    db = scenes.DeformableBatch(n_meshes=64, seed=5)
    assert len(db.F) == 20_000 and db.n_bones == 64 and len(db.tF) == 315 * 315 * 2
    with Collider(leaf_max=4) as gpu:
        ids = np.array([gpu.register_skinned(db.P, db.F, db.joints, db.weights, db.n_bones) for _ in range(64)],
                       np.uint32)
        terrain = gpu.register_mesh(db.tP.astype(np.float32), db.tF)
        bones = db.bone_matrices()
        gpu.skin_refit(ids, bones)
        q = np.zeros(64, QUERY_DTYPE)
        q["mesh_a"], q["mesh_b"], q["rel"] = ids, terrain, scenes.identity16()
        got = gpu.query_batch(q, MODE_BATCHED)
        soups, gperm = [], []
        for k, m in enumerate(ids):  # the GPU's skinned triangles in input order: what the oracle builds from
            tri = gpu.export_triangles(int(m))
            perm = gpu.export_tree(int(m))["perm"]
            inp = np.empty_like(tri)
            inp[perm] = tri
            want_pos = scenes.DeformableBatch.skin_reference(db.P, db.joints, db.weights, bones[k])[db.F]
            assert np.allclose(inp.reshape(-1, 3, 3), want_pos, rtol=0, atol=2e-5 * (1 + np.abs(want_pos).max()))
            soups.append(inp)
            gperm.append(perm)
        tperm = gpu.export_tree(int(terrain))["perm"]
    tsoup = scenes.soup(db.tP, db.tF)
    case("deformable batch: oracle trees and exact queries")
    trees = _trees(orc, soups + [tsoup])
    res = orc.batch(trees, np.arange(64, dtype=np.uint32), np.full(64, 64, np.uint32),
                    np.tile(scenes.identity16(), (64, 1)), mode=1, nthreads=THREADS)
    operm = [t.export()["perm"] for t in trees]
    # both sides as (query, input triangle of A, input triangle of B)
    g = got.copy()
    w = res.contacts.copy()
    for k in range(64):
        sg, sw = g["query"] == k, w["query"] == k
        g["tri_a"][sg] = gperm[k][g["tri_a"][sg]]
        w["tri_a"][sw] = operm[k][w["tri_a"][sw]]
    g["tri_b"] = tperm[g["tri_b"]]
    w["tri_b"] = operm[64][w["tri_b"]]
    g, w = sort_contacts(g), sort_contacts(w)
    assert len(w) > 1000 and len(np.unique(w["query"])) > 32
    assert_same_keys(g, w)
    A = np.stack([soups[k][t] for k, t in zip(w["query"], w["tri_a"])]).reshape(-1, 3, 3)
    B = tsoup[w["tri_b"]].reshape(-1, 3, 3)
    checked, worst = check_batched_geometry(g, w, A, B)
    print(f"[deformable] {len(w)} contacts over {len(np.unique(w['query']))} meshes, {checked} segments checked, "
          f"worst relative error {worst:.3g}")
