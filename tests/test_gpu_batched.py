"""Batched path on the GPU against the exact-semantics oracle.

The batched path evaluates every orientation predicate with its exact sign
(float32 filter, float64 fallback, exact expansion), and its traversal is
conservative, so its colliding pair set must equal the oracle's exact set
(bignum orientation, conservative double traversal) with zero mismatches.
"""
import numpy as np
import pytest

from conftest import case, golden
from contacts_check import (MIN_SIN, SEG_TOL, TreePositions, assert_same_keys, check_batched_geometry,
                            corners_from_trees, crossing_sin, sort_contacts)
from inmyroom_vulkan_amd import MODE_BATCHED, MODE_PLAIN, scenes

pytestmark = pytest.mark.gpu


def _keys(c):
    return {(int(x["query"]), int(x["tri_a"]), int(x["tri_b"])) for x in c}


def test_tritri_exact_stream_matches_oracle(gpu, orc):
    g = golden("moller_ref.npz")
    a1, b1, _ = scenes.config4(100000, seed=51)
    a2, b2, _ = scenes.config4(100000, seed=52, span=1.0)
    a9 = np.concatenate([g["a9"], a1, a2])
    b9 = np.concatenate([g["b9"], b1, b2])
    case(f"tritri exact n={len(a9)}")
    f, seg = gpu.tritri(a9, b9, MODE_BATCHED)
    want = orc.tritri_exact(a9, b9)
    assert np.array_equal(f, want)
    cnt = gpu.counters()
    assert cnt["fp64_fallbacks"] > 0  # the near-degenerate cases do reach the fallback
    # every intersecting pair's segment against the float64 segment of the
    # same float corners, where the crossing edges are conditioned
    hit = np.flatnonzero(f == 1)
    A, B = a9[hit].reshape(-1, 3, 3), b9[hit].reshape(-1, 3, 3)
    cond = np.minimum(crossing_sin(A, B), crossing_sin(B, A)) >= MIN_SIN
    s64 = orc.segment_f64(a9[hit], b9[hit]).astype(np.float64)
    g = seg[hit].astype(np.float64)
    err = np.minimum(np.abs(g - s64).max(1), np.abs(g - np.c_[s64[:, 3:], s64[:, :3]]).max(1))
    scale = np.maximum(np.abs(A).reshape(len(A), -1).max(1), np.abs(B).reshape(len(B), -1).max(1))
    assert cond.mean() > 0.8 and len(hit) > 10_000
    bad = np.flatnonzero(cond & (err > SEG_TOL * scale))
    assert len(bad) == 0, (len(bad), hit[bad[:5]])


def _exact_oracle(orc, soups, qa, qb, rel, L):
    trees = [orc.tree(s, L) for s in soups]
    res = orc.batch(trees, qa, qb, rel, mode=1)
    res.tree_pos = TreePositions(soups, trees)
    return res


def _check_exact(got, want, qa, qb, rel):
    """Pair set equal to the exact oracle's; flags, normals and segments per pair."""
    g, w = sort_contacts(got), sort_contacts(want.contacts)
    assert_same_keys(g, w)
    A, B = corners_from_trees(w, want.tree_pos, qa, qb, rel)
    check_batched_geometry(g, w, A, B)


def test_batched_queries_config1(gpu, orc):
    mesh, rel = scenes.config1(300, seed=1)
    case("batched queries config1 n=300")
    mid = gpu.register_soups([mesh])[0]
    q = gpu.make_queries(mid, mid, rel)
    got = gpu.query_batch(q, MODE_BATCHED)
    z = np.zeros(len(rel), np.uint32)
    want = _exact_oracle(orc, [mesh], z, z, rel, 4)
    _check_exact(got, want, z, z, rel)
    assert len(got) > 100
    # same pair set as the plain path away from degenerate contacts
    plain = gpu.query_batch(q, MODE_PLAIN)
    diff = _keys(got) ^ _keys(plain)
    assert len(diff) <= max(3, len(plain) // 500)


def test_batched_queries_uneven_subset(gpu, orc):
    soups, qa, qb, rel, _ = scenes.config3(128, seed=3, freq_cap=16)
    case(f"batched queries config3 subset nq={len(qa)}")
    ids = np.array(gpu.register_soups(soups), np.uint32)
    q = np.zeros(len(qa), dtype=gpu.make_queries(0, 0, rel[:1]).dtype)
    q["mesh_a"], q["mesh_b"], q["rel"] = ids[qa], ids[qb], rel
    got = gpu.query_batch(q, MODE_BATCHED)
    want = _exact_oracle(orc, soups, qa, qb, rel, 4)
    _check_exact(got, want, qa, qb, rel)
    assert len(got) > 100


def test_batched_self_contact_deferred_tests(gpu, orc):
    """A mesh against itself at the identity: neighbouring triangles share
    vertices, so their predicates are exactly zero and every such test goes
    from the float32 pass through the float64 rerun to the exact one; there
    are more of them than the rerun lists first hold, so the run is repeated
    with longer lists."""
    P, F = scenes.icosphere(32)
    soup = P.astype(np.float32)[F].reshape(-1, 9)
    case(f"batched self contact at identity, {len(soup)} triangles")
    mid = gpu.register_soups([soup])[0]
    rel = scenes.identity16()[None]
    got = gpu.query_batch(gpu.make_queries(mid, mid, rel), MODE_BATCHED)
    cnt = gpu.counters()
    z = np.zeros(1, np.uint32)
    want = _exact_oracle(orc, [soup], z, z, rel, 4)
    g, w = sort_contacts(got), sort_contacts(want.contacts)
    assert_same_keys(g, w)
    # touching neighbours: flags and normals per pair; their segments are points or edges
    A, B = corners_from_trees(w, want.tree_pos, z, z, rel)
    check_batched_geometry(g, w, A, B, min_checked=0.0)
    assert cnt["exact_fallbacks"] > 65536
    assert len(got) > 3 * len(soup)  # each triangle touches its edge and vertex neighbours


@pytest.mark.parametrize("tuning", [
    dict(split_levels=0, stack_entries=512, pop_width=64, waves_per_cu=2),   # no split; spills in LDS stack
    dict(split_levels=4, stack_entries=2048, pop_width=16, waves_per_cu=4),
    dict(split_levels=12, stack_entries=1024, pop_width=64, waves_per_cu=8),
    dict(split_levels=6, stack_entries=512, pop_width=1, waves_per_cu=1),
    dict(split_levels=2, stack_entries=512, pop_width=64, waves_per_cu=4, work_sharing=0),
    # idle wavefronts take donated node pairs; a tiny spin budget leaves donations to the second launch
    dict(split_levels=2, stack_entries=512, pop_width=64, waves_per_cu=4, work_sharing=1),
    dict(split_levels=0, stack_entries=1024, pop_width=32, waves_per_cu=2, work_sharing=1, spin_polls=1),
    dict(split_levels=10, stack_entries=512, pop_width=64, waves_per_cu=8, work_sharing=1, spin_polls=8),
    # deep, selective splitting: only pairs over split_tris triangles go on to the next level
    dict(split_levels=24, split_tris=64, stack_entries=512, pop_width=64, waves_per_cu=4),
    dict(split_levels=40, split_tris=1, stack_entries=512, pop_width=64, waves_per_cu=4),
    # one batch of node pairs in flight per wavefront, against the two-batch default
    dict(split_levels=16, split_tris=256, stack_entries=512, pop_width=64, waves_per_cu=6, pipeline=1),
    dict(split_levels=0, stack_entries=2048, pop_width=16, waves_per_cu=2, pipeline=2),
    dict(split_levels=6, stack_entries=1024, pop_width=3, waves_per_cu=10, pipeline=2),
    # split levels one per launch, and two per launch with an odd count (the last level alone)
    dict(split_levels=16, split_tris=256, stack_entries=512, waves_per_cu=6, split_depth=1),
    dict(split_levels=15, split_tris=64, stack_entries=512, waves_per_cu=6, split_depth=2),
    dict(split_levels=1, stack_entries=512, waves_per_cu=6, split_depth=2),
])
def test_batched_tuning_variants_agree(gpu, tuning):
    soups, qa, qb, rel, _ = scenes.config3(64, seed=5, freq_cap=24)
    case(f"batched tuning variant {tuning}")
    ids = np.array(gpu.register_soups(soups), np.uint32)
    q = np.zeros(len(qa), dtype=gpu.make_queries(0, 0, rel[:1]).dtype)
    q["mesh_a"], q["mesh_b"], q["rel"] = ids[qa], ids[qb], rel
    gpu.set_tuning()
    base = _keys(gpu.query_batch(q, MODE_BATCHED))
    base_cnt = gpu.counters()
    try:
        gpu.set_tuning(**tuning)
        got = _keys(gpu.query_batch(q, MODE_BATCHED))
        cnt = gpu.counters()
    finally:
        gpu.set_tuning()
    assert got == base
    assert cnt["leaf_pairs"] == base_cnt["leaf_pairs"] and cnt["tri_tests"] == base_cnt["tri_tests"]
    assert cnt["sat_tests"] == base_cnt["sat_tests"]  # every node pair tested exactly once, whichever launch takes it


@pytest.mark.parametrize("tuning", [
    dict(split_levels=16, split_tris=512, stack_entries=1024, waves_per_cu=8, work_sharing=3, handoff_pct=10,
         resplit_levels=2),  # a tiny budget: most of the work is handed on and split again
    dict(split_levels=0, stack_entries=512, waves_per_cu=4, work_sharing=3, handoff_pct=100, resplit_levels=4),
    dict(split_levels=8, stack_entries=512, waves_per_cu=2, work_sharing=3, handoff_pct=50, resplit_levels=16),
    dict(split_levels=6, stack_entries=512, waves_per_cu=4, work_sharing=3, handoff_pct=30, resplit_levels=0),
])
def test_batched_budgeted_pass_agrees(gpu, tuning):
    """Budgeted traversal: each run's budget comes from the run before, so
    the later runs hand work on; every run tests every node pair once."""
    soups, qa, qb, rel, _ = scenes.config3(64, seed=5, freq_cap=24)
    case(f"budgeted pass {tuning}")
    ids = np.array(gpu.register_soups(soups), np.uint32)
    q = np.zeros(len(qa), dtype=gpu.make_queries(0, 0, rel[:1]).dtype)
    q["mesh_a"], q["mesh_b"], q["rel"] = ids[qa], ids[qb], rel
    gpu.set_tuning()
    base = _keys(gpu.query_batch(q, MODE_BATCHED))
    base_cnt = gpu.counters()
    try:
        for use_graph in (0, 0, 1, 1):
            gpu.set_tuning(**tuning, use_graph=use_graph)
            got = _keys(gpu.query_batch(q, MODE_BATCHED))
            cnt = gpu.counters()
            assert got == base
            for k in ("sat_tests", "leaf_pairs", "tri_tests"):
                assert cnt[k] == base_cnt[k], k
    finally:
        gpu.set_tuning()


def test_batched_graph_replays_match_eager(gpu):
    soups, qa, qb, rel, _ = scenes.config3(96, seed=7, freq_cap=20)
    case("batched graph replays")
    ids = np.array(gpu.register_soups(soups), np.uint32)
    q = np.zeros(len(qa), dtype=gpu.make_queries(0, 0, rel[:1]).dtype)
    q["mesh_a"], q["mesh_b"], q["rel"] = ids[qa], ids[qb], rel
    gpu.set_tuning()
    eager = _keys(gpu.query_batch(q, MODE_BATCHED))
    try:
        gpu.set_tuning(use_graph=1)
        gpu.stage(q)
        for rep in range(4):
            gpu.run_staged(MODE_BATCHED)
            got = _keys(gpu.fetch())
            assert got == eager, f"replay {rep}"
            # eager work of another kind between replays
            gpu.query_batch(q[:8], MODE_PLAIN)
            gpu.stage(q)
    finally:
        gpu.set_tuning()
