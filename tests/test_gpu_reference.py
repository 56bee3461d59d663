"""Reference-semantics mode on the GPU: the plain path's tests (float SAT,
the reference's box transform and descent measure, the Möller test with its
snap) run through the batched machinery (split levels, persistent
traversal, dense narrow phase).  Every contact field and every count must
equal the plain path's and the oracle's reference-semantics results, bit for
bit (OBBtree.cpp:414-477, CreateUncollideRays.cpp:74-115)."""
import numpy as np
import pytest

from conftest import case
from contacts_check import assert_plain_bit_exact, sort_contacts
from inmyroom_vulkan_amd import ENTITY_DTYPE, MODE_PLAIN, MODE_REFERENCE, scenes

pytestmark = pytest.mark.gpu

COUNTS = ("sat_tests", "leaf_pairs", "tri_tests", "contacts", "ub_cases")


def _queries(gpu, soups, qa, qb, rel):
    ids = np.array(gpu.register_soups(soups), np.uint32)
    q = np.zeros(len(qa), dtype=gpu.make_queries(0, 0, rel[:1]).dtype)
    q["mesh_a"], q["mesh_b"], q["rel"] = ids[qa], ids[qb], rel
    return q


def test_reference_mode_matches_plain_and_oracle(gpu, orc):
    soups, qa, qb, rel, _ = scenes.config3(160, seed=3, freq_cap=16)
    case(f"reference mode, config3 subset nq={len(qa)}")
    q = _queries(gpu, soups, qa, qb, rel)
    plain = sort_contacts(gpu.query_batch(q, MODE_PLAIN))
    pc = gpu.counters()
    ref = sort_contacts(gpu.query_batch(q, MODE_REFERENCE))
    rc = gpu.counters()
    assert_plain_bit_exact(ref, plain)
    for k in COUNTS:
        assert rc[k] == pc[k], k
    trees = [orc.tree(s, 4) for s in soups]
    res = orc.batch(trees, qa.astype(np.uint32), qb.astype(np.uint32), rel, mode=0)
    assert_plain_bit_exact(ref, sort_contacts(res.contacts))
    assert (rc["sat_tests"], rc["leaf_pairs"], rc["tri_tests"], rc["ub_cases"]) == (
        res.total_sat, res.total_leafpairs, res.total_tritests, res.total_ub)
    assert len(ref) > 100


@pytest.mark.parametrize("tuning", [
    dict(pipeline=1),
    dict(split_depth=1, split_levels=9),
    dict(split_levels=2, stack_entries=512, pop_width=16, waves_per_cu=2),
    dict(split_levels=24, split_tris=1, stack_entries=1024, waves_per_cu=4),
    dict(split_levels=4, stack_entries=512, waves_per_cu=4, work_sharing=1, spin_polls=8),
    dict(split_levels=8, stack_entries=512, waves_per_cu=2, work_sharing=3, handoff_pct=30, resplit_levels=2),
])
def test_reference_mode_tuning_variants(gpu, tuning):
    """Whichever launch takes a node pair, the pair's test, its children and
    its leaf pairs are the same: contacts and counts never change."""
    soups, qa, qb, rel, _ = scenes.config3(64, seed=5, freq_cap=24)
    case(f"reference mode tuning {tuning}")
    q = _queries(gpu, soups, qa, qb, rel)
    plain = sort_contacts(gpu.query_batch(q, MODE_PLAIN))
    pc = gpu.counters()
    try:
        gpu.set_tuning(**tuning)
        got = sort_contacts(gpu.query_batch(q, MODE_REFERENCE))
        rc = gpu.counters()
    finally:
        gpu.set_tuning()
    assert_plain_bit_exact(got, plain)
    for k in COUNTS:
        assert rc[k] == pc[k], k


def test_reference_mode_graph_replays(gpu):
    soups, qa, qb, rel, _ = scenes.config3(96, seed=7, freq_cap=20)
    case("reference mode graph replays")
    q = _queries(gpu, soups, qa, qb, rel)
    want = sort_contacts(gpu.query_batch(q, MODE_PLAIN))
    try:
        gpu.set_tuning(use_graph=1)
        gpu.stage(q)
        for rep in range(4):
            gpu.run_staged(MODE_REFERENCE)
            assert_plain_bit_exact(sort_contacts(gpu.fetch()), want)
            gpu.query_batch(q[:8], MODE_PLAIN)  # eager work of another kind between replays
            gpu.stage(q)
    finally:
        gpu.set_tuning()


def test_reference_mode_frame_512_entities(gpu, orc):
    """A 512-entity scene frame (the scene workload) in reference mode against
    the oracle's reference-semantics frame: same pairs, same contacts."""
    sc = scenes.Scene(n_entities=512, seed=2)
    case("reference mode frame, 512 entities")
    ids = gpu.register_soups(sc.soups)
    trees = [orc.tree(s, 4) for s in sc.soups]
    for frame in range(2):
        ents = sc.entities(ids, ENTITY_DTYPE)
        pairs, cons = gpu.frame(ents, MODE_REFERENCE)
        res = orc.frame(trees, sc.mesh_of, ents["current"], sc.should_callback, mode=0)
        want = {}
        for c in res.contacts:
            a, b = res.pairs[c["query"]]
            want[(int(a), int(b), int(c["tri_a"]), int(c["tri_b"]))] = c
        got = {}
        for p in pairs:
            for c in cons[p["contact_begin"]: p["contact_begin"] + p["contact_count"]]:
                got[(int(p["first"]), int(p["second"]), int(c["tri_a"]), int(c["tri_b"]))] = c
        assert set(got) == set(want), frame
        for k, c in want.items():
            g = got[k]
            for f in ("flags", "s", "t", "na", "nb"):
                assert g[f].tobytes() == c[f].tobytes(), (k, f)
        keys = [(int(p["first"]), int(p["second"])) for p in pairs]
        assert keys == sorted(keys) and all(p["contact_count"] > 0 for p in pairs)
        assert len(want) > 100
        sc.step()


def test_reference_mode_empty_and_tiny(gpu):
    soups, qa, qb, rel, _ = scenes.config3(8, seed=9, freq_cap=12)
    q = _queries(gpu, soups, qa, qb, rel)
    case("reference mode, empty and one-query batches")
    assert len(gpu.query_batch(q[:0], MODE_REFERENCE)) == 0
    for i in range(len(q)):
        got = sort_contacts(gpu.query_batch(q[i:i + 1], MODE_REFERENCE))
        want = sort_contacts(gpu.query_batch(q[i:i + 1], MODE_PLAIN))
        assert_plain_bit_exact(got, want)
