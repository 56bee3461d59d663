"""Broad phase and whole frames on the GPU against the oracle."""
import numpy as np
import pytest

from conftest import case
from inmyroom_vulkan_amd import ENTITY_DTYPE, MODE_BATCHED, MODE_PLAIN, scenes

pytestmark = pytest.mark.gpu


def _scene(n, seed, cube=25.0):
    return scenes.Scene(n_entities=n, seed=seed, cube=cube, torus_grid=(32, 16))


def test_broad_phase_matches_oracle(gpu, orc):
    sc = _scene(300, 21)
    sc.should_callback[::7] = 0  # a pair is kept only if one of its entities wants callbacks
    case("broad phase n=300")
    ids = gpu.register_soups(sc.soups)
    trees = [orc.tree(s, 4) for s in sc.soups]
    for frame in range(3):
        ents = sc.entities(ids, ENTITY_DTYPE)
        got = [tuple(p) for p in gpu.broad_phase(ents).tolist()]
        want = [tuple(p) for p in orc.frame(trees, sc.mesh_of, ents["current"], sc.should_callback, mode=0).pairs.tolist()]
        assert got == want, frame  # same pairs, same orientation, same (first, second) order
        assert len(want) > 5
        sc.step()


def test_broad_phase_crowded_matches_oracle(gpu, orc):
    """240 unit boxes in a 2 m cube: an entity overlaps up to ~200 later ones
    along U, so the sweep tests them in several chunks of 64."""
    case("broad phase, 240 boxes in a 2 m cube")
    box = scenes.box_soup()
    mid = gpu.register_soups([box])[0]
    tree = orc.tree(box, 4)
    rng = np.random.default_rng(5)
    n = 240
    ents = np.zeros(n, ENTITY_DTYPE)
    ents["current"] = ents["previous"] = scenes.pose_matrices(rng.uniform(-1, 1, (n, 3)), scenes.random_quaternions(rng, n))
    ents["mesh_id"] = mid
    ents["entity"] = np.arange(n)
    ents["should_callback"] = 1
    ents["should_callback"][::5] = 0
    got = [tuple(p) for p in gpu.broad_phase(ents).tolist()]
    res = orc.frame([tree], np.zeros(n, np.uint32), ents["current"], ents["should_callback"], mode=0)
    want = [tuple(p) for p in res.pairs.tolist()]
    assert got == want
    later = np.bincount(np.array(want)[:, 0], minlength=n)
    assert later.max() > 128  # some entity's pairs span three chunks or more


def test_frame_plain_matches_oracle(gpu, orc):
    sc = _scene(200, 22, cube=18.0)
    case("frame plain n=200")
    ids = gpu.register_soups(sc.soups)
    trees = [orc.tree(s, 4) for s in sc.soups]
    for frame in range(3):
        ents = sc.entities(ids, ENTITY_DTYPE)
        pairs, cons = gpu.frame(ents, MODE_PLAIN)
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
            assert g["s"].tobytes() == c["s"].tobytes() and g["t"].tobytes() == c["t"].tobytes()
            assert g["flags"] == c["flags"] and g["na"].tobytes() == c["na"].tobytes()
        # pairs without contacts are dropped, the rest are in (first, second) order
        keys = [(int(p["first"]), int(p["second"])) for p in pairs]
        assert keys == sorted(keys) and all(p["contact_count"] > 0 for p in pairs)
        assert len(want) > 0
        sc.step()


def test_frame_batched_matches_exact_oracle(gpu, orc):
    sc = _scene(200, 23, cube=18.0)
    case("frame batched n=200")
    ids = gpu.register_soups(sc.soups)
    trees = [orc.tree(s, 4) for s in sc.soups]
    ents = sc.entities(ids, ENTITY_DTYPE)
    pairs, cons = gpu.frame(ents, MODE_BATCHED)
    res = orc.frame(trees, sc.mesh_of, ents["current"], sc.should_callback, mode=1)
    want = {(int(res.pairs[c["query"]][0]), int(res.pairs[c["query"]][1]), int(c["tri_a"]), int(c["tri_b"]))
            for c in res.contacts}
    got = set()
    for p in pairs:
        for c in cons[p["contact_begin"]: p["contact_begin"] + p["contact_count"]]:
            got.add((int(p["first"]), int(p["second"]), int(c["tri_a"]), int(c["tri_b"])))
    assert got == want and len(want) > 0


def test_frame_batched_graph_replays(orc):
    """Frames through one captured graph: the pair count changes from frame
    to frame (the kernels read it from the device), every replay against the
    exact oracle, with eager frames in between."""
    from inmyroom_vulkan_amd import Collider

    sc = _scene(200, 24, cube=18.0)
    case("frame batched n=200 through graph replays")
    with Collider(leaf_max=4) as c:
        c.set_tuning(use_graph=1)
        ids = c.register_soups(sc.soups)
        trees = [orc.tree(s, 4) for s in sc.soups]
        counts = set()
        for frame in range(6):
            if frame == 3:
                c.set_tuning(use_graph=0)  # an eager frame between replays
            if frame == 4:
                c.set_tuning(use_graph=1)
            ents = sc.entities(ids, ENTITY_DTYPE)
            pairs, cons = c.frame(ents, MODE_BATCHED)
            res = orc.frame(trees, sc.mesh_of, ents["current"], sc.should_callback, mode=1)
            want = {(int(res.pairs[q["query"]][0]), int(res.pairs[q["query"]][1]), int(q["tri_a"]), int(q["tri_b"]))
                    for q in res.contacts}
            got = set()
            for p in pairs:
                for q in cons[p["contact_begin"]: p["contact_begin"] + p["contact_count"]]:
                    got.add((int(p["first"]), int(p["second"]), int(q["tri_a"]), int(q["tri_b"])))
            assert got == want and len(want) > 0, frame
            counts.add(c.counters()["broad_pairs"])
            for _ in range(3):
                sc.step()
        assert len(counts) > 1  # the replays saw different pair counts


def test_frame_many_contacts_sorted(gpu, orc):
    """The 512-entity scene: tens of thousands of contacts go through the
    multi-workgroup radix sort; they come back grouped by pair in (first,
    second) order and by (tri_a, tri_b) within a pair, and as a set equal
    the exact oracle's."""
    sc = scenes.Scene(n_entities=512, seed=2)
    case("frame batched, 512 entities")
    ids = gpu.register_soups(sc.soups)
    trees = [orc.tree(s, 4) for s in sc.soups]
    ents = sc.entities(ids, ENTITY_DTYPE)
    pairs, cons = gpu.frame(ents, MODE_BATCHED)
    assert len(cons) > 4096
    keys = [(int(p["first"]), int(p["second"])) for p in pairs]
    assert keys == sorted(keys) and len(set(keys)) == len(keys)
    begin = 0
    for p in pairs:
        assert p["contact_begin"] == begin
        c = cons[begin: begin + p["contact_count"]]
        tk = [(int(x["tri_a"]), int(x["tri_b"])) for x in c]
        assert tk == sorted(tk) and len(set(tk)) == len(tk)
        begin += int(p["contact_count"])
    assert begin == len(cons)
    res = orc.frame(trees, sc.mesh_of, ents["current"], sc.should_callback, mode=1)
    want = {(int(res.pairs[c["query"]][0]), int(res.pairs[c["query"]][1]), int(c["tri_a"]), int(c["tri_b"]))
            for c in res.contacts}
    got = {(int(p["first"]), int(p["second"]), int(x["tri_a"]), int(x["tri_b"]))
           for p in pairs for x in cons[p["contact_begin"]: p["contact_begin"] + p["contact_count"]]}
    assert got == want


def test_frame_small_inputs(gpu):
    case("frame with fewer than two entities")
    ids = gpu.register_soups([scenes.box_soup()])
    e = np.zeros(1, ENTITY_DTYPE)
    e["current"] = scenes.identity16()
    e["mesh_id"] = ids[0]
    pairs, cons = gpu.frame(e, MODE_PLAIN)
    assert len(pairs) == 0 and len(cons) == 0
    assert len(gpu.broad_phase(e[:0])) == 0


def _sparse_boxes(n, side, seed, mid):
    rng = np.random.default_rng(seed)
    ents = np.zeros(n, ENTITY_DTYPE)
    ents["current"] = ents["previous"] = scenes.pose_matrices(rng.uniform(0, side, (n, 3)),
                                                              scenes.random_quaternions(rng, n))
    ents["mesh_id"] = mid
    ents["entity"] = np.arange(n)
    ents["should_callback"] = 1
    return ents


def test_broad_phase_over_4096_entities_keeps_pair_order(gpu, orc):
    """4,500 unit boxes in a 100 m cube.  The pair sort's key (first << 16 |
    second) needs 29 bits above 4,096 entities; with at most 2,048 pairs it
    runs as the one-pass rank sort, whose digit mask must then cover all 32
    bits.  Pairs come out in (first, second) order, equal to the oracle's,
    and the same on a second run."""
    case("broad phase, 4,500 sparse boxes (29-bit pair keys)")
    box = scenes.box_soup()
    mid = gpu.register_soups([box])[0]
    tree = orc.tree(box, 4)
    ents = _sparse_boxes(4500, 100.0, 31, mid)
    got = [tuple(p) for p in gpu.broad_phase(ents).tolist()]
    res = orc.frame([tree], np.zeros(len(ents), np.uint32), ents["current"], ents["should_callback"], mode=0)
    want = [tuple(p) for p in res.pairs.tolist()]
    assert 16 < len(want) <= 2048, len(want)
    assert max(a for a, _ in want) >= 4096  # keys above 2^28 take part
    assert got == want
    assert [tuple(p) for p in gpu.broad_phase(ents).tolist()] == got


def test_frame_over_4096_entities_matches_oracle(gpu, orc):
    """The same sparse scene as one plain frame: pairs in (first, second)
    order, contacts bit-exact against the oracle."""
    case("frame plain, 4,500 sparse boxes")
    box = scenes.box_soup()
    mid = gpu.register_soups([box])[0]
    tree = orc.tree(box, 4)
    ents = _sparse_boxes(4500, 100.0, 32, mid)
    pairs, cons = gpu.frame(ents, MODE_PLAIN)
    res = orc.frame([tree], np.zeros(len(ents), np.uint32), ents["current"], ents["should_callback"], mode=0)
    keys = [(int(p["first"]), int(p["second"])) for p in pairs]
    assert keys == sorted(keys) and len(keys) > 0
    want = {}
    for c in res.contacts:
        a, b = res.pairs[c["query"]]
        want[(int(a), int(b), int(c["tri_a"]), int(c["tri_b"]))] = c
    got = {}
    for p in pairs:
        for c in cons[p["contact_begin"]: p["contact_begin"] + p["contact_count"]]:
            got[(int(p["first"]), int(p["second"]), int(c["tri_a"]), int(c["tri_b"]))] = c
    assert set(got) == set(want)
    for k, c in want.items():
        assert got[k]["s"].tobytes() == c["s"].tobytes() and got[k]["t"].tobytes() == c["t"].tobytes()


def test_broad_phase_ends_the_last_run(gpu):
    """cd_broad_phase overwrites the relative matrices and entity copy the
    last frame's rays and response would read: after it, both refuse."""
    from inmyroom_vulkan_amd import CollideError

    case("rays and response refused after a standalone broad phase")
    sc = _scene(120, 25, cube=12.0)
    ids = gpu.register_soups(sc.soups)
    ents = sc.entities(ids, ENTITY_DTYPE)
    pairs, cons = gpu.frame(ents, MODE_PLAIN)
    assert len(cons) > 0
    gpu.uncollide_rays()  # valid straight after the frame
    sc.step()
    gpu.broad_phase(sc.entities(ids, ENTITY_DTYPE))
    for call in (gpu.uncollide_rays, gpu.shoot_rays):
        with pytest.raises(CollideError):
            call()
