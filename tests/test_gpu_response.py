"""Response on the GPU (cd_shoot_rays) against the oracle
(ShootUncollideRays.cpp:14-179, Ray.cpp:13-236, CollisionDetection.cpp:70-99).

Each ray's passes follow the reference's operations; the pair's force sum
runs over its rays in the order they are listed, which for the reference
follows an unordered_map walk (CreateUncollideRays.cpp:138) that neither the
GPU nor the oracle reproduces, so pair deltas are compared within 1e-5
relative.  Which pairs are shot, and whose share is zero, must match exactly.
"""
import math

import numpy as np
import pytest

from conftest import case
from inmyroom_vulkan_amd import ENTITY_DTYPE, MODE_BATCHED, MODE_PLAIN, MODE_REFERENCE, CollideError, Collider, scenes

pytestmark = pytest.mark.gpu

DEG = np.float32(0.017453292519943295)
COS_START = np.float32(math.cos(np.float32(40 * DEG)))
COS_FINISH = np.float32(math.cos(np.float32(65 * DEG)))


def _tmat(x, y=0.0, z=0.0):
    m = np.eye(4, dtype=np.float32)
    m[3, :3] = (x, y, z)
    return m.reshape(-1)


def _compare(got, ids, d):
    assert len(got) == len(ids)
    want = {(int(i[0]), int(i[1])): (int(i[2]), d[k]) for k, i in enumerate(ids)}
    for r in got:
        shot, dd = want[(int(r["first"]), int(r["second"]))]
        assert int(r["shot"]) == shot
        for j, name in enumerate(("delta", "delta_first", "delta_second")):
            scale = max(1e-6, float(np.abs(dd[0]).max()))
            np.testing.assert_allclose(r[name], dd[j], rtol=1e-5, atol=1e-5 * scale, err_msg=name)


def test_response_two_spheres(gpu, orc):
    case("response, two spheres")
    P, F = scenes.icosphere(8)
    P = P.astype(np.float32)
    vn = (P / np.linalg.norm(P, axis=1, keepdims=True)).astype(np.float32)
    mid = gpu.register_mesh(P, F)
    gpu.set_mesh_normals(mid, vn)
    t = orc.tree(P[F].reshape(-1, 9), 4)
    t.set_attrs(vn[F].reshape(-1, 9), F)
    I = scenes.identity16()
    for dist, prev2 in ((1.6, 2.1), (1.75, 1.75), (1.3, 1.1)):
        e = np.zeros(2, ENTITY_DTYPE)
        e["current"] = [I, _tmat(dist, 0.1, -0.05)]
        e["previous"] = [_tmat(-0.2), _tmat(prev2, 0.1, -0.05)]
        e["mesh_id"] = mid
        e["should_callback"] = 1
        gpu.frame(e, MODE_PLAIN)
        gpu.uncollide_rays()
        got = gpu.shoot_rays()
        ids, d, _ = orc.shoot_frame([t], [0, 0], e["current"], e["previous"], e["should_callback"], COS_START,
                                    COS_FINISH, np.float32(1.01))
        _compare(got, ids, d)
        assert got[0]["shot"] == 1
        n = np.linalg.norm(got[0]["delta"])
        assert got[0]["delta"][0] / n > 0.98  # the second is pushed away along +x


def test_response_scene_frames(gpu, orc):
    sc = scenes.Scene(n_entities=96, seed=2, cube=12.0)
    case("response, 96-entity scene")
    ids = gpu.register_soups(sc.soups)
    trees = [orc.tree(s, 4) for s in sc.soups]
    total = shot = 0
    for frame in range(3):
        ents = sc.entities(ids, ENTITY_DTYPE)
        gpu.frame(ents, MODE_PLAIN)
        gpu.uncollide_rays()
        got = gpu.shoot_rays()
        want_ids, want_d, _ = orc.shoot_frame(trees, sc.mesh_of, ents["current"], ents["previous"],
                                              ents["should_callback"], COS_START, COS_FINISH, np.float32(1.01))
        _compare(got, want_ids, want_d)
        total += len(got)
        shot += int(got["shot"].sum())
        sc.step()
    assert total > 10 and shot > 0


def test_response_batched_mode_invariants(gpu):
    """After an exact-semantics frame the response has the same shape: a
    push apart along the centre line, shared by the entity that moved."""
    case("response after a batched frame")
    P, F = scenes.icosphere(8)
    P = P.astype(np.float32)
    mid = gpu.register_mesh(P, F)
    gpu.set_mesh_normals(mid, (P / np.linalg.norm(P, axis=1, keepdims=True)).astype(np.float32))
    e = np.zeros(2, ENTITY_DTYPE)
    e["current"] = [scenes.identity16(), _tmat(1.6)]
    e["previous"] = [scenes.identity16(), _tmat(2.2)]
    e["mesh_id"] = mid
    e["should_callback"] = 1
    gpu.frame(e, MODE_BATCHED)
    gpu.uncollide_rays()
    r = gpu.shoot_rays()
    assert len(r) == 1 and r[0]["shot"] == 1
    np.testing.assert_array_equal(r[0]["delta_first"], 0.0)
    np.testing.assert_array_equal(r[0]["delta_second"], r[0]["delta"])
    assert 0.2 < r[0]["delta"][0] < 0.55


def test_shoot_needs_frame_and_rays(gpu):
    c = Collider(leaf_max=4)
    try:
        with pytest.raises(CollideError):
            c.shoot_rays()  # no frame yet
        mid = c.register_soups([scenes.box_soup()])[0]
        e = np.zeros(2, ENTITY_DTYPE)
        e["current"] = [scenes.identity16(), _tmat(0.5)]
        e["previous"] = e["current"]
        e["mesh_id"] = mid
        e["should_callback"] = 1
        c.frame(e, MODE_PLAIN)
        with pytest.raises(CollideError):
            c.shoot_rays()  # rays not made yet
        c.uncollide_rays()
        r = c.shoot_rays()
        assert len(r) >= 0 and (len(r) == 0 or r["shot"].sum() == 0)  # nothing moved
    finally:
        c.close()


@pytest.mark.parametrize("mode,step_graph", [(MODE_PLAIN, 0), (MODE_REFERENCE, 0), (MODE_BATCHED, 0),
                                             (MODE_REFERENCE, 2), (MODE_BATCHED, 2)])
def test_step_equals_frame_rays_and_shoot(mode, step_graph):
    """cd_step (contacts and rays kept on the device; for the batched modes
    one graph per step, or with step_graph 2 the same kernels launched one by
    one) gives the records the three calls give, bit for bit, frame after
    frame, and leaves the state the three calls leave: rays and shots taken
    after it give the same records again.  Its first step outgrows the
    initial pair, contact and ray capacities of the on-device step and runs
    again with larger ones."""
    sc = scenes.Scene(n_entities=96, seed=2, cube=12.0)
    case(f"cd_step vs cd_frame + cd_uncollide_rays + cd_shoot_rays, mode {mode}, step_graph {step_graph}")
    gpu = Collider(leaf_max=4)  # a fresh context: its step starts at the initial capacities
    ids = gpu.register_soups(sc.soups)
    gpu.set_tuning(step_graph=step_graph)
    try:
        shot = 0
        for frame in range(3):
            ents = sc.entities(ids, ENTITY_DTYPE)
            gpu.frame(ents, mode)
            pr, rays = gpu.uncollide_rays()
            want = gpu.shoot_rays()
            got = gpu.step(ents, mode)
            assert got.tobytes() == want.tobytes(), frame
            assert gpu.shoot_rays().tobytes() == want.tobytes(), frame
            pr2, rays2 = gpu.uncollide_rays()
            assert pr2.tobytes() == pr.tobytes() and rays2.tobytes() == rays.tobytes(), frame
            shot += int(got["shot"].sum())
            sc.step()
        assert shot > 0
        e = sc.entities(ids, ENTITY_DTYPE)[:1]
        assert len(gpu.step(e, mode)) == 0  # one entity: no pairs
        far = sc.entities(ids, ENTITY_DTYPE)[:2]
        far["current"][1][12] += 1e4
        assert len(gpu.step(far, mode)) == 0  # two entities far apart
    finally:
        gpu.close()


@pytest.mark.parametrize("step_graph", [0, 2])
def test_step_pair_beyond_per_pair_order(step_graph):
    """An icosphere inside a copy of itself with seeded radial noise: the two
    surfaces cross all over, one pair with more contacts than the step's
    per-pair contact order ranks (2,048).  The step runs again with the radix
    sorts and still gives the three calls' records, bit for bit."""
    case(f"cd_step, one pair beyond the per-pair contact order, step_graph {step_graph}")
    P, F = scenes.icosphere(12)
    rng = np.random.default_rng(7)
    Q = P * (1.0 + rng.uniform(-0.02, 0.02, (len(P), 1)))
    gpu = Collider(leaf_max=4)
    try:
        gpu.set_tuning(step_graph=step_graph)
        a = gpu.register_mesh(P.astype(np.float32), F)
        b = gpu.register_mesh(Q.astype(np.float32), F)
        e = np.zeros(2, ENTITY_DTYPE)
        e["current"] = [scenes.identity16(), scenes.identity16()]
        e["previous"] = [_tmat(-0.2), scenes.identity16()]
        e["mesh_id"] = [a, b]
        e["should_callback"] = 1
        for mode in (MODE_REFERENCE, MODE_BATCHED):
            pairs, _ = gpu.frame(e, mode)
            assert len(pairs) == 1 and int(pairs[0]["contact_count"]) > 2048, pairs
            gpu.uncollide_rays()
            want = gpu.shoot_rays()
            got = gpu.step(e, mode)
            assert len(got) == 1 and got.tobytes() == want.tobytes(), mode
    finally:
        gpu.close()


@pytest.mark.parametrize("mode", [MODE_REFERENCE, MODE_BATCHED])
def test_step_crowded_scene_grows_its_capacities(mode):
    """40 entities packed into a 4 m cube: far more pairs than the step's
    first pair capacity (the entity count), and many contacts and rays.  The
    step runs again with larger capacities and gives the three calls'
    records, bit for bit; the next frame replays the grown step."""
    sc = scenes.Scene(n_entities=40, seed=5, cube=4.0, torus_grid=(48, 12))
    case(f"cd_step, crowded scene, mode {mode}")
    gpu = Collider(leaf_max=4)
    try:
        ids = gpu.register_soups(sc.soups)
        for frame in range(2):
            ents = sc.entities(ids, ENTITY_DTYPE)
            pairs, _ = gpu.frame(ents, mode)
            gpu.uncollide_rays()
            want = gpu.shoot_rays()
            got = gpu.step(ents, mode)
            assert got.tobytes() == want.tobytes(), frame
            assert gpu.counters()["broad_pairs"] > 64, frame
            sc.step()
    finally:
        gpu.close()
