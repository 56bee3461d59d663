"""Trees deeper than the LDS stacks hold, on the GPU against the oracle.

A unit icosphere (1,280 triangles) with a chain of 50 thin triangles whose
distances double along x: every centre split of a box holding the chain
peels its far end off, so the sphere's leaves sit some 60 levels down
(the oracle's tree, built on the CPU, says how deep).  A pair of such meshes
needs about 122 stack entries in the plain walk (its LDS stack holds 64) and
a ray through one needs about 62 (its LDS stack holds 48): both walks then
keep their stacks in HBM.  Results must be the oracle's: contacts and counts
bit for bit in the plain and reference modes, the exact pair set in the
batched mode, and response deltas within 1e-5."""
import math

import numpy as np
import pytest

from conftest import case
from contacts_check import assert_plain_bit_exact, sort_contacts
from inmyroom_vulkan_amd import ENTITY_DTYPE, MODE_BATCHED, MODE_PLAIN, MODE_REFERENCE, Collider, scenes

pytestmark = pytest.mark.gpu

DEG = np.float32(0.017453292519943295)
COS_START = np.float32(math.cos(np.float32(40 * DEG)))
COS_FINISH = np.float32(math.cos(np.float32(65 * DEG)))


def deep_soup(n_chain=50):
    sphere = scenes.unit_icosphere_soup(8)
    x = 2.0 * 2.0 ** np.arange(n_chain)
    chain = np.zeros((n_chain, 9), np.float32)
    chain[:, 0] = x
    chain[:, 3], chain[:, 4] = 1.1 * x, 1e-3
    chain[:, 6], chain[:, 8] = 1.05 * x, 1e-3
    return np.concatenate([sphere.astype(np.float32), chain])


def _rel(t, axis, angle):
    q = np.r_[np.sin(angle / 2) * np.asarray(axis) / np.linalg.norm(axis), np.cos(angle / 2)]
    return scenes.pose_matrices(np.asarray([t], float), q[None]).astype(np.float32)[0]


@pytest.fixture(scope="module")
def deep(orc):
    soup = deep_soup()
    tree = orc.tree(soup, 4)
    gpu = Collider(leaf_max=4)
    mid = gpu.register_soups([soup])[0]
    yield soup, tree, gpu, mid
    gpu.close()


def test_deep_tree_is_deep_and_matches_the_oracle(deep):
    soup, tree, gpu, mid = deep
    case(f"deep tree: depth {tree.depth}, {len(soup)} triangles")
    assert tree.depth >= 60
    assert gpu.mesh_info(mid)["depth"] == tree.depth
    assert 2 * tree.depth + 2 > 64 and tree.depth + 2 > 48  # past both LDS stacks


def test_deep_tree_queries_all_modes(deep, orc):
    soup, tree, gpu, mid = deep
    rel = np.stack([_rel((1.3, 0.2, -0.1), (0, 0, 1), 0.3), _rel((0.9, -0.4, 0.3), (1, 1, 0), 1.1),
                    _rel((0.2, 1.5, 0.0), (0, 1, 0), 2.0), _rel((0.0, 0.0, 0.0), (1, 0, 0), 0.7)])
    case("deep tree: 4 self queries in plain, reference and batched modes")
    q = gpu.make_queries(mid, mid, rel)
    z = np.zeros(len(rel), np.uint32)
    want = orc.batch([tree], z, z, rel, mode=0)
    plain = sort_contacts(gpu.query_batch(q, MODE_PLAIN))
    pc = gpu.counters()
    assert_plain_bit_exact(plain, sort_contacts(want.contacts))
    assert (pc["sat_tests"], pc["leaf_pairs"], pc["tri_tests"]) == (want.total_sat, want.total_leafpairs,
                                                                     want.total_tritests)
    ref = sort_contacts(gpu.query_batch(q, MODE_REFERENCE))
    assert_plain_bit_exact(ref, plain)
    exact = orc.batch([tree], z, z, rel, mode=1)
    got = {(int(c["query"]), int(c["tri_a"]), int(c["tri_b"])) for c in gpu.query_batch(q, MODE_BATCHED)}
    assert got == {(int(c["query"]), int(c["tri_a"]), int(c["tri_b"])) for c in exact.contacts}
    assert len(plain) > 100


def test_deep_tree_frame_rays_and_response(deep, orc):
    soup, tree, gpu, mid = deep
    case("deep tree: frame, uncollide rays and response of two overlapping entities")
    e = np.zeros(2, ENTITY_DTYPE)
    e["current"] = [scenes.identity16(), _rel((1.6, 0.1, -0.05), (0, 0, 1), 0.2).reshape(-1)]
    e["previous"] = [scenes.identity16(), _rel((2.1, 0.1, -0.05), (0, 0, 1), 0.2).reshape(-1)]
    e["mesh_id"] = mid
    e["should_callback"] = 1
    for mode in (MODE_PLAIN, MODE_REFERENCE):
        pairs, cons = gpu.frame(e, mode)
        assert len(pairs) == 1 and len(cons) > 10
        _, rays = gpu.uncollide_rays()
        assert len(rays) > 10
        got = gpu.shoot_rays()
        ids, d, _ = orc.shoot_frame([tree], [0, 0], e["current"], e["previous"], e["should_callback"], COS_START,
                                    COS_FINISH, np.float32(1.01))
        assert len(got) == len(ids) == 1 and int(got[0]["shot"]) == int(ids[0, 2]) == 1
        for j, name in enumerate(("delta", "delta_first", "delta_second")):
            scale = max(1e-6, float(np.abs(d[0, 0]).max()))
            np.testing.assert_allclose(got[0][name], d[0, j], rtol=1e-5, atol=1e-5 * scale, err_msg=name)
