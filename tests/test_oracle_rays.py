"""The oracle's ray generation (CreateUncollideRays.cpp:60-200), pinned by
hand-computed cases and invariants.

The reference has no tests for this step and glm is absent, so the pins are
closed forms: crossing triangles whose contact segment, not-outside
vertices, fallback point and averages follow from the geometry, plus the
vertex de-duplication and the invariants of rays on a sphere.
"""
import numpy as np
import pytest

from inmyroom_vulkan_amd import scenes

I16 = scenes.identity16()


def _one(orc, tris_a, tris_b, vid_a=None, vid_b=None, nrm_a=None, nrm_b=None, m16=I16):
    ta = orc.tree(np.asarray(tris_a, np.float32).reshape(-1, 9), 4)
    tb = orc.tree(np.asarray(tris_b, np.float32).reshape(-1, 9), 4)
    ta.set_attrs(nrm_a, vid_a)
    tb.set_attrs(nrm_b, vid_b)
    r = orc.rays_queries([ta, tb], np.array([0], np.uint32), np.array([1], np.uint32), np.asarray(m16)[None])
    return r


# A in z = 0 (normal +z); B in x = 0 (normal +x).  A's vertex 0 (x = -1) is the
# only one not outside B's plane; B's vertices 0 and 1 (z = -1) are not
# outside A's plane.  The contact segment is x = 0, z = 0, y in [-1/2, -1/6].
A1 = [[-1, -0.5, 0], [2, -0.5, 0], [2, 0.5, 0]]
B1 = [[0, -1, -1], [0, 1, -1], [0, 0, 1]]


def test_rays_crossing_triangles(orc):
    r = _one(orc, [A1], [B1])
    assert r.pairs[0, 2] == 1 and r.pairs[0, 3] == 2
    fa, fb = r.split(0)
    np.testing.assert_array_equal(fa[0, 0], np.float32(A1[0]))
    np.testing.assert_array_equal(fa[0, 1], [0, 0, -1])  # minus A's face normal
    got_b = sorted(map(tuple, fb[:, 0].tolist()))
    assert got_b == [(0.0, -1.0, -1.0), (0.0, 1.0, -1.0)]
    assert np.all(fb[:, 1] == np.float32([-1, 0, 0]))
    np.testing.assert_allclose(r.avg[0, 0], A1[0], rtol=0, atol=0)
    np.testing.assert_allclose(r.avg[0, 1], [0, 0, -1], rtol=0, atol=1e-7)


def test_rays_second_side_transformed(orc):
    """The second entity's rays are in the first's space; its average goes back to its own space."""
    rng = np.random.default_rng(3)
    q = scenes.random_quaternions(rng, 1)
    M = scenes.pose_matrices(np.array([[0.25, -0.5, 0.125]]), q)[0]
    Mi = np.linalg.inv(M.reshape(4, 4).T.astype(np.float64))
    # B given in its own space such that M * B = B1
    Bown = [(Mi @ np.array(p + [1.0]))[:3] for p in B1]
    r = _one(orc, [A1], [Bown], m16=M)
    fa, fb = r.split(0)
    assert len(fa) == 1 and len(fb) == 2
    np.testing.assert_allclose(np.sort(fb[:, 0], axis=0), np.sort(np.float32(B1[:2]), axis=0), atol=1e-5)
    np.testing.assert_allclose(fb[:, 1], np.tile([-1, 0, 0], (2, 1)), atol=1e-5)
    own_avg = np.mean(np.float64(Bown[:2]), axis=0)
    np.testing.assert_allclose(r.avg[0, 1], own_avg, atol=1e-5)


def test_rays_fallback_to_weighted_point(orc):
    """A crosses B1 (x = 0, normal +x) and B2 (x = 1, normal -x): vertex 0 is outside
    B2's plane and vertices 1, 2 outside B1's, so no vertex is inside and one ray
    starts at the weighted mean of the two segment midpoints."""
    B2 = [[1, 1, -1], [1, -1, -1], [1, 0, 1]]  # normal -x
    r = _one(orc, [A1], [B1, B2])
    fa, fb = r.split(0)
    assert len(fa) == 1
    # segments: x = 0, y in [-1/2, -1/6] (length 1/3); x = 1, y in [-1/2, 1/6] (length 2/3)
    p1 = np.array([0.0, (-0.5 - 1 / 6) / 2, 0.0])
    p2 = np.array([1.0, (-0.5 + 1 / 6) / 2, 0.0])
    w1, w2 = 1 / 3, 2 / 3
    want = (w1 * p1 + w2 * p2) / (w1 + w2)
    np.testing.assert_allclose(fa[0, 0], want, atol=1e-6)
    np.testing.assert_allclose(fa[0, 1], [0, 0, -1], atol=1e-7)  # barycentric normal of a flat triangle
    # B1 keeps its vertices 0, 1; B2's vertices 0, 1 (z = -1) are not outside A's plane either
    assert len(fb) == 4


def test_rays_deduplicate_shared_vertices(orc):
    """Two A triangles sharing vertex 0 both keep it inside: one ray for it."""
    A2 = [[-1, -0.5, 0], [2, 0.5, 0], [-0.5, 0.5, 0]]  # shares A1's vertex 0
    vid = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)  # corner 0 is vertex 0 in both
    r = _one(orc, [A1, A2], [B1], vid_a=vid)
    fa, _ = r.split(0)
    origins = [tuple(o) for o in fa[:, 0].tolist()]
    assert origins.count((-1.0, -0.5, 0.0)) == 1
    # without shared ids (a soup) the vertex appears once per triangle
    r2 = _one(orc, [A1, A2], [B1])
    fa2, _ = r2.split(0)
    assert [tuple(o) for o in fa2[:, 0].tolist()].count((-1.0, -0.5, 0.0)) == 2


def test_rays_zero_weight_leaf_is_skipped(orc):
    """Triangles touching at a single point give a zero-length segment: the
    candidate is null (IsNull) and produces no ray."""
    A = [[0, 0, 0], [1, 0, 0], [0, 1, 0]]
    B = [[0, 0, 0], [-1, 0, 0.5], [0, -1, -0.5]]  # meets A only at the origin
    flags, seg = orc.tritri(np.float32(A).reshape(1, 9), np.float32(B).reshape(1, 9))
    assert flags[0] == 1 and not seg.any()  # a contact, of zero length
    r = _one(orc, [A], [B])
    assert r.pairs[0, 2] == 0 and r.pairs[0, 3] == 0


def test_rays_sphere_invariants(orc):
    soup, mats = scenes.config1(60, seed=1, freq=6)
    P, F = scenes.icosphere(6)
    P = P.astype(np.float32)
    t = orc.tree(P[F].reshape(-1, 9), 4)
    vn = (P / np.linalg.norm(P, axis=1, keepdims=True)).astype(np.float32)
    t.set_attrs(vn[F].reshape(-1, 9), F)
    r = orc.rays_queries([t], np.zeros(60, np.uint32), np.zeros(60, np.uint32), mats)
    assert r.pairs[:, 2].sum() > 50
    for i in range(60):
        fa, fb = r.split(i)
        assert (len(fa) > 0) == (len(fb) > 0)
        if not len(fa):
            continue
        np.testing.assert_allclose(np.linalg.norm(fa[:, 1], axis=1), 1.0, atol=1e-6)
        # vertex rays of A start on the unit sphere and point inward along the radius
        on = np.abs(np.linalg.norm(fa[:, 0], axis=1) - 1.0) < 1e-6
        np.testing.assert_allclose(fa[on, 1], -fa[on, 0], atol=1e-5)
        # each vertex at most once per side
        assert len({tuple(o) for o in fa[on, 0].tolist()}) == int(on.sum())
        np.testing.assert_allclose(r.avg[i, 0], fa[:, 0].astype(np.float64).mean(axis=0), atol=1e-5)
