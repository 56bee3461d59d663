"""Pins of the oracle (CPU restatement of the reference collision path).

Each function of the oracle is checked against something independent of
it: the reference's own code compiled from /root/reference (live when the
reference is mounted, through the golden vectors in tests/golden otherwise),
closed forms, exact rational arithmetic, or brute force.
"""
from fractions import Fraction

import numpy as np
import pytest

from conftest import golden
from inmyroom_vulkan_amd import scenes
from oracle import RefPieces

F32_EPS = np.float32(np.finfo(np.float32).eps)


# ---------------------------------------------------------------- eig3 ----
def test_eig3_matches_reference_golden(orc):
    g = golden("eig3_ref.npz")
    for A, V, d in zip(g["A"], g["V"], g["d"]):
        V1, d1 = orc.eig3(A)
        assert V1.tobytes() == V.tobytes() and d1.tobytes() == d.tobytes()


@pytest.mark.skipif(not RefPieces.available(), reason="reference pieces not built (needs /root/reference)")
def test_eig3_matches_reference_live(orc):
    ref = RefPieces()
    rng = np.random.default_rng(5)
    for _ in range(3000):
        a = rng.normal(size=(3, 3)) * rng.uniform(1e-3, 1e3)
        a = a + a.T
        assert orc.eig3(a)[0].tobytes() == ref.eig3(a)[0].tobytes()


def test_eig3_closed_forms(orc):
    V, d = orc.eig3(np.diag([3.0, 1.0, 2.0]))
    assert list(d) == [1.0, 2.0, 3.0]  # ascending
    assert np.array_equal(np.abs(V), np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0]], float))  # eigenvectors in columns
    rng = np.random.default_rng(1)
    for _ in range(200):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        lam = np.sort(rng.normal(size=3))
        A = q @ np.diag(lam) @ q.T
        V, d = orc.eig3(A)
        assert np.allclose(d, lam, atol=1e-12)
        assert np.allclose(V.T @ V, np.eye(3), atol=1e-12)
        assert np.allclose(A @ V, V @ np.diag(d), atol=1e-12)


# ------------------------------------------------------------- Möller ----
def test_moller_matches_reference_golden(orc):
    g = golden("moller_ref.npz")
    f, s = orc.tritri(g["a9"], g["b9"])
    ok = ~g["ub"]
    assert np.array_equal(f[ok] & 3, g["flags"][ok])
    assert np.array_equal(((f & 4) != 0), g["ub"])
    keep = ok & ((g["flags"] & 3) == 1)
    assert np.array_equal(s[keep].view(np.uint32), g["seg"][keep].view(np.uint32))


@pytest.mark.skipif(not RefPieces.available(), reason="reference pieces not built (needs /root/reference)")
def test_moller_matches_reference_live(orc):
    ref = RefPieces()
    a9, b9, _ = scenes.config4(30000, seed=21)
    ua, ub_, _ = scenes.config4(10000, seed=22, span=1.0)
    a9, b9 = np.concatenate([a9, ua]), np.concatenate([b9, ub_])
    f1, s1 = orc.tritri(a9, b9)
    f2, s2 = ref.tritri(a9, b9)
    ok = (f1 & 4) == 0
    assert np.array_equal(f1[ok] & 3, f2[ok])
    keep = ok & ((f2 & 3) == 1)
    assert np.array_equal(s1[keep].view(np.uint32), s2[keep].view(np.uint32))


def test_moller_closed_forms(orc):
    T0 = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], np.float32).reshape(1, 9)
    cross_ = np.array([[0, 0, -1], [0, 0, 1], [0, 2, 0]], np.float32).reshape(1, 9)
    f, s = orc.tritri(T0, cross_)
    assert f[0] == 1
    ends = {tuple(s[0, :3]), tuple(s[0, 3:])}
    assert ends == {(0.0, 0.0, 0.0), (0.0, 1.0, 0.0)}  # T0 ∩ {x=0} = y in [-1,1]; T1 ∩ {z=0} = y in [0,1]
    apart = cross_.copy().reshape(3, 3)
    apart[:, 0] = 2.0
    assert orc.tritri(T0, apart.reshape(1, 9))[0][0] == 0
    # coplanar: overlapping -> hit + coplanar; disjoint -> coplanar only
    inner = np.array([[-0.2, -0.5, 0], [0.2, -0.5, 0], [0, 0.2, 0]], np.float32).reshape(1, 9)
    far = inner.reshape(3, 3).copy()
    far[:, 0] += 5
    assert orc.tritri(T0, inner)[0][0] == 3
    assert orc.tritri(T0, far.reshape(1, 9))[0][0] == 2
    # the 1e-6 snap acts on N1.U + d1 with |N1| = 4 here: a vertex 2e-7 above
    # the plane (distance 8e-7) counts as on it, one 5e-7 above (2e-6) does not
    touch = np.array([[0, 0, 2e-7], [0, 0, 1], [0, 1, 1]], np.float32).reshape(1, 9)
    lift = np.array([[0, 0, 5e-7], [0, 0, 1], [0, 1, 1]], np.float32).reshape(1, 9)
    assert orc.tritri(T0, touch)[0][0] & 1 == 1
    assert orc.tritri(T0, lift)[0][0] & 1 == 0


# ------------------------------------------------------ exact predicates ----
def _orient_fraction(a, b, c, d):
    A = [Fraction(float(a[i])) - Fraction(float(c[i])) for i in range(3)]
    B = [Fraction(float(b[i])) - Fraction(float(c[i])) for i in range(3)]
    D = [Fraction(float(d[i])) - Fraction(float(c[i])) for i in range(3)]
    cx = A[1] * B[2] - B[1] * A[2]
    cy = A[2] * B[0] - B[2] * A[0]
    cz = A[0] * B[1] - B[0] * A[1]
    v = cx * D[0] + cy * D[1] + cz * D[2]
    return (v > 0) - (v < 0)


def _near_degenerate_quads(rng, n):
    out = []
    for _ in range(n):
        a, b, c = rng.uniform(-1e3, 1e3, size=(3, 3)).astype(np.float32)
        s, t = rng.uniform(-1, 2, size=2)
        d = (a + s * (b - a) + t * (c - a)).astype(np.float32)  # on the plane up to rounding
        k = rng.integers(0, 3)
        if k == 1:
            d = np.nextafter(d, np.float32(np.inf))
        elif k == 2:
            d = a.copy()
        out.append((a, b, c, d))
    return out


def test_orient_exact_vs_fractions(orc):
    rng = np.random.default_rng(3)
    quads = _near_degenerate_quads(rng, 600)
    quads += [tuple(rng.uniform(-1, 1, size=(4, 3)).astype(np.float32)) for _ in range(200)]
    tiny = [tuple((rng.uniform(-1, 1, size=(4, 3)) * 10.0 ** rng.uniform(-40, 30, size=(4, 1))).astype(np.float32))
            for _ in range(200)]
    for a, b, c, d in quads + tiny:
        want = _orient_fraction(a, b, c, d)
        assert orc.orient_exact(a, b, c, d) == want
        assert orc.orient_exact(a, b, c, d, bignum_only=True) == want


def _seg_tri_fraction(s, e, a, b, c):
    """Closed segment vs closed triangle, exact (Fractions)."""
    F = lambda v: [Fraction(float(x)) for x in v]
    s, e, a, b, c = F(s), F(e), F(a), F(b), F(c)
    sub = lambda p, q: [p[i] - q[i] for i in range(3)]
    crs = lambda p, q: [p[1] * q[2] - p[2] * q[1], p[2] * q[0] - p[0] * q[2], p[0] * q[1] - p[1] * q[0]]
    dt = lambda p, q: p[0] * q[0] + p[1] * q[1] + p[2] * q[2]
    n = crs(sub(b, a), sub(c, a))
    os_, oe = dt(n, sub(s, a)), dt(n, sub(e, a))
    if (os_ > 0 and oe > 0) or (os_ < 0 and oe < 0):
        return False

    def inside(p):
        return (dt(crs(sub(b, a), sub(p, a)), n) >= 0 and dt(crs(sub(c, b), sub(p, b)), n) >= 0
                and dt(crs(sub(a, c), sub(p, c)), n) >= 0)

    if os_ == 0 and oe == 0:  # segment in the plane: endpoint inside or crossing an edge
        if inside(s) or inside(e):
            return True

        def seg_seg(p1, p2, q1, q2):
            d1 = dt(crs(sub(p2, p1), sub(q1, p1)), n)
            d2 = dt(crs(sub(p2, p1), sub(q2, p1)), n)
            d3 = dt(crs(sub(q2, q1), sub(p1, q1)), n)
            d4 = dt(crs(sub(q2, q1), sub(p2, q1)), n)
            if ((d1 > 0 and d2 < 0) or (d1 < 0 and d2 > 0)) and ((d3 > 0 and d4 < 0) or (d3 < 0 and d4 > 0)):
                return True
            return False

        return seg_seg(s, e, a, b) or seg_seg(s, e, b, c) or seg_seg(s, e, c, a)
    t = os_ / (os_ - oe)
    p = [s[i] + t * (e[i] - s[i]) for i in range(3)]
    return inside(p)


def _tri_tri_fraction(A, B):
    """Exact closed-triangle intersection for non-coplanar pairs: some edge of one meets the other."""
    A = A.reshape(3, 3)
    B = B.reshape(3, 3)
    for (P, Q) in ((A, B), (B, A)):
        for i in range(3):
            if _seg_tri_fraction(P[i], P[(i + 1) % 3], Q[0], Q[1], Q[2]):
                return True
    return False


def test_tritri_exact_vs_bruteforce(orc):
    rng = np.random.default_rng(8)
    a9, b9, kind = scenes.config4(1500, seed=9, span=1.0)
    ex = orc.tritri_exact(a9, b9)
    for i in range(len(a9)):
        if ex[i] == 2:
            # coplanar: all of one triangle's vertices on the other's plane, exactly
            A, B = a9[i].reshape(3, 3), b9[i].reshape(3, 3)
            on = all(_orient_fraction(A[0], A[1], A[2], B[k]) == 0 for k in range(3)) or \
                all(_orient_fraction(B[0], B[1], B[2], A[k]) == 0 for k in range(3))
            assert on
            continue
        assert bool(ex[i]) == _tri_tri_fraction(a9[i], b9[i]), (i, kind[i])
    assert (ex == 1).sum() > 100 and (ex == 0).sum() > 100
    del rng


def _segment_fraction(a9, b9):
    """Exact intersection segment of two triangles by clipping the planes'
    common line against both triangles (Cyrus-Beck), in Fractions: the line
    P(t) = P0 + t D, D = n1 x n2, P0 on both planes with D . P0 = 0; each
    edge keeps the side of its triangle's third vertex."""
    F = Fraction
    A = [[F(float(x)) for x in a9[3 * k:3 * k + 3]] for k in range(3)]
    B = [[F(float(x)) for x in b9[3 * k:3 * k + 3]] for k in range(3)]
    sub = lambda p, q: [p[i] - q[i] for i in range(3)]
    dot = lambda p, q: p[0] * q[0] + p[1] * q[1] + p[2] * q[2]
    crs = lambda p, q: [p[1] * q[2] - p[2] * q[1], p[2] * q[0] - p[0] * q[2], p[0] * q[1] - p[1] * q[0]]
    n1, n2 = crs(sub(A[1], A[0]), sub(A[2], A[0])), crs(sub(B[1], B[0]), sub(B[2], B[0]))
    D = crs(n1, n2)
    rows, rhs = [n1, n2, D], [dot(n1, A[0]), dot(n2, B[0]), F(0)]
    det3 = lambda m: dot(m[0], crs(m[1], m[2]))
    den = det3(rows)
    cols = [[rows[r][c] for r in range(3)] for c in range(3)]
    P0 = []
    for c in range(3):
        mc = [list(col) for col in cols]
        mc[c] = rhs
        P0.append(det3([[mc[cc][r] for cc in range(3)] for r in range(3)]) / den)
    lo, hi = None, None
    for T, n in ((A, n1), (B, n2)):
        for i in range(3):
            j, k = (i + 1) % 3, (i + 2) % 3
            m = crs(n, sub(T[j], T[i]))
            if dot(m, sub(T[k], T[i])) < 0:
                m = [-x for x in m]
            a, b = dot(m, D), -dot(m, sub(P0, T[i]))  # need a t >= b
            if a > 0:
                lo = b / a if lo is None else max(lo, b / a)
            elif a < 0:
                hi = b / a if hi is None else min(hi, b / a)
    pt = lambda t: [float(P0[i] + t * D[i]) for i in range(3)]
    return np.array(pt(lo) + pt(hi))


def test_segment_f64_closed_form_and_exact_clip(orc):
    """The exact mode's contact segments are checked against segment_f64:
    a hand-computed crossing, then exact clipping on random intersecting pairs."""
    a = np.array([-1, -1, 0, 2, -1, 0, -1, 2, 0], np.float32)   # z = 0, hypotenuse x + y = 1
    b = np.array([0, -5, -1, 0, 5, -1, 0, 0, 1], np.float32)    # x = 0, crosses z = 0 at y = +-2.5
    seg = orc.segment_f64(a[None], b[None])[0]
    assert seg.tolist() in ([0, -1, 0, 0, 1, 0], [0, 1, 0, 0, -1, 0])
    a9, b9, _ = scenes.config4(3000, seed=19, span=1.0)
    ex = orc.tritri_exact(a9, b9)
    hit = np.flatnonzero(ex == 1)[:150]
    assert len(hit) >= 100
    got = orc.segment_f64(a9[hit], b9[hit])
    for i, g in zip(hit, got):
        want = _segment_fraction(a9[i], b9[i])
        scale = max(1.0, np.abs(want).max())
        err = min(np.abs(g - want).max(), np.abs(g - np.r_[want[3:], want[:3]]).max())
        assert err <= 1e-6 * scale, (i, g, want)


def test_exact_vs_moller_agree_away_from_degeneracy(orc):
    a9, b9, kind = scenes.config4(20000, seed=31)
    f, _ = orc.tritri(a9, b9)
    ex = orc.tritri_exact(a9, b9)
    rnd = kind == 0
    moller = (f & 1) & ((f & 2) == 0)
    assert np.array_equal(moller[rnd].astype(bool), (ex[rnd] == 1))


# ------------------------------------------------------------- glm math ----
def test_mat_inverse_and_mul(orc):
    I = scenes.identity16()
    assert np.array_equal(orc.mat_inverse(I), I)  # equal values; the cofactor signs leave some zeros as -0
    rng = np.random.default_rng(4)
    M = scenes.pose_matrices(rng.uniform(-3, 3, (50, 3)), scenes.random_quaternions(rng, 50), rng.uniform(0.5, 3, 50))
    for m in M:
        inv = orc.mat_inverse(m)
        m64 = m.astype(np.float64).reshape(4, 4).T  # row-major view of the column-major array
        assert np.allclose(inv.astype(np.float64).reshape(4, 4).T, np.linalg.inv(m64), atol=1e-5)
        prod = orc.mat_mul(inv, m).astype(np.float64).reshape(4, 4).T
        assert np.allclose(prod, np.eye(4), atol=1e-5)
    a, b = M[0], M[1]
    ab = orc.mat_mul(a, b).astype(np.float64).reshape(4, 4).T
    want = a.astype(np.float64).reshape(4, 4).T @ b.astype(np.float64).reshape(4, 4).T
    assert np.allclose(ab, want, atol=1e-5)
    assert not np.allclose(ab, want.T, atol=1e-3)  # a transposed operand would be caught


def test_sap_axes_orthonormal(orc):
    ax = orc.sap_axes().astype(np.float64)
    assert np.allclose(ax @ ax.T, np.eye(3), atol=1e-6)
    assert np.allclose(ax[0], np.array([0.8, -0.2, 0.0]) / np.hypot(0.8, 0.2), atol=1e-7)


# ------------------------------------------------------------ OBB / SAT ----
def _box_contains(box, pts, tol):
    c, S = box[:3].astype(np.float64), box[3:].reshape(3, 3).astype(np.float64)
    coef = np.linalg.solve(S.T, (pts.astype(np.float64) - c).T).T
    return np.all(np.abs(coef) <= 1 + tol)


def test_box_from_points(orc):
    rng = np.random.default_rng(6)
    for _ in range(100):
        p = (rng.normal(size=(60, 3)) * rng.uniform(0.05, 5, size=3)).astype(np.float32)
        b = orc.box_from_points(p)
        S = b[3:].reshape(3, 3).astype(np.float64)
        Sn = S / np.linalg.norm(S, axis=1, keepdims=True)
        assert np.allclose(Sn @ Sn.T, np.eye(3), atol=1e-5)
        assert _box_contains(b, p, 1e-5)
    eps = np.float32(np.finfo(np.float32).eps)
    one = orc.box_from_points(np.array([[1, 2, 3]] * 6, np.float32))
    assert list(one) == [1, 2, 3, eps, 0, 0, 0, eps, 0, 0, 0, eps]
    assert list(orc.box_from_points(np.zeros((0, 3), np.float32))) == [0, 0, 0, eps, 0, 0, 0, eps, 0, 0, 0, eps]
    # an axis-aligned box's points give axis-aligned sides with half-extents + FLT_EPSILON
    cube = np.array([[x, y, z] for x in (-1, 1) for y in (-2, 2) for z in (-3, 3)], np.float32)
    b = orc.box_from_points(cube)
    half = sorted(np.linalg.norm(b[3:].reshape(3, 3), axis=1))
    assert np.allclose(half, [1 + eps, 2 + eps, 3 + eps], rtol=0, atol=1e-6)


def _lp_overlap(a, b):
    from scipy.optimize import linprog
    ca, Sa = a[:3].astype(float), a[3:].reshape(3, 3).astype(float)
    cb, Sb = b[:3].astype(float), b[3:].reshape(3, 3).astype(float)
    A_eq = np.concatenate([Sa.T, -Sb.T], axis=1)
    res = linprog(np.zeros(6), A_eq=A_eq, b_eq=cb - ca, bounds=[(-1, 1)] * 6, method="highs")
    return res.status == 0


def test_sat_matches_exact_overlap(orc):
    rng = np.random.default_rng(7)
    agree = 0
    for _ in range(300):
        boxes = []
        for _ in range(2):
            q, _r = np.linalg.qr(rng.normal(size=(3, 3)))
            ext = rng.uniform(0.2, 2, size=3)
            boxes.append(np.concatenate([rng.uniform(-2, 2, 3), (q * ext[:, None]).reshape(9)]).astype(np.float32))
        s = orc.sat(boxes[0], boxes[1])
        # skip near-tangent pairs: shrink/grow both boxes by 1e-3 and require the same exact verdict
        shrink = [np.concatenate([x[:3], x[3:] * 0.999]) for x in boxes]
        grow = [np.concatenate([x[:3], x[3:] * 1.001]) for x in boxes]
        e1, e2 = _lp_overlap(*shrink), _lp_overlap(*grow)
        if e1 != e2:
            continue
        assert s == e1
        agree += 1
    assert agree > 250


# --------------------------------------------------------------- trees ----
def _tree_checks(tree, tri9, L):
    ex = tree.export()
    T = len(tri9)
    perm = ex["perm"]
    assert sorted(perm.tolist()) == list(range(T))
    links, obbs = ex["links"], ex["obbs"]
    n = len(links)
    tri = tri9[perm]
    seen_nodes = {0} if n else set()
    covered = np.zeros(T, int)
    slack = tree.slack
    assert slack < 1e-5
    for i in range(n):
        for s in range(2):
            cnt, link = int(links[i, 2 + s]), int(links[i, s])
            box = obbs[i, 12 * s: 12 * s + 12]
            if cnt == 0:
                assert i < link < n
                seen_nodes.add(link)
            else:
                assert 1 <= cnt <= L
                covered[link: link + cnt] += 1
                assert _box_contains(box, tri[link: link + cnt].reshape(-1, 3), 1e-4)
    if n:
        assert seen_nodes == set(range(n))
        assert np.all(covered == 1)
    else:
        assert T <= L
    assert _box_contains(ex["root"], tri9.reshape(-1, 3), 1e-4) if T else True


@pytest.mark.parametrize("L", [1, 4])
def test_tree_invariants(orc, L):
    rng = np.random.default_rng(10)
    meshes = [scenes.unit_icosphere_soup(4), scenes.box_soup(), scenes.torus_soup(rng, 16, 8),
              rng.normal(size=(57, 9)).astype(np.float32), scenes.box_soup()[:1],
              np.repeat(np.ones((1, 9), np.float32), 9, axis=0)]  # all points identical
    for m in meshes:
        t = orc.tree(m, L)
        _tree_checks(t, m, L)


# ------------------------------------------------------------ traversal ----
def _transform(m16, tri9):
    M = m16.astype(np.float32).reshape(4, 4)  # column-major: M[c, r]
    p = tri9.reshape(-1, 3)
    out = np.empty_like(p)
    for r in range(3):
        a, b, c, d = M[0, r] * p[:, 0], M[1, r] * p[:, 1], M[2, r] * p[:, 2], M[3, r] * np.float32(1)
        out[:, r] = (a + b) + (c + d)
    return out.reshape(-1, 9)


def test_traversal_against_all_pairs(orc):
    mesh = scenes.unit_icosphere_soup(3)  # 180 triangles
    soup, rel = mesh, scenes.config1(40, seed=5)[1]
    for L in (1, 4):
        tree = orc.tree(soup, L)
        tri = soup[tree.export()["perm"]]
        res = orc.batch([tree], np.zeros(len(rel), np.uint32), np.zeros(len(rel), np.uint32), rel, mode=0)
        got = {(int(c["query"]), int(c["tri_a"]), int(c["tri_b"])) for c in res.contacts}
        want = set()
        exact_want = set()
        for q, m in enumerate(rel):
            tb = _transform(m, tri)
            A = np.repeat(tri, len(tri), axis=0)
            B = np.tile(tb, (len(tri), 1))
            f, _ = orc.tritri(A, B)
            hit = np.flatnonzero((f & 1) & ((f & 2) == 0))
            want |= {(q, int(h // len(tri)), int(h % len(tri))) for h in hit}
            ex = orc.tritri_exact(A, B)
            exact_want |= {(q, int(h // len(tri)), int(h % len(tri))) for h in np.flatnonzero(ex == 1)}
        assert got <= want
        assert len(want - got) <= max(2, len(want) // 200)  # only pairs cut by a boundary SAT rounding
        rex = orc.batch([tree], np.zeros(len(rel), np.uint32), np.zeros(len(rel), np.uint32), rel, mode=1)
        got_ex = {(int(c["query"]), int(c["tri_a"]), int(c["tri_b"])) for c in rex.contacts}
        assert got_ex == exact_want


# ------------------------------------------------------------ broad phase ----
def test_sap_against_brute_force(orc):
    rng = np.random.default_rng(12)
    n = 120
    soups = [scenes.box_soup(), scenes.unit_icosphere_soup(2)]
    trees = [orc.tree(s, 4) for s in soups]
    mesh_of = rng.integers(0, 2, n).astype(np.uint32)
    cur = scenes.pose_matrices(rng.uniform(0, 12, (n, 3)), scenes.random_quaternions(rng, n), rng.uniform(0.5, 3, n))
    cb = (rng.random(n) < 0.8).astype(np.uint8)
    res = orc.frame(trees, mesh_of, cur, cb, mode=0)
    got = {tuple(p) for p in res.pairs.tolist()}
    axes = orc.sap_axes()
    iv = np.zeros((n, 3, 2), np.float32)
    for i in range(n):
        box = orc.xform_box(cur[i], trees[mesh_of[i]].export()["root"])
        c, S = box[:3], box[3:].reshape(3, 3)
        for k in range(3):
            a = axes[k]
            cproj = np.float32(np.float32(c[0] * a[0]) + np.float32(c[1] * a[1])) + np.float32(c[2] * a[2])
            s = [abs(np.float32(np.float32(a[0] * S[j, 0]) + np.float32(a[1] * S[j, 1])) + np.float32(a[2] * S[j, 2]))
                 for j in range(3)]
            tot = np.float32(np.float32(s[0] + s[1]) + s[2])
            iv[i, k] = (np.float32(cproj - tot), np.float32(cproj + tot))
    want = set()
    for i in range(n):
        for j in range(i + 1, n):
            if not (cb[i] or cb[j]):
                continue
            if all(iv[i, k, 1] >= iv[j, k, 0] and iv[j, k, 1] >= iv[i, k, 0] for k in range(3)):
                first = i if (iv[i, 0, 0], i) < (iv[j, 0, 0], j) else j
                want.add((first, j if first == i else i))
    assert got == want
    assert len(want) > 10
