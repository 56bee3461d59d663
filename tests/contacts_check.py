"""Per-contact comparison of GPU contacts with the oracle's (test helper).

Contacts are matched by (query, tri_a, tri_b) after sorting both sides.

* Plain mode: every field bit for bit (the reference's arithmetic).
* Batched (exact) mode: the pair set equal to the exact oracle's; flags and
  normals bit for bit (the reference's float contact filter on the same float
  corners); the segment against the oracle's float64 segment, for every pair
  whose crossing edges meet the other triangle's plane at an angle of at
  least MIN_SIN (below that the float32 segment ends are ill-conditioned:
  their error grows like eps * scale / sin), within SEG_TOL * scale, scale
  the largest coordinate magnitude of the pair's six corners.
"""
import numpy as np

SEG_TOL = 1e-5
MIN_SIN = 1e-2


def sort_contacts(c):
    order = np.lexsort((c["tri_b"], c["tri_a"], c["query"]))
    return c[order]


def keys(c):
    return np.stack([c["query"], c["tri_a"], c["tri_b"]], 1).astype(np.int64)


def assert_same_keys(got, want):
    kg, kw = keys(got), keys(want)
    if len(kg) != len(kw) or not np.array_equal(kg, kw):
        sg = set(map(tuple, kg.tolist()))
        sw = set(map(tuple, kw.tolist()))
        raise AssertionError(f"pair sets differ: {len(sg - sw)} extra, {len(sw - sg)} missing "
                             f"(of {len(sw)}); e.g. {sorted(sg ^ sw)[:5]}")


def assert_plain_bit_exact(got, want):
    """Both sorted; every field identical."""
    assert_same_keys(got, want)
    for f in ("flags", "s", "t", "na", "nb"):
        assert got[f].tobytes() == want[f].tobytes(), f


def xform_f32(m16, p):
    """glm's mat4 * vec4(p, 1) in float32: (m0 x + m4 y) + (m8 z + m12), per row.
    m16: (n, 16) column-major, p: (n, 3)."""
    m = m16.astype(np.float32)
    p = p.astype(np.float32)
    out = np.empty_like(p)
    one = np.float32(1.0)
    for r in range(3):
        a = m[:, 0 + r] * p[:, 0]
        b = m[:, 4 + r] * p[:, 1]
        c = m[:, 8 + r] * p[:, 2]
        d = m[:, 12 + r] * one
        out[:, r] = (a + b) + (c + d)
    return out


def crossing_sin(X, Y):
    """Smallest sine of the angle between an edge of triangle X that crosses
    triangle Y's plane and that plane (float64); 1 where no edge crosses
    strictly (the triangle touches the plane at a corner or edge)."""
    X = X.astype(np.float64)
    Y = Y.astype(np.float64)
    n = np.cross(Y[:, 1] - Y[:, 0], Y[:, 2] - Y[:, 0])
    nl = np.linalg.norm(n, axis=1)
    d = np.einsum("nkc,nc->nk", X - Y[:, None, 0], n)
    best = np.ones(len(X))
    for k in range(3):
        j = (k + 1) % 3
        cross = d[:, k] * d[:, j] < 0
        el = np.linalg.norm(X[:, j] - X[:, k], axis=1)
        s = np.abs(d[:, k] - d[:, j]) / np.maximum(nl * el, 1e-300)
        best = np.where(cross, np.minimum(best, s), best)
    return best


def check_batched_geometry(got, want, A, B, min_checked=0.9):
    """got, want: sorted contacts with equal keys; A, B: (n, 3, 3) float32
    corners of each contact's triangles in A's space (B already moved).
    Returns (pairs checked, largest relative segment error)."""
    assert len(got) == len(want) == len(A) == len(B)
    for f in ("flags", "na", "nb"):
        bad = np.flatnonzero(np.any((got[f] != want[f]).reshape(len(got), -1), axis=1))
        assert len(bad) == 0, (f, len(bad), got[bad[:3]], want[bad[:3]])
    if len(got) == 0:
        return 0, 0.0
    cond = np.minimum(crossing_sin(A, B), crossing_sin(B, A)) >= MIN_SIN
    scale = np.maximum(np.abs(A).reshape(len(A), -1).max(1), np.abs(B).reshape(len(B), -1).max(1))
    scale = np.maximum(scale, 1e-30).astype(np.float64)
    gs, gt = got["s"].astype(np.float64), got["t"].astype(np.float64)
    ws, wt = want["s"].astype(np.float64), want["t"].astype(np.float64)
    same = np.maximum(np.abs(gs - ws).max(1), np.abs(gt - wt).max(1))
    swap = np.maximum(np.abs(gs - wt).max(1), np.abs(gt - ws).max(1))
    rel = np.minimum(same, swap) / scale
    worst = float(rel[cond].max()) if cond.any() else 0.0
    bad = np.flatnonzero(cond & (rel > SEG_TOL))
    assert len(bad) == 0, (f"{len(bad)} of {int(cond.sum())} conditioned segments off by more than {SEG_TOL} "
                           f"relative (worst {worst:.3g})", got[bad[:3]], want[bad[:3]])
    assert cond.mean() >= min_checked, f"only {cond.mean():.3f} of the pairs are conditioned"
    return int(cond.sum()), worst


def corners_from_trees(contacts, tree_pos, qa, qb, rel):
    """Each contact's triangles in A's space: tree_pos[m] is mesh m's
    tree-order triangles (n, 9); qa, qb the queries' meshes; rel their
    matrices (B's corners moved by xform_f32, as the narrow phase moves them)."""
    q = contacts["query"].astype(np.int64)
    ma, mb = np.asarray(qa)[q], np.asarray(qb)[q]
    A = np.empty((len(q), 9), np.float32)
    Bl = np.empty((len(q), 9), np.float32)
    for m in np.unique(ma):
        sel = ma == m
        A[sel] = tree_pos[int(m)][contacts["tri_a"][sel]]
    for m in np.unique(mb):
        sel = mb == m
        Bl[sel] = tree_pos[int(m)][contacts["tri_b"][sel]]
    B = xform_f32(np.repeat(np.asarray(rel, np.float32).reshape(-1, 16)[q], 3, axis=0), Bl.reshape(-1, 3))
    return A.reshape(-1, 3, 3), B.reshape(-1, 3, 3)


class TreePositions(dict):
    """Mesh -> tree-order triangles, from the oracle's trees, on demand."""

    def __init__(self, soups, trees):
        super().__init__()
        self.soups, self.trees = soups, trees

    def __missing__(self, m):
        v = self.soups[m][self.trees[m].export()["perm"]]
        self[m] = v
        return v
