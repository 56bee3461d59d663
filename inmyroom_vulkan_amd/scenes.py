"""Synthetic, seeded workloads for the collision path.

Meshes are float32 triangle soups (T, 9); poses are column-major float32
4x4 matrices (glm layout, translation in elements 12..14).

  config 1  two unit icospheres of 1,280 triangles, 1,000 relative poses
  config 2  512 rigid entities (boxes / icospheres / noisy tori), 100 frames
  config 3  4,096 mesh-pair queries, triangle counts log-uniform in [12, 65,536]
  config 4  16M triangle pairs: random, near-coplanar, touching
"""
from __future__ import annotations

import functools
import math

import numpy as np

# --------------------------------------------------------------------------
# meshes
# --------------------------------------------------------------------------
_PHI = (1.0 + 5.0 ** 0.5) / 2.0
_ICO_V = np.array([[-1, _PHI, 0], [1, _PHI, 0], [-1, -_PHI, 0], [1, -_PHI, 0], [0, -1, _PHI], [0, 1, _PHI],
                   [0, -1, -_PHI], [0, 1, -_PHI], [_PHI, 0, -1], [_PHI, 0, 1], [-_PHI, 0, -1], [-_PHI, 0, 1]],
                  np.float64)
_ICO_F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                   [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                   [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)


@functools.lru_cache(maxsize=128)
def icosphere(freq: int) -> tuple[np.ndarray, np.ndarray]:
    """Closed geodesic icosphere of frequency `freq`: 20*freq^2 triangles,
    10*freq^2 + 2 vertices on the unit sphere (float64 directions, int32 faces).
    Edge points are computed once per edge, so neighbouring faces share them."""
    n = int(freq)
    assert n >= 1
    base = _ICO_V / np.linalg.norm(_ICO_V, axis=1, keepdims=True)
    verts = [v for v in base]
    edge_ids: dict[tuple[int, int], list[int]] = {}

    def edge_point(a: int, b: int, k: int) -> int:  # k steps from a towards b
        lo, hi = (a, b) if a < b else (b, a)
        kk = k if a < b else n - k
        if (lo, hi) not in edge_ids:
            ids = []
            for s in range(1, n):
                verts.append(base[lo] + (base[hi] - base[lo]) * (s / n))
                ids.append(len(verts) - 1)
            edge_ids[(lo, hi)] = ids
        return edge_ids[(lo, hi)][kk - 1]

    faces = []
    for A, B, Cc in _ICO_F:
        grid = {}
        for i in range(n + 1):
            for j in range(n + 1 - i):
                if (i, j) == (0, 0):
                    idx = A
                elif (i, j) == (n, 0):
                    idx = B
                elif (i, j) == (0, n):
                    idx = Cc
                elif j == 0:
                    idx = edge_point(A, B, i)
                elif i == 0:
                    idx = edge_point(A, Cc, j)
                elif i + j == n:
                    idx = edge_point(B, Cc, j)
                else:
                    verts.append(base[A] + (base[B] - base[A]) * (i / n) + (base[Cc] - base[A]) * (j / n))
                    idx = len(verts) - 1
                grid[(i, j)] = idx
        for i in range(n):
            for j in range(n - i):
                faces.append((grid[(i, j)], grid[(i + 1, j)], grid[(i, j + 1)]))
                if i + j < n - 1:
                    faces.append((grid[(i + 1, j)], grid[(i + 1, j + 1)], grid[(i, j + 1)]))
    V = np.array(verts, np.float64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    F = np.array(faces, np.int32)
    assert len(F) == 20 * n * n and len(V) == 10 * n * n + 2
    return V, F


def soup(V: np.ndarray, F: np.ndarray) -> np.ndarray:
    """Indexed mesh -> (T, 9) float32 triangle soup (glTF 'triangles' mode)."""
    return np.ascontiguousarray(np.asarray(V, np.float32)[F].reshape(-1, 9))


def unit_icosphere_soup(freq: int = 8) -> np.ndarray:
    V, F = icosphere(freq)
    return soup(V, F)


def displaced_icosphere(freq: int, amp: float, rng: np.random.Generator) -> np.ndarray:
    """Icosphere with smooth radial displacement r = 1 + amp * f(dir), |f| <= 1."""
    V, F = icosphere(freq)
    k = 4
    w = rng.normal(size=(k, 3))
    w *= rng.uniform(1.0, 4.0, size=(k, 1)) / np.linalg.norm(w, axis=1, keepdims=True)
    ph = rng.uniform(0, 2 * np.pi, size=k)
    a = rng.uniform(0.5, 1.0, size=k)
    f = (np.sin(V @ w.T + ph) * a).sum(axis=1) / a.sum()
    return soup(V * (1.0 + amp * f)[:, None], F)


def box_soup(half: float = 0.5) -> np.ndarray:
    c = np.array([[x, y, z] for x in (-half, half) for y in (-half, half) for z in (-half, half)], np.float64)
    F = np.array([[0, 2, 3], [0, 3, 1], [4, 5, 7], [4, 7, 6], [0, 1, 5], [0, 5, 4], [2, 6, 7], [2, 7, 3],
                  [0, 4, 6], [0, 6, 2], [1, 3, 7], [1, 7, 5]], np.int32)
    return soup(c, F)


def torus_soup(rng: np.random.Generator, nu: int = 128, nv: int = 32, R: float = 1.0, r: float = 0.3,
               sigma: float = 0.02) -> np.ndarray:
    u = np.arange(nu) * (2 * np.pi / nu)
    v = np.arange(nv) * (2 * np.pi / nv)
    U, Vv = np.meshgrid(u, v, indexing="ij")
    rr = r + rng.normal(0.0, sigma, size=U.shape)
    P = np.stack([(R + rr * np.cos(Vv)) * np.cos(U), (R + rr * np.cos(Vv)) * np.sin(U), rr * np.sin(Vv)], -1)
    P = P.reshape(-1, 3)
    i, j = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    a = i * nv + j
    b = ((i + 1) % nu) * nv + j
    c = ((i + 1) % nu) * nv + (j + 1) % nv
    d = i * nv + (j + 1) % nv
    F = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3), np.stack([a, c, d], -1).reshape(-1, 3)]).astype(np.int32)
    return soup(P, F)


# --------------------------------------------------------------------------
# poses
# --------------------------------------------------------------------------
def random_quaternions(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform unit quaternions (Shoemake 1992), (w, x, y, z)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    return np.stack([b * np.cos(2 * np.pi * u3), a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                     b * np.sin(2 * np.pi * u3)], -1)


def quat_to_mat3(q: np.ndarray) -> np.ndarray:
    q = q / np.linalg.norm(q, axis=-1, keepdims=True)
    w, x, y, z = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1),
    ], -2)


def pose_matrices(t: np.ndarray, q: np.ndarray, s: np.ndarray | float = 1.0) -> np.ndarray:
    """T * R * S as column-major float32 (n, 16)."""
    t = np.asarray(t, np.float64).reshape(-1, 3)
    R = quat_to_mat3(np.asarray(q, np.float64).reshape(-1, 4))
    s = np.broadcast_to(np.asarray(s, np.float64), (len(t),))
    M = np.zeros((len(t), 4, 4), np.float64)
    M[:, :3, :3] = R * s[:, None, None]
    M[:, :3, 3] = t
    M[:, 3, 3] = 1.0
    # row-major M[r, c] -> column-major flat index c*4 + r
    return np.ascontiguousarray(M.transpose(0, 2, 1).reshape(-1, 16).astype(np.float32))


# --------------------------------------------------------------------------
# config 1: plain, CPU-runnable
# --------------------------------------------------------------------------
def config1(n_poses: int = 1000, seed: int = 1, freq: int = 8):
    """Two unit icospheres (20*freq^2 triangles; freq 8 -> 1,280), n_poses
    relative poses: t ~ U[-2.5, 2.5]^3, uniform rotation.  Entity A sits at the
    identity, so the relative matrix is B's pose."""
    rng = np.random.default_rng(seed)
    t = rng.uniform(-2.5, 2.5, size=(n_poses, 3))
    q = random_quaternions(rng, n_poses)
    return unit_icosphere_soup(freq), pose_matrices(t, q)


# --------------------------------------------------------------------------
# config 2: scene of rigid entities
# --------------------------------------------------------------------------
def quat_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    aw, ax, ay, az = a[..., 0], a[..., 1], a[..., 2], a[..., 3]
    bw, bx, by, bz = b[..., 0], b[..., 1], b[..., 2], b[..., 3]
    return np.stack([aw * bw - ax * bx - ay * by - az * bz, aw * bx + ax * bw + ay * bz - az * by,
                     aw * by - ax * bz + ay * bw + az * bx, aw * bz + ax * by - ay * bx + az * bw], -1)


class Scene:
    """Rigid entities drawn from three mesh kinds, integrated at fixed dt."""

    def __init__(self, n_entities: int = 512, seed: int = 2, cube: float = 50.0, dt: float = 1.0 / 60.0,
                 torus_grid: tuple[int, int] = (128, 32)):
        rng = np.random.default_rng(seed)
        self.soups = [box_soup(), unit_icosphere_soup(8), torus_soup(rng, *torus_grid)]
        n = int(n_entities)
        self.mesh_of = rng.integers(0, 3, size=n).astype(np.uint32)
        self.scale = rng.uniform(0.5, 3.0, size=n)
        self.x = rng.uniform(0.0, cube, size=(n, 3))
        self.q = random_quaternions(rng, n)
        self.v = rng.normal(0.0, 2.0, size=(n, 3))
        self.w = rng.normal(0.0, 1.0, size=(n, 3))
        self.dt = dt
        self.should_callback = np.ones(n, np.uint8)
        self.prev = self.matrices()

    def matrices(self) -> np.ndarray:
        return pose_matrices(self.x, self.q, self.scale)

    def step(self):
        """x += v dt; q <- q * exp(w dt / 2), renormalised."""
        self.prev = self.matrices()
        self.x = self.x + self.v * self.dt
        ang = np.linalg.norm(self.w, axis=1) * self.dt / 2
        axis = self.w / np.maximum(np.linalg.norm(self.w, axis=1, keepdims=True), 1e-30)
        dq = np.concatenate([np.cos(ang)[:, None], axis * np.sin(ang)[:, None]], axis=1)
        q = quat_mul(self.q, dq)
        self.q = q / np.linalg.norm(q, axis=1, keepdims=True)

    def entities(self, mesh_ids, dtype) -> np.ndarray:
        e = np.zeros(len(self.mesh_of), dtype)
        e["current"] = self.matrices()
        e["previous"] = self.prev
        e["mesh_id"] = np.asarray(mesh_ids, np.uint32)[self.mesh_of]
        e["entity"] = np.arange(len(self.mesh_of), dtype=np.uint16)
        e["should_callback"] = self.should_callback
        return e


# --------------------------------------------------------------------------
# config 3: 4,096 uneven mesh-pair queries
# --------------------------------------------------------------------------
def config3_sizes(n_queries: int = 4096, seed: int = 3, tmin: float = 12.0, tmax: float = 65536.0):
    """Frequencies of the 2*n_queries meshes: T ~ logU[tmin, tmax] mapped to
    a geodesic icosphere n = max(1, round(sqrt(T/20))), i.e. T = 20 n^2."""
    rng = np.random.default_rng(seed)
    T = np.exp(rng.uniform(np.log(tmin), np.log(tmax), size=2 * n_queries))
    return np.maximum(1, np.rint(np.sqrt(T / 20.0))).astype(np.int64)


def config3(n_queries: int = 4096, seed: int = 3, amp: float = 0.1, overlap_p: float = 0.3,
            freq_cap: int | None = None):
    """Returns (soups, qa, qb, rel, overlap_flag).  Mesh 2i and 2i+1 belong to
    query i.  freq_cap (tests only) bounds the mesh size."""
    freqs = config3_sizes(n_queries, seed)
    if freq_cap is not None:
        freqs = np.minimum(freqs, freq_cap)
    rng = np.random.default_rng(seed + 1000)
    soups = [displaced_icosphere(int(f), amp, rng) for f in freqs]
    radii = np.array([float(np.sqrt((s.reshape(-1, 3).astype(np.float64) ** 2).sum(1)).max()) for s in soups])
    overlap = rng.random(n_queries) < overlap_p
    rsum = radii[0::2] + radii[1::2]
    dist = np.where(overlap, rng.uniform(0.5, 1.0, n_queries), rng.uniform(1.2, 2.0, n_queries)) * rsum
    d = rng.normal(size=(n_queries, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rel = pose_matrices(d * dist[:, None], random_quaternions(rng, n_queries))
    qa = np.arange(0, 2 * n_queries, 2, dtype=np.uint32)
    qb = qa + 1
    return soups, qa, qb, rel, overlap


# --------------------------------------------------------------------------
# config 5: deformable batch against a terrain
# --------------------------------------------------------------------------
def torus_indexed(nu: int, nv: int, R: float = 1.0, r: float = 0.3):
    u = np.arange(nu) * (2 * np.pi / nu)
    v = np.arange(nv) * (2 * np.pi / nv)
    U, Vv = np.meshgrid(u, v, indexing="ij")
    P = np.stack([(R + r * np.cos(Vv)) * np.cos(U), (R + r * np.cos(Vv)) * np.sin(U), r * np.sin(Vv)], -1).reshape(-1, 3)
    i, j = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    a = i * nv + j
    b = ((i + 1) % nu) * nv + j
    c = ((i + 1) % nu) * nv + (j + 1) % nv
    d = i * nv + (j + 1) % nv
    F = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3), np.stack([a, c, d], -1).reshape(-1, 3)]).astype(np.uint32)
    return P, F, U.reshape(-1)


def heightfield(n: int = 316, size: float = 100.0, amp: float = 5.0, octaves: int = 6, seed: int = 5):
    """n x n vertex grid over [0, size]^2 with fBm heights in [-amp, amp]:
    2 (n-1)^2 triangles."""
    rng = np.random.default_rng(seed)
    xs = np.linspace(0.0, size, n)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    h = np.zeros_like(X)
    a = 1.0
    for o in range(octaves):
        res = 2 ** (o + 2)
        g = rng.uniform(-1, 1, size=(res + 1, res + 1))
        fx, fy = X / size * res, Y / size * res
        ix, iy = np.minimum(fx.astype(int), res - 1), np.minimum(fy.astype(int), res - 1)
        tx, ty = fx - ix, fy - iy
        tx, ty = tx * tx * (3 - 2 * tx), ty * ty * (3 - 2 * ty)
        v = (g[ix, iy] * (1 - tx) * (1 - ty) + g[ix + 1, iy] * tx * (1 - ty) + g[ix, iy + 1] * (1 - tx) * ty
             + g[ix + 1, iy + 1] * tx * ty)
        h += a * v
        a *= 0.5
    h *= amp / np.abs(h).max()
    P = np.stack([X, Y, h], -1).reshape(-1, 3)
    i, j = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    a0, b0, c0, d0 = i * n + j, (i + 1) * n + j, (i + 1) * n + j + 1, i * n + j + 1
    F = np.concatenate([np.stack([a0, b0, c0], -1).reshape(-1, 3), np.stack([a0, c0, d0], -1).reshape(-1, 3)]).astype(np.uint32)
    return P, F, h


def rotation_about(axis: np.ndarray, angle: np.ndarray) -> np.ndarray:
    axis = axis / np.linalg.norm(axis, axis=-1, keepdims=True)
    h = angle / 2
    return quat_to_mat3(np.concatenate([np.cos(h)[..., None], axis * np.sin(h)[..., None]], -1))


class DeformableBatch:
    """n_meshes skinned tori (nu x nv grid, 2 nu nv triangles) with n_bones
    bones on the ring; each vertex takes its 4 nearest bones with
    Dirichlet(1,1,1,1) weights; every frame each bone turns about its rest
    point by a random angle in [0, max_deg] degrees.  The tori sit on an fBm
    heightfield terrain."""

    def __init__(self, n_meshes: int = 64, nu: int = 100, nv: int = 100, n_bones: int = 64, seed: int = 5,
                 terrain_n: int = 316, terrain_size: float = 100.0, max_deg: float = 30.0):
        rng = np.random.default_rng(seed)
        self.rng = rng
        self.P, self.F, ring = torus_indexed(nu, nv)
        self.n_meshes, self.n_bones, self.max_deg = n_meshes, n_bones, max_deg
        theta = np.arange(n_bones) * (2 * np.pi / n_bones)
        self.bone_pos = np.stack([np.cos(theta), np.sin(theta), np.zeros(n_bones)], -1)
        d = np.abs(((ring[:, None] - theta[None, :]) + np.pi) % (2 * np.pi) - np.pi)
        self.joints = np.argsort(d, axis=1)[:, :4].astype(np.uint16)
        self.weights = rng.dirichlet(np.ones(4), size=len(self.P)).astype(np.float32)
        self.tP, self.tF, h = heightfield(terrain_n, terrain_size, 5.0, seed=seed + 1)
        # world placement: uniform scale 1-2, on the terrain, random heading, half sunk
        self.scale = rng.uniform(1.0, 2.0, n_meshes)
        xy = rng.uniform(10.0, terrain_size - 10.0, size=(n_meshes, 2))
        gi = np.clip(np.rint(xy / terrain_size * (terrain_n - 1)).astype(int), 0, terrain_n - 1)
        z = h[gi[:, 0], gi[:, 1]]
        self.world_t = np.concatenate([xy, z[:, None]], 1)
        self.world_q = random_quaternions(rng, n_meshes)

    def bone_matrices(self) -> np.ndarray:
        """(n_meshes, n_bones, 16) column-major: World * T(b) * R * T(-b)."""
        nm, nb = self.n_meshes, self.n_bones
        axis = self.rng.normal(size=(nm, nb, 3))
        ang = np.radians(self.rng.uniform(0.0, self.max_deg, size=(nm, nb)))
        R = rotation_about(axis, ang)
        Wr = quat_to_mat3(self.world_q) * self.scale[:, None, None]
        M = np.zeros((nm, nb, 4, 4))
        b = self.bone_pos[None]
        M[..., :3, :3] = Wr[:, None] @ R
        local_t = b - (R @ b[..., None])[..., 0]  # T(b) R T(-b) translation
        M[..., :3, 3] = (Wr[:, None] @ local_t[..., None])[..., 0] + self.world_t[:, None]
        M[..., 3, 3] = 1.0
        return np.ascontiguousarray(M.transpose(0, 1, 3, 2).reshape(nm, nb, 16).astype(np.float32))

    @staticmethod
    def skin_reference(P, joints, weights, bones) -> np.ndarray:
        """float64 linear blend skinning, for tests."""
        B = bones.astype(np.float64).reshape(-1, 4, 4).transpose(0, 2, 1)  # row-major
        ph = np.concatenate([P, np.ones((len(P), 1))], 1)
        out = np.zeros((len(P), 3))
        for k in range(4):
            out += weights[:, k:k + 1].astype(np.float64) * (B[joints[:, k].astype(int)] @ ph[..., None])[:, :3, 0]
        return out


# --------------------------------------------------------------------------
# config 4: precision stress
# --------------------------------------------------------------------------
def config4(n_pairs: int = 1 << 24, seed: int = 4, span: float = 1e3):
    """(a9, b9, kind) float32; kind 0 random, 1 near-coplanar (offset <= 1e-7
    relative), 2 touching (a shared vertex or edge point moved by 1 ulp)."""
    rng = np.random.default_rng(seed)
    n = int(n_pairs)
    kind = np.zeros(n, np.int8)
    kind[: n // 10] = 1
    kind[n // 10: n // 5] = 2
    rng.shuffle(kind)
    A = rng.uniform(-span, span, size=(n, 3, 3))
    B = rng.uniform(-span, span, size=(n, 3, 3))
    # near-coplanar: B in A's plane, then offset along the unit normal
    m = kind == 1
    k = int(m.sum())
    if k:
        a = A[m]
        e1, e2 = a[:, 1] - a[:, 0], a[:, 2] - a[:, 0]
        nrm = np.cross(e1, e2)
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        ab = rng.uniform(-0.5, 1.2, size=(k, 3, 2))
        pts = a[:, None, 0] + ab[..., :1] * e1[:, None] + ab[..., 1:] * e2[:, None]
        off = rng.uniform(-1e-7, 1e-7, size=(k, 1, 1)) * span
        B[m] = pts + off * nrm[:, None]
    A32 = A.astype(np.float32)
    B32 = B.astype(np.float32)
    # touching: B's first vertex is one of A's vertices or a point of one of A's edges, +-1 ulp
    m = kind == 2
    k = int(m.sum())
    if k:
        a = A32[m].astype(np.float64)
        which = rng.integers(0, 3, size=k)
        onedge = rng.random(k) < 0.5
        t = rng.uniform(0, 1, size=k)
        v0 = a[np.arange(k), which]
        v1 = a[np.arange(k), (which + 1) % 3]
        p = np.where(onedge[:, None], v0 + t[:, None] * (v1 - v0), v0).astype(np.float32)
        step = rng.integers(-1, 2, size=(k, 3))
        p = np.where(step > 0, np.nextafter(p, np.float32(np.inf)), np.where(step < 0, np.nextafter(p, np.float32(-np.inf)), p))
        b = B32[m]
        b[:, 0] = p
        # keep B small and next to the contact point so it crosses A's plane or touches it
        b[:, 1] = (p + rng.uniform(-span / 10, span / 10, size=(k, 3))).astype(np.float32)
        b[:, 2] = (p + rng.uniform(-span / 10, span / 10, size=(k, 3))).astype(np.float32)
        B32[m] = b
    return A32.reshape(n, 9), B32.reshape(n, 9), kind


def identity16() -> np.ndarray:
    return np.eye(4, dtype=np.float32).T.reshape(16).copy()


def config4_count_default() -> int:
    return 1 << 24


def log_uniform_mean(tmin: float = 12.0, tmax: float = 65536.0) -> float:
    return (tmax - tmin) / math.log(tmax / tmin)
