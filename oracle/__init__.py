"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the oracle.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The shipped
library (``inmyroom_vulkan_amd``) never imports, links or calls it.

The oracle is a CPU restatement of inMyRoom_vulkan's collision path (see
``oracle.cpp``).  ``_ref/`` holds the two pieces of the reference that compile
on their own (``eig3/eig3.cpp`` and the Möller section of
``inMyRoom_vulkan/src/Geometry/Triangle.cpp``), used to pin the restatement.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_ROOT = "/root/reference"

CONTACT_DTYPE = np.dtype([
    ("query", "<u4"), ("tri_a", "<u4"), ("tri_b", "<u4"), ("flags", "<u4"),
    ("s", "<f4", 3), ("t", "<f4", 3), ("na", "<f4", 3), ("nb", "<f4", 3),
])
assert CONTACT_DTYPE.itemsize == 64

_F = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_D = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_U32 = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_I32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_U64 = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_U8 = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def build(ref: bool | None = None, fast: bool = True) -> None:
    """Compile liboracle.so (and the reference pieces when the reference is mounted)."""
    targets = ["liboracle.so"] + (["liboracle_fast.so"] if fast else [])
    if ref is None:
        ref = os.path.isdir(REF_ROOT)
    if ref:
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def build_native(out: str) -> str:
    """Compile the oracle with the reference's release flags (-O3 -ffast-math
    -march=native, inMyRoom_vulkan/CMakeLists.txt:30) for this host into
    `out`; the CPU baseline of the benchmark, built where it is measured."""
    subprocess.run(["make", "-s", "-C", HERE, "native", f"NATIVE_OUT={os.path.abspath(out)}"], check=True)
    return os.path.abspath(out)


NATIVE_FLAGS = "-std=c++17 -O3 -ffast-math -march=native"
STRICT_FLAGS = "-std=c++17 -O2 -ffp-contract=off -fno-fast-math"


def _load(name: str) -> C.CDLL:
    path = name if os.path.isabs(name) else os.path.join(HERE, name)
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built; run oracle.build() or `make -C oracle`")
    return C.CDLL(path)


class Oracle:
    def __init__(self, fast: bool = False, path: str | None = None):
        lib = self.lib = _load(path or ("liboracle_fast.so" if fast else "liboracle.so"))
        vp = C.c_void_p
        lib.orc_eig3.argtypes = [_D, _D, _D]
        lib.orc_box_from_points.argtypes = [_F, C.c_int64, _F]
        lib.orc_sort3_desc.argtypes = [_F, C.POINTER(C.c_int32)]
        lib.orc_segment_f64.argtypes = [_F, _F, C.c_int64, _F]
        lib.orc_sat.argtypes = [_F, _F]
        lib.orc_sat.restype = C.c_int
        lib.orc_surface.argtypes = [_F]
        lib.orc_surface.restype = C.c_float
        lib.orc_xform_box.argtypes = [_F, _F, _F]
        lib.orc_mat_inverse.argtypes = [_F, _F]
        lib.orc_node_globals.argtypes = [_F, _U32, C.c_int64, _F]
        lib.orc_node_globals.restype = C.c_int
        lib.orc_mat_mul.argtypes = [_F, _F, _F]
        lib.orc_sap_axes.argtypes = [_F]
        lib.orc_tritri.argtypes = [_F, _F, C.c_int64, _I32, C.c_void_p]
        lib.orc_tritri_exact.argtypes = [_F, _F, C.c_int64, _I32]
        lib.orc_orient_exact.argtypes = [_F, _F, _F, _F]
        lib.orc_orient_exact.restype = C.c_int
        lib.orc_orient_bignum.argtypes = [_F, _F, _F, _F]
        lib.orc_orient_bignum.restype = C.c_int
        lib.orc_bignum_calls.restype = C.c_uint64
        lib.orc_tree_build.argtypes = [_F, C.c_int64, C.c_int]
        lib.orc_tree_build.restype = vp
        lib.orc_tree_free.argtypes = [vp]
        for f in ("orc_tree_nodes", "orc_tree_tris"):
            getattr(lib, f).argtypes = [vp]
            getattr(lib, f).restype = C.c_int64
        lib.orc_tree_depth.argtypes = [vp]
        lib.orc_tree_depth.restype = C.c_int32
        lib.orc_tree_slack.argtypes = [vp]
        lib.orc_tree_slack.restype = C.c_double
        lib.orc_tree_export.argtypes = [vp, _F, _F, _U32, _U32]
        lib.orc_batch_run.argtypes = [C.POINTER(vp), _U32, _U32, _F, C.c_int64, C.c_int, C.c_int, C.c_int]
        lib.orc_batch_run.restype = vp
        lib.orc_frame_run.argtypes = [C.POINTER(vp), _U32, _F, _U8, C.c_int64, C.c_int]
        lib.orc_frame_run.restype = vp
        lib.orc_result_free.argtypes = [vp]
        lib.orc_result_count.argtypes = [vp, C.c_int]
        lib.orc_result_count.restype = C.c_int64
        lib.orc_result_contacts.argtypes = [vp, vp]
        lib.orc_result_leafpairs.argtypes = [vp, _U32, C.c_void_p]
        lib.orc_result_pairs.argtypes = [vp, _U32]
        lib.orc_result_per_query.argtypes = [vp, _U64]
        lib.orc_tree_set_attrs.argtypes = [vp, C.c_void_p, C.c_void_p]
        lib.orc_rays_queries.argtypes = [C.POINTER(vp), _U32, _U32, _F, C.c_int64]
        lib.orc_rays_queries.restype = vp
        lib.orc_rays_frame.argtypes = [C.POINTER(vp), _U32, _F, _U8, C.c_int64]
        lib.orc_rays_frame.restype = vp
        lib.orc_rays_count.argtypes = [vp, C.c_int]
        lib.orc_rays_count.restype = C.c_int64
        lib.orc_rays_get.argtypes = [vp, _U32, _F, _F]
        lib.orc_rays_free.argtypes = [vp]
        lib.orc_ray_triangle.argtypes = [_F, _F, _F]
        lib.orc_ray_box.argtypes = [_F, _F, _F]
        lib.orc_ray_box.restype = C.c_int
        lib.orc_ray_tree.argtypes = [vp, _F, _F, _F, _F]
        lib.orc_shoot_frame.argtypes = [C.POINTER(vp), _U32, _F, _F, _U8, C.c_int64, C.c_float, C.c_float,
                                        C.c_float]
        lib.orc_shoot_frame.restype = vp
        lib.orc_shoot_get.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_float), C.POINTER(C.c_uint64)]
        lib.orc_shoot_get.restype = C.c_int64
        lib.orc_shoot_free.argtypes = [vp]

    # -- small functions -------------------------------------------------
    def eig3(self, A):
        A = np.ascontiguousarray(A, np.float64).reshape(9)
        V = np.zeros(9, np.float64)
        d = np.zeros(3, np.float64)
        self.lib.orc_eig3(A, V, d)
        return V.reshape(3, 3), d

    def box_from_points(self, pts):
        pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
        out = np.zeros(12, np.float32)
        self.lib.orc_box_from_points(pts.reshape(-1), len(pts), out)
        return out

    def segment_f64(self, a9, b9):
        """float64 intersection segments (n, 6) of intersecting, non-coplanar pairs."""
        a9 = np.ascontiguousarray(a9, np.float32).reshape(-1, 9)
        b9 = np.ascontiguousarray(b9, np.float32).reshape(-1, 9)
        out = np.zeros((len(a9), 6), np.float32)
        if len(a9):
            self.lib.orc_segment_f64(a9.reshape(-1), b9.reshape(-1), len(a9), out.reshape(-1))
        return out

    def sat(self, a12, b12) -> bool:
        return bool(self.lib.orc_sat(np.ascontiguousarray(a12, np.float32), np.ascontiguousarray(b12, np.float32)))

    def surface(self, b12) -> np.float32:
        return np.float32(self.lib.orc_surface(np.ascontiguousarray(b12, np.float32)))

    def xform_box(self, m16, b12):
        out = np.zeros(12, np.float32)
        self.lib.orc_xform_box(np.ascontiguousarray(m16, np.float32).reshape(16), np.ascontiguousarray(b12, np.float32), out)
        return out

    def mat_inverse(self, m16):
        out = np.zeros(16, np.float32)
        self.lib.orc_mat_inverse(np.ascontiguousarray(m16, np.float32).reshape(16), out)
        return out

    def mat_mul(self, a16, b16):
        out = np.zeros(16, np.float32)
        self.lib.orc_mat_mul(np.ascontiguousarray(a16, np.float32).reshape(16),
                             np.ascontiguousarray(b16, np.float32).reshape(16), out)
        return out

    def node_globals(self, trs13, parent):
        """GetGlobalMatrix down a hierarchy: trs13 (n, 13) = scale, rotation
        (x, y, z, w), translation, global translation; parent (n,) with ~0 for
        roots.  Returns (n, 16) column-major float32."""
        t = np.ascontiguousarray(trs13, np.float32).reshape(-1, 13)
        p = np.ascontiguousarray(parent, np.uint32).reshape(-1)
        out = np.zeros((len(t), 16), np.float32)
        if self.lib.orc_node_globals(t, p, len(t), out) != 0:
            raise ValueError("cyclic node hierarchy")
        return out

    def sap_axes(self):
        out = np.zeros(9, np.float32)
        self.lib.orc_sap_axes(out)
        return out.reshape(3, 3)

    def tritri(self, a9, b9):
        """Reference-semantics Möller test: flags (bit0 hit, bit1 coplanar, bit2 UB case), segments."""
        a9 = np.ascontiguousarray(a9, np.float32).reshape(-1, 9)
        b9 = np.ascontiguousarray(b9, np.float32).reshape(-1, 9)
        n = len(a9)
        flags = np.zeros(n, np.int32)
        seg = np.zeros((n, 6), np.float32)
        self.lib.orc_tritri(a9.reshape(-1), b9.reshape(-1), n, flags, seg.ctypes.data)
        return flags, seg

    def tritri_exact(self, a9, b9):
        """Exact semantics: 1 intersect, 0 disjoint, 2 coplanar."""
        a9 = np.ascontiguousarray(a9, np.float32).reshape(-1, 9)
        b9 = np.ascontiguousarray(b9, np.float32).reshape(-1, 9)
        out = np.zeros(len(a9), np.int32)
        self.lib.orc_tritri_exact(a9.reshape(-1), b9.reshape(-1), len(a9), out)
        return out

    def orient_exact(self, a, b, c, d, bignum_only: bool = False) -> int:
        f = self.lib.orc_orient_bignum if bignum_only else self.lib.orc_orient_exact
        cv = lambda x: np.ascontiguousarray(x, np.float32).reshape(3)
        return int(f(cv(a), cv(b), cv(c), cv(d)))

    # -- trees and queries -------------------------------------------------
    def tree(self, tri9, leaf_max: int = 4) -> "OracleTree":
        return OracleTree(self, tri9, leaf_max)

    def batch(self, trees, qa, qb, mats, mode: int = 0, nthreads: int = 1, keep_leafpairs: bool = False):
        handles = (C.c_void_p * len(trees))(*[t.h for t in trees])
        qa = np.ascontiguousarray(qa, np.uint32)
        qb = np.ascontiguousarray(qb, np.uint32)
        mats = np.ascontiguousarray(mats, np.float32).reshape(-1, 16)
        r = self.lib.orc_batch_run(handles, qa, qb, mats.reshape(-1), len(qa), mode, nthreads, int(keep_leafpairs))
        return OracleResult(self, r, nq=len(qa))

    def frame(self, trees, mesh_of, cur16, should_cb, mode: int = 0):
        handles = (C.c_void_p * len(trees))(*[t.h for t in trees])
        n = len(mesh_of)
        r = self.lib.orc_frame_run(handles, np.ascontiguousarray(mesh_of, np.uint32),
                                   np.ascontiguousarray(cur16, np.float32).reshape(-1),
                                   np.ascontiguousarray(should_cb, np.uint8), n, mode)
        return OracleResult(self, r, nq=None)


    def rays_queries(self, trees, qa, qb, mats):
        """CDentriesUncollideRays of each query (CreateUncollideRays.cpp:60-200)."""
        handles = (C.c_void_p * len(trees))(*[t.h for t in trees])
        mats = np.ascontiguousarray(mats, np.float32).reshape(-1, 16)
        r = self.lib.orc_rays_queries(handles, np.ascontiguousarray(qa, np.uint32),
                                      np.ascontiguousarray(qb, np.uint32), mats.reshape(-1), len(mats))
        return OracleRays(self, r)

    def rays_frame(self, trees, mesh_of, cur16, should_cb):
        """Pairs of a frame that produce rays, with their rays (CollisionDetection.cpp:38-67)."""
        handles = (C.c_void_p * len(trees))(*[t.h for t in trees])
        r = self.lib.orc_rays_frame(handles, np.ascontiguousarray(mesh_of, np.uint32),
                                    np.ascontiguousarray(cur16, np.float32).reshape(-1),
                                    np.ascontiguousarray(should_cb, np.uint8), len(mesh_of))
        return OracleRays(self, r)


    # -- response: shooting the uncollide rays (ShootUncollideRays.cpp, Ray.cpp) --
    def ray_triangle(self, direction, tri9):
        """(hit, backface, distance, bary_u, bary_v) of a ray from the origin."""
        out = np.zeros(6, np.float32)
        self.lib.orc_ray_triangle(np.ascontiguousarray(direction, np.float32), np.ascontiguousarray(tri9, np.float32),
                                  out)
        return bool(out[0]), bool(out[1]), float(out[2]), float(out[3]), float(out[4])

    def ray_box(self, direction, box12):
        """(hit, t_enter, t_exit) of a ray from the origin through a box (centre, half-axes)."""
        out = np.zeros(2, np.float32)
        hit = self.lib.orc_ray_box(np.ascontiguousarray(direction, np.float32), np.ascontiguousarray(box12, np.float32),
                                   out)
        return bool(hit), float(out[0]), float(out[1])

    def ray_tree(self, tree, m16, origin, direction):
        """Closest hit (hit, backface, distance, tree-order triangle or -1, u, v) of a ray on a moved tree."""
        out = np.zeros(8, np.float32)
        self.lib.orc_ray_tree(tree.h, np.ascontiguousarray(m16, np.float32).reshape(-1),
                              np.ascontiguousarray(origin, np.float32), np.ascontiguousarray(direction, np.float32), out)
        return bool(out[0]), bool(out[1]), float(out[2]), int(out[3]), float(out[4]), float(out[5])

    def shoot_frame(self, trees, mesh_of, cur16, prev16, should_cb, cos_start, cos_finish, bias):
        """Per pair with rays: ids (first, second, shot) and deltas (world delta,
        first's share, second's share) (CollisionDetection.cpp:70-99), plus
        (passes, successful passes)."""
        handles = (C.c_void_p * len(trees))(*[t.h for t in trees])
        h = self.lib.orc_shoot_frame(handles, np.ascontiguousarray(mesh_of, np.uint32),
                                     np.ascontiguousarray(cur16, np.float32).reshape(-1),
                                     np.ascontiguousarray(prev16, np.float32).reshape(-1),
                                     np.ascontiguousarray(should_cb, np.uint8), len(mesh_of), cos_start, cos_finish, bias)
        n = int(self.lib.orc_shoot_get(h, None, None, None))
        ids = np.zeros((n, 3), np.uint32)
        d = np.zeros((n, 3, 3), np.float32)
        st = np.zeros(2, np.uint64)
        self.lib.orc_shoot_get(h, ids.ctypes.data_as(C.POINTER(C.c_uint32)), d.ctypes.data_as(C.POINTER(C.c_float)),
                               st.ctypes.data_as(C.POINTER(C.c_uint64)))
        self.lib.orc_shoot_free(h)
        return ids, d, (int(st[0]), int(st[1]))


class OracleRays:
    """pairs: (first, second, n_from_first, n_from_second); avg: (n, 2, 3);
    rays: per pair the first's rays then the second's, (origin, direction)."""

    def __init__(self, orc: Oracle, h):
        lib = orc.lib
        n, nr = int(lib.orc_rays_count(h, 0)), int(lib.orc_rays_count(h, 1))
        self.pairs = np.zeros((n, 4), np.uint32)
        self.avg = np.zeros((n, 2, 3), np.float32)
        self.rays = np.zeros((nr, 2, 3), np.float32)
        if n:
            lib.orc_rays_get(h, self.pairs.reshape(-1), self.avg.reshape(-1), self.rays.reshape(-1))
        lib.orc_rays_free(h)

    def split(self, i: int):
        """Rays from the first and from the second entity of pair i."""
        begin = int(self.pairs[:i, 2:].sum())
        nf, ns = int(self.pairs[i, 2]), int(self.pairs[i, 3])
        return self.rays[begin:begin + nf], self.rays[begin + nf:begin + nf + ns]


class OracleTree:
    def __init__(self, orc: Oracle, tri9, leaf_max: int):
        self.orc = orc
        self.tri9 = np.ascontiguousarray(tri9, np.float32).reshape(-1, 9)
        self.h = orc.lib.orc_tree_build(self.tri9.reshape(-1), len(self.tri9), leaf_max)
        self.leaf_max = leaf_max

    def __del__(self):
        h, self.h = getattr(self, "h", None), None
        if h:
            self.orc.lib.orc_tree_free(h)

    @property
    def nodes(self) -> int:
        return int(self.orc.lib.orc_tree_nodes(self.h))

    @property
    def depth(self) -> int:
        return int(self.orc.lib.orc_tree_depth(self.h))

    @property
    def slack(self) -> float:
        return float(self.orc.lib.orc_tree_slack(self.h))

    def set_attrs(self, normals9=None, vid3=None):
        """Corner normals (input order, 9 floats per triangle; None = face normals) and
        index-buffer vertex ids (3 per triangle; None = 3 t + k) for ray generation."""
        self._nrm = None if normals9 is None else np.ascontiguousarray(normals9, np.float32).reshape(-1, 9)
        self._vid = None if vid3 is None else np.ascontiguousarray(vid3, np.uint32).reshape(-1, 3)
        self.orc.lib.orc_tree_set_attrs(self.h, None if self._nrm is None else self._nrm.ctypes.data,
                                        None if self._vid is None else self._vid.ctypes.data)

    def export(self):
        n = self.nodes
        t = int(self.orc.lib.orc_tree_tris(self.h))
        root = np.zeros(12, np.float32)
        obbs = np.zeros((max(n, 1), 24), np.float32)
        links = np.zeros((max(n, 1), 4), np.uint32)
        perm = np.zeros(max(t, 1), np.uint32)
        self.orc.lib.orc_tree_export(self.h, root, obbs.reshape(-1), links.reshape(-1), perm)
        return {"root": root, "obbs": obbs[:n], "links": links[:n], "perm": perm[:t]}


class OracleResult:
    def __init__(self, orc: Oracle, h, nq):
        self.orc, self.h, self.nq = orc, h, nq
        lib = orc.lib
        nc = lib.orc_result_count(h, 0)
        self.contacts = np.zeros(nc, CONTACT_DTYPE)
        if nc:
            lib.orc_result_contacts(h, self.contacts.ctypes.data)
        nl = lib.orc_result_count(h, 1)
        self.leafpairs = np.zeros((nl, 4), np.uint32)
        self.lp_query = np.zeros(nl, np.uint32)
        if nl:
            lib.orc_result_leafpairs(h, self.leafpairs.reshape(-1), self.lp_query.ctypes.data)
        npairs = lib.orc_result_count(h, 2)
        self.pairs = np.zeros((npairs, 2), np.uint32)
        if npairs:
            lib.orc_result_pairs(h, self.pairs.reshape(-1))
        nrows = nq if nq is not None else npairs
        self.per_query = np.zeros((nrows, 4), np.uint64)
        if nrows:
            lib.orc_result_per_query(h, self.per_query.reshape(-1))
        self.total_sat = int(lib.orc_result_count(h, 3))
        self.total_leafpairs = int(lib.orc_result_count(h, 4))
        self.total_tritests = int(lib.orc_result_count(h, 5))
        self.total_ub = int(lib.orc_result_count(h, 6))
        lib.orc_result_free(h)
        self.h = None


class RefPieces:
    """The two reference pieces compiled from /root/reference (oracle/_ref)."""

    EIG3 = "_Z19eigen_decompositionPA3_dS0_Pd"
    MOLLER = "_Z32tri_tri_intersect_with_isectlinePfS_S_S_S_S_PiS_S_"

    def __init__(self):
        self.eig = _load(os.path.join("_ref", "libref_eig3.so"))
        self.mol = _load(os.path.join("_ref", "libref_moller.so"))
        self._eig = getattr(self.eig, self.EIG3)
        self._eig.argtypes = [_D, _D, _D]
        self._mol = getattr(self.mol, self.MOLLER)
        P = C.POINTER(C.c_float)
        self._mol.argtypes = [P, P, P, P, P, P, C.POINTER(C.c_int), P, P]
        self._mol.restype = C.c_int

    @staticmethod
    def available() -> bool:
        return all(os.path.exists(os.path.join(HERE, "_ref", f)) for f in ("libref_eig3.so", "libref_moller.so"))

    def eig3(self, A):
        A = np.ascontiguousarray(A, np.float64).reshape(9).copy()
        V = np.zeros(9, np.float64)
        d = np.zeros(3, np.float64)
        self._eig(A, V, d)
        return V.reshape(3, 3), d

    def tritri(self, a9, b9):
        a9 = np.ascontiguousarray(a9, np.float32).reshape(-1, 9)
        b9 = np.ascontiguousarray(b9, np.float32).reshape(-1, 9)
        n = len(a9)
        flags = np.zeros(n, np.int32)
        seg = np.zeros((n, 6), np.float32)
        P = C.POINTER(C.c_float)
        for i in range(n):
            a = a9[i].copy()
            b = b9[i].copy()
            s = np.zeros(3, np.float32)
            t = np.zeros(3, np.float32)
            cop = C.c_int(0)
            pa = a.ctypes.data_as(P)
            pb = b.ctypes.data_as(P)
            f4 = C.sizeof(C.c_float)
            args = [C.cast(C.c_void_p(a.ctypes.data + 3 * k * f4), P) for k in range(3)]
            args += [C.cast(C.c_void_p(b.ctypes.data + 3 * k * f4), P) for k in range(3)]
            hit = self._mol(*args, C.byref(cop), s.ctypes.data_as(P), t.ctypes.data_as(P))
            del pa, pb
            flags[i] = (1 if hit else 0) | (2 if cop.value else 0)
            if hit and not cop.value:
                seg[i, :3] = s
                seg[i, 3:] = t
        return flags, seg
