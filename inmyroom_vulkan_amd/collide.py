"""ctypes binding of the HIP collision library (include/mi355x_collide.h).

The library is the product; this module only marshals numpy arrays across
the C ABI.  Loading fails loudly when the HIP library is missing: there is
no CPU fallback anywhere behind this interface.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmi355x_collide.so")

CD_OK = 0
CD_E_INVALID = -1
CD_E_CAPACITY = -2
CD_E_HIP = -3
CD_E_STACK = -4
CD_E_NOMEM = -5
CD_E_FORMAT = -6
CD_NO_MESH = 0xFFFFFFFF
MODE_PLAIN = 0
MODE_BATCHED = 1
MODE_REFERENCE = 2  # the plain mode's tests and results on the batched mode's machinery

# Symbols the header declares (checked by the CPU tests).
EXPORTS = (
    "cd_create", "cd_destroy", "cd_last_error", "cd_register_mesh", "cd_register_soups", "cd_mesh_info_get",
    "cd_export_tree", "cd_tritri_batch", "cd_query_batch", "cd_query_leafpairs", "cd_frame", "cd_broad_phase",
    "cd_stage_queries", "cd_run_staged", "cd_fetch_contacts", "cd_get_counters", "cd_get_timings",
    "cd_set_tuning", "cd_get_tuning", "cd_export_triangles", "cd_register_skinned", "cd_skin_refit",
    "cd_uncollide_rays", "cd_set_mesh_normals", "cd_shoot_rays", "cd_gltf_open", "cd_gltf_close",
    "cd_gltf_mesh_count", "cd_gltf_mesh_info_get", "cd_gltf_primitive_get", "cd_register_gltf",
    "cd_node_global_matrices", "cd_gltf_node_count", "cd_gltf_nodes", "cd_tritri_batch_noalloc", "cd_step",
)
CD_NOALLOC_MAX_PAIRS = 1 << 20

CONTACT_DTYPE = np.dtype([
    ("query", "<u4"), ("tri_a", "<u4"), ("tri_b", "<u4"), ("flags", "<u4"),
    ("s", "<f4", 3), ("t", "<f4", 3), ("na", "<f4", 3), ("nb", "<f4", 3),
])
QUERY_DTYPE = np.dtype([("mesh_a", "<u4"), ("mesh_b", "<u4"), ("rel", "<f4", 16)])
ENTITY_DTYPE = np.dtype([
    ("current", "<f4", 16), ("previous", "<f4", 16), ("mesh_id", "<u4"), ("entity", "<u2"),
    ("should_callback", "u1"), ("pad", "u1"),
])
PAIR_DTYPE = np.dtype([("first", "<u4"), ("second", "<u4"), ("contact_begin", "<u4"), ("contact_count", "<u4")])
RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("direction", "<f4", 3)])
PAIR_RAYS_DTYPE = np.dtype([
    ("query", "<u4"), ("first", "<u4"), ("second", "<u4"), ("n_first", "<u4"), ("n_second", "<u4"), ("pad", "<u4"),
    ("begin_first", "<u8"), ("begin_second", "<u8"), ("avg_first", "<f4", 3), ("avg_second", "<f4", 3),
])
PAIR_RESPONSE_DTYPE = np.dtype([
    ("first", "<u4"), ("second", "<u4"), ("shot", "<u4"), ("pad", "<u4"), ("delta", "<f4", 3),
    ("delta_first", "<f4", 3), ("delta_second", "<f4", 3), ("pad2", "<f4", 3),
])
NODE_DTYPE = np.dtype([
    ("local_scale", "<f4", 3), ("local_rotation", "<f4", 4), ("local_translation", "<f4", 3),
    ("global_translation", "<f4", 3), ("parent", "<u4"), ("pad", "<u4", 2),
])
CD_NO_PARENT = 0xFFFFFFFF
assert CONTACT_DTYPE.itemsize == 64 and QUERY_DTYPE.itemsize == 72 and ENTITY_DTYPE.itemsize == 136
assert NODE_DTYPE.itemsize == 64
assert RAY_DTYPE.itemsize == 24 and PAIR_RAYS_DTYPE.itemsize == 64 and PAIR_RESPONSE_DTYPE.itemsize == 64


class cd_shoot_params(C.Structure):
    _fields_ = [("smoothstep_start_rads", C.c_float), ("smoothstep_finish_rads", C.c_float),
                ("distance_bias", C.c_float)]


class cd_config(C.Structure):
    _fields_ = [("leaf_max", C.c_uint32), ("device", C.c_uint32), ("reserved", C.c_uint32 * 6)]


class cd_mesh_info(C.Structure):
    _fields_ = [("ntri", C.c_uint32), ("nnodes", C.c_uint32), ("depth", C.c_uint32), ("leaf_max", C.c_uint32),
                ("root", C.c_float * 12), ("slack", C.c_float), ("radius", C.c_float)]


class cd_counters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("sat_tests", "leaf_pairs", "tri_tests", "contacts", "fp64_fallbacks",
                                            "exact_fallbacks", "ub_cases", "broad_pairs")]


class cd_timings(C.Structure):
    _fields_ = [("total_ms", C.c_float), ("traverse_ms", C.c_float), ("narrow_ms", C.c_float),
                ("broad_ms", C.c_float), ("launches", C.c_uint32), ("pad", C.c_uint32)]


class cd_tuning(C.Structure):
    _fields_ = [("split_levels", C.c_uint32), ("stack_entries", C.c_uint32), ("pop_width", C.c_uint32),
                ("waves_per_cu", C.c_uint32), ("use_graph", C.c_uint32), ("work_sharing", C.c_uint32),
                ("spin_polls", C.c_uint32), ("split_tris", C.c_uint32), ("handoff_pct", C.c_uint32),
                ("resplit_levels", C.c_uint32), ("pipeline", C.c_uint32), ("split_depth", C.c_uint32),
                ("step_graph", C.c_uint32)]


class cd_gltf_mesh_info(C.Structure):
    _fields_ = [("nprim", C.c_uint32), ("with_normals", C.c_uint32), ("without_normals", C.c_uint32),
                ("pad", C.c_uint32), ("nvert", C.c_uint64), ("ntri", C.c_uint64)]


class cd_gltf_primitive(C.Structure):
    _fields_ = [("gltf_mode", C.c_int32), ("mode", C.c_uint32), ("excluded", C.c_uint32), ("nvert", C.c_uint32),
                ("nindex", C.c_uint64), ("ntri", C.c_uint64), ("positions", C.POINTER(C.c_float)),
                ("normals", C.POINTER(C.c_float)), ("indices", C.POINTER(C.c_uint32))]


class CollideError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{msg} (status {code})")
        self.code = code


_lib = None


def load_library() -> C.CDLL:
    """Load the built HIP library; raise if it is not there."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"HIP collision library not built: {LIB_PATH} (run python -m inmyroom_vulkan_amd.build)")
    lib = C.CDLL(LIB_PATH)
    vp, u32, u64 = C.c_void_p, C.c_uint32, C.c_uint64
    P = C.POINTER
    lib.cd_create.argtypes = [P(vp), P(cd_config)]
    lib.cd_destroy.argtypes = [vp]
    lib.cd_last_error.argtypes = [vp]
    lib.cd_last_error.restype = C.c_char_p
    lib.cd_register_mesh.argtypes = [vp, vp, u32, vp, u32, P(u32)]
    lib.cd_register_soups.argtypes = [vp, vp, vp, u32, P(u32)]
    lib.cd_mesh_info_get.argtypes = [vp, u32, P(cd_mesh_info)]
    lib.cd_export_tree.argtypes = [vp, u32, vp, vp, vp, vp]
    lib.cd_tritri_batch.argtypes = [vp, vp, vp, u64, u32, vp, vp]
    lib.cd_tritri_batch_noalloc.argtypes = [vp, vp, vp, u64, u32, vp, vp]
    lib.cd_query_batch.argtypes = [vp, vp, u32, u32, vp, u64, P(u64)]
    lib.cd_query_leafpairs.argtypes = [vp, vp, u32, vp, u64, P(u64)]
    lib.cd_frame.argtypes = [vp, vp, u32, u32, vp, u32, P(u32), vp, u64, P(u64)]
    lib.cd_broad_phase.argtypes = [vp, vp, u32, vp, u32, P(u32)]
    lib.cd_stage_queries.argtypes = [vp, vp, u32]
    lib.cd_run_staged.argtypes = [vp, u32, P(u64)]
    lib.cd_fetch_contacts.argtypes = [vp, vp, u64, P(u64)]
    lib.cd_uncollide_rays.argtypes = [vp, vp, u32, P(u32), vp, u64, P(u64)]
    lib.cd_shoot_rays.argtypes = [vp, P(cd_shoot_params), vp, u32, P(u32)]
    lib.cd_step.argtypes = [vp, vp, u32, u32, P(cd_shoot_params), vp, u32, P(u32)]
    lib.cd_set_mesh_normals.argtypes = [vp, u32, vp, u32]
    lib.cd_get_counters.argtypes = [vp, P(cd_counters)]
    lib.cd_get_timings.argtypes = [vp, P(cd_timings)]
    lib.cd_set_tuning.argtypes = [vp, P(cd_tuning)]
    lib.cd_export_triangles.argtypes = [vp, u32, vp]
    lib.cd_register_skinned.argtypes = [vp, vp, u32, vp, u32, vp, vp, u32, P(u32)]
    lib.cd_skin_refit.argtypes = [vp, vp, u32, vp]
    lib.cd_get_tuning.argtypes = [vp, P(cd_tuning)]
    lib.cd_gltf_open.argtypes = [C.c_char_p, P(vp), C.c_char_p, u32]
    lib.cd_gltf_close.argtypes = [vp]
    lib.cd_gltf_mesh_count.argtypes = [vp, P(u32)]
    lib.cd_gltf_mesh_info_get.argtypes = [vp, u32, P(cd_gltf_mesh_info)]
    lib.cd_gltf_primitive_get.argtypes = [vp, u32, u32, P(cd_gltf_primitive)]
    lib.cd_register_gltf.argtypes = [vp, vp, vp, u32]
    lib.cd_node_global_matrices.argtypes = [vp, vp, u32, vp]
    lib.cd_gltf_node_count.argtypes = [vp, P(u32)]
    lib.cd_gltf_nodes.argtypes = [vp, vp, vp, vp, u32]
    for name in EXPORTS:
        if name != "cd_last_error":
            getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


class GltfDocument:
    """A parsed .gltf / .glb file (cd_gltf_open); host memory only, no GPU needed."""

    def __init__(self, path: str):
        self.lib = load_library()
        h = C.c_void_p()
        err = C.create_string_buffer(1024)
        rc = self.lib.cd_gltf_open(os.fsencode(path), C.byref(h), err, len(err))
        if rc != CD_OK:
            raise CollideError(rc, f"cd_gltf_open: {err.value.decode(errors='replace')}")
        self.h = h
        self.path = path

    def close(self):
        if getattr(self, "h", None):
            self.lib.cd_gltf_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __len__(self) -> int:
        n = C.c_uint32()
        if self.lib.cd_gltf_mesh_count(self.h, C.byref(n)) != CD_OK:
            raise CollideError(CD_E_INVALID, "cd_gltf_mesh_count")
        return n.value

    def mesh_info(self, mesh: int) -> dict:
        info = cd_gltf_mesh_info()
        if self.lib.cd_gltf_mesh_info_get(self.h, mesh, C.byref(info)) != CD_OK:
            raise CollideError(CD_E_INVALID, f"cd_gltf_mesh_info_get: no mesh {mesh}")
        return {k: getattr(info, k) for k in ("nprim", "with_normals", "without_normals", "nvert", "ntri")}

    def nodes(self):
        """(NODE_DTYPE array, glTF mesh per node or -1, collides per node) as the nodes are imported."""
        n = C.c_uint32()
        if self.lib.cd_gltf_node_count(self.h, C.byref(n)) != CD_OK:
            raise CollideError(CD_E_INVALID, "cd_gltf_node_count")
        nodes = np.zeros(n.value, NODE_DTYPE)
        mesh = np.zeros(n.value, np.int32)
        coll = np.zeros(n.value, np.uint32)
        rc = self.lib.cd_gltf_nodes(self.h, _ptr(nodes), _ptr(mesh), _ptr(coll), n.value)
        if rc != CD_OK:
            raise CollideError(rc, "cd_gltf_nodes")
        return nodes, mesh, coll.astype(bool)

    def primitives(self, mesh: int) -> list[dict]:
        """The mesh's primitives in concatenation order, as numpy copies."""
        out = []
        for k in range(self.mesh_info(mesh)["nprim"]):
            p = cd_gltf_primitive()
            if self.lib.cd_gltf_primitive_get(self.h, mesh, k, C.byref(p)) != CD_OK:
                raise CollideError(CD_E_INVALID, f"cd_gltf_primitive_get: mesh {mesh} primitive {k}")
            arr = lambda ptr, n, dt: (np.ctypeslib.as_array(ptr, (n,)).astype(dt, copy=True) if n and ptr
                                      else np.zeros(0, dt))
            rec = {"gltf_mode": p.gltf_mode, "mode": p.mode, "excluded": bool(p.excluded), "nvert": p.nvert,
                   "ntri": p.ntri, "positions": None, "normals": None, "indices": None}
            if not p.excluded:
                rec["positions"] = arr(p.positions, 3 * p.nvert, np.float32).reshape(-1, 3)
                rec["indices"] = arr(p.indices, p.nindex, np.uint32)
                if p.normals:
                    rec["normals"] = arr(p.normals, 3 * p.nvert, np.float32).reshape(-1, 3)
            out.append(rec)
        return out


def _grown(n: int) -> int:
    """A capacity for n records with a quarter more room, so that the next
    slightly larger frame does not call again."""
    return int(n) + int(n) // 4


class Collider:
    """One library context (cd_ctx) on one GPU."""

    def __init__(self, leaf_max: int = 4, device: int = 0):
        self.lib = load_library()
        cfg = cd_config(leaf_max=leaf_max, device=device)
        h = C.c_void_p()
        rc = self.lib.cd_create(C.byref(h), C.byref(cfg))
        if rc != CD_OK:
            raise CollideError(rc, "cd_create failed (no HIP device?)")
        self.h = h
        self.leaf_max = leaf_max
        self._caps = {}  # output capacities of the last calls: a frame like the last needs no second call
        self._bufs = {}  # output buffers kept across calls

    def close(self):
        if getattr(self, "h", None):
            self.lib.cd_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int, what: str):
        if rc != CD_OK:
            raise CollideError(rc, f"{what}: {self.lib.cd_last_error(self.h).decode()}")

    # -- meshes ---------------------------------------------------------------
    def register_mesh(self, positions, indices) -> int:
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        idx = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
        mid = C.c_uint32()
        self._check(self.lib.cd_register_mesh(self.h, _ptr(pos), len(pos), _ptr(idx), len(idx), C.byref(mid)),
                    "cd_register_mesh")
        return mid.value

    def register_soups(self, soups) -> list[int]:
        """soups: list of (T, 9) float32 triangle arrays.  Returns their mesh ids."""
        counts = np.array([len(s) for s in soups], np.uint32)
        tri = np.ascontiguousarray(np.concatenate([np.asarray(s, np.float32).reshape(-1, 9) for s in soups])
                                   if len(soups) else np.zeros((0, 9), np.float32))
        first = C.c_uint32()
        self._check(self.lib.cd_register_soups(self.h, _ptr(tri), _ptr(counts), len(counts), C.byref(first)),
                    "cd_register_soups")
        return list(range(first.value, first.value + len(soups)))

    def register_gltf(self, doc: GltfDocument) -> np.ndarray:
        """Trees of every mesh of a parsed document; CD_NO_MESH where a mesh has no collision triangles."""
        ids = np.zeros(max(1, len(doc)), np.uint32)
        self._check(self.lib.cd_register_gltf(self.h, doc.h, _ptr(ids), len(ids)), "cd_register_gltf")
        return ids[: len(doc)]

    def node_globals(self, nodes: np.ndarray) -> np.ndarray:
        """Global matrices (n, 16) of a node hierarchy (NODE_DTYPE records)."""
        nd = np.ascontiguousarray(nodes, NODE_DTYPE)
        out = np.zeros((len(nd), 16), np.float32)
        self._check(self.lib.cd_node_global_matrices(self.h, _ptr(nd), len(nd), _ptr(out)), "cd_node_global_matrices")
        return out

    def mesh_info(self, mesh: int) -> dict:
        info = cd_mesh_info()
        self._check(self.lib.cd_mesh_info_get(self.h, mesh, C.byref(info)), "cd_mesh_info_get")
        return {"ntri": info.ntri, "nnodes": info.nnodes, "depth": info.depth, "leaf_max": info.leaf_max,
                "root": np.array(info.root, np.float32), "slack": info.slack, "radius": info.radius}

    def export_tree(self, mesh: int) -> dict:
        info = self.mesh_info(mesh)
        n, t = info["nnodes"], info["ntri"]
        root = np.zeros(12, np.float32)
        obbs = np.zeros((max(n, 1), 24), np.float32)
        links = np.zeros((max(n, 1), 4), np.uint32)
        perm = np.zeros(max(t, 1), np.uint32)
        self._check(self.lib.cd_export_tree(self.h, mesh, _ptr(root), _ptr(obbs), _ptr(links), _ptr(perm)),
                    "cd_export_tree")
        return {"root": root, "obbs": obbs[:n], "links": links[:n], "perm": perm[:t], "depth": info["depth"]}

    def export_triangles(self, mesh: int) -> np.ndarray:
        n = self.mesh_info(mesh)["ntri"]
        out = np.zeros((max(n, 1), 9), np.float32)
        self._check(self.lib.cd_export_triangles(self.h, mesh, _ptr(out)), "cd_export_triangles")
        return out[:n]

    def register_skinned(self, positions, indices, joints, weights, nbones: int) -> int:
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        idx = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
        jt = np.ascontiguousarray(joints, np.uint16).reshape(-1, 4)
        wt = np.ascontiguousarray(weights, np.float32).reshape(-1, 4)
        mid = C.c_uint32()
        self._check(self.lib.cd_register_skinned(self.h, _ptr(pos), len(pos), _ptr(idx), len(idx), _ptr(jt), _ptr(wt),
                                                 nbones, C.byref(mid)), "cd_register_skinned")
        return mid.value

    def skin_refit(self, mesh_ids, bone_matrices):
        ids = np.ascontiguousarray(mesh_ids, np.uint32)
        bones = np.ascontiguousarray(bone_matrices, np.float32)
        self._check(self.lib.cd_skin_refit(self.h, _ptr(ids), len(ids), _ptr(bones)), "cd_skin_refit")

    # -- queries --------------------------------------------------------------
    def tritri(self, a9, b9, mode: int = MODE_PLAIN):
        a9 = np.ascontiguousarray(a9, np.float32).reshape(-1, 9)
        b9 = np.ascontiguousarray(b9, np.float32).reshape(-1, 9)
        n = len(a9)
        flags = np.zeros(n, np.int32)
        seg = np.zeros((n, 6), np.float32)
        self._check(self.lib.cd_tritri_batch(self.h, _ptr(a9), _ptr(b9), n, mode, _ptr(flags), _ptr(seg)),
                    "cd_tritri_batch")
        return flags, seg

    @staticmethod
    def make_queries(mesh_a, mesh_b, rel) -> np.ndarray:
        rel = np.asarray(rel, np.float32).reshape(-1, 16)
        q = np.zeros(len(rel), QUERY_DTYPE)
        q["mesh_a"] = mesh_a
        q["mesh_b"] = mesh_b
        q["rel"] = rel
        return q

    def query_batch(self, queries: np.ndarray, mode: int = MODE_PLAIN, cap: int | None = None) -> np.ndarray:
        queries = np.ascontiguousarray(queries, QUERY_DTYPE)
        cap = cap if cap is not None else 1 << 16
        while True:
            out = np.empty(cap, CONTACT_DTYPE)  # the call fills the prefix that is returned
            n = C.c_uint64()
            rc = self.lib.cd_query_batch(self.h, _ptr(queries), len(queries), mode, _ptr(out), cap, C.byref(n))
            if rc == CD_E_CAPACITY:
                cap = int(n.value)
                continue
            self._check(rc, "cd_query_batch")
            return out[: n.value]

    def query_leafpairs(self, queries: np.ndarray) -> np.ndarray:
        queries = np.ascontiguousarray(queries, QUERY_DTYPE)
        cap = 1 << 16
        while True:
            out = np.zeros((cap, 4), np.uint32)
            n = C.c_uint64()
            rc = self.lib.cd_query_leafpairs(self.h, _ptr(queries), len(queries), _ptr(out), cap, C.byref(n))
            if rc == CD_E_CAPACITY:
                cap = int(n.value)
                continue
            self._check(rc, "cd_query_leafpairs")
            return out[: n.value]

    def stage(self, queries: np.ndarray):
        queries = np.ascontiguousarray(queries, QUERY_DTYPE)
        self._check(self.lib.cd_stage_queries(self.h, _ptr(queries), len(queries)), "cd_stage_queries")

    def run_staged(self, mode: int = MODE_PLAIN) -> int:
        n = C.c_uint64()
        self._check(self.lib.cd_run_staged(self.h, mode, C.byref(n)), "cd_run_staged")
        return int(n.value)

    def fetch(self) -> np.ndarray:
        n = C.c_uint64()
        rc = self.lib.cd_fetch_contacts(self.h, None, 0, C.byref(n))
        if rc not in (CD_OK, CD_E_CAPACITY):
            self._check(rc, "cd_fetch_contacts")
        out = np.zeros(int(n.value), CONTACT_DTYPE)
        self._check(self.lib.cd_fetch_contacts(self.h, _ptr(out), len(out), C.byref(n)), "cd_fetch_contacts")
        return out

    def frame(self, entities: np.ndarray, mode: int = MODE_PLAIN):
        ents = np.ascontiguousarray(entities, ENTITY_DTYPE)
        pcap, ccap = self._caps.get("frame", (1 << 12, 1 << 16))
        while True:
            pairs = np.empty(pcap, PAIR_DTYPE)  # the call fills the prefix that is returned
            cons = np.empty(ccap, CONTACT_DTYPE)
            npairs, ncons = C.c_uint32(), C.c_uint64()
            rc = self.lib.cd_frame(self.h, _ptr(ents), len(ents), mode, _ptr(pairs), pcap, C.byref(npairs),
                                   _ptr(cons), ccap, C.byref(ncons))
            if rc == CD_E_CAPACITY:
                pcap = max(pcap, _grown(npairs.value))
                ccap = max(ccap, _grown(ncons.value))
                continue
            self._check(rc, "cd_frame")
            self._caps["frame"] = (pcap, ccap)
            return pairs[: npairs.value], cons[: ncons.value]

    def uncollide_rays(self):
        """Rays of the last run's contacts: (pairs [PAIR_RAYS_DTYPE], rays [RAY_DTYPE])."""
        pcap, rcap = self._caps.get("rays", (1 << 10, 1 << 14))
        while True:
            pairs = np.empty(pcap, PAIR_RAYS_DTYPE)  # the call fills the prefix that is returned
            rays = np.empty(rcap, RAY_DTYPE)
            npairs, nrays = C.c_uint32(), C.c_uint64()
            rc = self.lib.cd_uncollide_rays(self.h, _ptr(pairs), pcap, C.byref(npairs), _ptr(rays), rcap,
                                            C.byref(nrays))
            if rc == CD_E_CAPACITY:
                pcap = max(pcap, _grown(npairs.value))
                rcap = max(rcap, _grown(nrays.value))
                continue
            self._check(rc, "cd_uncollide_rays")
            self._caps["rays"] = (pcap, rcap)
            return pairs[: npairs.value], rays[: nrays.value]

    def shoot_rays(self, start_rads=None, finish_rads=None, bias=None):
        """Response of the last frame's pairs with rays (PAIR_RESPONSE_DTYPE);
        after frame() and uncollide_rays().  No arguments: the reference's
        smoother-step angles (40, 65 degrees) and distance bias (1.01)."""
        p = None
        if start_rads is not None or finish_rads is not None or bias is not None:
            p = C.byref(cd_shoot_params(math.radians(40.0) if start_rads is None else start_rads,
                                        math.radians(65.0) if finish_rads is None else finish_rads,
                                        1.01 if bias is None else bias))
        cap = 1 << 10
        while True:
            out = np.zeros(cap, PAIR_RESPONSE_DTYPE)
            n = C.c_uint32()
            rc = self.lib.cd_shoot_rays(self.h, p, _ptr(out), cap, C.byref(n))
            if rc == CD_E_CAPACITY:
                cap = max(cap, int(n.value))
                continue
            self._check(rc, "cd_shoot_rays")
            return out[: n.value]

    def step(self, entities: np.ndarray, mode: int = MODE_REFERENCE, start_rads=None, finish_rads=None, bias=None):
        """One whole scene step (cd_step): frame, uncollide rays and response
        in one call, contacts and rays kept on the device; the responses of
        the pairs with rays (PAIR_RESPONSE_DTYPE), as shoot_rays() gives them."""
        ents = np.ascontiguousarray(entities, ENTITY_DTYPE)
        p = None
        if start_rads is not None or finish_rads is not None or bias is not None:
            p = C.byref(cd_shoot_params(math.radians(40.0) if start_rads is None else start_rads,
                                        math.radians(65.0) if finish_rads is None else finish_rads,
                                        1.01 if bias is None else bias))
        cap = self._caps.get("step", 1 << 10)
        out = self._bufs.get("step")
        if out is None or len(out) < cap:
            out = self._bufs["step"] = np.zeros(cap, PAIR_RESPONSE_DTYPE)
        n = C.c_uint32()
        rc = self.lib.cd_step(self.h, _ptr(ents), len(ents), mode, p, _ptr(out), cap, C.byref(n))
        if rc == CD_E_CAPACITY:  # the frame and its rays are done: a larger buffer finishes the step
            cap = _grown(int(n.value))
            out = self._bufs["step"] = np.zeros(cap, PAIR_RESPONSE_DTYPE)
            rc = self.lib.cd_shoot_rays(self.h, p, _ptr(out), cap, C.byref(n))
            self._caps["step"] = cap
        self._check(rc, "cd_step")
        return out[: n.value].copy()

    def set_mesh_normals(self, mesh: int, normals: np.ndarray) -> None:
        n = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
        self._check(self.lib.cd_set_mesh_normals(self.h, int(mesh), _ptr(n), len(n)), "cd_set_mesh_normals")

    def broad_phase(self, entities: np.ndarray) -> np.ndarray:
        ents = np.ascontiguousarray(entities, ENTITY_DTYPE)
        cap = max(1024, 4 * len(ents))
        while True:
            out = np.zeros((cap, 2), np.uint32)
            n = C.c_uint32()
            rc = self.lib.cd_broad_phase(self.h, _ptr(ents), len(ents), _ptr(out), cap, C.byref(n))
            if rc == CD_E_CAPACITY:
                cap = int(n.value)
                continue
            self._check(rc, "cd_broad_phase")
            return out[: n.value]

    def counters(self) -> dict:
        c = cd_counters()
        self._check(self.lib.cd_get_counters(self.h, C.byref(c)), "cd_get_counters")
        return {f: getattr(c, f) for f, _ in cd_counters._fields_}

    def set_tuning(self, split_levels=0, stack_entries=0, pop_width=0, waves_per_cu=0, use_graph=0, work_sharing=0,
                   spin_polls=0, split_tris=0, handoff_pct=0, resplit_levels=0, pipeline=0, split_depth=0,
                   step_graph=0):
        t = cd_tuning(split_levels=split_levels, stack_entries=stack_entries, pop_width=pop_width,
                      waves_per_cu=waves_per_cu, use_graph=use_graph, work_sharing=work_sharing,
                      spin_polls=spin_polls, split_tris=split_tris, handoff_pct=handoff_pct,
                      resplit_levels=resplit_levels, pipeline=pipeline, split_depth=split_depth,
                      step_graph=step_graph)
        self._check(self.lib.cd_set_tuning(self.h, C.byref(t)), "cd_set_tuning")

    def tuning(self) -> dict:
        t = cd_tuning()
        self._check(self.lib.cd_get_tuning(self.h, C.byref(t)), "cd_get_tuning")
        return {f: getattr(t, f) for f, _ in cd_tuning._fields_}

    def timings(self) -> dict:
        t = cd_timings()
        self._check(self.lib.cd_get_timings(self.h, C.byref(t)), "cd_get_timings")
        return {f: getattr(t, f) for f, _ in cd_timings._fields_ if f != "pad"}
