"""TEST INFRASTRUCTURE ONLY: the oracle's restatement of the reference's
glTF mesh ingest, in numpy.

Only ``tests/`` may import this module, and only as the checker; the shipped
library never does.  It restates, from inMyRoom_vulkan/src:

  - tinygltf's reading of the document (the tinygltf submodule of the
    reference, a fork of syoyo/tinygltf, is an empty directory at an
    unrecorded commit; what is used of it is the glTF 2.0 data model:
    buffers, buffer views, accessors, meshes, primitives, attributes) and
    PrimitiveInitializationData's reads of POSITION, NORMAL and the index
    list (Graphics/Meshes/PrimitivesOfMeshes.cpp:43-137, 528-552), through
    GetAccessorBeginEndPtrs (:695-752), which takes count x element size
    bytes from accessor.byteOffset + bufferView.byteOffset and never reads
    byteStride;
  - the primitive filter of GetPrimitiveOBBtreeData (:637-671): skinned
    (JOINTS_0) or position-morphed primitives give nothing;
  - the concatenation order of MeshesOfNodes::AddMeshesOfModel
    (Graphics/Meshes/MeshesOfNodes.cpp:41-43), as the insertion sort that
    libstdc++'s std::sort runs on lists of up to 16 elements, step by step;
  - Triangle::CreateIndicesTriplets for every draw mode
    (Geometry/Triangle.cpp:10-62) and CreateTriangleList (:259-280).

Where the reference's result is undefined the oracle takes the library's
choice, so that the two can be compared there at all: a primitive without
indices gets 0 .. nvert-1 (the reference sizes that list by its 4-float
position array, PrimitivesOfMeshes.cpp:660-663), and strips and fans of
fewer than two indices give no triangles (size() - 2 wraps around).
"""
from __future__ import annotations

import base64
import json
import os
import struct
import urllib.parse

import numpy as np

POINTS, LINES, LINE_LOOP, LINE_STRIP, TRIANGLES, TRIANGLE_STRIP, TRIANGLE_FAN = range(7)
_COMP = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4}


def read_document(path: str):
    """(json dict, list of buffer bytes) of a .gltf or .glb file."""
    raw = open(path, "rb").read()
    base = os.path.dirname(path)
    bin_chunk = None
    if raw[:4] == b"glTF":
        _, _, total = struct.unpack_from("<III", raw, 0)
        off, doc = 12, None
        while off + 8 <= total:
            n, kind = struct.unpack_from("<II", raw, off)
            chunk = raw[off + 8: off + 8 + n]
            if doc is None:
                doc = json.loads(chunk.decode("utf-8"))
            elif kind == 0x004E4942 and bin_chunk is None:
                bin_chunk = chunk
            off += 8 + n
    else:
        doc = json.loads(raw.decode("utf-8"))
    bufs = []
    for i, b in enumerate(doc.get("buffers", [])):
        uri = b.get("uri")
        if uri is None:
            data = bin_chunk[: b["byteLength"]]
        elif uri.startswith("data:"):
            data = base64.b64decode(uri.split(",", 1)[1])
        else:
            data = open(os.path.join(base, urllib.parse.unquote(uri)), "rb").read()
        bufs.append(data)
    return doc, bufs


def accessor(doc, bufs, index: int) -> np.ndarray:
    """count x ncomp elements, packed from the accessor's start
    (PrimitivesOfMeshes.cpp:695-752)."""
    a = doc["accessors"][index]
    bv = doc["bufferViews"][a["bufferView"]]
    dt = np.dtype(_COMP[a["componentType"]]).newbyteorder("<")
    nc = _NCOMP[a["type"]]
    start = bv.get("byteOffset", 0) + a.get("byteOffset", 0)
    buf = bufs[bv["buffer"]]
    return np.frombuffer(buf, dt, count=a["count"] * nc, offset=start).reshape(a["count"], nc)


def _flip(v: np.ndarray) -> np.ndarray:
    """(x, y, z) -> (x, -y, -z) (PrimitivesOfMeshes.cpp:78-80, 129-131)."""
    return (v.astype(np.float32) * np.float32([1, -1, -1])).astype(np.float32)


def primitive(doc, bufs, p: dict) -> dict:
    mode = p.get("mode", -1)
    draw = TRIANGLES if mode == -1 else (LINE_STRIP if mode == LINE_LOOP else mode)
    attrs = p["attributes"]
    targets = p.get("targets", [])
    excluded = "JOINTS_0" in attrs or (len(targets) > 0 and "POSITION" in targets[0])
    out = {"gltf_mode": mode, "mode": draw, "excluded": excluded}
    if excluded:
        return out
    pos = _flip(accessor(doc, bufs, attrs["POSITION"]))
    nrm = _flip(accessor(doc, bufs, attrs["NORMAL"])) if "NORMAL" in attrs else None
    if "indices" in p:
        idx = accessor(doc, bufs, p["indices"])[:, 0].astype(np.uint32)
    else:
        idx = np.arange(len(pos), dtype=np.uint32)
    out.update(positions=pos, normals=nrm, indices=idx)
    return out


def insertion_order(gltf_modes) -> list[int]:
    """libstdc++'s __insertion_sort under the comparator of
    MeshesOfNodes.cpp:42 (lhs is triangles-mode or has no mode), run element
    by element: an element the comparator puts before the first goes to the
    front, any other is inserted by a backward scan that stops at the first
    element it is not less than."""
    less = lambda lhs, rhs: gltf_modes[lhs] == TRIANGLES or gltf_modes[lhs] == -1
    seq = list(range(len(gltf_modes)))
    for i in range(1, len(seq)):
        val = seq[i]
        if less(val, seq[0]):
            seq[1: i + 1] = seq[0:i]
            seq[0] = val
        else:
            j = i
            while less(val, seq[j - 1]):
                seq[j] = seq[j - 1]
                j -= 1
            seq[j] = val
    return seq


def triplets(mode: int, n: int) -> np.ndarray:
    """Index-list positions of each triangle (Triangle.cpp:14-56)."""
    rows = []
    if mode == POINTS:
        rows = [(i, i, i) for i in range(n)]
    elif mode == LINES:
        rows = [(2 * i, 2 * i, 2 * i + 1) for i in range(n // 2)]
    elif mode == LINE_STRIP:
        rows = [(i, i, i + 1) for i in range(max(0, n - 1))]
    elif mode == TRIANGLES:
        rows = [(3 * i, 3 * i + 1, 3 * i + 2) for i in range(n // 3)]
    elif mode == TRIANGLE_STRIP:
        rows = [(i, i + (1 + i % 2), i + (2 - i % 2)) for i in range(max(0, n - 2))]
    elif mode == TRIANGLE_FAN:
        rows = [(i + 1, i + 2, 0) for i in range(max(0, n - 2))]
    return np.array(rows, np.int64).reshape(-1, 3)


def load(path: str) -> list[list[dict]]:
    """Per glTF mesh, its primitives in concatenation order."""
    doc, bufs = read_document(path)
    meshes = []
    for m in doc.get("meshes", []):
        prims = [primitive(doc, bufs, p) for p in m["primitives"]]
        order = insertion_order([p["gltf_mode"] for p in prims])
        meshes.append([prims[i] for i in order])
    return meshes


def mesh_triangles(prims: list[dict]):
    """CreateTriangleList over the included primitives, concatenated
    (PrimitivesOfMeshes.cpp:840-863): positions (T, 9), vertex normals
    (T, 9) or None, and index triplets (T, 3) as the reference keeps them
    (raw per-primitive indices) and with each primitive's first-vertex
    offset within the mesh added."""
    P, N, I, Io = [], [], [], []
    vbase = 0
    for p in prims:
        if p["excluded"]:
            continue
        t = triplets(p["mode"], len(p["indices"]))
        ids = p["indices"][t] if len(t) else np.zeros((0, 3), np.uint32)
        P.append(p["positions"][ids].reshape(-1, 9))
        if p["normals"] is not None:
            N.append(p["normals"][ids].reshape(-1, 9))
        I.append(ids.astype(np.uint32))
        Io.append((ids + vbase).astype(np.uint32))
        vbase += len(p["positions"])
    if not P:
        z = np.zeros((0, 9), np.float32)
        return z, None, np.zeros((0, 3), np.uint32), np.zeros((0, 3), np.uint32)
    normals = np.concatenate(N) if len(N) == len(P) else None
    return np.concatenate(P), normals, np.concatenate(I), np.concatenate(Io)


# ---------------------------------------------------------------------------
# Nodes: GameImporter::CreatePositionInitMap (src/GameImporter.cpp:503-552)
# then NodeDataCompEntity::CreateComponentEntityByMap
# (ECS/GeneralCompEntities/NodeDataCompEntity.cpp:19-61), in float32 scalar
# operations.  glm (an empty submodule of the reference) supplies
# glm::decompose, glm::normalize and the quaternion dot; they are restated
# from glm 0.9.9's published source order, so the matrix-node path and the
# normalization are pinned by round trips and invariants, not by glm itself.
# ---------------------------------------------------------------------------
_f = np.float32


def _mat_mul(A, B):
    R = [_f(0)] * 16
    for c in range(4):
        for r in range(4):
            t0 = _f(A[0 * 4 + r] * B[c * 4 + 0])
            t1 = _f(A[1 * 4 + r] * B[c * 4 + 1])
            t2 = _f(A[2 * 4 + r] * B[c * 4 + 2])
            t3 = _f(A[3 * 4 + r] * B[c * 4 + 3])
            R[c * 4 + r] = _f(_f(_f(t0 + t1) + t2) + t3)
    return R


def _dot3(a, b):
    return _f(_f(_f(a[0] * b[0]) + _f(a[1] * b[1])) + _f(a[2] * b[2]))


def _cross3(a, b):
    return [_f(_f(a[1] * b[2]) - _f(b[1] * a[2])), _f(_f(a[2] * b[0]) - _f(b[2] * a[0])),
            _f(_f(a[0] * b[1]) - _f(b[0] * a[1]))]


def quat_normalize(q):
    """glm::normalize of (x, y, z, w): dot = (w w + x x) + (y y + z z)."""
    x, y, z, w = (_f(v) for v in q)
    d = _f(_f(_f(w * w) + _f(x * x)) + _f(_f(y * y) + _f(z * z)))
    ln = _f(np.sqrt(d))
    if not ln > 0:
        return [_f(0), _f(0), _f(0), _f(1)]
    inv = _f(_f(1) / ln)
    return [_f(x * inv), _f(y * inv), _f(z * inv), _f(w * inv)]


def decompose(M):
    """glm::decompose of a float32 4x4 (column-major list): (scale, (x, y, z, w), translation) or None."""
    eps = _f(1.1920928955078125e-7)
    L = [_f(v) for v in M]
    if abs(L[15]) < eps:
        return None
    w33 = L[15]
    L = [_f(v / w33) for v in L[:15]] + [_f(w33 / w33)]
    rows = [L[0:3], L[4:7], L[8:11]]
    if abs(_dot3(rows[0], _cross3(rows[1], rows[2]))) < eps:
        return None
    t = [L[12], L[13], L[14]]
    length = lambda v: _f(np.sqrt(_dot3(v, v)))
    unit = lambda v: [_f(c / length(v)) for c in v]
    comb = lambda a, b, s: [_f(a[i] + _f(b[i] * s)) for i in range(3)]
    scale = [length(rows[0])]
    rows[0] = unit(rows[0])
    sz = _dot3(rows[0], rows[1])
    rows[1] = comb(rows[1], rows[0], _f(-sz))
    scale.append(length(rows[1]))
    rows[1] = unit(rows[1])
    sy = _dot3(rows[0], rows[2])
    rows[2] = comb(rows[2], rows[0], _f(-sy))
    sx = _dot3(rows[1], rows[2])
    rows[2] = comb(rows[2], rows[1], _f(-sx))
    scale.append(length(rows[2]))
    rows[2] = unit(rows[2])
    if _dot3(rows[0], _cross3(rows[1], rows[2])) < 0:
        scale = [_f(s * _f(-1)) for s in scale]
        rows = [[_f(c * _f(-1)) for c in r] for r in rows]
    q = [_f(0)] * 4
    trace = _f(_f(rows[0][0] + rows[1][1]) + rows[2][2])
    if trace > 0:
        root = _f(np.sqrt(_f(trace + _f(1))))
        q[3] = _f(_f(0.5) * root)
        root = _f(_f(0.5) / root)
        q[0] = _f(root * _f(rows[1][2] - rows[2][1]))
        q[1] = _f(root * _f(rows[2][0] - rows[0][2]))
        q[2] = _f(root * _f(rows[0][1] - rows[1][0]))
    else:
        nxt = [1, 2, 0]
        i = 0
        if rows[1][1] > rows[0][0]:
            i = 1
        if rows[2][2] > rows[i][i]:
            i = 2
        j = nxt[i]
        k = nxt[j]
        root = _f(np.sqrt(_f(_f(_f(rows[i][i] - rows[j][j]) - rows[k][k]) + _f(1))))
        q[i] = _f(_f(0.5) * root)
        root = _f(_f(0.5) / root)
        q[j] = _f(root * _f(rows[i][j] + rows[j][i]))
        q[k] = _f(root * _f(rows[i][k] + rows[k][i]))
        q[3] = _f(root * _f(rows[j][k] - rows[k][j]))
    return scale, q, t


def nodes(path: str):
    """(trs13 (n, 13) float32, parent (n,) uint32 with ~0 for roots, mesh (n,), collides (n,))."""
    doc, _ = read_document(path)
    nodes_ = doc.get("nodes", [])
    n = len(nodes_)
    trs = np.zeros((n, 13), np.float32)
    parent = np.full(n, 0xFFFFFFFF, np.uint32)
    mesh = np.full(n, -1, np.int32)
    coll = np.zeros(n, bool)
    meshes = doc.get("meshes", [])

    def mesh_flags(m):
        prims = meshes[m]["primitives"]
        skinned = any("JOINTS_0" in p["attributes"] or "WEIGHTS_0" in p["attributes"] for p in prims)
        morphed = any(len(p.get("targets", [])) > 0 and len(p["targets"][0]) > 0 for p in prims)
        return skinned, morphed

    F = [_f(1), _f(0), _f(0), _f(0), _f(0), _f(-1), _f(0), _f(0), _f(0), _f(0), _f(-1), _f(0),
         _f(0), _f(0), _f(0), _f(1)]
    for i, nd in enumerate(nodes_):
        scale, rot, tr = [_f(1)] * 3, [_f(0), _f(0), _f(0), _f(1)], [_f(0)] * 3
        if len(nd.get("matrix", [])) == 16:
            A = [_f(v) for v in nd["matrix"]]
            out = decompose(_mat_mul(_mat_mul(F, A), F))
            if out is not None:
                scale, rot, tr = out
        else:
            if "scale" in nd:
                scale = [_f(v) for v in nd["scale"]]
            # tinygltf holds these as doubles: the negations act on doubles (an integer 0 gives -0.0)
            if "rotation" in nd:
                x, y, z, w = (float(v) for v in nd["rotation"])
                rot = [_f(x), _f(-y), _f(-z), _f(w)]
            if "translation" in nd:
                x, y, z = (float(v) for v in nd["translation"])
                tr = [_f(x), _f(-y), _f(-z)]
        rot = quat_normalize(rot)
        trs[i, 0:3], trs[i, 3:7], trs[i, 7:10] = scale, rot, tr
        if "mesh" in nd:
            mesh[i] = nd["mesh"]
            skinned, morphed = mesh_flags(nd["mesh"])
            coll[i] = not skinned and not morphed
        for ch in nd.get("children", []):
            parent[ch] = i
    return trs, parent, mesh, coll
