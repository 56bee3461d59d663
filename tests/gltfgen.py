"""Writes small glTF 2.0 documents for the ingest tests: meshes of primitives
with chosen draw modes, index types, layouts and containers (.gltf with an
external buffer, .gltf with a base64 data URI, .glb)."""
from __future__ import annotations

import base64
import json
import os
import struct

import numpy as np


def prim(positions, indices=None, normals=None, mode=None, index_type=5125, joints=False, morph=False,
         stride=None):
    return {"positions": np.asarray(positions, np.float32).reshape(-1, 3),
            "indices": None if indices is None else np.asarray(indices, np.int64).reshape(-1),
            "normals": None if normals is None else np.asarray(normals, np.float32).reshape(-1, 3),
            "mode": mode, "index_type": index_type, "joints": joints, "morph": morph, "stride": stride}


def document(meshes, container="gltf", uri_name="buf.bin"):
    """(json dict, binary buffer) for meshes = [[prim, ...], ...]."""
    blob = bytearray()
    views, accs = [], []

    def add(arr: np.ndarray, comp: int, typ: str, stride=None, minmax=False) -> int:
        while len(blob) % 4:
            blob.append(0)
        raw = arr.tobytes()
        view = {"buffer": 0, "byteOffset": len(blob), "byteLength": len(raw)}
        if stride is not None:
            view["byteStride"] = stride
        blob.extend(raw)
        views.append(view)
        n = arr.shape[0]
        acc = {"bufferView": len(views) - 1, "componentType": comp, "count": int(n), "type": typ}
        if minmax:
            acc["min"] = [float(x) for x in arr.min(axis=0)]
            acc["max"] = [float(x) for x in arr.max(axis=0)]
        accs.append(acc)
        return len(accs) - 1

    out_meshes = []
    for mesh in meshes:
        ps = []
        for p in mesh:
            attrs = {"POSITION": add(p["positions"], 5126, "VEC3", p["stride"], minmax=True)}
            if p["normals"] is not None:
                attrs["NORMAL"] = add(p["normals"], 5126, "VEC3")
            if p["joints"]:
                attrs["JOINTS_0"] = add(np.zeros((len(p["positions"]), 4), np.uint16), 5123, "VEC4")
            q = {"attributes": attrs}
            if p["indices"] is not None:
                dt = {5121: np.uint8, 5123: np.uint16, 5125: np.uint32}[p["index_type"]]
                q["indices"] = add(p["indices"].astype(dt).reshape(-1, 1), p["index_type"], "SCALAR")
            if p["mode"] is not None:
                q["mode"] = p["mode"]
            if p["morph"]:
                q["targets"] = [{"POSITION": add(np.zeros_like(p["positions"]), 5126, "VEC3")}]
            ps.append(q)
        out_meshes.append({"primitives": ps})
    doc = {"asset": {"version": "2.0"}, "meshes": out_meshes, "accessors": accs, "bufferViews": views,
           "buffers": [{"byteLength": len(blob)}]}
    if container == "gltf":
        doc["buffers"][0]["uri"] = uri_name
    elif container == "datauri":
        doc["buffers"][0]["uri"] = "data:application/octet-stream;base64," + base64.b64encode(bytes(blob)).decode()
    return doc, bytes(blob)


def write(path_stem: str, meshes, container="gltf") -> str:
    """Writes the document; returns the path of the .gltf / .glb."""
    name = os.path.basename(path_stem) + ".bin"
    doc, blob = document(meshes, container, uri_name=name)
    if container == "glb":
        js = json.dumps(doc).encode()
        js += b" " * (-len(js) % 4)
        binc = blob + b"\0" * (-len(blob) % 4)
        total = 12 + 8 + len(js) + 8 + len(binc)
        data = struct.pack("<III", 0x46546C67, 2, total) + struct.pack("<II", len(js), 0x4E4F534A) + js
        data += struct.pack("<II", len(binc), 0x004E4942) + binc
        path = path_stem + ".glb"
        open(path, "wb").write(data)
        return path
    if container == "gltf":
        open(path_stem + ".bin", "wb").write(blob)
    path = path_stem + ".gltf"
    json.dump(doc, open(path, "w"))
    return path


def grid_strip(n: int):
    """n vertices zig-zagging along x in the z = 0 plane: (i // 2, i % 2, 0)."""
    i = np.arange(n)
    return np.stack([i // 2, i % 2, np.zeros(n)], axis=1).astype(np.float32)


def polygon(n: int):
    """A centre vertex then n vertices of a regular polygon, counter-clockwise in z = 0."""
    a = 2 * np.pi * np.arange(n) / n
    ring = np.stack([np.cos(a), np.sin(a), np.zeros(n)], axis=1)
    return np.concatenate([np.zeros((1, 3)), ring]).astype(np.float32)


def mixed_meshes(rng):
    """Seven meshes over every draw mode, index type and primitive filter."""
    P = rng.normal(size=(40, 3)).astype(np.float32)
    N = (P / np.linalg.norm(P, axis=1, keepdims=True)).astype(np.float32)
    idx = rng.integers(0, 40, 60)
    return [
        [prim(P, idx, N)],                                          # triangles, mode absent
        [prim(P, idx[:31], N, mode=5), prim(P, idx, N, mode=4),  # strip, triangles, fan, triangles:
         prim(P, idx[:17], N, mode=6), prim(P, idx[:9], N)],     # the order the sort leaves
        [prim(P, idx[:12], mode=0, index_type=5123), prim(P, idx[:13], mode=1),
         prim(P, idx[:8], mode=2), prim(P, idx[:8], mode=3)],    # points, lines, loop -> strip, strip
        [prim(P, None, N, mode=4), prim(P[:7], None, N[:7], mode=5)],  # no indices: 0 .. n-1
        [prim(P, idx, N, joints=True), prim(P, idx, N, morph=True), prim(P, idx[:6], N)],
        [prim(P, idx, N, joints=True)],                             # nothing left to collide
        [prim(P, idx[:1], N, mode=5), prim(P, [], N, mode=6)],   # too short for a triangle
    ]
