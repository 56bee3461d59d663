"""Writes the glTF fixtures under tests/golden/gltf/ from the scenes the
reference ships (run in a container where /root/reference is mounted):

  apple/scene.gltf, apple/scene.bin     testGames/SnakeGame/Apple, copied as is
                                        (one mesh, 804 triangles, byteStride
                                        views equal to the element size)
  table_room.gltf, table_room.bin       testGames/SnakeGame/Table_room: 46
                                        meshes (4 skinned), 156 nodes
  environment.gltf, environment.bin     testGames/Sponza/environmentTest: 3
                                        meshes, 4 nodes

The last two keep what the mesh ingest and the node transforms read (every
mesh and primitive, POSITION, NORMAL, the index lists, JOINTS_0 of skinned
primitives, nodes, scenes, skins without their matrices) with the
accessors' min / max, each accessor repacked into a buffer view of its own.
Textures, materials, animations and the other attributes are dropped.

    python tests/golden/make_gltf_fixtures.py
"""
from __future__ import annotations

import json
import os
import shutil

REF = "/root/reference/inMyRoom_vulkan/testGames"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gltf")
KEEP_ATTRS = ("POSITION", "NORMAL", "JOINTS_0")
SIZES = {5120: 1, 5121: 1, 5122: 2, 5123: 2, 5125: 4, 5126: 4}
NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT4": 16}


def trim(src: str, dst_stem: str) -> None:
    doc = json.load(open(src))
    base = os.path.dirname(src)
    bufs = [open(os.path.join(base, b["uri"]), "rb").read() for b in doc["buffers"]]
    out_bin = bytearray()
    views, accs, remap = [], [], {}

    def keep(ai: int) -> int:
        if ai in remap:
            return remap[ai]
        a = dict(doc["accessors"][ai])
        bv = doc["bufferViews"][a["bufferView"]]
        n = a["count"] * NCOMP[a["type"]] * SIZES[a["componentType"]]
        start = bv.get("byteOffset", 0) + a.get("byteOffset", 0)
        while len(out_bin) % 4:
            out_bin.append(0)
        views.append({"buffer": 0, "byteOffset": len(out_bin), "byteLength": n})
        out_bin.extend(bufs[bv["buffer"]][start: start + n])
        a["bufferView"] = len(views) - 1
        a.pop("byteOffset", None)
        accs.append(a)
        remap[ai] = len(accs) - 1
        return remap[ai]

    meshes = []
    for m in doc["meshes"]:
        prims = []
        for p in m["primitives"]:
            q = {"attributes": {k: keep(v) for k, v in p["attributes"].items() if k in KEEP_ATTRS}}
            if "indices" in p:
                q["indices"] = keep(p["indices"])
            if "mode" in p:
                q["mode"] = p["mode"]
            if "targets" in p:
                q["targets"] = [{k: keep(v) for k, v in t.items() if k == "POSITION"} for t in p["targets"]]
            prims.append(q)
        meshes.append({"name": m.get("name", ""), "primitives": prims})
    out = {"asset": doc["asset"], "scene": doc.get("scene", 0), "scenes": doc["scenes"],
           "nodes": doc["nodes"], "meshes": meshes, "accessors": accs, "bufferViews": views,
           "buffers": [{"uri": os.path.basename(dst_stem) + ".bin", "byteLength": len(out_bin)}]}
    if "skins" in doc:
        out["skins"] = [{k: v for k, v in s.items() if k != "inverseBindMatrices"} for s in doc["skins"]]
    for n in out["nodes"]:
        n.pop("camera", None)
    json.dump(out, open(dst_stem + ".gltf", "w"), separators=(",", ":"))
    open(dst_stem + ".bin", "wb").write(bytes(out_bin))


def main() -> None:
    os.makedirs(os.path.join(OUT, "apple"), exist_ok=True)
    for f in ("scene.gltf", "scene.bin"):
        shutil.copyfile(os.path.join(REF, "SnakeGame/Apple", f), os.path.join(OUT, "apple", f))
    trim(os.path.join(REF, "SnakeGame/Table_room/scene.gltf"), os.path.join(OUT, "table_room"))
    trim(os.path.join(REF, "Sponza/environmentTest/environment.gltf"), os.path.join(OUT, "environment"))


if __name__ == "__main__":
    main()
