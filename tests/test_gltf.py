"""glTF mesh ingest on the CPU: the oracle's restatement pinned by the files
themselves and by library routines, and the library's parser (host code
behind cd_gltf_open, no GPU involved) against the oracle.

Pins of the oracle (oracle/gltf.py):
  - every POSITION accessor carries min / max (glTF 2.0 requires them): the
    decoded, flipped vertices must reproduce them exactly, which fails for a
    wrong offset, count, component type or flip;
  - decoded NORMALs are unit vectors (a misplaced read is not);
  - triangle counts of the shipped scenes, from their index accessors;
  - the primitive order against libstdc++'s std::sort itself, compiled here
    with the reference's comparator, over every mode sequence up to length 7;
  - each draw mode's triplets against geometry: a strip over a planar grid
    and a fan over a convex polygon are consistently wound and tile the
    area exactly; points and lines give the degenerate triplets.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import gltfgen as gg
from conftest import GOLDEN
from oracle import gltf as og

FIXTURES = {
    "apple": os.path.join(GOLDEN, "gltf", "apple", "scene.gltf"),
    "table_room": os.path.join(GOLDEN, "gltf", "table_room.gltf"),
    "environment": os.path.join(GOLDEN, "gltf", "environment.gltf"),
}
ORIGINALS = {
    "apple": "/root/reference/inMyRoom_vulkan/testGames/SnakeGame/Apple/scene.gltf",
    "table_room": "/root/reference/inMyRoom_vulkan/testGames/SnakeGame/Table_room/scene.gltf",
    "environment": "/root/reference/inMyRoom_vulkan/testGames/Sponza/environmentTest/environment.gltf",
}


def _all_paths():
    out = list(FIXTURES.items())
    out += [(k + "_original", p) for k, p in ORIGINALS.items() if os.path.exists(p)]
    return out


@pytest.mark.parametrize("name,path", list(FIXTURES.items()))
def test_oracle_decode_reproduces_accessor_bounds(name, path):
    doc, bufs = og.read_document(path)
    checked = 0
    for m in doc["meshes"]:
        for p in m["primitives"]:
            a = doc["accessors"][p["attributes"]["POSITION"]]
            dec = og.primitive(doc, bufs, p)
            if dec["excluded"]:
                continue
            pos = dec["positions"]
            lo, hi = np.float32(a["min"]), np.float32(a["max"])
            # (x, y, z) -> (x, -y, -z): the flipped axes swap their bounds
            np.testing.assert_array_equal(pos.min(axis=0), [lo[0], -hi[1], -hi[2]])
            np.testing.assert_array_equal(pos.max(axis=0), [hi[0], -lo[1], -lo[2]])
            if dec["normals"] is not None:
                ln = np.linalg.norm(dec["normals"].astype(np.float64), axis=1)
                assert np.abs(ln - 1).max() < 1e-3
            assert dec["indices"].max() < len(pos)
            checked += 1
    assert checked > 0


def test_oracle_triangle_counts():
    counts = {}
    for name, path in FIXTURES.items():
        meshes = og.load(path)
        counts[name] = [len(og.mesh_triangles(m)[0]) for m in meshes]
    assert counts["apple"] == [804]
    assert counts["environment"] == [8448, 8448, 8448]
    tr = counts["table_room"]
    assert len(tr) == 46 and sum(tr) == 38694 and tr.count(0) == 4  # 4 skinned meshes give none
    doc, _ = og.read_document(FIXTURES["table_room"])
    every = sum(doc["accessors"][p["indices"]]["count"] // 3 for m in doc["meshes"] for p in m["primitives"])
    assert every == 41422


SORT_PROGRAM = r'''
#include <algorithm>
#include <cstdio>
#include <vector>
// For every sequence of the modes {-1, 0, 4, 5} up to length 7, the order
// std::sort leaves under a comparator that looks at its left operand only.
int main() {
    const int alpha[4] = {-1, 0, 4, 5};
    for (int n = 0; n <= 7; ++n) {
        int total = 1;
        for (int k = 0; k < n; ++k) total *= 4;
        for (int code = 0; code < total; ++code) {
            std::vector<std::pair<int, int>> v;
            for (int k = 0, c = code; k < n; ++k, c /= 4) v.push_back({alpha[c % 4], k});
            std::sort(v.begin(), v.end(), [](const std::pair<int, int>& l, const std::pair<int, int>&) {
                return l.first == 4 || l.first == -1;
            });
            std::printf("%d %d:", n, code);
            for (auto& e : v) std::printf(" %d", e.second);
            std::printf("\n");
        }
    }
    return 0;
}
'''


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ for libstdc++'s std::sort")
def test_primitive_order_matches_std_sort(tmp_path):
    src = tmp_path / "sort.cpp"
    src.write_text(SORT_PROGRAM)
    exe = tmp_path / "sort"
    subprocess.run(["g++", "-O1", "-std=c++17", str(src), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines()
    alpha = [-1, 0, 4, 5]
    checked = 0
    for line in lines:
        head, _, rest = line.partition(":")
        n, code = map(int, head.split())
        modes = [alpha[(code // 4 ** k) % 4] for k in range(n)]
        assert og.insertion_order(modes) == [int(x) for x in rest.split()], modes
        checked += 1
    assert checked == sum(4 ** n for n in range(8))


def _area_normal(P, tri_idx):
    a, b, c = P[tri_idx[:, 0]], P[tri_idx[:, 1]], P[tri_idx[:, 2]]
    return np.cross(b - a, c - a).astype(np.float64) / 2


def test_oracle_triplets_geometry():
    # a strip over a unit-wide grid: n - 2 triangles, all wound the same way, tiling the rectangle
    for n in (3, 4, 7, 10):
        P = gg.grid_strip(n).astype(np.float64)
        t = og.triplets(og.TRIANGLE_STRIP, n)
        assert len(t) == n - 2
        an = _area_normal(P, t)
        assert np.all(an[:, 2] == an[0, 2]) and abs(an[0, 2]) == 0.5
        assert np.allclose(an[:, :2], 0)
    # a fan over a regular polygon (centre first): counter-clockwise, the polygon's area
    for k in (3, 5, 12):
        P = gg.polygon(k).astype(np.float64)
        # the centre, then the ring and its first vertex again, as a fan index list
        idx = np.array([1 + i for i in range(k)] + [1], np.int64)
        idx = np.concatenate([[0], idx])
        t = og.triplets(og.TRIANGLE_FAN, len(idx))
        assert len(t) == len(idx) - 2 == k
        an = _area_normal(P, idx[t])
        assert np.all(an[:, 2] > 0)
        assert abs(an[:, 2].sum() - 0.5 * k * np.sin(2 * np.pi / k)) < 1e-6  # float32 vertices
    t = og.triplets(og.TRIANGLES, 10)
    assert np.array_equal(t, np.arange(9).reshape(3, 3))
    assert np.array_equal(og.triplets(og.POINTS, 4), np.repeat(np.arange(4), 3).reshape(4, 3))
    assert np.array_equal(og.triplets(og.LINES, 5), [[0, 0, 1], [2, 2, 3]])
    assert np.array_equal(og.triplets(og.LINE_STRIP, 3), [[0, 0, 1], [1, 1, 2]])
    for mode in (og.TRIANGLE_STRIP, og.TRIANGLE_FAN, og.LINE_STRIP):
        assert len(og.triplets(mode, 0)) == 0


def _same_as_oracle(path):
    from inmyroom_vulkan_amd import GltfDocument

    want = og.load(path)
    with GltfDocument(path) as doc:
        assert len(doc) == len(want)
        for m, wprims in enumerate(want):
            got = doc.primitives(m)
            info = doc.mesh_info(m)
            assert len(got) == len(wprims) == info["nprim"]
            for g, w in zip(got, wprims):
                assert (g["gltf_mode"], g["mode"], g["excluded"]) == (w["gltf_mode"], w["mode"], w["excluded"])
                if w["excluded"]:
                    continue
                assert g["positions"].tobytes() == w["positions"].tobytes()
                assert (g["normals"] is None) == (w["normals"] is None)
                if w["normals"] is not None:
                    assert g["normals"].tobytes() == w["normals"].tobytes()
                assert np.array_equal(g["indices"], w["indices"])
                assert g["ntri"] == len(og.triplets(w["mode"], len(w["indices"])))
            assert info["ntri"] == len(og.mesh_triangles(wprims)[0])


@pytest.mark.parametrize("name,path", _all_paths())
def test_parser_matches_oracle_on_shipped_scenes(name, path):
    _same_as_oracle(path)


@pytest.mark.parametrize("container", ["gltf", "datauri", "glb"])
def test_parser_matches_oracle_on_synthetic_documents(tmp_path, container):
    rng = np.random.default_rng(7)
    path = gg.write(str(tmp_path / "doc"), gg.mixed_meshes(rng), container)
    _same_as_oracle(path)
    from inmyroom_vulkan_amd import GltfDocument

    with GltfDocument(path) as doc:
        # per mode: triangles n / 3, strip and fan n - 2, points n, lines n / 2, line strip n - 1
        want = [60 // 3, (31 - 2) + 60 // 3 + (17 - 2) + 9 // 3, 12 + 13 // 2 + (8 - 1) + (8 - 1), 40 // 3 + (7 - 2),
                6 // 3, 0, 0]
        assert [doc.mesh_info(m)["ntri"] for m in range(len(doc))] == want


def _expect_format_error(path, needle):
    from inmyroom_vulkan_amd import GltfDocument
    from inmyroom_vulkan_amd.collide import CD_E_FORMAT, CollideError

    with pytest.raises(CollideError) as e:
        GltfDocument(path)
    assert e.value.code == CD_E_FORMAT and needle in str(e.value), str(e.value)


def test_parser_rejects(tmp_path):
    import json

    P = np.eye(3, dtype=np.float32)
    ok = [[gg.prim(P, [0, 1, 2])]]
    cases = {
        "stride": ([[gg.prim(P, [0, 1, 2], stride=16)]], None, "interleaved"),
        "u8": ([[gg.prim(P, [0, 1, 2], index_type=5121)]], None, "component type 5121"),
        "range": ([[gg.prim(P, [0, 1, 3])]], None, "out of range"),
    }
    for name, (meshes, _, needle) in cases.items():
        _expect_format_error(gg.write(str(tmp_path / name), meshes), needle)

    def edited(name, fn):
        doc, blob = gg.document(ok, "gltf", uri_name=name + ".bin")
        fn(doc)
        open(tmp_path / (name + ".bin"), "wb").write(blob)
        path = tmp_path / (name + ".gltf")
        json.dump(doc, open(path, "w"))
        return str(path)

    _expect_format_error(edited("sparse", lambda d: d["accessors"][0].update(sparse={"count": 0})), "sparse")
    _expect_format_error(edited("short", lambda d: d["buffers"][0].update(byteLength=999)), "byteLength")
    _expect_format_error(edited("mode", lambda d: d["meshes"][0]["primitives"][0].update(mode=9)), "invalid mode")
    _expect_format_error(edited("nopos", lambda d: d["meshes"][0]["primitives"][0]["attributes"].pop("POSITION")),
                         "POSITION")
    _expect_format_error(edited("vec2", lambda d: d["accessors"][0].update(type="VEC2")), "expected type VEC3")
    _expect_format_error(edited("past", lambda d: d["accessors"][0].update(count=4)), "runs past")
    _expect_format_error(edited("huge", lambda d: d["accessors"][0].update(count=2 ** 45)), "count too large")
    _expect_format_error(edited("nonint", lambda d: d["accessors"][0].update(count=1e300)), "not an integer")
    bad = tmp_path / "bad.gltf"
    bad.write_text('{"meshes": [1, 2,]}')
    _expect_format_error(str(bad), "JSON")
    _expect_format_error(str(tmp_path / "missing.gltf"), "cannot read")
