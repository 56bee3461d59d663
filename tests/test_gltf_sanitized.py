"""The glTF reader (host code behind cd_gltf_open) under AddressSanitizer and
UndefinedBehaviorSanitizer on the CPU: every shipped scene, the synthetic
documents of each container, every reject case of the parser tests, and a
seeded fuzz pass over byte- and field-level mutants of valid documents.  A
sanitizer report aborts the harness, and the test stops at the first one,
naming the file.  No GPU is involved."""
import base64
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import gltfgen as gg
from conftest import GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "inmyroom_vulkan_amd", "csrc")
SCENES = [os.path.join(GOLDEN, "gltf", "apple", "scene.gltf"), os.path.join(GOLDEN, "gltf", "table_room.gltf"),
          os.path.join(GOLDEN, "gltf", "environment.gltf")]
SAN_ENV = dict(os.environ, ASAN_OPTIONS="abort_on_error=1:detect_leaks=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not found")
    out = str(tmp_path_factory.mktemp("san") / "gltf_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I", CSRC,
                    os.path.join(ROOT, "tests", "native", "gltf_check.cpp"), "-o", out], check=True)
    return out


def run(harness, paths):
    """One line per path; raises with the sanitizer's report on an abort."""
    r = subprocess.run([harness, *paths], capture_output=True, text=True, env=SAN_ENV, timeout=300)
    lines = r.stdout.splitlines()
    if r.returncode != 0:
        bad = paths[len(lines)] if len(lines) < len(paths) else "?"
        raise AssertionError(f"harness exit {r.returncode} on {bad}\n{r.stdout[-2000:]}\n{r.stderr[-6000:]}")
    assert len(lines) == len(paths)
    return lines


def test_sanitized_reader_on_documents_and_rejects(harness, tmp_path):
    rng = np.random.default_rng(7)
    docs = list(SCENES) + [gg.write(str(tmp_path / f"mixed_{c}"), gg.mixed_meshes(rng), c)
                           for c in ("gltf", "datauri", "glb")]
    for line in run(harness, docs):
        assert line.startswith("ok "), line
    P = np.eye(3, dtype=np.float32)
    ok = [[gg.prim(P, [0, 1, 2])]]
    rejects = {gg.write(str(tmp_path / "stride"), [[gg.prim(P, [0, 1, 2], stride=16)]]): "interleaved",
               gg.write(str(tmp_path / "u8"), [[gg.prim(P, [0, 1, 2], index_type=5121)]]): "component type 5121",
               gg.write(str(tmp_path / "range"), [[gg.prim(P, [0, 1, 3])]]): "out of range"}

    def edited(name, fn):
        doc, blob = gg.document(ok, "gltf", uri_name=name + ".bin")
        fn(doc)
        open(tmp_path / (name + ".bin"), "wb").write(blob)
        path = tmp_path / (name + ".gltf")
        json.dump(doc, open(path, "w"))
        return str(path)

    rejects[edited("sparse", lambda d: d["accessors"][0].update(sparse={"count": 0}))] = "sparse"
    rejects[edited("short", lambda d: d["buffers"][0].update(byteLength=999))] = "byteLength"
    rejects[edited("mode", lambda d: d["meshes"][0]["primitives"][0].update(mode=9))] = "invalid mode"
    rejects[edited("nopos", lambda d: d["meshes"][0]["primitives"][0]["attributes"].pop("POSITION"))] = "POSITION"
    rejects[edited("vec2", lambda d: d["accessors"][0].update(type="VEC2"))] = "expected type VEC3"
    rejects[edited("past", lambda d: d["accessors"][0].update(count=4))] = "runs past"
    rejects[edited("huge", lambda d: d["accessors"][0].update(count=2 ** 45))] = "count too large"
    rejects[edited("nonint", lambda d: d["accessors"][0].update(count=1e300))] = "not an integer"
    bad = tmp_path / "bad.gltf"
    bad.write_text('{"meshes": [1, 2,]}')
    rejects[str(bad)] = "JSON"
    rejects[str(tmp_path / "missing.gltf")] = "cannot read"
    for (path, needle), line in zip(rejects.items(), run(harness, list(rejects))):
        assert line.startswith("err ") and needle in line, (path, line)


# Values a mutated JSON field takes: boundaries of the integer and float
# ranges the reader converts, wrong types, and indices of things that do not
# exist.
ODD_VALUES = [-1, 0, 1, 3, 4, 2 ** 31 - 1, 2 ** 31, 2 ** 32 + 1, 2 ** 53 + 1, 2 ** 63, -2 ** 63, 10 ** 30, 1e308,
              -1e-308, 0.5, -0.0, float("nan"), "x", "", None, True, [], {}, [1, 2, 3], {"a": 1}]


def _paths(node, prefix=()):
    """Every (path, value) of a JSON tree."""
    yield prefix, node
    if isinstance(node, dict):
        for k, v in node.items():
            yield from _paths(v, prefix + (k,))
    elif isinstance(node, list):
        for i, v in enumerate(node):
            yield from _paths(v, prefix + (i,))


def _set(doc, path, value):
    d = doc
    for k in path[:-1]:
        d = d[k]
    d[path[-1]] = value


def _delete(doc, path):
    d = doc
    for k in path[:-1]:
        d = d[k]
    del d[path[-1]]


def _mutant_json(rng, doc):
    doc = json.loads(json.dumps(doc))
    paths = [p for p, _ in _paths(doc) if p]
    for _ in range(int(rng.integers(1, 4))):
        p = paths[int(rng.integers(len(paths)))]
        try:
            if rng.random() < 0.2:
                _delete(doc, p)
            else:
                _set(doc, p, ODD_VALUES[int(rng.integers(len(ODD_VALUES)))])
        except (KeyError, IndexError, TypeError):
            pass
    return json.dumps(doc, allow_nan=True).encode()


def _mutant_bytes(rng, data: bytes):
    b = bytearray(data)
    kind = int(rng.integers(4))
    if kind == 0 and len(b):  # truncate
        return bytes(b[: int(rng.integers(len(b)))])
    if kind == 1:  # flip bytes
        for _ in range(int(rng.integers(1, 8))):
            if b:
                b[int(rng.integers(len(b)))] ^= 1 << int(rng.integers(8))
        return bytes(b)
    if kind == 2:  # insert random bytes
        at = int(rng.integers(len(b) + 1))
        return bytes(b[:at] + rng.integers(0, 256, int(rng.integers(1, 16)), dtype=np.uint8).tobytes() + b[at:])
    at = int(rng.integers(len(b) + 1))  # delete a span
    return bytes(b[:at] + b[at + int(rng.integers(1, 32)):])


def test_fuzz_reader_stops_at_first_sanitizer_report(harness, tmp_path):
    """2,400 mutants (JSON fields, JSON text, the external buffer, data: URIs
    and .glb bytes) of valid documents, run in batches; every one must load
    or be rejected with a message, under both sanitizers."""
    rng = np.random.default_rng(2024)
    base = tmp_path / "base"
    base.mkdir()
    gltf_path = gg.write(str(base / "doc"), gg.mixed_meshes(np.random.default_rng(3)), "gltf")
    uri_path = gg.write(str(base / "uri"), gg.mixed_meshes(np.random.default_rng(4)), "datauri")
    glb_path = gg.write(str(base / "bin"), gg.mixed_meshes(np.random.default_rng(5)), "glb")
    gltf_doc = json.load(open(gltf_path))
    uri_doc = json.load(open(uri_path))
    glb_bytes = open(glb_path, "rb").read()
    bin_name = gltf_doc["buffers"][0]["uri"]
    bin_bytes = open(os.path.join(base, bin_name), "rb").read()
    apple = json.load(open(SCENES[0]))
    apple_bin = os.path.join(os.path.dirname(SCENES[0]), apple["buffers"][0]["uri"])
    work = tmp_path / "mutants"
    work.mkdir()
    shutil.copy(os.path.join(base, bin_name), work / bin_name)
    shutil.copy(apple_bin, work / os.path.basename(apple_bin))
    n_ok = n_err = 0
    for batch in range(12):
        paths = []
        for i in range(200):
            stem = work / f"m{batch}_{i}"
            kind = int(rng.integers(6))
            if kind == 0:  # a JSON field of the external-buffer document
                data, ext = _mutant_json(rng, gltf_doc), ".gltf"
            elif kind == 1:  # the JSON text itself
                data, ext = _mutant_bytes(rng, json.dumps(gltf_doc).encode()), ".gltf"
            elif kind == 2:  # the external buffer (the document then names the mutant)
                own = dict(gltf_doc, buffers=[dict(gltf_doc["buffers"][0], uri=stem.name + ".bin")])
                open(str(stem) + ".bin", "wb").write(_mutant_bytes(rng, bin_bytes))
                data, ext = json.dumps(own).encode(), ".gltf"
            elif kind == 3:  # a data: URI's base64 text
                u = json.loads(json.dumps(uri_doc))
                head, _, body = u["buffers"][0]["uri"].partition(",")
                u["buffers"][0]["uri"] = head + "," + _mutant_bytes(rng, body.encode()).decode("latin-1")
                data, ext = json.dumps(u).encode(), ".gltf"
            elif kind == 4:  # .glb bytes (header, chunk lengths, JSON, buffer)
                data, ext = _mutant_bytes(rng, glb_bytes), ".glb"
            else:  # a field of a shipped scene (nodes, skins, materials too)
                data, ext = _mutant_json(rng, apple), ".gltf"
            path = str(stem) + ext
            open(path, "wb").write(data)
            paths.append(path)
        for line in run(harness, paths):
            assert line.startswith(("ok ", "err ")), line
            n_ok += line.startswith("ok ")
            n_err += line.startswith("err ")
    assert n_ok > 100 and n_err > 500  # both outcomes are exercised
    print(f"[fuzz] {n_ok} mutants loaded, {n_err} rejected, no sanitizer report")
