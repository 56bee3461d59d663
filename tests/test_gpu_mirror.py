"""The C++ mirror (include/CollisionDetection.hpp) end to end on the GPU:
ExecuteCollisionDetection through ray shooting, the ancestor fan-out
(CollisionDetection.cpp:101-121) and MakeCallbacks (:124-134)."""
import os
import subprocess

import pytest

from conftest import case

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r'''
#include "CollisionDetection.hpp"
#include <cstdio>
#include <map>
int main() {
    // a cube of side 2 (12 triangles, outward winding)
    const float P[8 * 3] = {-1,-1,-1, 1,-1,-1, 1,1,-1, -1,1,-1, -1,-1,1, 1,-1,1, 1,1,1, -1,1,1};
    const uint32_t I[12 * 3] = {0,2,1, 0,3,2, 4,5,6, 4,6,7, 0,1,5, 0,5,4, 2,3,7, 2,7,6, 1,2,6, 1,6,5, 0,4,7, 0,7,3};
    mi355x::CollisionDetection cd(4, MODE);
    cd.SetKeepDetails(KEEP);
    const mi355x::OBBtree tree(cd.RegisterMesh(P, 8, I, 12));
    // entities 5 and 7 share ancestor 2: callbacks go to 5 and 7 only
    cd.SetAncestorsProvider([](mi355x::Entity e) { return std::vector<mi355x::Entity>{2, e}; });
    std::map<int, std::vector<mi355x::CollisionCallbackData>> got;
    cd.SetCallbackSink([&](const mi355x::CallbackEntityDataPairs& v) {
        for (const auto& p : v) got[p.first] = p.second;
    });
    auto entry = [&](mi355x::Entity ent, float x, float prev_x, float y, float z) {
        mi355x::CollisionDetectionEntry e{};
        for (int i = 0; i < 16; ++i) e.currentGlobalMatrix.m[i] = e.previousGlobalMatrix.m[i] = (i % 5 == 0) ? 1.f : 0.f;
        e.currentGlobalMatrix[3][0] = x;
        e.currentGlobalMatrix[3][1] = e.previousGlobalMatrix[3][1] = y;
        e.currentGlobalMatrix[3][2] = e.previousGlobalMatrix[3][2] = z;
        e.previousGlobalMatrix[3][0] = prev_x;
        e.OBBtree_ptr = &tree;
        e.shouldCallback = true;
        e.entity = ent;
        return e;
    };
    cd.Reset();
    cd.AddCollisionDetectionEntry(entry(5, 0.f, 0.f, 0.f, 0.f));         // static
    cd.AddCollisionDetectionEntry(entry(7, 1.5f, 2.5f, 0.25f, 0.35f));  // moved in, overlapping by 0.5 in x
    cd.ExecuteCollisionDetection();
    for (const auto& r : cd.GetCollidingPairs())
        std::printf("pair %d %d shot %d delta %.6f %.6f %.6f\n", r.first, r.second, int(r.shot), r.delta[0], r.delta[1],
                    r.delta[2]);
    for (const auto& kv : got)
        for (const auto& c : kv.second)
            std::printf("callback %d family %d other %d %.6f %.6f %.6f\n", kv.first, c.familyEntity,
                        c.collideWithEntity, c.deltaVector[0], c.deltaVector[1], c.deltaVector[2]);
    // output buffers kept across frames: a frame without pairs, then the first frame again
    for (int again = 0; again < 2; ++again) {
        cd.Reset();
        cd.AddCollisionDetectionEntry(entry(5, 0.f, 0.f, 0.f, 0.f));
        cd.AddCollisionDetectionEntry(again ? entry(7, 1.5f, 2.5f, 0.25f, 0.35f) : entry(7, 9.f, 10.f, 0.25f, 0.35f));
        cd.ExecuteCollisionDetection();
        for (const auto& r : cd.GetCollidingPairs())
            std::printf("again%d %d %d shot %d delta %.6f %.6f %.6f\n", again, r.first, r.second, int(r.shot),
                        r.delta[0], r.delta[1], r.delta[2]);
    }
    return 0;
}
'''


def _oracle_delta(orc):
    """The same frame through the oracle: the world-space delta of the pair."""
    import math

    import numpy as np

    P = np.float32([-1, -1, -1, 1, -1, -1, 1, 1, -1, -1, 1, -1, -1, -1, 1, 1, -1, 1, 1, 1, 1, -1, 1, 1]).reshape(-1, 3)
    I = np.uint32([0, 2, 1, 0, 3, 2, 4, 5, 6, 4, 6, 7, 0, 1, 5, 0, 5, 4, 2, 3, 7, 2, 7, 6, 1, 2, 6, 1, 6, 5,
                   0, 4, 7, 0, 7, 3]).reshape(-1, 3)
    t = orc.tree(P[I].reshape(-1, 9), 4)
    t.set_attrs(None, I)

    def m(x, y, z):
        a = np.eye(4, dtype=np.float32)
        a[3, :3] = (x, y, z)
        return a.reshape(-1)

    deg = np.float32(0.017453292519943295)
    cs, cf = np.float32(math.cos(np.float32(40 * deg))), np.float32(math.cos(np.float32(65 * deg)))
    cur = np.stack([m(0, 0, 0), m(1.5, 0.25, 0.35)])
    prev = np.stack([m(0, 0, 0), m(2.5, 0.25, 0.35)])
    ids, d, _ = orc.shoot_frame([t], [0, 0], cur, prev, np.ones(2, np.uint8), cs, cf, np.float32(1.01))
    assert len(ids) == 1 and ids[0, 2] == 1
    return d[0, 0]


@pytest.mark.parametrize("mode,keep", [("CD_MODE_PLAIN", 1), ("CD_MODE_REFERENCE", 0), ("CD_MODE_REFERENCE", 1)])
def test_mirror_frame_to_callbacks(tmp_path, orc, mode, keep):
    """keep 0: one cd_step per ExecuteCollisionDetection; 1: the three calls
    with every pair's contacts and rays kept."""
    case(f"C++ mirror, two cubes to callbacks, {mode}, details {keep}")
    from inmyroom_vulkan_amd import collide
    src = tmp_path / "mirror.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "mirror"
    libdir = os.path.dirname(collide.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", f"-DMODE={mode}", f"-DKEEP={keep}", "-I", os.path.join(ROOT, "include"),
                    str(src), "-L", libdir, "-lmi355x_collide", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=60).stdout
    lines = out.split("\n")
    pairs = [l.split() for l in lines if l.startswith("pair")]
    cbs = [l.split() for l in lines if l.startswith("callback")]
    assert len(pairs) == 1 and pairs[0][1:3] == ["5", "7"] and pairs[0][4] == "1"
    delta = [float(v) for v in pairs[0][6:9]]
    want = _oracle_delta(orc)
    for g, w in zip(delta, want):
        assert abs(g - w) <= 1e-5 * max(1.0, abs(w)) + 2e-6  # printed with 6 decimals
    by_target = {int(c[1]): c for c in cbs}
    assert set(by_target) == {5, 7}  # the shared ancestor 2 gets nothing (CollisionDetection.cpp:101-121)
    c5, c7 = by_target[5], by_target[7]
    assert (c5[3], c5[5]) == ("5", "7") and (c7[3], c7[5]) == ("7", "5")
    assert all(abs(float(v)) == 0.0 for v in c5[6:9])  # the static cube takes no share
    assert [float(v) for v in c7[6:9]] == delta  # the moving cube takes all of it
    assert not [l for l in lines if l.startswith("again0")]  # no pairs
    again = [l.split() for l in lines if l.startswith("again1")]
    assert len(again) == 1 and again[0][1:] == pairs[0][1:]  # the first frame again, through the kept buffers


GLTF_PROGRAM = r'''
#include "CollisionDetection.hpp"
#include <cstdio>
int main(int argc, char** argv) {
    mi355x::CollisionDetection cd(4, CD_MODE_PLAIN);
    mi355x::MeshesOfNodes meshes(cd);
    meshes.AddMeshesOfModel(argv[1]);  // Apple: one mesh
    // taken before the next model is added: the mirror's meshes never move
    const mi355x::OBBtree* apple_tree = &meshes.GetMeshInfo(0).boundBoxTree;
    meshes.AddMeshesOfModel(argv[2]);  // Table_room: 46 meshes, 4 skinned
    const size_t apple = meshes.GetMeshIndexOffsetOfModel(argv[1]);
    const size_t room = meshes.GetMeshIndexOffsetOfModel(argv[2]);
    std::printf("meshes %zu apple %zu room %zu\n", meshes.size(), apple, room);
    int none = 0;
    for (size_t i = room; i < meshes.size(); ++i) none += meshes.GetMeshInfo(i).boundBoxTree.meshId == CD_NO_MESH;
    std::printf("without trees %d\n", none);
    auto entry = [&](mi355x::Entity ent, float x) {
        mi355x::CollisionDetectionEntry e{};
        for (int i = 0; i < 16; ++i) e.currentGlobalMatrix.m[i] = e.previousGlobalMatrix.m[i] = (i % 5 == 0) ? 1.f : 0.f;
        e.currentGlobalMatrix[3][0] = x;
        e.previousGlobalMatrix[3][0] = x + 0.5f;
        e.OBBtree_ptr = &meshes.GetMeshInfo(apple).boundBoxTree;
        if (e.OBBtree_ptr != apple_tree) std::printf("moved\n");
        e.shouldCallback = true;
        e.entity = ent;
        return e;
    };
    cd.Reset();
    cd.AddCollisionDetectionEntry(entry(1, 0.f));
    cd.AddCollisionDetectionEntry(entry(2, 1.2f));   // overlaps the first
    cd.AddCollisionDetectionEntry(entry(3, 10.f));   // far away
    cd.ExecuteCollisionDetection();
    for (const auto& r : cd.GetCollidingPairs()) std::printf("pair %d %d shot %d\n", r.first, r.second, int(r.shot));
    return 0;
}
'''


def test_mirror_meshes_of_nodes_gltf(tmp_path):
    case("C++ mirror, MeshesOfNodes from glTF files")
    from conftest import GOLDEN
    from inmyroom_vulkan_amd import collide
    src = tmp_path / "gltf.cpp"
    src.write_text(GLTF_PROGRAM)
    exe = tmp_path / "gltf"
    libdir = os.path.dirname(collide.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-L", libdir,
                    "-lmi355x_collide", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    apple = os.path.join(GOLDEN, "gltf", "apple", "scene.gltf")
    room = os.path.join(GOLDEN, "gltf", "table_room.gltf")
    out = subprocess.run([str(exe), apple, room], check=True, capture_output=True, text=True, timeout=60).stdout
    lines = out.split("\n")
    assert lines[0] == "meshes 47 apple 0 room 1"
    assert lines[1] == "without trees 4"
    pairs = [l.split() for l in lines if l.startswith("pair")]
    assert [p[1:3] for p in pairs] == [["1", "2"]] and pairs[0][4] == "1"
    assert "moved" not in lines
