"""The C++ mirror against the engine's own types and call pattern.

A harness restates the engine's shapes: glm-style matrices and vectors,
`CollisionDetectionEntry` / `CollisionCallbackData` as ECStypes.h:149-163
declares them, an `OBBtree` that owns heap storage, `MeshesOfNodes` keeping
`std::vector<MeshInfo>` grown by emplace_back one model at a time
(MeshesOfNodes.cpp:29-64), and an ECS wrapper with an entities handler and
components (CollisionDetection.cpp:106-134).  The engine's own lines compile
unchanged against `mi355x::BasicCollisionDetection<CollisionDetectionEntry,
CollisionCallbackData>`: the construction of Engine.cpp:50 and the entry of
ModelCollisionCompEntity.cpp:66-76.

* CPU: the tree registry.  Meshes are registered by index while models are
  added; a tree address taken before the last model has moved by the end
  (the hazard of keying addresses during loading), every address bound after
  the last model resolves to its mesh's id, and the stale one is refused.
* GPU: two cubes through the engine-typed mirror, the callbacks reaching the
  components through the ECS wrapper, against the oracle's delta.
"""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT, case

HARNESS = r'''
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace glm {  // the shapes the mirror relies on: m[column][row], v[k]
struct vec3 {
    float x = 0.f, y = 0.f, z = 0.f;
    vec3() = default;
    vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    float& operator[](int i) { return (&x)[i]; }
    const float& operator[](int i) const { return (&x)[i]; }
};
struct vec4 {
    float x = 0.f, y = 0.f, z = 0.f, w = 0.f;
    float& operator[](int i) { return (&x)[i]; }
    const float& operator[](int i) const { return (&x)[i]; }
};
struct mat4 {
    vec4 c[4];
    explicit mat4(float d = 1.f) { for (int i = 0; i < 4; ++i) c[i][i] = d; }
    vec4& operator[](int i) { return c[i]; }
    const vec4& operator[](int i) const { return c[i]; }
};
typedef mat4 mat4x4;
}  // namespace glm

typedef uint16_t Entity;  // ECStypes.h:18
class OBBtree {           // owns its nodes and triangles on the heap, like the engine's
public:
    std::vector<float> triangles_position;
};
struct CollisionDetectionEntry {  // ECStypes.h:149-156
    glm::mat4x4 currentGlobalMatrix;
    glm::mat4x4 previousGlobalMatrix;
    const class OBBtree* OBBtree_ptr;
    bool shouldCallback;
    Entity entity;
};
struct CollisionCallbackData {  // ECStypes.h:158-163
    Entity familyEntity;
    Entity collideWithEntity;
    glm::vec3 deltaVector = glm::vec3(0.f, 0.f, 0.f);
};
class ComponentBaseClass {  // ComponentBaseClass.h:25
public:
    virtual ~ComponentBaseClass() = default;
    virtual void CollisionCallback(const std::vector<std::pair<Entity, std::vector<CollisionCallbackData>>>&) {}
};
class EntitiesHandler {
public:
    std::vector<Entity> GetEntityAncestors(Entity e) const { return {2, e}; }  // 5 and 7 share ancestor 2
};
class ECSwrapper {
public:
    EntitiesHandler* GetEntitiesHandler() { return &handler; }
    const std::map<int, ComponentBaseClass*>& GetComponentIDtoComponentBaseClassMap() const { return comps; }
    EntitiesHandler handler;
    std::map<int, ComponentBaseClass*> comps;
};

#include "CollisionDetection.hpp"
typedef mi355x::BasicCollisionDetection<CollisionDetectionEntry, CollisionCallbackData> CollisionDetection;

struct MeshInfo {  // MeshesOfNodes.h:26-29
    std::vector<size_t> primitivesIndex;
    OBBtree boundBoxTree;
};
class MeshesOfNodes {  // one model after another into a growing vector
public:
    void AddModel(size_t nmeshes, std::vector<uint32_t> ids, mi355x::TreeRegistry& reg) {
        for (size_t m = 0; m < nmeshes; ++m) {
            MeshInfo info;
            info.boundBoxTree.triangles_position.assign(9 * (m + 1), float(m));
            reg.SetMeshIdOfIndex(meshes.size(), ids[m]);  // by index: stable while the vector grows
            meshes.emplace_back(std::move(info));          // MeshesOfNodes.cpp:62
        }
    }
    const MeshInfo& GetMeshInfo(size_t i) const { return meshes[i]; }
    size_t size() const { return meshes.size(); }
    std::vector<MeshInfo> meshes;
};

static int cpu_part() {
    mi355x::TreeRegistry reg;
    MeshesOfNodes mon;
    uint32_t next = 100;
    mon.AddModel(3, {next, next + 1, next + 2}, reg);
    next += 3;
    const OBBtree* early = &mon.GetMeshInfo(0).boundBoxTree;  // keyed before the other models: goes stale
    for (int model = 0; model < 6; ++model) {
        std::vector<uint32_t> ids;
        for (int k = 0; k < 5; ++k) ids.push_back(next++);
        mon.AddModel(5, ids, reg);
    }
    std::printf("moved %d\n", int(early != &mon.GetMeshInfo(0).boundBoxTree));
    for (size_t i = 0; i < mon.size(); ++i) reg.Bind(&mon.GetMeshInfo(i).boundBoxTree, i);  // after the last model
    int bad = 0;
    for (size_t i = 0; i < mon.size(); ++i)
        if (reg.MeshId(&mon.GetMeshInfo(i).boundBoxTree) != 100 + i) ++bad;
    std::printf("resolved %zu bad %d\n", mon.size(), bad);
    bool refused = false;
    if (early != &mon.GetMeshInfo(0).boundBoxTree) {
        try {
            (void)reg.MeshId(early);
        } catch (const std::exception&) {
            refused = true;
        }
    }
    std::printf("stale refused %d\n", int(refused));
    return 0;
}

// ---- the engine's own lines, against the engine-typed mirror (needs a GPU) ----
struct Sink : ComponentBaseClass {
    std::map<int, std::vector<CollisionCallbackData>> got;
    void CollisionCallback(const std::vector<std::pair<Entity, std::vector<CollisionCallbackData>>>& v) override {
        for (const auto& p : v) got[p.first] = p.second;
    }
};

static void AddCollisionDetectionEntryToVector(const glm::mat4x4& this_frame_global_matrix,
                                               const glm::mat4x4& previous_frame_global_matrix, bool shouldCallback,
                                               const MeshesOfNodes* meshesOfNodes_ptr, size_t meshIndex, Entity thisEntity,
                                               CollisionDetection* collisionDetection_ptr) {
    // ModelCollisionCompEntity.cpp:66-76
    CollisionDetectionEntry this_collisionDetectionEntry;
    this_collisionDetectionEntry.currentGlobalMatrix = this_frame_global_matrix;
    this_collisionDetectionEntry.previousGlobalMatrix = previous_frame_global_matrix;
    this_collisionDetectionEntry.shouldCallback = shouldCallback;
    this_collisionDetectionEntry.OBBtree_ptr = &(meshesOfNodes_ptr->GetMeshInfo(meshIndex).boundBoxTree);
    this_collisionDetectionEntry.entity = thisEntity;

    collisionDetection_ptr->AddCollisionDetectionEntry(this_collisionDetectionEntry);
}

static int gpu_part() {
    std::unique_ptr<ECSwrapper> ECSwrapper_uptr = std::make_unique<ECSwrapper>();
    Sink sink;
    ECSwrapper_uptr->comps[1] = &sink;
    ECSwrapper_uptr->comps[2] = nullptr;  // MakeCallbacks skips empty component slots
    std::unique_ptr<CollisionDetection> collisionDetection_uptr;
    collisionDetection_uptr = std::make_unique<CollisionDetection>(ECSwrapper_uptr.get());  // Engine.cpp:50

    // a cube of side 2 (12 triangles, outward winding), registered at load
    const float P[8 * 3] = {-1,-1,-1, 1,-1,-1, 1,1,-1, -1,1,-1, -1,-1,1, 1,-1,1, 1,1,1, -1,1,1};
    const uint32_t I[12 * 3] = {0,2,1, 0,3,2, 4,5,6, 4,6,7, 0,1,5, 0,5,4, 2,3,7, 2,7,6, 1,2,6, 1,6,5, 0,4,7, 0,7,3};
    MeshesOfNodes mon;
    mi355x::TreeRegistry& reg = collisionDetection_uptr->Trees();
    const uint32_t id = collisionDetection_uptr->RegisterMesh(P, 8, I, 12);
    mon.AddModel(1, {id}, reg);
    mon.AddModel(2, {id, id}, reg);  // more models move the first one's tree
    for (size_t i = 0; i < mon.size(); ++i) reg.Bind(&mon.GetMeshInfo(i).boundBoxTree, i);

    auto pose = [](float x, float y, float z) {
        glm::mat4 m(1.f);
        m[3][0] = x; m[3][1] = y; m[3][2] = z;
        return m;
    };
    collisionDetection_uptr->Reset();
    AddCollisionDetectionEntryToVector(pose(0, 0, 0), pose(0, 0, 0), true, &mon, 0, 5, collisionDetection_uptr.get());
    AddCollisionDetectionEntryToVector(pose(1.5f, 0.25f, 0.35f), pose(2.5f, 0.25f, 0.35f), true, &mon, 2, 7,
                                       collisionDetection_uptr.get());
    collisionDetection_uptr->ExecuteCollisionDetection();
    for (const auto& r : collisionDetection_uptr->GetCollidingPairs())
        std::printf("pair %d %d shot %d delta %.6f %.6f %.6f\n", r.first, r.second, int(r.shot), r.delta[0], r.delta[1],
                    r.delta[2]);
    for (const auto& kv : sink.got)
        for (const auto& c : kv.second)
            std::printf("callback %d family %d other %d %.6f %.6f %.6f\n", kv.first, c.familyEntity,
                        c.collideWithEntity, c.deltaVector.x, c.deltaVector.y, c.deltaVector.z);
    return 0;
}

int main(int argc, char** argv) { return (argc > 1 && std::string(argv[1]) == "gpu") ? gpu_part() : cpu_part(); }
'''


def _build(tmp_path):
    from inmyroom_vulkan_amd import collide

    src = tmp_path / "engine.cpp"
    src.write_text(HARNESS)
    exe = tmp_path / "engine"
    libdir = os.path.dirname(collide.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-L",
                    libdir, "-lmi355x_collide", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    return exe


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_engine_pattern_tree_registry(tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=60).stdout.split("\n")
    assert "moved 1" in out  # an address keyed during loading no longer names the tree
    assert "resolved 33 bad 0" in out  # bound after the last model: every entry finds its mesh
    assert "stale refused 1" in out


@pytest.mark.gpu
def test_engine_typed_mirror_to_callbacks(tmp_path, orc):
    from test_gpu_mirror import _oracle_delta

    case("engine-typed C++ mirror: two cubes to the components' callbacks")
    exe = _build(tmp_path)
    out = subprocess.run([str(exe), "gpu"], check=True, capture_output=True, text=True, timeout=60).stdout.split("\n")
    pairs = [l.split() for l in out if l.startswith("pair")]
    cbs = [l.split() for l in out if l.startswith("callback")]
    assert len(pairs) == 1 and pairs[0][1:3] == ["5", "7"] and pairs[0][4] == "1"
    delta = [float(v) for v in pairs[0][6:9]]
    for g, w in zip(delta, _oracle_delta(orc)):
        assert abs(g - w) <= 1e-5 * max(1.0, abs(w)) + 2e-6
    by_target = {int(c[1]): c for c in cbs}
    assert set(by_target) == {5, 7}  # the shared ancestor 2 gets nothing
    assert [float(v) for v in by_target[7][6:9]] == delta
    assert all(float(v) == 0.0 for v in by_target[5][6:9])
