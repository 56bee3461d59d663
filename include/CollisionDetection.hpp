// CollisionDetection.hpp — C++ host mirror of the reference's collision
// interface, implemented over the C ABI of mi355x_collide.h.
//
// Same member names, call order and argument meaning as
// inMyRoom_vulkan/include/CollisionDetection/CollisionDetection.h:11-19
// (Reset, AddCollisionDetectionEntry, ExecuteCollisionDetection), with the
// entry type of inMyRoom_vulkan/include/ECS/ECStypes.h:149-156 restated on
// plain floats (the OBBtree pointer becomes a registered mesh id).
// ExecuteCollisionDetection runs src/CollisionDetection/CollisionDetection.cpp:38-141
// on the GPU: phases 1-3 (broad, mid, narrow), the uncollide rays of
// CreateUncollideRays.cpp:60-200, the ray shooting of ShootUncollideRays.cpp
// with the movement split (:69-99), then the callback fan-out to entity
// ancestors (:101-121) and MakeCallbacks (:124-134).  The two engine
// services those last steps use are hooks: the entity's ancestor list
// (EntitiesHandler::GetEntityAncestors) and the components' CollisionCallback
// (ComponentBaseClass.h:25), which receives the callbacks as the reference
// hands them over.
//
// Error behaviour: the reference has no error returns; this mirror throws
// std::runtime_error with the library's message when a call fails (the C ABI
// itself never throws).
#pragma once

#include <algorithm>
#include <cstring>
#include <functional>
#include <iterator>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "mi355x_collide.h"

namespace mi355x {

typedef uint16_t Entity;  // ECStypes.h:18

struct CollisionDetectionEntry {  // ECStypes.h:149-156
    float currentGlobalMatrix[16];
    float previousGlobalMatrix[16];
    uint32_t meshId;  // replaces const OBBtree* OBBtree_ptr
    bool shouldCallback;
    Entity entity;
};

struct CollisionCallbackData {  // ECStypes.h:158-163
    Entity familyEntity;
    Entity collideWithEntity;
    float deltaVector[3] = {0.f, 0.f, 0.f};
};
typedef std::vector<std::pair<Entity, std::vector<CollisionCallbackData>>> CallbackEntityDataPairs;

struct CollidingPair {
    Entity first, second;          // oriented as SweepAndPrune.cpp:63 orients them
    uint32_t firstIndex, secondIndex;  // positions in the frame's entry list
    std::vector<cd_contact> contacts;  // in the first entity's model space
    // CDentriesUncollideRays (include/CollisionDetection/CreateUncollideRays.h:6-14);
    // every ray in the first entity's model space
    std::vector<cd_ray> rays_from_first_to_second, rays_from_second_to_first;
    float average_point_first_modelspace[3];
    float average_point_second_modelspace[3];
    // the response (CollisionDetection.cpp:70-99): zero when neither entity moved
    bool shot = false;
    float delta[3] = {0.f, 0.f, 0.f};
    CollisionCallbackData first_callback, second_callback;
};

class CollisionDetection {
public:
    explicit CollisionDetection(uint32_t leaf_max = 4, uint32_t mode = CD_MODE_PLAIN) : mode_(mode) {
        cd_config cfg{};
        cfg.leaf_max = leaf_max;
        check(cd_create(&ctx_, &cfg), "cd_create");
        cd_tuning t{};
        t.use_graph = 1;  // the batched pipeline is captured once and replayed every frame
        check(cd_set_tuning(ctx_, &t), "cd_set_tuning");
    }
    ~CollisionDetection() {
        if (ctx_) cd_destroy(ctx_);
    }
    CollisionDetection(const CollisionDetection&) = delete;
    CollisionDetection& operator=(const CollisionDetection&) = delete;

    // Mesh ingest: an indexed triangle list (glTF "triangles" mode); the OBB
    // tree is built on the GPU (OBBtree.cpp:321-358).
    uint32_t RegisterMesh(const float* positions, uint32_t nvert, const uint32_t* indices, uint32_t ntri) {
        uint32_t id = 0;
        check(cd_register_mesh(ctx_, positions, nvert, indices, ntri, &id), "cd_register_mesh");
        return id;
    }

    // Vertex normals for a mesh's rays (Triangle.cpp:214-234); without them the
    // face normals are used, as for a glTF mesh without normals.
    void SetMeshNormals(uint32_t mesh, const float* normals, uint32_t nvert) {
        check(cd_set_mesh_normals(ctx_, mesh, normals, nvert), "cd_set_mesh_normals");
    }

    void Reset() { entries_.clear(); }  // CollisionDetection.cpp:28-31

    void AddCollisionDetectionEntry(const CollisionDetectionEntry in_collisionDetectionEntry) {  // :33-36
        cd_entity e{};
        std::memcpy(e.current, in_collisionDetectionEntry.currentGlobalMatrix, sizeof e.current);
        std::memcpy(e.previous, in_collisionDetectionEntry.previousGlobalMatrix, sizeof e.previous);
        e.mesh_id = in_collisionDetectionEntry.meshId;
        e.entity = in_collisionDetectionEntry.entity;
        e.should_callback = in_collisionDetectionEntry.shouldCallback ? 1 : 0;
        entries_.push_back(e);
    }

    void ExecuteCollisionDetection() {  // :38-67, phases 1-3
        results_.clear();
        if (entries_.size() < 2) return;
        // The output buffers keep the largest sizes seen, with a quarter more
        // room (they never shrink): a frame that fits them takes one call; a
        // larger one is called again.
        uint32_t np = 0;
        uint64_t nc = 0;
        int rc = cd_frame(ctx_, entries_.data(), uint32_t(entries_.size()), mode_, pairs_.data(), uint32_t(pairs_.size()),
                          &np, contacts_.data(), contacts_.size(), &nc);
        if (rc == CD_E_CAPACITY) {
            pairs_.resize(std::max<size_t>(pairs_.size(), np + np / 4));
            contacts_.resize(std::max<size_t>(contacts_.size(), nc + nc / 4));
            rc = cd_frame(ctx_, entries_.data(), uint32_t(entries_.size()), mode_, pairs_.data(), uint32_t(pairs_.size()),
                          &np, contacts_.data(), contacts_.size(), &nc);
        }
        check(rc, "cd_frame");
        // phase 3 (CreateUncollideRays), on the contacts the frame left on the device
        uint32_t nrp = 0;
        uint64_t nr = 0;
        rc = cd_uncollide_rays(ctx_, ray_pairs_.data(), uint32_t(ray_pairs_.size()), &nrp, rays_.data(), rays_.size(), &nr);
        if (rc == CD_E_CAPACITY) {
            ray_pairs_.resize(std::max<size_t>(ray_pairs_.size(), nrp + nrp / 4));
            rays_.resize(std::max<size_t>(rays_.size(), nr + nr / 4));
            rc = cd_uncollide_rays(ctx_, ray_pairs_.data(), uint32_t(ray_pairs_.size()), &nrp, rays_.data(), rays_.size(),
                                   &nr);
        }
        check(rc, "cd_uncollide_rays");
        size_t k = 0;  // both lists follow the frame's pair order
        for (uint32_t i = 0; i < nrp; ++i) {
            const cd_pair_rays& rp = ray_pairs_[i];
            while (k < np && (pairs_[k].first != rp.first || pairs_[k].second != rp.second)) ++k;
            CollidingPair r;
            r.firstIndex = rp.first;
            r.secondIndex = rp.second;
            r.first = entries_[rp.first].entity;
            r.second = entries_[rp.second].entity;
            if (k < np)
                r.contacts.assign(contacts_.begin() + pairs_[k].contact_begin,
                                  contacts_.begin() + pairs_[k].contact_begin + pairs_[k].contact_count);
            r.rays_from_first_to_second.assign(rays_.begin() + rp.begin_first, rays_.begin() + rp.begin_first + rp.n_first);
            r.rays_from_second_to_first.assign(rays_.begin() + rp.begin_second,
                                               rays_.begin() + rp.begin_second + rp.n_second);
            std::memcpy(r.average_point_first_modelspace, rp.avg_first, sizeof rp.avg_first);
            std::memcpy(r.average_point_second_modelspace, rp.avg_second, sizeof rp.avg_second);
            results_.push_back(std::move(r));
        }
        if (results_.empty()) {
            MakeCallbacks({});
            return;
        }
        // ray shooting and the movement split (:69-99), one record per pair with rays
        std::vector<cd_pair_response> resp(results_.size());
        uint32_t nresp = 0;
        check(cd_shoot_rays(ctx_, nullptr, resp.data(), uint32_t(resp.size()), &nresp), "cd_shoot_rays");
        std::unordered_map<Entity, std::vector<CollisionCallbackData>> callbacks_to_be_made;
        for (size_t i = 0; i < results_.size() && i < nresp; ++i) {
            CollidingPair& r = results_[i];
            r.shot = resp[i].shot != 0;
            std::memcpy(r.delta, resp[i].delta, sizeof r.delta);
            r.first_callback.familyEntity = r.first;
            r.first_callback.collideWithEntity = r.second;
            std::memcpy(r.first_callback.deltaVector, resp[i].delta_first, sizeof r.first_callback.deltaVector);
            r.second_callback.familyEntity = r.second;
            r.second_callback.collideWithEntity = r.first;
            std::memcpy(r.second_callback.deltaVector, resp[i].delta_second, sizeof r.second_callback.deltaVector);
            // the callback goes to every ancestor the two entities do not share (:101-121)
            const std::vector<Entity> fa = Ancestors(r.first), sa = Ancestors(r.second);
            for (size_t k = 0; k != fa.size(); ++k)
                if (k >= sa.size() || fa[k] != sa[k]) callbacks_to_be_made[fa[k]].emplace_back(r.first_callback);
            for (size_t k = 0; k != sa.size(); ++k)
                if (k >= fa.size() || fa[k] != sa[k]) callbacks_to_be_made[sa[k]].emplace_back(r.second_callback);
        }
        MakeCallbacks(std::move(callbacks_to_be_made));
    }

    // Engine hooks for the last steps (CollisionDetection.cpp:101-134).
    // Without an ancestor provider an entity is its own only ancestor.
    void SetAncestorsProvider(std::function<std::vector<Entity>(Entity)> f) { ancestors_ = std::move(f); }
    void SetCallbackSink(std::function<void(const CallbackEntityDataPairs&)> f) { sink_ = std::move(f); }

    const std::vector<CollidingPair>& GetCollidingPairs() const { return results_; }
    cd_ctx* Context() const { return ctx_; }

private:
    std::vector<Entity> Ancestors(Entity e) const { return ancestors_ ? ancestors_(e) : std::vector<Entity>{e}; }

    // CollisionDetection::MakeCallbacks (:124-134): the map becomes a vector
    // of (entity, callbacks) pairs handed to the components.
    void MakeCallbacks(std::unordered_map<Entity, std::vector<CollisionCallbackData>>&& callbacks_to_be_made) const {
        CallbackEntityDataPairs v(std::make_move_iterator(callbacks_to_be_made.begin()),
                                  std::make_move_iterator(callbacks_to_be_made.end()));
        if (sink_) sink_(v);
    }

    void check(int rc, const char* what) const {
        if (rc != CD_OK) throw std::runtime_error(std::string(what) + ": " + (ctx_ ? cd_last_error(ctx_) : "no context"));
    }

    cd_ctx* ctx_ = nullptr;
    uint32_t mode_;
    std::vector<cd_entity> entries_;
    std::vector<cd_pair> pairs_;
    std::vector<cd_contact> contacts_;
    std::vector<cd_pair_rays> ray_pairs_;
    std::vector<cd_ray> rays_;
    std::vector<CollidingPair> results_;
    std::function<std::vector<Entity>(Entity)> ancestors_;
    std::function<void(const CallbackEntityDataPairs&)> sink_;
};

// The collision part of MeshesOfNodes (include/Graphics/Meshes/MeshesOfNodes.h:
// 28, 48-51): AddMeshesOfModel gives every mesh of a glTF model its tree
// (src/Graphics/Meshes/MeshesOfNodes.cpp:29-64), GetMeshIndexOffsetOfModel
// (:66-73) and GetMeshInfo(i).boundBoxTree name it.  A model is named by its
// path where the reference keys it by its tinygltf::Model; boundBoxTree is
// the mesh id entries carry (CD_NO_MESH for a mesh with no collision
// triangles: skinned or morphed only, which the reference never attaches a
// collision component to, src/GameImporter.cpp:320-321).
class MeshesOfNodes {
public:
    struct MeshInfo {
        uint32_t boundBoxTree = CD_NO_MESH;
    };

    explicit MeshesOfNodes(CollisionDetection& collisionDetection) : cd_(collisionDetection) {}

    void AddMeshesOfModel(const std::string& path) {
        cd_gltf* doc = nullptr;
        char err[512];
        if (cd_gltf_open(path.c_str(), &doc, err, sizeof err) != CD_OK)
            throw std::runtime_error("cd_gltf_open: " + std::string(err));
        uint32_t n = 0;
        cd_gltf_mesh_count(doc, &n);
        std::vector<uint32_t> ids(n);
        const int rc = cd_register_gltf(cd_.Context(), doc, ids.data(), n);
        cd_gltf_close(doc);
        if (rc != CD_OK) throw std::runtime_error(std::string("cd_register_gltf: ") + cd_last_error(cd_.Context()));
        modelToMeshIndexOffset_.emplace(path, meshes_.size());
        for (uint32_t id : ids) meshes_.push_back(MeshInfo{id});
    }

    size_t GetMeshIndexOffsetOfModel(const std::string& path) const {
        auto it = modelToMeshIndexOffset_.find(path);
        if (it == modelToMeshIndexOffset_.end()) throw std::runtime_error("model not added: " + path);
        return it->second;
    }

    const MeshInfo& GetMeshInfo(size_t this_mesh_index) const { return meshes_.at(this_mesh_index); }
    size_t size() const { return meshes_.size(); }

private:
    CollisionDetection& cd_;
    std::vector<MeshInfo> meshes_;
    std::unordered_map<std::string, size_t> modelToMeshIndexOffset_;
};

}  // namespace mi355x
