// CollisionDetection.hpp — C++ host mirror of the reference's collision
// interface, implemented over the C ABI of mi355x_collide.h.
//
// Same member names, call order and argument meaning as
// inMyRoom_vulkan/include/CollisionDetection/CollisionDetection.h:11-19
// (CollisionDetection(ECSwrapper*), Reset, AddCollisionDetectionEntry,
// ExecuteCollisionDetection).  BasicCollisionDetection takes the engine's own
// CollisionDetectionEntry and CollisionCallbackData (ECStypes.h:149-163) as
// template arguments, so the engine's callers compile unchanged: the entry's
// `const OBBtree* OBBtree_ptr` (set at ModelCollisionCompEntity.cpp:73) is
// mapped to the GPU mesh by TreeRegistry, and the ECS wrapper's entities
// handler and components receive the callbacks as the reference hands them
// over.  mi355x::CollisionDetection instantiates it on plain-float
// restatements of those types for hosts without the engine.
// ExecuteCollisionDetection runs src/CollisionDetection/CollisionDetection.cpp:38-141
// on the GPU: phases 1-3 (broad, mid, narrow), the uncollide rays of
// CreateUncollideRays.cpp:60-200, the ray shooting of ShootUncollideRays.cpp
// with the movement split (:69-99), then the callback fan-out to entity
// ancestors (:101-121) and MakeCallbacks (:124-134).
//
// Error behaviour: the reference has no error returns; this mirror throws
// std::runtime_error with the library's message when a call fails (the C ABI
// itself never throws).
#pragma once

#include <algorithm>
#include <cstring>
#include <deque>
#include <functional>
#include <iterator>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <utility>
#include <vector>

#include "mi355x_collide.h"

namespace mi355x {

typedef uint16_t Entity;  // ECStypes.h:18

// With this header alone, an entry's `const OBBtree*` points to a handle of a
// registered GPU tree (MeshesOfNodes::MeshInfo::boundBoxTree below).  In the
// engine the pointer is the engine's own OBBtree (MeshesOfNodes.h:28), which
// TreeRegistry maps to its mesh id.
class OBBtree {
public:
    OBBtree() = default;
    explicit OBBtree(uint32_t mesh_id) : meshId(mesh_id) {}
    uint32_t meshId = CD_NO_MESH;
};

// Column-major 4x4 with glm's access pattern: m[column][row].
struct Mat4 {
    float m[16];
    float* operator[](int column) { return m + 4 * column; }
    const float* operator[](int column) const { return m + 4 * column; }
};
struct Vec3 {
    float v[3] = {0.f, 0.f, 0.f};
    float& operator[](int k) { return v[k]; }
    const float& operator[](int k) const { return v[k]; }
};

struct CollisionDetectionEntry {  // ECStypes.h:149-156
    Mat4 currentGlobalMatrix;
    Mat4 previousGlobalMatrix;
    const OBBtree* OBBtree_ptr = nullptr;
    bool shouldCallback = false;
    Entity entity = 0;
};

struct CollisionCallbackData {  // ECStypes.h:158-163
    Entity familyEntity;
    Entity collideWithEntity;
    Vec3 deltaVector;
};

// Engine tree object -> GPU mesh id.  The engine keeps its meshes in a
// std::vector grown by emplace_back while models are added
// (MeshesOfNodes.cpp:62, MeshesOfNodes.h:62), so a tree's address moves until
// the last model is in; MeshesOfNodes::GetMeshInfo asserts that the meshes
// were flashed (MeshesOfNodes.h:48), after which the vector never grows.
// Register each mesh by its index at load, then, once the list is final,
// bind the addresses: Bind(&GetMeshInfo(i).boundBoxTree, i).
class TreeRegistry {
public:
    void SetMeshIdOfIndex(size_t mesh_index, uint32_t mesh_id) {
        if (idOfIndex_.size() <= mesh_index) idOfIndex_.resize(mesh_index + 1, CD_NO_MESH);
        idOfIndex_[mesh_index] = mesh_id;
    }
    void Bind(const void* tree, size_t mesh_index) {
        if (mesh_index >= idOfIndex_.size()) throw std::runtime_error("TreeRegistry::Bind: mesh index never registered");
        idOfTree_[tree] = idOfIndex_[mesh_index];
    }
    void Clear() { idOfTree_.clear(); }
    uint32_t MeshId(const OBBtree* t) const {  // the handle carries its id
        if (!t) throw std::runtime_error("collision entry without a tree");
        return t->meshId;
    }
    template <class Tree> uint32_t MeshId(const Tree* t) const {
        auto it = idOfTree_.find(static_cast<const void*>(t));
        if (it == idOfTree_.end())
            throw std::runtime_error("collision entry names a tree that was never bound (bind after the last model)");
        return it->second;
    }

private:
    std::vector<uint32_t> idOfIndex_;
    std::unordered_map<const void*, uint32_t> idOfTree_;
};

// One colliding entity pair of the last ExecuteCollisionDetection.
template <class EntityT, class CallbackData>
struct BasicCollidingPair {
    EntityT first, second;                 // oriented as SweepAndPrune.cpp:63 orients them
    uint32_t firstIndex, secondIndex;      // positions in the frame's entry list
    std::vector<cd_contact> contacts;      // in the first entity's model space
    // CDentriesUncollideRays (include/CollisionDetection/CreateUncollideRays.h:6-14);
    // every ray in the first entity's model space
    std::vector<cd_ray> rays_from_first_to_second, rays_from_second_to_first;
    float average_point_first_modelspace[3];
    float average_point_second_modelspace[3];
    // the response (CollisionDetection.cpp:70-99): zero when neither entity moved
    bool shot = false;
    float delta[3] = {0.f, 0.f, 0.f};
    CallbackData first_callback{}, second_callback{};
};

// CollisionDetection (CollisionDetection.h:11-19) over the engine's own entry
// and callback types: an Entry has currentGlobalMatrix / previousGlobalMatrix
// (m[column][row] access, glm::mat4 or Mat4), OBBtree_ptr, shouldCallback and
// entity; a CallbackData has familyEntity, collideWithEntity and deltaVector
// (v[k] access, glm::vec3 or Vec3).  In the engine:
//   using CollisionDetection = mi355x::BasicCollisionDetection<CollisionDetectionEntry, CollisionCallbackData>;
template <class Entry = CollisionDetectionEntry, class CallbackData = CollisionCallbackData>
class BasicCollisionDetection {
public:
    using EntityT = typename std::decay<decltype(std::declval<Entry&>().entity)>::type;
    using CallbackEntityDataPairs = std::vector<std::pair<EntityT, std::vector<CallbackData>>>;
    using CollidingPair = BasicCollidingPair<EntityT, CallbackData>;

    explicit BasicCollisionDetection(uint32_t leaf_max = 4, uint32_t mode = CD_MODE_REFERENCE) : mode_(mode) {
        cd_config cfg{};
        cfg.leaf_max = leaf_max;
        check(cd_create(&ctx_, &cfg), "cd_create");
        cd_tuning t{};
        t.use_graph = 1;  // the batched pipeline is captured once and replayed every frame
        check(cd_set_tuning(ctx_, &t), "cd_set_tuning");
    }

    // The reference's constructor (CollisionDetection.h:14): the engine's ECS
    // gives the ancestors of an entity (CollisionDetection.cpp:106-107) and the
    // components whose CollisionCallback receives the callbacks (:124-134).
    template <class ECSwrapperT, class = decltype(std::declval<ECSwrapperT&>().GetEntitiesHandler())>
    explicit BasicCollisionDetection(ECSwrapperT* ecs, uint32_t mode = CD_MODE_REFERENCE, uint32_t leaf_max = 4)
        : BasicCollisionDetection(leaf_max, mode) {
        ancestors_ = [ecs](EntityT e) {
            const auto a = ecs->GetEntitiesHandler()->GetEntityAncestors(e);
            return std::vector<EntityT>(a.begin(), a.end());
        };
        sink_ = [ecs](const CallbackEntityDataPairs& v) {
            for (const auto& this_compID_component_ptr_pair : ecs->GetComponentIDtoComponentBaseClassMap())
                if (this_compID_component_ptr_pair.second != nullptr)
                    this_compID_component_ptr_pair.second->CollisionCallback(v);
        };
    }
    ~BasicCollisionDetection() {
        if (ctx_) cd_destroy(ctx_);
    }
    BasicCollisionDetection(const BasicCollisionDetection&) = delete;
    BasicCollisionDetection& operator=(const BasicCollisionDetection&) = delete;

    // Mesh ingest: an indexed triangle list (glTF "triangles" mode); the OBB
    // tree is built on the GPU (OBBtree.cpp:321-358).
    uint32_t RegisterMesh(const float* positions, uint32_t nvert, const uint32_t* indices, uint32_t ntri) {
        uint32_t id = 0;
        check(cd_register_mesh(ctx_, positions, nvert, indices, ntri, &id), "cd_register_mesh");
        return id;
    }
    // A triangle soup (9 floats per triangle), as GetOBBtreeAndReset hands
    // its triangles to OBBtree (PrimitivesOfMeshes.cpp:840-863).
    uint32_t RegisterTriangles(const float* tri9, uint32_t ntri) {
        uint32_t id = 0;
        check(cd_register_soups(ctx_, tri9, &ntri, 1, &id), "cd_register_soups");
        return id;
    }

    // Vertex normals for a mesh's rays (Triangle.cpp:214-234); without them the
    // face normals are used, as for a glTF mesh without normals.
    void SetMeshNormals(uint32_t mesh, const float* normals, uint32_t nvert) {
        check(cd_set_mesh_normals(ctx_, mesh, normals, nvert), "cd_set_mesh_normals");
    }

    // Engine trees: see TreeRegistry.
    TreeRegistry& Trees() { return trees_; }

    void Reset() { entries_.clear(); }  // CollisionDetection.cpp:28-31

    void AddCollisionDetectionEntry(const Entry in_collisionDetectionEntry) {  // :33-36
        const Entry& src = in_collisionDetectionEntry;
        cd_entity e{};
        for (int c = 0; c < 4; ++c)
            for (int r = 0; r < 4; ++r) {
                e.current[4 * c + r] = float(src.currentGlobalMatrix[c][r]);
                e.previous[4 * c + r] = float(src.previousGlobalMatrix[c][r]);
            }
        e.mesh_id = trees_.MeshId(src.OBBtree_ptr);
        e.entity = uint16_t(src.entity);
        e.should_callback = src.shouldCallback ? 1 : 0;
        entries_.push_back(e);
    }

    void ExecuteCollisionDetection() {  // :38-141
        results_.clear();
        if (entries_.size() < 2) return;
        if (!keep_details_) {
            // one call for the whole step; contacts and rays stay on the device
            uint32_t nresp = 0;
            int rc = cd_step(ctx_, entries_.data(), uint32_t(entries_.size()), mode_, nullptr, resp_.data(),
                             uint32_t(resp_.size()), &nresp);
            if (rc == CD_E_CAPACITY) {  // the frame and its rays are done: shoot into a larger buffer
                resp_.resize(std::max<size_t>(resp_.size(), nresp + nresp / 4));
                rc = cd_shoot_rays(ctx_, nullptr, resp_.data(), uint32_t(resp_.size()), &nresp);
            }
            check(rc, "cd_step");
            for (uint32_t i = 0; i < nresp; ++i) {
                CollidingPair r;
                r.firstIndex = resp_[i].first;
                r.secondIndex = resp_[i].second;
                r.first = EntityT(entries_[resp_[i].first].entity);
                r.second = EntityT(entries_[resp_[i].second].entity);
                results_.push_back(std::move(r));
            }
            Respond(resp_.data(), nresp);
            return;
        }
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
            r.first = EntityT(entries_[rp.first].entity);
            r.second = EntityT(entries_[rp.second].entity);
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
        resp_.resize(std::max(resp_.size(), results_.size()));
        uint32_t nresp = 0;
        check(cd_shoot_rays(ctx_, nullptr, resp_.data(), uint32_t(resp_.size()), &nresp), "cd_shoot_rays");
        Respond(resp_.data(), nresp);
    }

    // Keep every pair's contacts and rays in GetCollidingPairs (three calls
    // per step, the contacts and rays copied to the host); off by default,
    // when the step is one cd_step and a pair holds its entities, shot flag
    // and deltas.
    void SetKeepDetails(bool keep) { keep_details_ = keep; }

    // Engine hooks for the last steps (CollisionDetection.cpp:101-134), for a
    // host without the engine's ECS.  Without an ancestor provider an entity
    // is its own only ancestor.
    void SetAncestorsProvider(std::function<std::vector<EntityT>(EntityT)> f) { ancestors_ = std::move(f); }
    void SetCallbackSink(std::function<void(const CallbackEntityDataPairs&)> f) { sink_ = std::move(f); }

    const std::vector<CollidingPair>& GetCollidingPairs() const { return results_; }
    cd_ctx* Context() const { return ctx_; }

private:
    // The responses of results_'s pairs, in order: shot flags, deltas, the
    // two callback records of each pair sent to every ancestor the two
    // entities do not share (:101-121), then MakeCallbacks.
    void Respond(const cd_pair_response* resp, uint32_t nresp) {
        std::unordered_map<EntityT, std::vector<CallbackData>> callbacks_to_be_made;
        for (size_t i = 0; i < results_.size() && i < nresp; ++i) {
            CollidingPair& r = results_[i];
            r.shot = resp[i].shot != 0;
            std::memcpy(r.delta, resp[i].delta, sizeof r.delta);
            r.first_callback.familyEntity = r.first;
            r.first_callback.collideWithEntity = r.second;
            r.second_callback.familyEntity = r.second;
            r.second_callback.collideWithEntity = r.first;
            for (int c = 0; c < 3; ++c) {
                r.first_callback.deltaVector[c] = resp[i].delta_first[c];
                r.second_callback.deltaVector[c] = resp[i].delta_second[c];
            }
            // the callback goes to every ancestor the two entities do not share (:101-121)
            const std::vector<EntityT> fa = Ancestors(r.first), sa = Ancestors(r.second);
            for (size_t a = 0; a != fa.size(); ++a)
                if (a >= sa.size() || fa[a] != sa[a]) callbacks_to_be_made[fa[a]].emplace_back(r.first_callback);
            for (size_t a = 0; a != sa.size(); ++a)
                if (a >= fa.size() || fa[a] != sa[a]) callbacks_to_be_made[sa[a]].emplace_back(r.second_callback);
        }
        MakeCallbacks(std::move(callbacks_to_be_made));
    }

    std::vector<EntityT> Ancestors(EntityT e) const { return ancestors_ ? ancestors_(e) : std::vector<EntityT>{e}; }

    // CollisionDetection::MakeCallbacks (:124-134): the map becomes a vector
    // of (entity, callbacks) pairs handed to the components.
    void MakeCallbacks(std::unordered_map<EntityT, std::vector<CallbackData>>&& callbacks_to_be_made) const {
        CallbackEntityDataPairs v(std::make_move_iterator(callbacks_to_be_made.begin()),
                                  std::make_move_iterator(callbacks_to_be_made.end()));
        if (sink_) sink_(v);
    }

    void check(int rc, const char* what) const {
        if (rc != CD_OK) throw std::runtime_error(std::string(what) + ": " + (ctx_ ? cd_last_error(ctx_) : "no context"));
    }

    cd_ctx* ctx_ = nullptr;
    uint32_t mode_;
    TreeRegistry trees_;
    std::vector<cd_entity> entries_;
    std::vector<cd_pair> pairs_;
    std::vector<cd_contact> contacts_;
    std::vector<cd_pair_rays> ray_pairs_;
    std::vector<cd_ray> rays_;
    std::vector<cd_pair_response> resp_ = std::vector<cd_pair_response>(64);
    bool keep_details_ = false;
    std::vector<CollidingPair> results_;
    std::function<std::vector<EntityT>(EntityT)> ancestors_;
    std::function<void(const CallbackEntityDataPairs&)> sink_;
};

using CollisionDetection = BasicCollisionDetection<>;
using CallbackEntityDataPairs = CollisionDetection::CallbackEntityDataPairs;
using CollidingPair = CollisionDetection::CollidingPair;

// The collision part of MeshesOfNodes (include/Graphics/Meshes/MeshesOfNodes.h:
// 28, 48-51): AddMeshesOfModel gives every mesh of a glTF model its tree
// (src/Graphics/Meshes/MeshesOfNodes.cpp:29-64), GetMeshIndexOffsetOfModel
// (:66-73) and GetMeshInfo(i).boundBoxTree name it.  A model is named by its
// path where the reference keys it by its tinygltf::Model; boundBoxTree is an
// OBBtree handle whose meshId is the GPU tree (CD_NO_MESH for a mesh with no
// collision triangles: skinned or morphed only, which the reference never
// attaches a collision component to, src/GameImporter.cpp:320-321), so
// `entry.OBBtree_ptr = &GetMeshInfo(i).boundBoxTree` works as in the engine.
class MeshesOfNodes {
public:
    struct MeshInfo {
        OBBtree boundBoxTree;  // a handle: boundBoxTree.meshId is the GPU tree
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
        for (uint32_t id : ids) meshes_.push_back(MeshInfo{OBBtree(id)});
    }

    size_t GetMeshIndexOffsetOfModel(const std::string& path) const {
        auto it = modelToMeshIndexOffset_.find(path);
        if (it == modelToMeshIndexOffset_.end()) throw std::runtime_error("model not added: " + path);
        return it->second;
    }

    // the address stays valid while models are added (a deque never moves its
    // elements when it grows at the end), so entries may take it at any time
    const MeshInfo& GetMeshInfo(size_t this_mesh_index) const { return meshes_.at(this_mesh_index); }
    size_t size() const { return meshes_.size(); }

private:
    CollisionDetection& cd_;
    std::deque<MeshInfo> meshes_;
    std::unordered_map<std::string, size_t> modelToMeshIndexOffset_;
};

}  // namespace mi355x
