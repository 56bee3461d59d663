// CollisionDetection.hpp — C++ host mirror of the reference's collision
// interface, implemented over the C ABI of mi355x_collide.h.
//
// Same member names, call order and argument meaning as
// inMyRoom_vulkan/include/CollisionDetection/CollisionDetection.h:11-19
// (Reset, AddCollisionDetectionEntry, ExecuteCollisionDetection), with the
// entry type of inMyRoom_vulkan/include/ECS/ECStypes.h:149-156 restated on
// plain floats (the OBBtree pointer becomes a registered mesh id).
// ExecuteCollisionDetection runs phases 1-3 of
// src/CollisionDetection/CollisionDetection.cpp:38-67 on the GPU and keeps the
// colliding entity pairs with their contacts; ray shooting and callbacks
// (CollisionDetection.cpp:69-141) stay with the engine.
//
// Error behaviour: the reference has no error returns; this mirror throws
// std::runtime_error with the library's message when a call fails (the C ABI
// itself never throws).
#pragma once

#include <cstring>
#include <stdexcept>
#include <string>
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

struct CollidingPair {
    Entity first, second;          // oriented as SweepAndPrune.cpp:63 orients them
    uint32_t firstIndex, secondIndex;  // positions in the frame's entry list
    std::vector<cd_contact> contacts;  // in the first entity's model space
};

class CollisionDetection {
public:
    explicit CollisionDetection(uint32_t leaf_max = 4, uint32_t mode = CD_MODE_PLAIN) : mode_(mode) {
        cd_config cfg{};
        cfg.leaf_max = leaf_max;
        check(cd_create(&ctx_, &cfg), "cd_create");
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
        uint32_t np = 0;
        uint64_t nc = 0;
        int rc = cd_frame(ctx_, entries_.data(), uint32_t(entries_.size()), mode_, nullptr, 0, &np, nullptr, 0, &nc);
        if (rc == CD_E_CAPACITY) {
            pairs_.resize(np);
            contacts_.resize(nc);
            rc = cd_frame(ctx_, entries_.data(), uint32_t(entries_.size()), mode_, pairs_.data(), np, &np,
                          contacts_.data(), nc, &nc);
        }
        check(rc, "cd_frame");
        pairs_.resize(np);
        contacts_.resize(nc);
        for (const cd_pair& p : pairs_) {
            CollidingPair r;
            r.firstIndex = p.first;
            r.secondIndex = p.second;
            r.first = entries_[p.first].entity;
            r.second = entries_[p.second].entity;
            r.contacts.assign(contacts_.begin() + p.contact_begin, contacts_.begin() + p.contact_begin + p.contact_count);
            results_.push_back(std::move(r));
        }
    }

    const std::vector<CollidingPair>& GetCollidingPairs() const { return results_; }
    cd_ctx* Context() const { return ctx_; }

private:
    void check(int rc, const char* what) const {
        if (rc != CD_OK) throw std::runtime_error(std::string(what) + ": " + (ctx_ ? cd_last_error(ctx_) : "no context"));
    }

    cd_ctx* ctx_ = nullptr;
    uint32_t mode_;
    std::vector<cd_entity> entries_;
    std::vector<cd_pair> pairs_;
    std::vector<cd_contact> contacts_;
    std::vector<CollidingPair> results_;
};

}  // namespace mi355x
