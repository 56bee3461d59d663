// TEST INFRASTRUCTURE ONLY — the oracle.  Nothing in the shipped library may
// include, link or call this code; tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py use it as the checker.
//
// CPU restatement of inMyRoom_vulkan's collision path (broad phase, OBB-tree
// build, OBB-tree traversal, Möller triangle-triangle test, contact output),
// written from the reference's behaviour.  Strict IEEE semantics: build with
// -O2 -ffp-contract=off and no fast-math, so every float/double operation is
// rounded exactly once, in the order written.
//
// glm (an un-vendored submodule of the reference, fork thesmallcreeper/glm,
// pinned commit unknown; public glm 0.9.9 semantics) is restated here as the
// handful of formulas the path uses.  Their expression order follows glm 0.9.9:
//   dot(a,b)        = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(a,b)      = (a.y*b.z - b.y*a.z, a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y)
//   length(v)       = sqrt(dot(v,v));  normalize(v) = v * (1 / sqrt(dot(v,v)))
//   mat4 * vec4     = (m0*x + m1*y) + (m2*z + m3*w)           (pairwise)
//   mat4 * mat4     = ((A0*b0 + A1*b1) + A2*b2) + A3*b3       (per column)
//   inverse(mat4)   = cofactor expansion, 1/det scaling
#pragma once
#include <cstdint>
#include <cstddef>
#include <vector>

namespace orc {

struct V3 { float x, y, z; };
struct D3 { double x, y, z; };
struct M4 { float m[16]; };  // column-major: m[c*4 + r]

// A "Paralgram" / OBB: centre plus three half-extent side vectors
// (inMyRoom_vulkan/include/Geometry/Paralgram.h:29-35).
struct Box { V3 c, u, v, w; };

// One flattened internal node (inMyRoom_vulkan/include/Geometry/OBBtree.h:61-75).
struct Node {
    Box obb[2];          // [0] = left child, [1] = right child
    uint32_t link[2];    // child node index (cnt==0) or triangle offset (cnt>0)
    uint16_t cnt[2];     // 0 = internal child, else triangles in the leaf
};

struct Tree {
    Box root;
    std::vector<Node> nodes;       // DFS preorder of internal nodes; node 0 = root
    std::vector<uint32_t> perm;    // tree-order slot -> input triangle index
    std::vector<float> pos;        // tree-order triangle positions, 9 floats each
    int depth = 0;                 // levels below the root (root leaf = 0)
    double slack = 0.0;            // max distance of a vertex outside a box that should contain it
    double radius = 0.0;           // max |vertex| in model space
    // Ray-generation attributes (Triangle.cpp:214-280), tree order; empty
    // until orc_tree_set_attrs: corner normals (9 floats per triangle, the
    // mesh's vertex normals or the face normal) and index-buffer vertex ids.
    std::vector<float> nrm;
    std::vector<uint32_t> vid;
};

struct LeafPair { uint32_t offA, cntA, offB, cntB; };

// Contact record: one colliding (intersecting, non-coplanar) triangle pair,
// expressed in the first entity's model space.
struct Contact {
    uint32_t query;
    uint32_t triA, triB;  // tree-order triangle slots
    uint32_t flags;       // bits 0-2: A vertex k outside B's plane; bits 3-5: B vertex k outside A's plane
    float s[3], t[3];     // intersection segment (source, target)
    float na[3], nb[3];   // unit face normals of A's and (transformed) B's triangle
};

struct Counters { uint64_t sat = 0, leafpairs = 0, tritests = 0, contacts = 0, ub = 0; };

// One uncollide ray (Ray.h): origin and direction, both in the first
// entity's model space.
struct RayRec { float o[3], d[3]; };

// CDentriesUncollideRays (CreateUncollideRays.h:6-14) for one entity pair.
struct PairRays {
    uint32_t first = 0, second = 0;
    std::vector<RayRec> from_first, from_second;
    float avg_first[3] = {0, 0, 0}, avg_second[3] = {0, 0, 0};
};

// Closest ray hit on a tree (Ray.h:17-24).
struct RayHit {
    bool hit = false, back = false;
    float dist = __builtin_huge_valf();
    uint32_t tri = 0xFFFFFFFFu;  // tree-order slot
    float bx = 0.f, by = 0.f;
};
// One entity pair's response (CollisionDetection.cpp:70-99): the uncollide
// delta in world space and its split between the two entities.
struct PairResponse {
    uint32_t first = 0, second = 0;
    bool shot = false;  // false when neither entity moved: all deltas zero
    float delta[3] = {0, 0, 0}, delta_first[3] = {0, 0, 0}, delta_second[3] = {0, 0, 0};
};
struct ShootStats { uint64_t passes = 0, successes = 0; };

bool ray_triangle(V3 dir, V3 v0, V3 v1, V3 v2, float& bx, float& by, float& dist, bool& back);
bool ray_box(V3 dir, const Box& b, float& tmin, float& tmax);
RayHit ray_tree(const Tree& T, const M4& M, V3 o, V3 d);
PairResponse respond(const Tree& first, const Tree& second, const M4& first_cur, const M4& first_prev,
                     const M4& second_cur, const M4& second_prev, const PairRays& pr, float cos_start,
                     float cos_finish, float bias, ShootStats* stats);

}  // namespace orc
