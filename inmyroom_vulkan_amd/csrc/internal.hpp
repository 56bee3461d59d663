// Internal layouts of the collision library (device arenas and records).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/mi355x_collide.h"
#include "geom.hpp"

namespace cdk {

// One child box of an internal OBB-tree node as the traversals read it: 64 B,
// four 16-byte loads from one half of a 128-byte line.  A node's two child
// records hold the content of the reference's OBBtreeNode
// (include/Geometry/OBBtree.h:61-75) plus each child box's surface, which the
// traversal's descent rule compares (OBBtree.cpp:426) and which is a pure
// function of the stored box.
constexpr uint32_t LEAF_BIT = 0x80000000u;
struct alignas(64) ChildRec {
    float box[12];   // centre, u, v, w
    float surf;      // surface(box)
    uint32_t link;   // internal child: mesh-local node index; leaf: mesh-local triangle offset (= toff)
    uint32_t tris;   // triangles under the child, LEAF_BIT set for a leaf
    uint32_t toff;   // the child's triangles: [toff, toff + tris) in tree order, mesh-local
};
static_assert(sizeof(ChildRec) == 64, "child record must stay 64 bytes");
__host__ __device__ inline bool child_leaf(const ChildRec& c) { return (c.tris & LEAF_BIT) != 0u; }
__host__ __device__ inline uint32_t child_tris(const ChildRec& c) { return c.tris & ~LEAF_BIT; }
// triangles of a leaf child, 0 for an internal one
__host__ __device__ inline uint32_t child_cnt(const ChildRec& c) { return child_leaf(c) ? child_tris(c) : 0u; }

// One internal OBB-tree node: its left and right child, 128 B (one L2 line).
struct alignas(128) NodeRec {
    ChildRec c[2];
};
static_assert(sizeof(NodeRec) == 128, "node record must stay one 128-byte line");

// Mesh table entry (device and host).  The root box is kept as a child
// record (link 0: the tree's node 0, or triangle offset 0 when the whole tree
// is one leaf), so a traversal reads a root exactly as it reads any child.
struct alignas(64) MeshRec {
    ChildRec root;        // root OBB (OBBtree.h:129)
    uint32_t node_base;   // first NodeRec of this mesh in the node arena
    uint32_t tri_base;    // first triangle of this mesh in the triangle arena (tree order)
    uint32_t ntri, nnodes;
    uint32_t depth;       // internal levels below the root
    float slack;          // max distance a vertex lies outside a box meant to hold it
    float radius;         // max |vertex|
    uint32_t pad[9];
};
static_assert(sizeof(MeshRec) == 128, "mesh record layout");

// Device copy of cd_query.
struct QueryRec {
    uint32_t mesh_a, mesh_b;
    float rel[16];
};
static_assert(sizeof(QueryRec) == sizeof(cd_query), "query layout");

// Per-query record of the batched path: relative matrix, the slack that
// makes its float32 separating-axis test conservative, and arena bases.
struct QueryPre {
    float m[16];
    float sigma;
    uint32_t mesh_a, mesh_b, node_base_a, node_base_b, tri_base_a, tri_base_b, pad;
};
static_assert(sizeof(QueryPre) == 96, "query record layout");

// The fixed sweep frame of the broad phase (CollisionDetection.cpp:9-11).
struct SapAxes {
    f3 u, v, w;
};
__host__ __device__ inline SapAxes sap_axes() {
    SapAxes a;
    a.u = normalize(mk(0.8f, -0.2f, 0.f));
    a.w = normalize(cross(a.u, mk(0.f, -1.f, 0.f)));
    a.v = normalize(cross(a.w, a.u));
    return a;
}

// One colliding triangle of one side of a query, aggregated over its
// contacts (TriangleCandidateRays, CreateUncollideRays.cpp:13-58).
struct RayCand {
    uint32_t query, tri;   // tree-order triangle slot of this side's mesh
    uint32_t mask, pad;    // vertex k not outside every plane it was tested against
    float w, wp[3];        // sum of weights, sum of weighted midpoints
};
static_assert(sizeof(RayCand) == 32, "candidate layout");

// One uncollide ray's two Hermann passes (the second shoots the reflected
// ray back): each pass's contribution to the summed force response and its
// push-out vector, as the response step accumulates them.
struct RayShot {
    float force[2][3], push[2][3];
    uint32_t ok;    // bit p: pass p succeeded
    uint32_t pair;  // query (entity pair) of the ray
};
// The box-only terms of the ray shooting's box test for one stored box
// (centre c, half-sides u, v, w): the three slab normals v x w, w x u, u x v,
// the box's projection radius on each, and the 1-norm sums of c and of
// u, v, w that bound the test's rounding.  Computed once per box (after each
// build or refit of its mesh) with the very operations the test would use.
struct alignas(64) RaySlab {
    float n[9];
    float h[3];
    float cr, sr;
    float pad[2];
};
static_assert(sizeof(RaySlab) == 64, "slab record layout");

constexpr int RAY_STACK_ENTRIES = 48;  // closest-hit traversal stack per lane (LDS): a tree's depth + 2
constexpr uint32_t RAY_LINK_BITS = 25;  // a stack entry packs a node or triangle index below 2^25 with a leaf count

// Counts of the one-call scene step, kept on the device so that no stage
// waits for the host (words of its count block): the ray generation's
// candidates and rays of each side (the first four words are its totals),
// the pairs the sweep found (also above its capacity) and the ones kept, and
// the contacts the later stages take, and a flag set when a pair has more
// contacts than the step's contact order handles.
constexpr uint32_t DC_CAND0 = 0, DC_RAYS0 = 1, DC_CAND1 = 2, DC_RAYS1 = 3;
constexpr uint32_t DC_SWEEP = 4, DC_NQ = 5, DC_NC = 6, DC_BIGSEG = 7, DC_WORDS = 16;

// A count read on the device: min(*n_dev, n), n being the capacity a launch
// (or a captured graph) was shaped for; without n_dev the count is n.
__device__ __forceinline__ uint32_t count_of(uint32_t n, const uint32_t* n_dev) { return n_dev ? min(*n_dev, n) : n; }

// A refit plan: small boxes (one thread each) and large boxes split into
// chunks of triangles (device arrays, cached per mesh list).
struct RefitPlan {
    const uint2* items;    // (mesh, traversal code) per box
    const uint2* ranges;   // (first triangle, triangles) per box, tree order, mesh-local
    const uint32_t* small; // boxes done whole by one thread
    uint32_t nsmall;
    const uint4* chunks;   // (box, first triangle, triangles, 0)
    uint32_t nchunks;
    const uint4* big;      // (box, first chunk, chunks, 0)
    uint32_t nbig;
};
// One skinned mesh in a skin/refit launch.
struct SkinRec {
    uint32_t mesh, vbase, nvert, nbones, bone_base, tri_base, ntri, vidx_base;
};

constexpr uint32_t ROOT_CODE = 0xFFFFFFFFu;  // traversal code of a tree's root
constexpr int PLAIN_STACK = 64;              // per-lane stack entries of the plain traversal

__host__ __device__ inline Box box_from12(const float* p) {
    Box b;
    b.c = mk(p[0], p[1], p[2]);
    b.u = mk(p[3], p[4], p[5]);
    b.v = mk(p[6], p[7], p[8]);
    b.w = mk(p[9], p[10], p[11]);
    return b;
}
__host__ __device__ inline void box_to12(const Box& b, float* p) {
    p[0] = b.c.x; p[1] = b.c.y; p[2] = b.c.z;
    p[3] = b.u.x; p[4] = b.u.y; p[5] = b.u.z;
    p[6] = b.v.x; p[7] = b.v.y; p[8] = b.v.z;
    p[9] = b.w.x; p[10] = b.w.y; p[11] = b.w.z;
}

// Growable device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need) {
        if (need <= bytes) return hipSuccess;
        const size_t bytes_before = bytes;
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            bytes = 0;
        }
        // at least 1.5x the old size: sizes that creep up from call to call
        // (contacts per frame) would otherwise free and allocate every call,
        // and hipFree waits for the device
        size_t want = need < 4096 ? 4096 : need;
        if (want < bytes_before + bytes_before / 2) want = bytes_before + bytes_before / 2;
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess) bytes = want;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};
// Pinned host memory, grown like DevBuf.
struct HostBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need) {
        if (need <= bytes) return hipSuccess;
        const size_t bytes_before = bytes;
        release();
        size_t want = need < 4096 ? 4096 : need;
        if (want < bytes_before + bytes_before / 2) want = bytes_before + bytes_before / 2;
        hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
        if (e == hipSuccess) bytes = want;
        else p = nullptr;
        return e;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
};
// A device buffer freed when it goes out of scope (temporaries).
struct ScopedBuf : DevBuf {
    ScopedBuf() = default;
    ScopedBuf(const ScopedBuf&) = delete;
    ScopedBuf& operator=(const ScopedBuf&) = delete;
    ~ScopedBuf() { release(); }
};

// One glTF primitive of the mesh ingest: its triangles' place in the output,
// its index list and its vertices (in the concatenated arrays), and the id of
// its first vertex within its mesh.
struct PrimSpan {
    uint32_t tri_off, ntri, idx_off, mode, vbase, vid_base, pad[2];
};
static_assert(sizeof(PrimSpan) == 32, "primitive span layout");
static_assert(sizeof(cd_node) == 64, "node layout");

// Copies of each mesh's slack maximum in the refit (see post_slack).
constexpr uint32_t SLACK_REPS = 64;

// Device-side counters of one run (zeroed by a fill kernel, never by a
// memset, so a run can be captured into a HIP graph).
struct RunCounters {
    unsigned long long sat, leafpairs, tritests, contacts, fp64, exact, ub, err;
    // leaf-pair slots handed out by the batched traversal; slots beyond the
    // leaf pairs it found hold empty entries (zero triangle counts)
    unsigned long long lpslots;
    // triangle tests the batched narrow phase left to its float64 and exact
    // reruns (list lengths; entries beyond the list capacity are not stored)
    unsigned long long deferred, deferred2;
};

// Build of a batch of triangle soups into the arenas (build.hip).
constexpr uint32_t BUILD_SCRATCH = 14;
constexpr size_t BUILD_KEEP = size_t(256) << 20;

struct BuildResult {
    std::vector<MeshRec> meshes;   // node_base / tri_base relative to the arenas
    std::string err;
};

}  // namespace cdk

// The context (opaque in the C ABI).
struct cd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t leaf_max = 4;
    std::string err;

    std::vector<cdk::MeshRec> meshes;      // host copy of the mesh table
    cdk::DevBuf d_meshes, d_nodes, d_tris, d_perm;
    size_t nodes_used = 0, tris_used = 0;
    // ray-generation attributes: leaf of every tree-order triangle (its leaf's
    // first slot), and per mesh {vertex-id base or ~0 (soup: 3 t + k),
    // normal base or ~0 (face normals)} with the arenas they index
    cdk::DevBuf d_leafof, d_attrs, d_vid, d_nrm;
    std::vector<uint2> attrs;
    std::vector<uint32_t> mesh_nvert;      // vertex ids per mesh (3 * ntri for a soup)
    size_t vid_used = 0, nrm_used = 0;
    // what the last run left on the device, for cd_uncollide_rays
    int last_run = 0;                      // 0 none, 1 staged queries, 2 frame
    bool contacts_sorted = false;          // sorted copy in d_work[W_CSORT]
    std::vector<uint2> staged_ids;         // per query: (mesh_a, mesh_b), or per frame pair: (first, second)
    std::vector<uint32_t> frame_mesh;      // mesh of each entity of the last cd_frame
    // rays of the last cd_uncollide_rays, still on the device, for cd_shoot_rays
    bool rays_ready = false;
    uint32_t ray_base1 = 0;                // first ray of the second sides
    uint64_t ray_total = 0;
    uint32_t ray_stride = 0;               // per-side spacing of the ray counts, begins and averages on the device
    std::vector<uint32_t> ray_counts;      // per query: rays of the first side, then (at +nq) of the second

    cdk::DevBuf d_queries, d_leafpairs, d_contacts, d_counters, d_tt_a, d_tt_b, d_tt_flags, d_tt_seg, d_tt_list;
    cdk::HostBuf h_stage;  // pinned staging of results copied back to the host
    cdk::DevBuf d_budget;  // node-pair tests of the last batched run (budgeted traversal pass)
    cdk::DevBuf d_nq;      // query count of the batched run, read by its kernels (graph replays keep it live)
    cdk::HostBuf h_nq;     // pinned staging of that count
    cdk::HostBuf h_pairs;  // pinned staging of a frame's broad-phase pairs
    cdk::DevBuf d_work[48];                 // scratch of the batched path, the broad phase, the rays and the step
    uint32_t staged_nq = 0;
    uint32_t plain_need = 0;  // stack entries the staged queries' plain walk needs (their trees' depths + 2)
    uint64_t staged_contacts = 0;
    uint64_t lp_cap = 0, contact_cap = 0, frontier_cap = 0;
    uint64_t defer_cap = 0;  // entries of each rerun list of the batched narrow phase
    uint64_t dump_cap = 0;   // entries of the budgeted traversal's hand-off and re-split buffers

    cd_counters counters{};
    cd_timings timings{};
    cd_tuning tuning{};
    int num_cu = 256;
    // skinned meshes: bind pose, joints, weights, tree-order vertex indices
    struct SkinHost { uint32_t vbase, nvert, nbones, vidx_base; };
    std::vector<std::pair<uint32_t, SkinHost>> skins;  // (mesh id, data)
    cdk::DevBuf d_bind, d_joints, d_weights, d_vidx, d_skinpos, d_bones, d_skinrec, d_refit_items, d_refit_bits,
        d_refit_scratch, d_refit_ranges, d_refit_plan, d_refit_partial;
    size_t skin_verts = 0, skin_vidx = 0;
    std::vector<uint32_t> refit_key;       // mesh list the cached refit items are for
    uint32_t refit_nitems = 0;
    cdk::RefitPlan refit_plan{};

    void* scratch = nullptr;               // this device's allocation-free scratch array (device address)
    std::string scratch_err;               // why it could not be resolved
    hipGraphExec_t graph = nullptr;        // captured batched pipeline
    std::vector<uint64_t> graph_key;       // sizes, tuning and every device pointer the graph captured
    // the frame and the scene step on the device: capacities (grown when a
    // run exceeds one), their captured graphs and keys, and the pinned block
    // their results land in
    // slab records of the ray shooting's box tests: two per internal node
    // (node arena order), one per mesh root; valid for a mesh until it is
    // refit
    cdk::DevBuf d_rayslab, d_rootslab;
    std::vector<uint8_t> slab_ok;
    // the tree build's work buffers, kept between builds that need at most
    // BUILD_KEEP bytes of them (a model load builds one document at a time)
    cdk::DevBuf build_scratch[cdk::BUILD_SCRATCH];
    uint32_t step_np_cap = 0, step_nc_cap = 0, step_ray_cap = 0;
    bool step_radix = false;  // a pair had more contacts than the per-pair order ranks: the radix sorts
    hipGraphExec_t step_graph = nullptr, frame_graph = nullptr;  // the whole step; cd_frame's part of it
    std::vector<uint64_t> step_key, frame_key;
    cdk::HostBuf h_step, h_ents;  // the run's result block; its entity records, staged for the graph's upload
    cdk::HostBuf h_ctr;           // a run's counters, read back
    hipEvent_t ev[8] = {};
};
