/* mi355x_collide.h — C ABI of the MI355X collision library.
 *
 * Drop-in for the CPU collision path of inMyRoom_vulkan:
 *   entity transforms in -> colliding entity pairs with contact segments,
 *   contact normals and vertex-outside flags out.
 * Plain pointers and sizes only; every function returns an int status
 * (CD_OK = 0, negative on error; cd_last_error() gives the message).  No
 * C++ exception crosses this boundary.  One cd_ctx per host thread; a
 * context is not internally synchronised.  All output buffers are owned by
 * the caller; the library keeps no caller pointer after a call returns.
 *
 * Reference paths are relative to the inMyRoom_vulkan repository root.
 */
#ifndef MI355X_COLLIDE_H
#define MI355X_COLLIDE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define CD_OK 0
#define CD_E_INVALID (-1)  /* bad argument (null pointer, unknown mesh, non-finite input, ...) */
#define CD_E_CAPACITY (-2) /* output buffer too small; the required count is written back */
#define CD_E_HIP (-3)      /* HIP runtime error */
#define CD_E_STACK (-4)    /* a traversal stack would overflow (tree deeper than supported) */
#define CD_E_NOMEM (-5)    /* device allocation failed */
#define CD_E_FORMAT (-6)   /* malformed file, or a layout the mesh ingest does not read */

#define CD_NO_MESH 0xFFFFFFFFu /* a glTF mesh without collision triangles */

/* ---- modes --------------------------------------------------------------- */
/* Plain: reference semantics bit for bit (Möller test with its 1e-6 snap,
 * float SAT), one lane per query.  Batched: persistent load-balanced
 * traversal, exact predicates (float filter, double fallback, exact
 * expansion as last resort), conservative SAT.  Reference: the plain mode's
 * tests on the batched mode's machinery (split levels, persistent
 * traversal, dense narrow phase): the plain mode's contacts and counts, bit
 * for bit, in another order, at the batched mode's speed.  The triangle-pair
 * stream (cd_tritri_batch) runs the plain test for it.                      */
#define CD_MODE_PLAIN 0u
#define CD_MODE_BATCHED 1u
#define CD_MODE_REFERENCE 2u

/* ---- records --------------------------------------------------------------- */
typedef struct cd_ctx cd_ctx;

typedef struct cd_config {
    uint32_t leaf_max;   /* triangles per OBB-tree leaf; 0 -> 4 (include/Geometry/OBBtree.h:49) */
    uint32_t device;     /* HIP device ordinal */
    uint32_t reserved[6];
} cd_config;

/* CollisionDetectionEntry (inMyRoom_vulkan/include/ECS/ECStypes.h:149-156):
 * the OBBtree pointer becomes a registered mesh id.                         */
typedef struct cd_entity {
    float current[16];       /* currentGlobalMatrix, column-major (glm) */
    float previous[16];      /* previousGlobalMatrix */
    uint32_t mesh_id;        /* from cd_register_mesh */
    uint16_t entity;         /* Entity (uint16_t, ECStypes.h:18) */
    uint8_t should_callback; /* shouldCallback */
    uint8_t pad;
} cd_entity;

/* One mesh-pair query: triangles of mesh_b are taken into mesh_a's space by
 * rel (OBBtreesCollision.cpp:15 computes rel = inverse(A) * B).             */
typedef struct cd_query {
    uint32_t mesh_a, mesh_b;
    float rel[16];
} cd_query;

/* One colliding triangle pair (intersecting and not coplanar,
 * CreateUncollideRays.cpp:88), in the first entity's model space.  tri_a /
 * tri_b are tree-order triangle slots (use cd_export_tree's permutation to
 * map them to input triangle indices).  flags bit k: A's vertex k is
 * outside B's plane; bit 3+k: B's vertex k is outside A's plane
 * (CreateUncollideRays.cpp:102-112).  na / nb: unit face normals.          */
typedef struct cd_contact {
    uint32_t query, tri_a, tri_b, flags;
    float s[3], t[3];
    float na[3], nb[3];
} cd_contact;

/* An overlapping entity pair from the broad phase, oriented as the
 * reference orients it (first = earlier on the sweep axis,
 * SweepAndPrune.cpp:63), with its range in the contact array.               */
typedef struct cd_pair {
    uint32_t first, second;
    uint32_t contact_begin, contact_count;
} cd_pair;

/* One uncollide ray (include/Geometry/Ray.h): origin and direction in the
 * first entity's model space, for the rays of either entity.              */
typedef struct cd_ray {
    float origin[3];
    float direction[3];
} cd_ray;

/* CDentriesUncollideRays (include/CollisionDetection/CreateUncollideRays.h:6-14)
 * of one query or frame pair: its rays sit at rays[begin_first ..
 * + n_first) (rays_from_first_to_second) and rays[begin_second .. +
 * n_second) (rays_from_second_to_first).                                   */
typedef struct cd_pair_rays {
    uint32_t query;          /* query index, or pair index of the last cd_frame */
    uint32_t first, second;  /* mesh ids (queries) or entity indices (cd_frame) */
    uint32_t n_first, n_second;
    uint32_t pad;
    uint64_t begin_first, begin_second;
    float avg_first[3];      /* average_point_first_modelspace */
    float avg_second[3];     /* average_point_second_modelspace (the second's model space) */
} cd_pair_rays;

/* Response parameters (ShootUncollideRays::ShootUncollideRays,
 * src/CollisionDetection/ShootUncollideRays.cpp:4-12; CollisionDetection.cpp:21-24
 * passes radians(40), radians(65), 1.01, which a NULL cd_shoot_params selects). */
typedef struct cd_shoot_params {
    float smoothstep_start_rads;   /* max_cos_from_force_response_smoothstep_start (an angle) */
    float smoothstep_finish_rads;  /* ..._finish */
    float distance_bias;           /* ray_distance_bias_multiplier */
} cd_shoot_params;

/* One entity pair's response (CollisionDetection.cpp:70-99). */
typedef struct cd_pair_response {
    uint32_t first, second;  /* entity indices of the last cd_frame */
    uint32_t shot;           /* 0: neither entity moved since the previous frame (:80-81), deltas zero */
    uint32_t pad;
    float delta[3];          /* ExecuteShootUncollideRays: the uncollide delta, world space */
    float delta_first[3];    /* CollisionCallbackData::deltaVector for the first entity: -delta * its share */
    float delta_second[3];   /* ... for the second: +delta * its share */
    float pad2[3];
} cd_pair_response;

typedef struct cd_mesh_info {
    uint32_t ntri, nnodes, depth, leaf_max;
    float root[12];          /* root OBB: centre, u, v, w half-extent sides */
    float slack;             /* max distance a vertex lies outside a box meant to hold it */
    float radius;            /* max |vertex| */
} cd_mesh_info;

typedef struct cd_counters {
    uint64_t sat_tests;       /* node-pair separating-axis tests */
    uint64_t leaf_pairs;      /* candidate leaf pairs */
    uint64_t tri_tests;       /* triangle-triangle tests */
    uint64_t contacts;        /* colliding pairs written */
    uint64_t fp64_fallbacks;  /* orientation predicates re-evaluated in double */
    uint64_t exact_fallbacks; /* ... and then in exact expansion arithmetic */
    uint64_t ub_cases;        /* Möller cases the reference leaves undefined */
    uint64_t broad_pairs;     /* entity pairs out of the broad phase */
} cd_counters;

typedef struct cd_timings {   /* device time of the last run, from HIP events */
    float total_ms;
    float traverse_ms;        /* plain: traversal kernel; batched: persistent kernel */
    float narrow_ms;          /* plain: narrow kernel */
    float broad_ms;
    uint32_t launches;        /* launches of the dominant kernel in the run */
    uint32_t pad;
} cd_timings;

/* Launch parameters of the batched path (no counterpart in the reference:
 * the reference has no GPU).  Chosen per workload by the autotuner
 * (inmyroom_vulkan_amd/autotune.py); zero fields mean "default".            */
typedef struct cd_tuning {
    uint32_t split_levels;   /* breadth-first levels before the persistent kernel (work-split granularity; 0: 16) */
    uint32_t stack_entries;  /* LDS stack entries per wavefront: 512, 1024 or 2048 (0: 512) */
    uint32_t pop_width;      /* node pairs a wavefront pops per step, 1..64 */
    uint32_t waves_per_cu;   /* persistent wavefronts per compute unit */
    uint32_t use_graph;      /* 1: capture the batched pipeline in a HIP graph and replay it */
    uint32_t work_sharing;   /* 1: wavefronts that run out of work take node pairs busy ones donate;
                                2: once half the grid is done, busy wavefronts hand their stacks to a
                                   second launch that shares them out;
                                3: budgeted pass: a wavefront hands its stack on after handoff_pct
                                   percent of the previous run's node-pair tests per wavefront; what
                                   is handed on is split again breadth first (resplit_levels) and
                                   shared out by a second launch */
    uint32_t spin_polls;     /* polls an idle wavefront waits for a donation before it leaves (0: 4096) */
    uint32_t split_tris;     /* a split level keeps splitting only node pairs over this many triangles
                                (0: 256; 1: all of them); the others go to the persistent kernel */
    uint32_t handoff_pct;    /* work_sharing 3: the budget, percent of the mean (0: 100) */
    uint32_t resplit_levels; /* work_sharing 3: breadth-first levels over what was handed on (0: 4) */
    uint32_t pipeline;       /* batches of node pairs a persistent wavefront keeps in flight, 1 or 2 (0: 2): with 2
                                it issues the next batch's loads before it tests the current one (work_sharing 0) */
    uint32_t split_depth;    /* split levels one launch runs, 1 or 2 (0: 2): with 2 a thread tests its pair and
                                then both of that pair's children */
    uint32_t step_graph;     /* cd_step and cd_frame (batched and reference modes): 0 or 1 replays the whole
                                step (the frame) as one captured HIP graph; 2 enqueues the same kernels one by
                                one, with events around the traversal and the narrow phase (cd_get_timings) */
} cd_tuning;

/* ---- context --------------------------------------------------------------- */
/* Replaces CollisionDetection::CollisionDetection (src/CollisionDetection/CollisionDetection.cpp:5-26). */
int cd_create(cd_ctx** out, const cd_config* cfg);
int cd_destroy(cd_ctx* ctx);
const char* cd_last_error(const cd_ctx* ctx);

/* ---- meshes -------------------------------------------------------------------- */
/* Builds the OBB tree of an indexed triangle list on the GPU.
 * Replaces OBBtree::OBBtree (src/Geometry/OBBtree.cpp:321-358) fed by
 * Triangle::CreateTriangleList in glTF "triangles" mode (src/Geometry/Triangle.cpp:259-280). */
int cd_register_mesh(cd_ctx* ctx, const float* positions, uint32_t nvert, const uint32_t* indices,
                     uint32_t ntri, uint32_t* mesh_id);
/* Many triangle soups at once (9 floats per triangle, tri_counts[i] triangles
 * for mesh i, concatenated); ids are first_id .. first_id + nmesh - 1.      */
int cd_register_soups(cd_ctx* ctx, const float* tri9, const uint32_t* tri_counts, uint32_t nmesh,
                      uint32_t* first_id);
int cd_mesh_info_get(cd_ctx* ctx, uint32_t mesh, cd_mesh_info* out);
/* Flattened tree as the reference stores it (include/Geometry/OBBtree.h:61-75,129-134):
 * node_obbs: nnodes x 24 floats (left box, right box); node_links: nnodes x 4
 * (left link, right link, left count, right count; count 0 = internal child);
 * perm: ntri tree-order slot -> input triangle index.                     */
int cd_export_tree(cd_ctx* ctx, uint32_t mesh, float* root12, float* node_obbs, uint32_t* node_links,
                   uint32_t* perm);

/* Tree-order triangle positions of a mesh (9 floats each; after a refit, the
 * skinned positions).  The reference keeps them in OBBtree::triangles_position
 * (include/Geometry/OBBtree.h:132).                                         */
int cd_export_triangles(cd_ctx* ctx, uint32_t mesh, float* tri9);

/* ---- glTF mesh ingest ---------------------------------------------------------------
 * A parsed .gltf / .glb document in host memory.  Parsing needs no GPU.
 * Replaces tinygltf's LoadASCIIFromFile / LoadBinaryFromFile as
 * GameImporter::LoadglTFmodel calls them (src/GameImporter.cpp:468-500) and
 * the POSITION / NORMAL / index reads of PrimitiveInitializationData
 * (src/Graphics/Meshes/PrimitivesOfMeshes.cpp:43-137, 528-552, 695-752):
 * positions and normals come out as (x, -y, -z); a line loop is read as a
 * line strip; skinned (JOINTS_0) and position-morphed primitives are marked
 * excluded and give no triangles (:637-671).  Interleaved (strided) views,
 * sparse accessors and index types other than unsigned short / unsigned int
 * are rejected with CD_E_FORMAT.                                            */
typedef struct cd_gltf cd_gltf;

typedef struct {
    uint32_t nprim;            /* primitives, excluded ones included */
    uint32_t with_normals;     /* included primitives with NORMAL */
    uint32_t without_normals;  /* included primitives without */
    uint32_t pad;
    uint64_t nvert;            /* vertices of the included primitives */
    uint64_t ntri;             /* collision triangles */
} cd_gltf_mesh_info;

typedef struct {
    int32_t gltf_mode;         /* as written; -1 when absent */
    uint32_t mode;             /* draw mode used (absent -> 4, line loop 2 -> line strip 3) */
    uint32_t excluded;         /* 1: skinned or morphed */
    uint32_t nvert;
    uint64_t nindex;
    uint64_t ntri;             /* triangles its mode makes (Triangle.cpp:10-62) */
    const float* positions;    /* nvert x 3; owned by the document */
    const float* normals;      /* nvert x 3, or NULL */
    const uint32_t* indices;   /* nindex (0 .. nvert-1 for a primitive without indices) */
} cd_gltf_primitive;

/* err (may be NULL) receives a message of at most err_cap - 1 bytes. */
int cd_gltf_open(const char* path, cd_gltf** doc, char* err, uint32_t err_cap);
int cd_gltf_close(cd_gltf* doc);
int cd_gltf_mesh_count(const cd_gltf* doc, uint32_t* n);
int cd_gltf_mesh_info_get(const cd_gltf* doc, uint32_t mesh, cd_gltf_mesh_info* out);
/* The k-th primitive of a mesh in the order its triangles are concatenated
 * (MeshesOfNodes::AddMeshesOfModel's sort, src/Graphics/Meshes/MeshesOfNodes.cpp:41-43). */
int cd_gltf_primitive_get(const cd_gltf* doc, uint32_t mesh, uint32_t k, cd_gltf_primitive* out);
/* Builds the collision tree of every mesh of the document in one GPU pass:
 * mesh_ids[m] (cap >= mesh count) receives glTF mesh m's id, or CD_NO_MESH
 * when it has no collision triangles.  Replaces MeshesOfNodes::AddMeshesOfModel
 * (MeshesOfNodes.cpp:29-64) -> PrimitivesOfMeshes::GetOBBtreeAndReset
 * (PrimitivesOfMeshes.cpp:840-863) -> Triangle::CreateTriangleList
 * (src/Geometry/Triangle.cpp:259-280) -> OBBtree::OBBtree.  Vertex ids (the
 * rays' de-duplication keys) are the index-list entries offset by the
 * primitive's first vertex in its mesh; vertex normals are the primitives'
 * NORMAL, or face normals for a mesh without them (a mesh that mixes the two
 * is rejected with CD_E_FORMAT).  cd_get_timings afterwards: broad_ms the
 * triangle assembly, traverse_ms the tree build, total_ms both.            */
int cd_register_gltf(cd_ctx* ctx, const cd_gltf* doc, uint32_t* mesh_ids, uint32_t cap);

/* ---- node transforms ----------------------------------------------------------------
 * A node's NodeData (include/ECS/GeneralCompEntities/NodeDataCompEntity.h:47-50). */
#define CD_NO_PARENT 0xFFFFFFFFu

typedef struct {
    float local_scale[3];
    float local_rotation[4];     /* unit quaternion (x, y, z, w) */
    float local_translation[3];
    float global_translation[3];
    uint32_t parent;             /* index in the same array, or CD_NO_PARENT */
    uint32_t pad[2];
} cd_node;                       /* 64 bytes */

/* Global matrices of a node hierarchy on the GPU, 16 floats (column-major)
 * per node: gT * parent_global * T * R * S, from each root down.  Replaces
 * NodeDataCompEntity::GetGlobalMatrix (src/ECS/GeneralCompEntities/
 * NodeDataCompEntity.cpp:63-72) as Early/LateNodeGlobalMatrixCompEntity::
 * Update apply it (EarlyNodeGlobalMatrixCompEntity.cpp:63-78).  Nodes may be
 * in any order; a chain deeper than 256 or a cycle is rejected.           */
int cd_node_global_matrices(cd_ctx* ctx, const cd_node* nodes, uint32_t n, float* global16);

/* The glTF nodes as the reference imports them (GameImporter::
 * CreatePositionInitMap, src/GameImporter.cpp:503-552, and NodeDataCompEntity::
 * CreateComponentEntityByMap, NodeDataCompEntity.cpp:19-61): TRS in
 * (x, -y, -z) space, a matrix node decomposed as glm::decompose does, the
 * rotation normalized; parents from the children lists.  mesh[i] is node i's
 * glTF mesh or -1; collides[i] is 1 when the node gets a collision component
 * (a mesh that is neither skinned nor morphed, GameImporter.cpp:320-326).
 * Any of nodes, mesh, collides may be NULL.                                 */
int cd_gltf_node_count(const cd_gltf* doc, uint32_t* n);
int cd_gltf_nodes(const cd_gltf* doc, cd_node* nodes, int32_t* mesh, uint32_t* collides, uint32_t cap);

/* ---- deformable meshes ---------------------------------------------------------- */
/* No counterpart in the reference: it gives trees only to non-skinned
 * primitives (src/Graphics/Meshes/PrimitivesOfMeshes.cpp:642) and never
 * refits.  A skinned mesh is registered in its bind pose (the tree is built
 * as cd_register_mesh builds it) with 4 joints and 4 weights per vertex.    */
int cd_register_skinned(cd_ctx* ctx, const float* bind_positions, uint32_t nvert, const uint32_t* indices,
                        uint32_t ntri, const uint16_t* joints4, const float* weights4, uint32_t nbones,
                        uint32_t* mesh_id);
/* Skin the listed meshes with their bone matrices (nbones x 16 floats per
 * mesh, concatenated in list order; each = joint matrix * inverse bind, the
 * product the skinning shader forms, shaders/dynamicMeshShader_glsl.comp:120-145)
 * and refit every box of their trees over the skinned triangles (same PCA
 * construction as the build; topology kept).                               */
int cd_skin_refit(cd_ctx* ctx, const uint32_t* mesh_ids, uint32_t n, const float* bone_matrices);

/* ---- queries ------------------------------------------------------------------- */
/* Triangle::IntersectTriangles (src/Geometry/Triangle.cpp:85-110) over n
 * pairs.  flags: bit0 hit, bit1 coplanar, bit2 undefined-in-reference case
 * (plain); bit0 hit, bit1 coplanar (batched = exact semantics).
 * seg6 (may be NULL): source, target when hit and not coplanar.           */
int cd_tritri_batch(cd_ctx* ctx, const float* a9, const float* b9, uint64_t n, uint32_t mode,
                    int32_t* flags, float* seg6);

/* No counterpart in the reference.  cd_tritri_batch for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): a9, b9, flags and seg6 (may be NULL) are
 * DEVICE pointers into memory of this runtime the caller owns, and the call
 * makes no allocation.  The exact (batched) mode's rerun lists and the run's
 * counters live in a fixed device array of the library, one per GPU, of
 * about 8 MiB (CD_NOALLOC_MAX_PAIRS pairs), which that GPU holds until the
 * process ends; n above CD_NOALLOC_MAX_PAIRS (either mode) is rejected with
 * CD_E_CAPACITY, and buffers shorter than n pairs with CD_E_INVALID.  Every
 * call on one GPU shares that array: calls that overlap on one GPU (two
 * contexts or two threads at once) race.  Synchronous, like the rest.     */
#define CD_NOALLOC_MAX_PAIRS (1u << 20)
int cd_tritri_batch_noalloc(cd_ctx* ctx, const float* a9, const float* b9, uint64_t n, uint32_t mode,
                            int32_t* flags, float* seg6);

/* Mesh-pair queries: OBBtreesCollision::ExecuteOBBtreesCollision
 * (src/CollisionDetection/OBBtreesCollision.cpp:7-18) followed by the
 * triangle loop of CreateUncollideRays::ExecuteCreateUncollideRays
 * (src/CollisionDetection/CreateUncollideRays.cpp:60-128).  Contacts of
 * query q carry query = q.  CD_E_CAPACITY writes the needed count to n_out. */
int cd_query_batch(cd_ctx* ctx, const cd_query* q, uint32_t nq, uint32_t mode, cd_contact* out,
                   uint64_t cap, uint64_t* n_out);

/* Candidate leaf pairs of the plain traversal (the mid-phase seam,
 * OBBtreesIntersectInfo, include/Geometry/OBBtree.h:8-19), for a host that
 * keeps its own narrow phase: 4 words each {query, offA, offB, cntA | cntB<<16}. */
int cd_query_leafpairs(cd_ctx* ctx, const cd_query* q, uint32_t nq, uint32_t* out4, uint64_t cap,
                       uint64_t* n_out);

/* Phases 1-3 of CollisionDetection::ExecuteCollisionDetection
 * (src/CollisionDetection/CollisionDetection.cpp:38-67): sweep-and-prune on
 * the fixed U/V/W frame, relative matrix per pair, traversal, triangle tests.
 * Pairs without contacts are dropped, as :52 and :63 drop them.  In the
 * batched and reference modes the frame is one launch of a captured HIP
 * graph (cd_step's first part, with its capacities) and one wait, then the
 * contacts are copied back; cd_get_timings reports its device time
 * (total_ms), and with cd_tuning.step_graph 2 the traversal (traverse_ms)
 * and the narrow phase (narrow_ms) too.                                    */
int cd_frame(cd_ctx* ctx, const cd_entity* entities, uint32_t n, uint32_t mode, cd_pair* pairs,
             uint32_t pair_cap, uint32_t* n_pairs, cd_contact* contacts, uint64_t contact_cap,
             uint64_t* n_contacts);

/* Rays of the last run's contacts (cd_frame, cd_query_batch or
 * cd_run_staged): CreateUncollideRays::ExecuteCreateUncollideRays
 * (src/CollisionDetection/CreateUncollideRays.cpp:60-200).  Per colliding
 * triangle: weight |s - t|, weighted midpoint and the vertices not outside
 * the other triangle's plane, merged over leaf pairs; then one ray per
 * such vertex (de-duplicated by vertex id per pair and side, direction
 * minus the vertex normal) or, when none is inside, one at the weighted
 * contact point with the barycentric normal.  Only pairs with rays are
 * listed (CollisionDetection.cpp:63).  Where several triangles claim one
 * vertex the reference keeps the first its unordered_map walk meets; here
 * the lowest tree-order triangle keeps it (this only matters for a mesh
 * with face normals).  CD_E_CAPACITY writes the needed counts back.       */
int cd_uncollide_rays(cd_ctx* ctx, cd_pair_rays* pairs, uint32_t pair_cap, uint32_t* n_pairs, cd_ray* rays,
                      uint64_t ray_cap, uint64_t* n_rays);

/* Response of the last cd_frame, from the rays cd_uncollide_rays made:
 * ShootUncollideRays::ExecuteShootUncollideRays
 * (src/CollisionDetection/ShootUncollideRays.cpp:14-179, Ray::IntersectOBBtree
 * src/Geometry/Ray.cpp:135-236) and the movement split of
 * CollisionDetection::ExecuteCollisionDetection (CollisionDetection.cpp:70-99,
 * PointMovementBetweenFrames :143-150).  Each ray is shot at the other
 * entity's tree and, when it leaves it through a back face before re-entering
 * its own, reflected and shot back; the pair's delta comes from the force
 * responses and push-out vectors of all its rays, and is split between the
 * entities by how far each one's average contact point moved since the
 * previous frame (cd_entity.previous).  One record per pair that has rays, in
 * cd_uncollide_rays' order.  glm::intersectRayTriangle's backface output
 * (a fork of glm that is not available) is taken as the negative-determinant
 * branch of glm 0.9.9's test.  The callbacks' fan-out to entity ancestors
 * (:101-121) needs the engine's entity tree and is left to the caller.
 * Params NULL: the reference's values.  CD_E_CAPACITY writes the count back;
 * it is also returned, with no count, for a pair whose mesh has 2^25
 * triangles or more (the traversal stack packs an index and a leaf count
 * into one word).  Trees deeper than 46 levels keep the traversal stacks in
 * device memory instead of LDS.  Afterwards cd_get_counters reports the
 * boxes tested (sat_tests) and ray-triangle tests (tri_tests) of the call,
 * and cd_get_timings the shooting kernel (traverse_ms) and, with the
 * response, the whole call's device time (total_ms). */
int cd_shoot_rays(cd_ctx* ctx, const cd_shoot_params* params, cd_pair_response* out, uint32_t cap,
                  uint32_t* n_out);

/* One whole step of CollisionDetection::ExecuteCollisionDetection
 * (src/CollisionDetection/CollisionDetection.cpp:38-99): cd_frame,
 * cd_uncollide_rays and cd_shoot_rays in one call, with the contacts and the
 * rays kept on the device; only the per-pair responses (the records
 * cd_shoot_rays gives) come back.  For hosts that need the deltas and not
 * the contacts or rays themselves (the engine's callbacks).  CD_E_CAPACITY
 * writes the count back with the frame and its rays done: cd_shoot_rays
 * with a larger buffer finishes the step.  Afterwards the context is as
 * after the three calls (cd_uncollide_rays and cd_shoot_rays may be called
 * again on the same frame).  In the batched and reference modes every count
 * stays on the device and the whole step is one launch of a captured HIP
 * graph (cd_tuning.step_graph) with one wait: its buffers are sized for
 * capacities of pairs, contacts and rays that grow when a step exceeds them
 * (that step runs again), so a context's first step may run more than once.
 * cd_get_counters then reports the run's counters (as cd_frame does) and
 * cd_get_timings the step's device time (total_ms); the plain mode runs the
 * three calls. */
int cd_step(cd_ctx* ctx, const cd_entity* entities, uint32_t n, uint32_t mode, const cd_shoot_params* params,
            cd_pair_response* out, uint32_t cap, uint32_t* n_out);

/* Vertex normals of a mesh for its rays (TriangleNormal::CreateTriangleNormalList,
 * src/Geometry/Triangle.cpp:214-234): nvert x 3 floats indexed like the
 * vertices given to cd_register_mesh (for a soup: 3 t + k).  A mesh without
 * them uses face normals, as the reference does for a mesh without normals. */
int cd_set_mesh_normals(cd_ctx* ctx, uint32_t mesh, const float* normals, uint32_t nvert);

/* Broad phase alone (src/CollisionDetection/SweepAndPrune.cpp:15-88):
 * oriented overlapping pairs, sorted by (first, second).                  */
int cd_broad_phase(cd_ctx* ctx, const cd_entity* entities, uint32_t n, uint32_t* pairs2, uint32_t cap,
                   uint32_t* n_pairs);

/* ---- device-resident runs (benchmarks, frame loops) ---------------------------- */
/* Copy queries into device memory owned by the context ... */
int cd_stage_queries(cd_ctx* ctx, const cd_query* q, uint32_t nq);
/* ... run them with inputs already in HBM; results stay on the device ... */
int cd_run_staged(cd_ctx* ctx, uint32_t mode, uint64_t* n_contacts);
/* ... and fetch them. */
int cd_fetch_contacts(cd_ctx* ctx, cd_contact* out, uint64_t cap, uint64_t* n_out);

int cd_get_counters(cd_ctx* ctx, cd_counters* out);
/* Tuning of the batched path (no counterpart in the reference). */
int cd_set_tuning(cd_ctx* ctx, const cd_tuning* t);
int cd_get_tuning(cd_ctx* ctx, cd_tuning* t);
int cd_get_timings(cd_ctx* ctx, cd_timings* out);

#ifdef __cplusplus
}
#endif

#endif /* MI355X_COLLIDE_H */
