// C ABI of the collision library (declared in include/mi355x_collide.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "gltf.hpp"
#include "kernels.hpp"

using namespace cdk;

namespace {

int fail(cd_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}
int hip_fail(cd_ctx* ctx, hipError_t e, const char* what) {
    return fail(ctx, CD_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define HCK(ctx, x)                                                 \
    do {                                                            \
        hipError_t e_ = (x);                                        \
        if (e_ != hipSuccess) return hip_fail((ctx), e_, #x);       \
    } while (0)

bool all_finite(const float* p, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

int upload_mesh_table(cd_ctx* ctx) {
    const size_t bytes = ctx->meshes.size() * sizeof(MeshRec);
    HCK(ctx, ctx->d_meshes.ensure(bytes + sizeof(MeshRec)));
    if (bytes) HCK(ctx, hipMemcpyAsync(ctx->d_meshes.p, ctx->meshes.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
    HCK(ctx, hipStreamSynchronize(ctx->stream));
    return CD_OK;
}

hipError_t grow_keep(DevBuf& b, size_t need, hipStream_t st);

int upload_attrs(cd_ctx* ctx) {
    const size_t bytes = ctx->attrs.size() * sizeof(uint2);
    HCK(ctx, ctx->d_attrs.ensure(bytes + sizeof(uint2)));
    if (bytes) HCK(ctx, hipMemcpyAsync(ctx->d_attrs.p, ctx->attrs.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
    HCK(ctx, hipStreamSynchronize(ctx->stream));
    return CD_OK;
}

// Index-buffer vertex ids of a registered mesh, in tree order (for the
// vertex de-duplication of the rays, CreateUncollideRays.cpp:146-152).
int store_vertex_ids(cd_ctx* ctx, uint32_t id, const uint32_t* indices, uint32_t ntri, uint32_t nvert) {
    hipStream_t st = ctx->stream;
    const MeshRec& m = ctx->meshes[id];
    const size_t n = size_t(ntri) * 3;
    if (ctx->vid_used + n >= 0xFFFFFFF0ull) return fail(ctx, CD_E_CAPACITY, "vertex-id arena exceeds 32-bit indexing");
    HCK(ctx, grow_keep(ctx->d_vid, (ctx->vid_used + n + 1) * sizeof(uint32_t), st));
    ScopedBuf tmp;
    HCK(ctx, tmp.ensure((n + 1) * sizeof(uint32_t)));
    if (n) HCK(ctx, hipMemcpyAsync(tmp.p, indices, n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    HCK(ctx, launch_tree_vidx(ctx->d_perm.as<uint32_t>() + m.tri_base, ntri, tmp.as<uint32_t>(), 0u,
                              ctx->d_vid.as<uint32_t>() + ctx->vid_used, st));
    HCK(ctx, hipStreamSynchronize(st));
    ctx->attrs[id].x = uint32_t(ctx->vid_used);
    ctx->mesh_nvert[id] = nvert;
    ctx->vid_used += n;
    return upload_attrs(ctx);
}

int register_soups(cd_ctx* ctx, const float* tri9, const uint32_t* counts, uint32_t nmesh, uint32_t* first_id) {
    uint64_t T = 0;
    for (uint32_t m = 0; m < nmesh; ++m) T += counts[m];
    if (T && !tri9) return fail(ctx, CD_E_INVALID, "null triangle array");
    if (!all_finite(tri9, size_t(T) * 9)) return fail(ctx, CD_E_INVALID, "non-finite vertex coordinate");
    for (uint32_t m = 0; m < nmesh; ++m)
        if (counts[m] > 0x7FFFFFFFu) return fail(ctx, CD_E_INVALID, "mesh too large");
    BuildResult r = build_soups(ctx, tri9, counts, nmesh);
    if (!r.err.empty()) return fail(ctx, CD_E_HIP, "tree build: " + r.err);
    if (first_id) *first_id = uint32_t(ctx->meshes.size());
    ctx->meshes.insert(ctx->meshes.end(), r.meshes.begin(), r.meshes.end());
    for (uint32_t m = 0; m < nmesh; ++m) {  // a soup: vertex ids 3 t + k, face normals
        ctx->attrs.push_back(make_uint2(~0u, ~0u));
        ctx->mesh_nvert.push_back(3u * counts[m]);
    }
    int rc = upload_mesh_table(ctx);
    return rc ? rc : upload_attrs(ctx);
}

// Also sets ctx->plain_need: the plain walk of a mesh pair holds at most its
// trees' depths + 2 node pairs on its stack.
int check_queries(cd_ctx* ctx, const QueryRec* q, uint32_t nq) {
    const uint32_t nm = uint32_t(ctx->meshes.size());
    uint32_t need = 0;
    for (uint32_t i = 0; i < nq; ++i) {
        if (q[i].mesh_a >= nm || q[i].mesh_b >= nm) return fail(ctx, CD_E_INVALID, "query names an unknown mesh");
        if (!all_finite(q[i].rel, 16)) return fail(ctx, CD_E_INVALID, "non-finite relative matrix");
        need = std::max(need, ctx->meshes[q[i].mesh_a].depth + ctx->meshes[q[i].mesh_b].depth + 2);
    }
    ctx->plain_need = need;
    return CD_OK;
}

// Work buffers (d_work slots): batched path, then broad phase / frame.
enum { W_QPRE = 0, W_FRONT0, W_FRONT1, W_SMALL, W_SPILL,
       W_ENTS, W_IV, W_K0, W_V0, W_K1, W_V1, W_PAIRS, W_SPAIRS, W_PCOUNT, W_CSORT, W_DUMP0, W_DEFER,
       W_RSTART, W_RPOS, W_RSUMS, W_RCAND, W_RDENSE, W_RCNT, W_RPOS2, W_RHKEY, W_RHVAL, W_RQ, W_RAVG, W_RTOT,
       W_RAYS, W_TASKS, W_SORTT, W_BUD0, W_BUD1, W_BUD2, W_HITS, W_SHOTS, W_RESP, W_PIDS, W_GSTACK,
       W_SCTR, W_STEPC, W_STEPOUT, W_SEGQ, W_CSEG };
static_assert(W_CSEG < sizeof(cd_ctx::d_work) / sizeof(DevBuf), "work slots");

RunCounters* dev_counters(cd_ctx* ctx) { return ctx->d_counters.as<RunCounters>(); }

// Through a pinned buffer: a copy into pageable memory is staged by the
// runtime and takes longer to complete.
int read_counters(cd_ctx* ctx, RunCounters& h) {
    HCK(ctx, ctx->h_ctr.ensure(sizeof h));
    HCK(ctx, hipMemcpyAsync(ctx->h_ctr.p, ctx->d_counters.p, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    HCK(ctx, hipStreamSynchronize(ctx->stream));
    std::memcpy(&h, ctx->h_ctr.p, sizeof h);
    return CD_OK;
}

// Plain run over the nq queries already in d_queries; contacts stay in d_contacts.
int run_plain(cd_ctx* ctx, uint32_t nq, uint64_t* n_contacts) {
    hipStream_t st = ctx->stream;
    const QueryRec* dq = ctx->d_queries.as<QueryRec>();
    const MeshRec* dm = ctx->d_meshes.as<MeshRec>();
    const NodeRec* dn = ctx->d_nodes.as<NodeRec>();
    if (ctx->lp_cap == 0) ctx->lp_cap = std::max<uint64_t>(1u << 20, 64ull * nq);
    if (ctx->contact_cap == 0) ctx->contact_cap = 1u << 20;
    RunCounters h{};
    for (int attempt = 0; attempt < 3; ++attempt) {
        HCK(ctx, ctx->d_leafpairs.ensure(ctx->lp_cap * sizeof(uint4)));
        HCK(ctx, ctx->d_contacts.ensure(ctx->contact_cap * sizeof(cd_contact)));
        HCK(ctx, launch_zero(dev_counters(ctx), st));
        HCK(ctx, hipEventRecord(ctx->ev[0], st));
        uint2* gstack = nullptr;  // trees deeper than the LDS stack holds: the stacks in HBM
        if (ctx->plain_need > uint32_t(PLAIN_STACK)) {
            HCK(ctx, ctx->d_work[W_GSTACK].ensure(size_t((nq + 63) / 64) * ctx->plain_need * 64 * sizeof(uint2)));
            gstack = ctx->d_work[W_GSTACK].as<uint2>();
        }
        HCK(ctx, launch_traverse_plain(dq, nq, dm, dn, ctx->d_leafpairs.as<uint4>(), ctx->lp_cap, dev_counters(ctx), st,
                                       ctx->plain_need, gstack));
        HCK(ctx, hipEventRecord(ctx->ev[1], st));
        int rc = read_counters(ctx, h);
        if (rc) return rc;
        if (h.err) return fail(ctx, CD_E_STACK, "plain traversal stack overflow");
        if (h.leafpairs > ctx->lp_cap) {
            ctx->lp_cap = h.leafpairs + h.leafpairs / 4;
            continue;
        }
        HCK(ctx, hipEventRecord(ctx->ev[2], st));
        HCK(ctx, launch_narrow_plain(ctx->d_leafpairs.as<uint4>(), h.leafpairs, ctx->leaf_max, dq, dm,
                                     ctx->d_tris.as<float>(), ctx->d_contacts.as<cd_contact>(), ctx->contact_cap,
                                     dev_counters(ctx), st));
        HCK(ctx, hipEventRecord(ctx->ev[3], st));
        rc = read_counters(ctx, h);
        if (rc) return rc;
        if (h.contacts > ctx->contact_cap) {
            ctx->contact_cap = h.contacts + h.contacts / 4;
            continue;
        }
        float t_trav = 0.f, t_narrow = 0.f;
        HCK(ctx, hipEventElapsedTime(&t_trav, ctx->ev[0], ctx->ev[1]));
        HCK(ctx, hipEventElapsedTime(&t_narrow, ctx->ev[2], ctx->ev[3]));
        ctx->timings = cd_timings{};
        ctx->timings.traverse_ms = t_trav;
        ctx->timings.narrow_ms = t_narrow;
        ctx->timings.total_ms = t_trav + t_narrow;
        ctx->timings.launches = 1;
        ctx->counters = cd_counters{};
        ctx->counters.sat_tests = h.sat;
        ctx->counters.leaf_pairs = h.leafpairs;
        ctx->counters.tri_tests = h.tritests;
        ctx->counters.contacts = h.contacts;
        ctx->counters.ub_cases = h.ub;
        *n_contacts = h.contacts;
        return CD_OK;
    }
    return fail(ctx, CD_E_NOMEM, "output sizing did not converge");
}

// nq_cap: the query capacity of the run (0: unknown, the large-batch defaults).
cd_tuning effective_tuning(const cd_ctx* ctx, uint32_t nq_cap = 0) {
    // defaults: the values the tuning sweeps pick with the pipelined traversal:
    // for the 4,096-query batch 16 levels, pairs over 256 triangles split
    // further, 512-entry stacks, 6 wavefronts per CU; for runs of 512 to
    // 2,048 queries (a 512-entity scene's ~1,100 pairs, 1,000 poses) 12
    // levels and 4 wavefronts.  Fewer queries keep 16 levels: their trees
    // must be split further to give the persistent kernel enough tasks (64
    // skinned meshes against a terrain: 0.32 ms against 0.40 with 12).
    cd_tuning t = ctx->tuning;
    const bool small = nq_cap >= 512 && nq_cap <= 2048;
    if (t.split_levels == 0) t.split_levels = small ? 12 : 16;
    if (t.split_levels > 64) t.split_levels = 64;
    if (t.split_tris == 0) t.split_tris = 256;
    if (t.stack_entries != 1024 && t.stack_entries != 2048) t.stack_entries = 512;
    if (t.pop_width == 0 || t.pop_width > 64) t.pop_width = 64;
    if (t.waves_per_cu == 0) t.waves_per_cu = small ? 4 : 6;
    if (t.spin_polls == 0) t.spin_polls = 4096;
    if (t.handoff_pct == 0) t.handoff_pct = 100;
    if (t.resplit_levels == 0) t.resplit_levels = 4;
    if (t.pipeline == 0 || t.pipeline > 2) t.pipeline = 2;
    if (t.split_depth == 0 || t.split_depth > 2) t.split_depth = 2;
    return t;
}

constexpr uint32_t SPILL_CAP = 16384;  // entries of each persistent wavefront's HBM spill area
constexpr uint32_t DQ_CAP = 1u << 20;  // node pairs the queue of the work-sharing traversal holds
// W_SMALL layout: [0..L] split-level counts; [TASK_COUNT] tasks the split
// levels hand to the persistent kernel; [HEAD0], [HEAD1] task heads of the
// traversal launch and of the launch that takes unclaimed donations;
// [DQ_CTL .. +3] donation queue: reserved, claimed, wavefronts idle or gone, pollers waiting;
// budgeted pass: [DUMP_COUNT] node pairs handed on, [RS_BASE + r] re-split
// level r's output, [TASK2_COUNT], [HEAD2] the second launch's tasks and head.
constexpr uint32_t TASK_COUNT = 70, HEAD0 = 72, HEAD1 = 73, DQ_CTL = 76;
constexpr uint32_t DUMP_COUNT = 80, RS_BASE = 81, TASK2_COUNT = 100, HEAD2 = 101;
constexpr uint32_t SMALL_WORDS = 104;
constexpr uint32_t MAX_RESPLIT = 16;


// Count table of the multi-workgroup radix sort for n keys.
hipError_t sort_table(cd_ctx* ctx, uint32_t n, uint32_t** table) {
    hipError_t e = ctx->d_work[W_SORTT].ensure(radix_table_words(n) * sizeof(uint32_t));
    *table = ctx->d_work[W_SORTT].as<uint32_t>();
    return e;
}

// Key bits of the contact fields query, tri_a, tri_b for nq queries.
void contact_field_bits(const cd_ctx* ctx, uint32_t nq, uint32_t bits[3]) {
    uint64_t max_tris = 1;
    for (const MeshRec& m : ctx->meshes) max_tris = std::max<uint64_t>(max_tris, m.ntri);
    bits[0] = key_bits_below(nq);
    bits[1] = bits[2] = key_bits_below(max_tris);
}

// Enqueue the whole batched pipeline on the stream (no host synchronisation
// inside, so it can be captured in a graph).
uint64_t frontier_cap(const cd_ctx* ctx, uint32_t nq, uint32_t L) {
    // worst case nq * 2^L, capped; a level that overflows sets an error bit and the run is repeated larger
    const uint64_t worst = (L >= 40) ? (uint64_t(1) << 40) : (uint64_t(std::max(nq, 1u)) << L);
    return std::max<uint64_t>(64, std::min<uint64_t>(worst, std::max<uint64_t>(ctx->frontier_cap, 1u << 20)));
}

// nq: the capacity the launch geometry is built for; the count itself is in d_nq
int enqueue_batched(cd_ctx* ctx, uint32_t nq, const cd_tuning& t, bool timed, bool ref) {
    hipStream_t st = ctx->stream;
    const uint32_t L = t.split_levels;
    const uint64_t fcap = frontier_cap(ctx, nq, L);
    const uint32_t nwaves = uint32_t(ctx->num_cu) * t.waves_per_cu;
    QueryPre* qp = ctx->d_work[W_QPRE].as<QueryPre>();
    uint3* F[2] = {ctx->d_work[W_FRONT0].as<uint3>(), ctx->d_work[W_FRONT1].as<uint3>()};
    unsigned int* small = ctx->d_work[W_SMALL].as<unsigned int>();
    uint4* dq = ctx->d_work[W_DUMP0].as<uint4>();
    const MeshRec* dm = ctx->d_meshes.as<MeshRec>();
    const NodeRec* dn = ctx->d_nodes.as<NodeRec>();
    RunCounters* ctr = dev_counters(ctx);
    const uint32_t grid = uint32_t(ctx->num_cu) * 8;
    if (t.work_sharing == 1)  // ready words of the donation queue start clear
        HCK(ctx, launch_fill_u32(reinterpret_cast<unsigned int*>(dq), 4 * DQ_CAP, 0u, st));
    if (timed) HCK(ctx, hipEventRecord(ctx->ev[0], st));
    HCK(ctx, launch_start(ctx->d_queries.as<QueryRec>(), nq, ctx->d_nq.as<unsigned int>(), dm, qp, F[0], small,
                          SMALL_WORDS, ctr, st));
    uint3* T = ctx->d_work[W_TASKS].as<uint3>();
    const uint32_t cap32 = uint32_t(std::min<uint64_t>(fcap, 0xFFFFFFF0ull));
    // level l holds at most nq * 2^l pairs: no more workgroups than that needs
    auto level_bound = [&](uint32_t l) {
        return l >= 32 ? uint64_t(cap32) : std::min<uint64_t>(cap32, uint64_t(std::max(nq, 1u)) << l);
    };
    // launch k takes frontier k from F[k & 1] (count small[k]); the last level sends every split to the task list
    uint32_t k = 0;
    for (uint32_t l = 0; l < L; ++k) {
        const uint32_t sgrid = uint32_t(ctx->num_cu) * split_blocks_per_cu(ref, t.split_depth == 2 && l + 1 < L);
        const uint32_t lgrid = uint32_t(std::min<uint64_t>(sgrid, (level_bound(l) + 255) / 256));
        const auto rule = [&](uint32_t level) { return level + 1 == L ? 0xFFFFFFFFu : t.split_tris; };
        if (t.split_depth == 2 && l + 1 < L) {
            HCK(ctx, launch_split_level2(F[k & 1], &small[k], F[(k + 1) & 1], &small[k + 1], cap32, qp, dm, dn,
                                         ctx->d_leafpairs.as<uint4>(), ctx->lp_cap, T, &small[TASK_COUNT], rule(l),
                                         rule(l + 1), ctr, lgrid, st, ref));
            l += 2;
        } else {
            HCK(ctx, launch_split_level(F[k & 1], &small[k], F[(k + 1) & 1], &small[k + 1], cap32, qp, dm, dn,
                                        ctx->d_leafpairs.as<uint4>(), ctx->lp_cap, T, &small[TASK_COUNT], rule(l), ctr,
                                        lgrid, st, ref));
            l += 1;
        }
    }
    if (timed) HCK(ctx, hipEventRecord(ctx->ev[1], st));
    const uint32_t share = t.work_sharing;
    unsigned long long* budget = ctx->d_budget.as<unsigned long long>();
    // budgeted pass: the node pairs handed on, the re-split levels' frontier and their tasks
    uint3* B3[3] = {ctx->d_work[W_BUD0].as<uint3>(), ctx->d_work[W_BUD1].as<uint3>(), ctx->d_work[W_BUD2].as<uint3>()};
    const uint32_t dcap = uint32_t(std::min<uint64_t>(ctx->dump_cap, 0xFFFFFFF0ull));
    HCK(ctx, launch_dfs(int(t.stack_entries), L ? T : F[0], nullptr, nullptr, &small[L ? TASK_COUNT : 0], cap32,
                        &small[HEAD0], qp, dm, dn, ctx->d_leafpairs.as<uint4>(), ctx->lp_cap,
                        ctx->d_work[W_SPILL].as<uint3>(), SPILL_CAP, t.pop_width, nwaves, dq, &small[DQ_CTL], DQ_CAP,
                        share, t.spin_polls, budget, t.handoff_pct, B3[0], &small[DUMP_COUNT], dcap, ctr, st,
                        t.pipeline == 2, ref));
    if (share == 1 || share == 2)  // donations nobody claimed, [claimed, reserved), or the stacks handed off
        HCK(ctx, launch_dfs(int(t.stack_entries), nullptr, dq, &small[DQ_CTL + 1], &small[DQ_CTL], DQ_CAP,
                            &small[HEAD1], qp, dm, dn, ctx->d_leafpairs.as<uint4>(), ctx->lp_cap,
                            ctx->d_work[W_SPILL].as<uint3>(), SPILL_CAP, t.pop_width, nwaves, dq, &small[DQ_CTL],
                            DQ_CAP, 0u, t.spin_polls, nullptr, 0u, nullptr, nullptr, 0u, ctr, st, false, ref));
    if (share == 3) {
        // what the budgeted launch handed on is split again breadth first and
        // shared out by a second launch over the new task list
        const uint32_t R = std::min(t.resplit_levels, MAX_RESPLIT);
        for (uint32_t r = 0; r < R; ++r)
            HCK(ctx, launch_split_level(B3[r & 1], &small[r ? RS_BASE + r : DUMP_COUNT], B3[(r + 1) & 1],
                                        &small[RS_BASE + r + 1], dcap, qp, dm, dn, ctx->d_leafpairs.as<uint4>(),
                                        ctx->lp_cap, B3[2], &small[TASK2_COUNT],
                                        r + 1 == R ? 0xFFFFFFFFu : t.split_tris, ctr, grid, st, ref));
        HCK(ctx, launch_dfs(int(t.stack_entries), R ? B3[2] : B3[0], nullptr, nullptr,
                            &small[R ? TASK2_COUNT : DUMP_COUNT], dcap, &small[HEAD2], qp, dm, dn,
                            ctx->d_leafpairs.as<uint4>(), ctx->lp_cap, ctx->d_work[W_SPILL].as<uint3>(), SPILL_CAP,
                            t.pop_width, nwaves, dq, &small[DQ_CTL], DQ_CAP, 0u, t.spin_polls, nullptr, 0u, nullptr,
                            nullptr, 0u, ctr, st, false, ref));
        HCK(ctx, launch_keep_budget(ctr, budget, st));
    }
    if (timed) HCK(ctx, hipEventRecord(ctx->ev[2], st));
    if (ref)
        HCK(ctx, launch_narrow_ref(ctx->d_leafpairs.as<uint4>(), &ctr->lpslots, ctx->lp_cap, qp, ctx->d_tris.as<float>(),
                                   ctx->d_contacts.as<cd_contact>(), ctx->contact_cap, ctr, uint32_t(ctx->num_cu), st));
    else
        HCK(ctx, launch_narrow_exact(ctx->d_leafpairs.as<uint4>(), &ctr->lpslots, ctx->lp_cap, qp,
                                     ctx->d_tris.as<float>(), ctx->d_work[W_HITS].as<uint4>(),
                                     ctx->d_contacts.as<cd_contact>(), ctx->contact_cap, ctx->d_work[W_DEFER].as<uint4>(),
                                     ctx->defer_cap, ctr, uint32_t(ctx->num_cu), st));
    if (timed) HCK(ctx, hipEventRecord(ctx->ev[3], st));
    return CD_OK;
}

// Buffers of a batched run for nq_cap queries (nq: the count, for the first
// sizing of the leaf-pair list); *fcap the frontier capacity they hold.
int ensure_batched(cd_ctx* ctx, uint32_t nq, uint32_t nq_cap, const cd_tuning& t, uint64_t* fcap) {
    const uint32_t nwaves = uint32_t(ctx->num_cu) * t.waves_per_cu;
    warm_launch_shapes();
    if (ctx->lp_cap == 0) ctx->lp_cap = std::max<uint64_t>(1u << 20, 64ull * nq);
    if (ctx->contact_cap == 0) ctx->contact_cap = 1u << 20;
    if (ctx->defer_cap == 0) ctx->defer_cap = 1u << 16;
    HCK(ctx, ctx->d_nq.ensure(sizeof(unsigned int)));
    HCK(ctx, ctx->h_nq.ensure(sizeof(unsigned int)));
    HCK(ctx, ctx->d_work[W_QPRE].ensure(size_t(nq_cap) * sizeof(QueryPre)));
    HCK(ctx, ctx->d_work[W_SMALL].ensure(SMALL_WORDS * sizeof(unsigned int)));
    HCK(ctx, ctx->d_work[W_SPILL].ensure(size_t(nwaves) * SPILL_CAP * sizeof(uint3)));
    if (t.work_sharing == 1 || t.work_sharing == 2)
        HCK(ctx, ctx->d_work[W_DUMP0].ensure(size_t(DQ_CAP) * sizeof(uint4)));
    if (!ctx->d_budget.p) {  // node-pair tests of the last run, kept across runs (none yet)
        HCK(ctx, ctx->d_budget.ensure(sizeof(unsigned long long)));
        HCK(ctx, hipMemsetAsync(ctx->d_budget.p, 0, sizeof(unsigned long long), ctx->stream));
    }
    *fcap = frontier_cap(ctx, nq_cap, t.split_levels);
    HCK(ctx, ctx->d_work[W_FRONT0].ensure(*fcap * sizeof(uint3)));
    HCK(ctx, ctx->d_work[W_FRONT1].ensure(*fcap * sizeof(uint3)));
    HCK(ctx, ctx->d_work[W_TASKS].ensure(*fcap * sizeof(uint3)));
    if (t.work_sharing == 3) {
        if (ctx->dump_cap == 0) ctx->dump_cap = 1u << 20;
        for (int w : {W_BUD0, W_BUD1, W_BUD2}) HCK(ctx, ctx->d_work[w].ensure(ctx->dump_cap * sizeof(uint3)));
    }
    HCK(ctx, ctx->d_leafpairs.ensure(ctx->lp_cap * sizeof(uint4)));
    HCK(ctx, ctx->d_contacts.ensure(ctx->contact_cap * sizeof(cd_contact)));
    HCK(ctx, ctx->d_work[W_HITS].ensure(ctx->contact_cap * sizeof(uint4)));
    HCK(ctx, ctx->d_work[W_DEFER].ensure(2 * ctx->defer_cap * sizeof(uint4)));
    return CD_OK;
}

// What a captured batched run depends on: sizes, tuning and the buffers.
void batched_key(const cd_ctx* ctx, uint32_t nq_cap, const cd_tuning& t, bool ref, uint64_t fcap,
                 std::vector<uint64_t>& key) {
    const uint32_t nwaves = uint32_t(ctx->num_cu) * t.waves_per_cu;
    const uint64_t k[] = {nq_cap, t.split_levels, t.stack_entries, t.pop_width, nwaves, t.work_sharing, t.spin_polls,
                          t.split_tris, t.handoff_pct, t.resplit_levels, t.pipeline, t.split_depth, uint64_t(ref),
                          ctx->lp_cap, ctx->contact_cap, ctx->defer_cap, ctx->dump_cap, uint64_t(ctx->leaf_max), fcap};
    key.insert(key.end(), std::begin(k), std::end(k));
    for (const DevBuf* b : {&ctx->d_queries, &ctx->d_meshes, &ctx->d_nodes, &ctx->d_tris, &ctx->d_leafpairs,
                            &ctx->d_contacts, &ctx->d_counters, &ctx->d_work[W_QPRE], &ctx->d_work[W_FRONT0],
                            &ctx->d_work[W_FRONT1], &ctx->d_work[W_SMALL], &ctx->d_work[W_SPILL],
                            &ctx->d_work[W_DUMP0], &ctx->d_work[W_TASKS], &ctx->d_work[W_DEFER], &ctx->d_budget,
                            &ctx->d_nq, &ctx->d_work[W_BUD0], &ctx->d_work[W_BUD1], &ctx->d_work[W_BUD2],
                            &ctx->d_work[W_HITS]})
        key.push_back(reinterpret_cast<uint64_t>(b->p));
}

// A batched run's counters: grows what ran out of room (*retry: run it
// again) or fails on an error the run cannot recover from.
int check_batched(cd_ctx* ctx, const RunCounters& h, uint64_t fcap, const cd_tuning& t, bool* retry) {
    *retry = false;
    if (h.err & 2u) {  // a split level or the hand-off ran out of room: repeat with twice as much
        ctx->frontier_cap = 2 * fcap;
        if (t.work_sharing == 3) ctx->dump_cap = 2 * std::max<uint64_t>(ctx->dump_cap, 1u << 20);
        *retry = true;
    }
    if (h.err & 8ull) return fail(ctx, CD_E_HIP, "batched traversal: a donated node pair was never published");
    if (h.err & ~2ull) return fail(ctx, CD_E_STACK, "batched traversal stack overflow (err " + std::to_string(h.err) + ")");
    if (h.lpslots > ctx->lp_cap) {
        ctx->lp_cap = h.lpslots + h.lpslots / 4;
        *retry = true;
    }
    if (h.contacts > ctx->contact_cap) {
        ctx->contact_cap = h.contacts + h.contacts / 4;
        *retry = true;
    }
    if (h.deferred > ctx->defer_cap || h.deferred2 > ctx->defer_cap) {  // rerun lists ran out of room
        const uint64_t need = std::max(h.deferred, h.deferred2);
        ctx->defer_cap = need + need / 4;
        *retry = true;
    }
    return CD_OK;
}

// The run's work counters as the API reports them.
void set_run_counters(cd_ctx* ctx, const RunCounters& h) {
    ctx->counters = cd_counters{};
    ctx->counters.sat_tests = h.sat;
    ctx->counters.leaf_pairs = h.leafpairs;
    ctx->counters.tri_tests = h.tritests;
    ctx->counters.contacts = h.contacts;
    ctx->counters.fp64_fallbacks = h.fp64;
    ctx->counters.exact_fallbacks = h.exact;
    ctx->counters.ub_cases = h.ub;
}

// Capture what enq enqueues on the stream as a graph in *exec (replacing
// the one there).
template <class F>
int capture_graph(cd_ctx* ctx, hipGraphExec_t* exec, F&& enq) {
    if (*exec) {
        (void)hipGraphExecDestroy(*exec);
        *exec = nullptr;
    }
    hipStream_t st = ctx->stream;
    hipGraph_t g = nullptr;
    HCK(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = enq();
    hipError_t ec = hipStreamEndCapture(st, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ec != hipSuccess) return hip_fail(ctx, ec, "hipStreamEndCapture");
    hipError_t ei = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return hip_fail(ctx, ei, "hipGraphInstantiate");
    return CD_OK;
}

// ref: reference semantics (the plain path's tests and results) on the
// batched machinery; else the exact-predicate batched path.
int run_batched(cd_ctx* ctx, uint32_t nq, uint64_t* n_contacts, bool ref) {
    hipStream_t st = ctx->stream;
    // the launch geometry (and a captured graph) is built for a power-of-two
    // capacity; the kernels read the count itself from d_nq, so frames whose
    // pair counts differ replay one graph
    uint32_t nq_cap = 64;
    while (nq_cap < nq) nq_cap *= 2;
    const cd_tuning t = effective_tuning(ctx, nq_cap);
    RunCounters h{};
    for (int attempt = 0; attempt < 8; ++attempt) {
        uint64_t fcap = 0;
        int rc = ensure_batched(ctx, nq, nq_cap, t, &fcap);
        if (rc) return rc;
        *static_cast<unsigned int*>(ctx->h_nq.p) = nq;
        HCK(ctx, hipMemcpyAsync(ctx->d_nq.p, ctx->h_nq.p, sizeof(unsigned int), hipMemcpyHostToDevice, st));
        if (t.use_graph) {
            std::vector<uint64_t> key;
            batched_key(ctx, nq_cap, t, ref, fcap, key);
            if (!(ctx->graph && key == ctx->graph_key)) {
                rc = capture_graph(ctx, &ctx->graph, [&] { return enqueue_batched(ctx, nq_cap, t, false, ref); });
                if (rc) return rc;
                ctx->graph_key = key;
            }
            HCK(ctx, hipEventRecord(ctx->ev[4], st));
            HCK(ctx, hipGraphLaunch(ctx->graph, st));
            HCK(ctx, hipEventRecord(ctx->ev[5], st));
        } else {
            rc = enqueue_batched(ctx, nq_cap, t, true, ref);
            if (rc) return rc;
        }
        rc = read_counters(ctx, h);
        if (rc) return rc;
        bool retry = false;
        rc = check_batched(ctx, h, fcap, t, &retry);
        if (rc) return rc;
        if (retry) continue;
        ctx->timings = cd_timings{};
        if (t.use_graph) {
            float total = 0.f;
            HCK(ctx, hipEventElapsedTime(&total, ctx->ev[4], ctx->ev[5]));
            ctx->timings.total_ms = total;
        } else {
            float t_split = 0.f, t_dfs = 0.f, t_narrow = 0.f;
            HCK(ctx, hipEventElapsedTime(&t_split, ctx->ev[0], ctx->ev[1]));
            HCK(ctx, hipEventElapsedTime(&t_dfs, ctx->ev[1], ctx->ev[2]));
            HCK(ctx, hipEventElapsedTime(&t_narrow, ctx->ev[2], ctx->ev[3]));
            ctx->timings.traverse_ms = t_split + t_dfs;
            ctx->timings.narrow_ms = t_narrow;
            ctx->timings.broad_ms = t_split;
            ctx->timings.total_ms = t_split + t_dfs + t_narrow;
        }
        ctx->timings.launches = 1;
        set_run_counters(ctx, h);
        *n_contacts = h.contacts;
        return CD_OK;
    }
    return fail(ctx, CD_E_NOMEM, "output sizing did not converge");
}

int run_queries_(cd_ctx* ctx, uint32_t nq, uint32_t mode, uint64_t* n_contacts);
int run_queries(cd_ctx* ctx, uint32_t nq, uint32_t mode, uint64_t* n_contacts) {
    ctx->last_run = 0;
    ctx->contacts_sorted = false;
    ctx->rays_ready = false;
    int rc = run_queries_(ctx, nq, mode, n_contacts);
    if (rc == CD_OK) ctx->last_run = 1;
    return rc;
}
int run_queries_(cd_ctx* ctx, uint32_t nq, uint32_t mode, uint64_t* n_contacts) {
    if (mode != CD_MODE_PLAIN && mode != CD_MODE_BATCHED && mode != CD_MODE_REFERENCE)
        return fail(ctx, CD_E_INVALID, "unknown mode");
    if (nq == 0) {  // nothing to launch (the kernels assume at least one query record)
        ctx->counters = cd_counters{};
        ctx->timings = cd_timings{};
        *n_contacts = 0;
        return CD_OK;
    }
    if (mode == CD_MODE_PLAIN) return run_plain(ctx, nq, n_contacts);
    if (mode == CD_MODE_BATCHED) return run_batched(ctx, nq, n_contacts, false);
    if (mode == CD_MODE_REFERENCE) return run_batched(ctx, nq, n_contacts, true);
    return fail(ctx, CD_E_INVALID, "unknown mode");
}

int stage(cd_ctx* ctx, const cd_query* q, uint32_t nq) {
    if (nq && !q) return fail(ctx, CD_E_INVALID, "null query array");
    const QueryRec* qr = reinterpret_cast<const QueryRec*>(q);
    int rc = check_queries(ctx, qr, nq);
    if (rc) return rc;
    HCK(ctx, ctx->d_queries.ensure(size_t(std::max<uint32_t>(nq, 1)) * sizeof(QueryRec)));
    if (nq) HCK(ctx, hipMemcpyAsync(ctx->d_queries.p, qr, size_t(nq) * sizeof(QueryRec), hipMemcpyHostToDevice, ctx->stream));
    HCK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->staged_nq = nq;
    ctx->staged_ids.resize(nq);
    for (uint32_t i = 0; i < nq; ++i) ctx->staged_ids[i] = make_uint2(qr[i].mesh_a, qr[i].mesh_b);
    ctx->last_run = 0;
    ctx->rays_ready = false;
    return CD_OK;
}

// Grow a device buffer, keeping its contents.
hipError_t grow_keep(DevBuf& b, size_t need, hipStream_t st) {
    if (need <= b.bytes) return hipSuccess;
    const size_t want = std::max(need, b.bytes * 2);
    void* np = nullptr;
    hipError_t e = hipMalloc(&np, want);
    if (e != hipSuccess) return e;
    if (b.p) {
        e = hipMemcpyAsync(np, b.p, b.bytes, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            (void)hipFree(np);
            return e;
        }
        (void)hipFree(b.p);
    }
    b.p = np;
    b.bytes = want;
    return hipSuccess;
}

// Sweep and prune on the GPU.  Leaves the oriented pairs, sorted by
// (first, second), in *hp and the matching queries (relative matrices) in
// d_queries.  With wait false the pairs are copied into h_pairs without
// waiting for them; finish_broad() takes them, and the broad phase's time,
// after a later synchronisation of the stream.
int check_entities(cd_ctx* ctx, const cd_entity* ents, uint32_t n) {
    if (!ents) return fail(ctx, CD_E_INVALID, "null entity array");
    if (n >= 0xFFFFu) return fail(ctx, CD_E_INVALID, "more entities than the 16-bit Entity type allows");
    for (uint32_t i = 0; i < n; ++i) {
        if (ents[i].mesh_id >= ctx->meshes.size()) return fail(ctx, CD_E_INVALID, "entity names an unknown mesh");
        if (!all_finite(ents[i].current, 16)) return fail(ctx, CD_E_INVALID, "non-finite entity matrix");
    }
    return CD_OK;
}

int broad_phase(cd_ctx* ctx, const cd_entity* ents, uint32_t n, uint32_t* np_out, std::vector<uint2>* hp,
                bool wait = true) {
    int rc = check_entities(ctx, ents, n);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    HCK(ctx, ctx->d_work[W_ENTS].ensure(size_t(n) * sizeof(cd_entity)));
    HCK(ctx, ctx->d_work[W_IV].ensure(size_t(n) * 6 * sizeof(float)));
    for (int w : {W_K0, W_V0, W_K1, W_V1}) HCK(ctx, ctx->d_work[w].ensure(size_t(n) * sizeof(uint32_t)));
    HCK(ctx, ctx->d_work[W_PCOUNT].ensure(sizeof(unsigned int)));
    HCK(ctx, hipMemcpyAsync(ctx->d_work[W_ENTS].p, ents, size_t(n) * sizeof(cd_entity), hipMemcpyHostToDevice, st));
    const cd_entity* de = ctx->d_work[W_ENTS].as<cd_entity>();
    uint32_t* k0 = ctx->d_work[W_K0].as<uint32_t>();
    uint32_t* v0 = ctx->d_work[W_V0].as<uint32_t>();
    uint32_t* k1 = ctx->d_work[W_K1].as<uint32_t>();
    uint32_t* v1 = ctx->d_work[W_V1].as<uint32_t>();
    unsigned int* cnt = ctx->d_work[W_PCOUNT].as<unsigned int>();
    uint32_t cap = std::max<uint32_t>(4 * n, 1024);
    uint32_t np = 0;
    HCK(ctx, hipEventRecord(ctx->ev[6], st));
    HCK(ctx, launch_sap_project(de, n, ctx->d_meshes.as<MeshRec>(), sap_axes(), ctx->d_work[W_IV].as<float>(), k0, v0, st));
    uint32_t* table = nullptr;
    HCK(ctx, sort_table(ctx, n, &table));
    HCK(ctx, launch_radix_sort(k0, v0, k1, v1, n, 32, table, st));
    for (int attempt = 0; attempt < 3; ++attempt) {
        HCK(ctx, ctx->d_work[W_PAIRS].ensure(size_t(cap) * sizeof(uint2)));
        HCK(ctx, launch_fill_u32(cnt, 1, 0u, st));
        HCK(ctx, launch_sap_sweep(v0, ctx->d_work[W_IV].as<float>(), de, n, ctx->d_work[W_PAIRS].as<uint2>(), cap, cnt, st));
        HCK(ctx, hipMemcpyAsync(&np, cnt, sizeof np, hipMemcpyDeviceToHost, st));
        HCK(ctx, hipStreamSynchronize(st));
        if (np <= cap) break;
        cap = np + np / 4;
    }
    if (np > cap) return fail(ctx, CD_E_NOMEM, "pair buffer sizing did not converge");
    if (np) {
        for (int w : {W_K0, W_V0, W_K1, W_V1}) HCK(ctx, ctx->d_work[w].ensure(size_t(np) * sizeof(uint32_t)));
        k0 = ctx->d_work[W_K0].as<uint32_t>();
        v0 = ctx->d_work[W_V0].as<uint32_t>();
        k1 = ctx->d_work[W_K1].as<uint32_t>();
        v1 = ctx->d_work[W_V1].as<uint32_t>();
        HCK(ctx, ctx->d_work[W_SPAIRS].ensure(size_t(np) * sizeof(uint2)));
        HCK(ctx, ctx->d_queries.ensure(size_t(np) * sizeof(QueryRec)));
        HCK(ctx, launch_pair_keys(ctx->d_work[W_PAIRS].as<uint2>(), np, k0, v0, st));
        HCK(ctx, sort_table(ctx, np, &table));
        HCK(ctx, launch_radix_sort(k0, v0, k1, v1, np, key_bits_below(uint64_t(n) << 16), table, st));
        HCK(ctx, launch_pair_queries(ctx->d_work[W_PAIRS].as<uint2>(), v0, np, de, ctx->d_queries.as<QueryRec>(),
                                     ctx->d_work[W_SPAIRS].as<uint2>(), st));
    }
    HCK(ctx, hipEventRecord(ctx->ev[7], st));
    ctx->timings = cd_timings{};
    *np_out = np;
    if (!wait) {
        HCK(ctx, ctx->h_pairs.ensure(size_t(np) * sizeof(uint2)));
        if (np)
            HCK(ctx, hipMemcpyAsync(ctx->h_pairs.p, ctx->d_work[W_SPAIRS].p, size_t(np) * sizeof(uint2),
                                    hipMemcpyDeviceToHost, st));
        return CD_OK;
    }
    hp->assign(np, make_uint2(0, 0));
    if (np) HCK(ctx, hipMemcpyAsync(hp->data(), ctx->d_work[W_SPAIRS].p, size_t(np) * sizeof(uint2), hipMemcpyDeviceToHost, st));
    HCK(ctx, hipStreamSynchronize(st));
    float ms = 0.f;
    HCK(ctx, hipEventElapsedTime(&ms, ctx->ev[6], ctx->ev[7]));
    ctx->timings.broad_ms = ms;
    return CD_OK;
}

// The pairs and time of a broad_phase(..., wait = false).
int finish_broad(cd_ctx* ctx, uint32_t np, std::vector<uint2>* hp, float* ms) {
    HCK(ctx, hipEventSynchronize(ctx->ev[7]));
    hp->resize(np);
    if (np) std::memcpy(hp->data(), ctx->h_pairs.p, size_t(np) * sizeof(uint2));
    HCK(ctx, hipEventElapsedTime(ms, ctx->ev[6], ctx->ev[7]));
    return CD_OK;
}

// ---- scratch of the allocation-free triangle-pair stream ----------------------
// cd_tritri_batch_noalloc keeps its run counters (first 256 bytes) and the
// exact mode's two rerun lists (64 + 2 n words) here.  Each GPU holds its own
// copy of the array at its own address, resolved once per device at context
// creation; in host code `scratch` names no device address, so only that
// resolved address is ever handed to kernels or copies.
constexpr size_t NOALLOC_COUNTER_BYTES = 256;
constexpr size_t NOALLOC_BYTES = NOALLOC_COUNTER_BYTES + (64 + 2 * size_t(CD_NOALLOC_MAX_PAIRS)) * sizeof(uint32_t);
static_assert(sizeof(RunCounters) <= NOALLOC_COUNTER_BYTES, "counters fit their slot");
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// The device address of `scratch` on `device`, which must be current.
hipError_t resolve_scratch(int device, void** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

// p points into one device allocation of this runtime with at least `bytes`
// bytes from p on.  A pointer that is not (host memory) fails the lookup,
// whose error would otherwise stay as the thread's last error and be
// reported by the next launch's hipGetLastError: it is cleared here.
bool device_span(const void* p, size_t bytes) {
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, const_cast<void*>(p)) != hipSuccess || !base) {
        (void)hipGetLastError();
        return false;
    }
    const uintptr_t b = reinterpret_cast<uintptr_t>(base), q = reinterpret_cast<uintptr_t>(p);
    return q >= b && q - b <= size && bytes <= size - (q - b);
}

}  // namespace

extern "C" {

int cd_create(cd_ctx** out, const cd_config* cfg) {
    if (!out) return CD_E_INVALID;
    *out = nullptr;
    cd_ctx* ctx = new (std::nothrow) cd_ctx();
    if (!ctx) return CD_E_NOMEM;
    if (cfg) {
        ctx->leaf_max = cfg->leaf_max ? cfg->leaf_max : 4;
        ctx->device = int(cfg->device);
    }
    if (ctx->leaf_max > 64) {
        delete ctx;
        return CD_E_INVALID;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ctx->device >= ndev) {
        delete ctx;
        return CD_E_HIP;
    }
    if (hipSetDevice(ctx->device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return CD_E_HIP;
    }
    for (auto& e : ctx->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete ctx;
            return CD_E_HIP;
        }
    if (ctx->d_counters.ensure(sizeof(RunCounters)) != hipSuccess) {
        delete ctx;
        return CD_E_NOMEM;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess && prop.multiProcessorCount > 0)
        ctx->num_cu = prop.multiProcessorCount;
    {  // the allocation-free entry point's scratch on this device (only that entry point needs it)
        const hipError_t e = resolve_scratch(ctx->device, &ctx->scratch);
        if (e != hipSuccess) {
            ctx->scratch = nullptr;
            ctx->scratch_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        }
    }
    *out = ctx;
    return CD_OK;
}

int cd_destroy(cd_ctx* ctx) {
    if (!ctx) return CD_E_INVALID;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (DevBuf* b : {&ctx->d_meshes, &ctx->d_nodes, &ctx->d_tris, &ctx->d_perm, &ctx->d_queries, &ctx->d_leafpairs,
                      &ctx->d_contacts, &ctx->d_counters, &ctx->d_tt_a, &ctx->d_tt_b, &ctx->d_tt_flags, &ctx->d_tt_seg,
                      &ctx->d_tt_list, &ctx->d_bind, &ctx->d_joints, &ctx->d_weights, &ctx->d_vidx, &ctx->d_skinpos, &ctx->d_bones,
                      &ctx->d_skinrec, &ctx->d_refit_items, &ctx->d_refit_bits, &ctx->d_refit_scratch,
                      &ctx->d_refit_ranges, &ctx->d_refit_plan, &ctx->d_refit_partial,
                      &ctx->d_leafof, &ctx->d_attrs, &ctx->d_vid, &ctx->d_nrm})
        b->release();
    for (auto& b : ctx->d_work) b.release();
    ctx->d_budget.release();
    ctx->d_nq.release();
    ctx->h_nq.release();
    ctx->h_stage.release();
    ctx->h_pairs.release();
    ctx->h_step.release();
    ctx->h_ents.release();
    ctx->h_ctr.release();
    ctx->d_rayslab.release();
    ctx->d_rootslab.release();
    for (auto& b : ctx->build_scratch) b.release();
    if (ctx->graph) (void)hipGraphExecDestroy(ctx->graph);
    if (ctx->step_graph) (void)hipGraphExecDestroy(ctx->step_graph);
    if (ctx->frame_graph) (void)hipGraphExecDestroy(ctx->frame_graph);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return CD_OK;
}

const char* cd_last_error(const cd_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int cd_register_mesh(cd_ctx* ctx, const float* positions, uint32_t nvert, const uint32_t* indices, uint32_t ntri,
                     uint32_t* mesh_id) {
    if (!ctx) return CD_E_INVALID;
    if (ntri && (!positions || !indices)) return fail(ctx, CD_E_INVALID, "null mesh arrays");
    std::vector<float> soup(size_t(ntri) * 9);
    for (uint32_t t = 0; t < ntri; ++t)
        for (int k = 0; k < 3; ++k) {
            const uint32_t vi = indices[size_t(t) * 3 + k];
            if (vi >= nvert) return fail(ctx, CD_E_INVALID, "index out of range");
            std::memcpy(&soup[size_t(t) * 9 + k * 3], positions + size_t(vi) * 3, 3 * sizeof(float));
        }
    uint32_t id = 0;
    int rc = register_soups(ctx, soup.data(), &ntri, 1, &id);
    if (rc) return rc;
    rc = store_vertex_ids(ctx, id, indices, ntri, nvert);
    if (rc) return rc;
    if (mesh_id) *mesh_id = id;
    return CD_OK;
}

int cd_set_mesh_normals(cd_ctx* ctx, uint32_t mesh, const float* normals, uint32_t nvert) {
    if (!ctx) return CD_E_INVALID;
    if (mesh >= ctx->meshes.size()) return fail(ctx, CD_E_INVALID, "unknown mesh");
    if (nvert != ctx->mesh_nvert[mesh]) return fail(ctx, CD_E_INVALID, "normal count differs from the mesh's vertex count");
    if (nvert && !normals) return fail(ctx, CD_E_INVALID, "null normals");
    if (!all_finite(normals, size_t(nvert) * 3)) return fail(ctx, CD_E_INVALID, "non-finite normal");
    (void)hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    const size_t n = size_t(nvert) * 3;
    if (ctx->nrm_used + n >= 0xFFFFFFF0ull) return fail(ctx, CD_E_CAPACITY, "normal arena exceeds 32-bit indexing");
    HCK(ctx, grow_keep(ctx->d_nrm, (ctx->nrm_used + n + 1) * sizeof(float), st));
    if (n) HCK(ctx, hipMemcpyAsync(ctx->d_nrm.as<float>() + ctx->nrm_used, normals, n * sizeof(float),
                                   hipMemcpyHostToDevice, st));
    HCK(ctx, hipStreamSynchronize(st));
    ctx->attrs[mesh].y = uint32_t(ctx->nrm_used);
    ctx->nrm_used += n;
    return upload_attrs(ctx);
}

int cd_register_soups(cd_ctx* ctx, const float* tri9, const uint32_t* tri_counts, uint32_t nmesh, uint32_t* first_id) {
    if (!ctx) return CD_E_INVALID;
    if (nmesh && !tri_counts) return fail(ctx, CD_E_INVALID, "null counts");
    (void)hipSetDevice(ctx->device);
    return register_soups(ctx, tri9, tri_counts, nmesh, first_id);
}

// ---- glTF mesh ingest ------------------------------------------------------------
struct cd_gltf {
    cdk::gltf::Document doc;
};

int cd_gltf_open(const char* path, cd_gltf** out, char* err, uint32_t err_cap) {
    auto report = [&](const std::string& m) {
        if (err && err_cap) {
            const size_t n = std::min<size_t>(m.size(), err_cap - 1);
            std::memcpy(err, m.data(), n);
            err[n] = 0;
        }
    };
    if (!path || !out) {
        report("null argument");
        return CD_E_INVALID;
    }
    *out = nullptr;
    cd_gltf* g = new (std::nothrow) cd_gltf;
    if (!g) {
        report("out of host memory");
        return CD_E_NOMEM;
    }
    std::string msg;
    bool ok;
    try {
        ok = cdk::gltf::load(path, g->doc, msg);
    } catch (const std::bad_alloc&) {
        ok = false;
        msg = "out of host memory";
    }
    if (!ok) {
        delete g;
        report(msg);
        return CD_E_FORMAT;
    }
    report("");
    *out = g;
    return CD_OK;
}

int cd_gltf_close(cd_gltf* doc) {
    delete doc;
    return CD_OK;
}

int cd_gltf_mesh_count(const cd_gltf* doc, uint32_t* n) {
    if (!doc || !n) return CD_E_INVALID;
    *n = uint32_t(doc->doc.meshes.size());
    return CD_OK;
}

int cd_gltf_mesh_info_get(const cd_gltf* doc, uint32_t mesh, cd_gltf_mesh_info* out) {
    if (!doc || !out || mesh >= doc->doc.meshes.size()) return CD_E_INVALID;
    const cdk::gltf::Mesh& m = doc->doc.meshes[mesh];
    out->nprim = uint32_t(m.prims.size());
    out->with_normals = m.with_normals;
    out->without_normals = m.without_normals;
    out->pad = 0;
    out->nvert = m.nvert;
    out->ntri = m.ntri;
    return CD_OK;
}

int cd_gltf_primitive_get(const cd_gltf* doc, uint32_t mesh, uint32_t k, cd_gltf_primitive* out) {
    if (!doc || !out || mesh >= doc->doc.meshes.size() || k >= doc->doc.meshes[mesh].prims.size()) return CD_E_INVALID;
    const cdk::gltf::Primitive& p = doc->doc.meshes[mesh].prims[k];
    out->gltf_mode = p.gltf_mode;
    out->mode = p.mode;
    out->excluded = p.excluded ? 1u : 0u;
    out->nvert = p.nvert;
    out->nindex = p.idx.size();
    out->ntri = p.ntri;
    out->positions = p.pos.empty() ? nullptr : p.pos.data();
    out->normals = p.nrm.empty() ? nullptr : p.nrm.data();
    out->indices = p.idx.empty() ? nullptr : p.idx.data();
    return CD_OK;
}

int cd_register_gltf(cd_ctx* ctx, const cd_gltf* doc, uint32_t* mesh_ids, uint32_t cap) {
    if (!ctx) return CD_E_INVALID;
    if (!doc) return fail(ctx, CD_E_INVALID, "null document");
    const auto& meshes = doc->doc.meshes;
    if (cap < meshes.size() || (!mesh_ids && !meshes.empty()))
        return fail(ctx, CD_E_CAPACITY, "mesh id array shorter than the document's " + std::to_string(meshes.size()) +
                                            " meshes");
    // what one tree build takes: every mesh with triangles, its primitives in order
    std::vector<uint32_t> reg, counts;
    std::vector<PrimSpan> spans;
    uint64_t T = 0, V = 0, N = 0;
    for (uint32_t m = 0; m < meshes.size(); ++m) {
        const auto& mesh = meshes[m];
        if (mesh.ntri == 0) continue;
        if (mesh.with_normals && mesh.without_normals)
            return fail(ctx, CD_E_FORMAT, "mesh " + std::to_string(m) +
                                              ": some primitives have NORMAL and some do not; the ray attributes "
                                              "need one kind per mesh");
        if (mesh.ntri > 0x7FFFFFFFull || mesh.nvert > 0x7FFFFFFFull)
            return fail(ctx, CD_E_INVALID, "mesh " + std::to_string(m) + " too large");
        uint64_t vid = 0;
        for (const auto& p : mesh.prims) {
            if (p.excluded) continue;
            if (p.ntri) {
                PrimSpan s{};
                s.tri_off = uint32_t(T);
                s.ntri = uint32_t(p.ntri);
                s.idx_off = uint32_t(N);
                s.mode = p.mode;
                s.vbase = uint32_t(V);
                s.vid_base = uint32_t(vid);
                spans.push_back(s);
            }
            T += p.ntri;
            V += p.nvert;
            N += p.idx.size();
            vid += p.nvert;
            if (T >= 0xFFFFFFF0ull || V >= 0x55555550ull || N >= 0xFFFFFFF0ull)
                return fail(ctx, CD_E_INVALID, "document too large for one ingest (32-bit indexing)");
        }
        reg.push_back(m);
        counts.push_back(uint32_t(mesh.ntri));
    }
    for (uint32_t m = 0; m < meshes.size(); ++m) mesh_ids[m] = CD_NO_MESH;
    if (reg.empty()) return CD_OK;
    (void)hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    // concatenated vertices and index lists of the included primitives
    std::vector<float> pos;
    std::vector<uint32_t> idx;
    pos.reserve(size_t(V) * 3);
    idx.reserve(size_t(N));
    for (uint32_t m : reg)
        for (const auto& p : meshes[m].prims) {
            if (p.excluded) continue;
            pos.insert(pos.end(), p.pos.begin(), p.pos.end());
            idx.insert(idx.end(), p.idx.begin(), p.idx.end());
        }
    ScopedBuf d_pos, d_idx, d_spans, d_soup, d_vid3;
    HCK(ctx, d_pos.ensure(pos.size() * sizeof(float) + 16));
    HCK(ctx, d_idx.ensure(idx.size() * sizeof(uint32_t) + 16));
    HCK(ctx, d_spans.ensure(spans.size() * sizeof(PrimSpan) + 16));
    HCK(ctx, d_soup.ensure(size_t(T) * 9 * sizeof(float) + 16));
    HCK(ctx, d_vid3.ensure(size_t(T) * 3 * sizeof(uint32_t) + 16));
    HCK(ctx, hipMemcpyAsync(d_pos.p, pos.data(), pos.size() * sizeof(float), hipMemcpyHostToDevice, st));
    HCK(ctx, hipMemcpyAsync(d_idx.p, idx.data(), idx.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    HCK(ctx, hipMemcpyAsync(d_spans.p, spans.data(), spans.size() * sizeof(PrimSpan), hipMemcpyHostToDevice, st));
    HCK(ctx, hipEventRecord(ctx->ev[0], st));
    HCK(ctx, launch_assemble(d_spans.as<PrimSpan>(), uint32_t(spans.size()), d_idx.as<uint32_t>(), d_pos.as<float>(),
                             uint32_t(T), d_soup.as<float>(), d_vid3.as<uint32_t>(), st));
    HCK(ctx, hipEventRecord(ctx->ev[1], st));
    BuildResult r = build_soups(ctx, d_soup.as<float>(), counts.data(), uint32_t(reg.size()), true);
    if (!r.err.empty()) return fail(ctx, CD_E_HIP, "tree build: " + r.err);
    HCK(ctx, hipEventRecord(ctx->ev[2], st));
    HCK(ctx, hipEventSynchronize(ctx->ev[2]));
    {  // assembly in broad_ms, the tree build in traverse_ms (device time between events)
        float t_asm = 0.f, t_build = 0.f;
        HCK(ctx, hipEventElapsedTime(&t_asm, ctx->ev[0], ctx->ev[1]));
        HCK(ctx, hipEventElapsedTime(&t_build, ctx->ev[1], ctx->ev[2]));
        ctx->timings = cd_timings{};
        ctx->timings.broad_ms = t_asm;
        ctx->timings.traverse_ms = t_build;
        ctx->timings.total_ms = t_asm + t_build;
        ctx->timings.launches = 1;
    }
    const uint32_t first = uint32_t(ctx->meshes.size());
    ctx->meshes.insert(ctx->meshes.end(), r.meshes.begin(), r.meshes.end());
    for (uint32_t m : reg) {  // plain soups (vertex ids 3 t + k) until their attributes are in
        ctx->attrs.push_back(make_uint2(~0u, ~0u));
        ctx->mesh_nvert.push_back(3u * uint32_t(meshes[m].ntri));
    }
    int rc = upload_mesh_table(ctx);
    if (!rc) rc = upload_attrs(ctx);
    if (rc) return rc;
    // vertex ids in tree order (one launch for the document), then the
    // vertex normals of the meshes that have them (one upload)
    const size_t nvid = size_t(T) * 3;
    if (ctx->vid_used + nvid >= 0xFFFFFFF0ull) return fail(ctx, CD_E_CAPACITY, "vertex-id arena exceeds 32-bit indexing");
    HCK(ctx, grow_keep(ctx->d_vid, (ctx->vid_used + nvid + 1) * sizeof(uint32_t), st));
    std::vector<uint32_t> prefix(reg.size() + 1, 0u);
    size_t nnrm = 0;
    for (size_t i = 0; i < reg.size(); ++i) {
        const MeshRec& mr = ctx->meshes[first + i];
        if (mr.tri_base != ctx->meshes[first].tri_base + prefix[i])
            return fail(ctx, CD_E_HIP, "tree build placed the document's meshes apart");
        prefix[i + 1] = prefix[i] + mr.ntri;
        if (meshes[reg[i]].with_normals) nnrm += size_t(meshes[reg[i]].nvert) * 3;
    }
    ScopedBuf d_prefix;
    HCK(ctx, d_prefix.ensure(prefix.size() * sizeof(uint32_t)));
    HCK(ctx, hipMemcpyAsync(d_prefix.p, prefix.data(), prefix.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    HCK(ctx, launch_tree_vidx_many(ctx->d_perm.as<uint32_t>() + ctx->meshes[first].tri_base, d_prefix.as<uint32_t>(),
                                   uint32_t(reg.size()), T, d_vid3.as<uint32_t>(),
                                   ctx->d_vid.as<uint32_t>() + ctx->vid_used, st));
    if (ctx->nrm_used + nnrm >= 0xFFFFFFF0ull) return fail(ctx, CD_E_CAPACITY, "normal arena exceeds 32-bit indexing");
    std::vector<float> nrm;
    nrm.reserve(nnrm);
    for (size_t i = 0; i < reg.size(); ++i) {
        const MeshRec& mr = ctx->meshes[first + i];
        const auto& mesh = meshes[reg[i]];
        uint2 attr = make_uint2(uint32_t(ctx->vid_used), ~0u);
        ctx->vid_used += size_t(mr.ntri) * 3;
        if (mesh.with_normals) {
            attr.y = uint32_t(ctx->nrm_used + nrm.size());
            const size_t at = nrm.size();
            for (const auto& p : mesh.prims)
                if (!p.excluded) nrm.insert(nrm.end(), p.nrm.begin(), p.nrm.end());
            nrm.resize(at + size_t(mesh.nvert) * 3, 0.f);
        }
        ctx->attrs[first + i] = attr;
        ctx->mesh_nvert[first + i] = uint32_t(mesh.nvert);
        mesh_ids[reg[i]] = first + uint32_t(i);
    }
    if (!nrm.empty()) {
        HCK(ctx, grow_keep(ctx->d_nrm, (ctx->nrm_used + nrm.size() + 1) * sizeof(float), st));
        HCK(ctx, hipMemcpyAsync(ctx->d_nrm.as<float>() + ctx->nrm_used, nrm.data(), nrm.size() * sizeof(float),
                                hipMemcpyHostToDevice, st));
        ctx->nrm_used += nrm.size();
    }
    HCK(ctx, hipStreamSynchronize(st));
    return upload_attrs(ctx);
}

int cd_gltf_node_count(const cd_gltf* doc, uint32_t* n) {
    if (!doc || !n) return CD_E_INVALID;
    *n = uint32_t(doc->doc.nodes.size());
    return CD_OK;
}

int cd_gltf_nodes(const cd_gltf* doc, cd_node* nodes, int32_t* mesh, uint32_t* collides, uint32_t cap) {
    if (!doc) return CD_E_INVALID;
    const auto& nv = doc->doc.nodes;
    if (cap < nv.size()) return CD_E_CAPACITY;
    for (size_t i = 0; i < nv.size(); ++i) {
        if (nodes) {
            cd_node& o = nodes[i];
            std::memset(&o, 0, sizeof o);
            std::memcpy(o.local_scale, nv[i].scale, sizeof o.local_scale);
            std::memcpy(o.local_rotation, nv[i].rotation, sizeof o.local_rotation);
            std::memcpy(o.local_translation, nv[i].translation, sizeof o.local_translation);
            o.parent = nv[i].parent;
        }
        if (mesh) mesh[i] = nv[i].mesh;
        if (collides) collides[i] = nv[i].collides ? 1u : 0u;
    }
    return CD_OK;
}

int cd_node_global_matrices(cd_ctx* ctx, const cd_node* nodes, uint32_t n, float* global16) {
    if (!ctx) return CD_E_INVALID;
    if (n && (!nodes || !global16)) return fail(ctx, CD_E_INVALID, "null node arrays");
    // depth of every node; parents out of range, cycles and chains deeper than 256 are rejected
    constexpr uint32_t MAX_DEPTH = 256;
    std::vector<uint32_t> depth(n, ~0u);
    for (uint32_t i = 0; i < n; ++i) {
        if (!all_finite(nodes[i].local_scale, 13)) return fail(ctx, CD_E_INVALID, "non-finite node transform");
        if (nodes[i].parent != CD_NO_PARENT && nodes[i].parent >= n)
            return fail(ctx, CD_E_INVALID, "node " + std::to_string(i) + ": parent out of range");
    }
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t a = i, steps = 0;
        while (depth[a] == ~0u && nodes[a].parent != CD_NO_PARENT) {
            a = nodes[a].parent;
            if (++steps > MAX_DEPTH) return fail(ctx, CD_E_INVALID, "node hierarchy deeper than 256 or cyclic");
        }
        uint32_t d = depth[a] == ~0u ? 0u : depth[a];
        if (depth[a] == ~0u) depth[a] = 0;
        // walk again from i, writing depths top-down
        std::vector<uint32_t> chain;
        for (uint32_t b = i; b != a; b = nodes[b].parent) chain.push_back(b);
        for (size_t k = chain.size(); k-- > 0;) {
            d += 1;
            if (d > MAX_DEPTH) return fail(ctx, CD_E_INVALID, "node hierarchy deeper than 256 or cyclic");
            depth[chain[k]] = d;
        }
    }
    if (n == 0) return CD_OK;
    (void)hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    ScopedBuf d_nodes, d_depth, d_out;
    HCK(ctx, d_nodes.ensure(size_t(n) * sizeof(cd_node)));
    HCK(ctx, d_depth.ensure(size_t(n) * sizeof(uint32_t)));
    HCK(ctx, d_out.ensure(size_t(n) * 16 * sizeof(float)));
    HCK(ctx, hipMemcpyAsync(d_nodes.p, nodes, size_t(n) * sizeof(cd_node), hipMemcpyHostToDevice, st));
    HCK(ctx, hipMemcpyAsync(d_depth.p, depth.data(), size_t(n) * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    HCK(ctx, launch_node_globals(d_nodes.as<cd_node>(), d_depth.as<uint32_t>(), n, d_out.as<float>(), st));
    HCK(ctx, hipMemcpyAsync(global16, d_out.p, size_t(n) * 16 * sizeof(float), hipMemcpyDeviceToHost, st));
    HCK(ctx, hipStreamSynchronize(st));
    return CD_OK;
}

int cd_mesh_info_get(cd_ctx* ctx, uint32_t mesh, cd_mesh_info* out) {
    if (!ctx || !out) return CD_E_INVALID;
    if (mesh >= ctx->meshes.size()) return fail(ctx, CD_E_INVALID, "unknown mesh");
    const MeshRec& m = ctx->meshes[mesh];
    out->ntri = m.ntri;
    out->nnodes = m.nnodes;
    out->depth = m.depth;
    out->leaf_max = ctx->leaf_max;
    std::memcpy(out->root, m.root.box, sizeof out->root);
    out->slack = m.slack;
    out->radius = m.radius;
    return CD_OK;
}

int cd_export_tree(cd_ctx* ctx, uint32_t mesh, float* root12, float* node_obbs, uint32_t* node_links, uint32_t* perm) {
    if (!ctx) return CD_E_INVALID;
    if (mesh >= ctx->meshes.size()) return fail(ctx, CD_E_INVALID, "unknown mesh");
    const MeshRec& m = ctx->meshes[mesh];
    if (root12) std::memcpy(root12, m.root.box, sizeof m.root.box);
    if (m.nnodes && (node_obbs || node_links)) {
        std::vector<NodeRec> recs(m.nnodes);
        HCK(ctx, hipMemcpyAsync(recs.data(), ctx->d_nodes.as<NodeRec>() + m.node_base, m.nnodes * sizeof(NodeRec),
                                hipMemcpyDeviceToHost, ctx->stream));
        HCK(ctx, hipStreamSynchronize(ctx->stream));
        for (uint32_t i = 0; i < m.nnodes; ++i) {
            if (node_obbs) {
                std::memcpy(node_obbs + size_t(i) * 24, recs[i].c[0].box, 12 * sizeof(float));
                std::memcpy(node_obbs + size_t(i) * 24 + 12, recs[i].c[1].box, 12 * sizeof(float));
            }
            if (node_links) {
                node_links[size_t(i) * 4 + 0] = recs[i].c[0].link;
                node_links[size_t(i) * 4 + 1] = recs[i].c[1].link;
                node_links[size_t(i) * 4 + 2] = child_cnt(recs[i].c[0]);
                node_links[size_t(i) * 4 + 3] = child_cnt(recs[i].c[1]);
            }
        }
    }
    if (perm && m.ntri) {
        HCK(ctx, hipMemcpyAsync(perm, ctx->d_perm.as<uint32_t>() + m.tri_base, m.ntri * sizeof(uint32_t),
                                hipMemcpyDeviceToHost, ctx->stream));
        HCK(ctx, hipStreamSynchronize(ctx->stream));
    }
    return CD_OK;
}

int cd_tritri_batch(cd_ctx* ctx, const float* a9, const float* b9, uint64_t n, uint32_t mode, int32_t* flags, float* seg6) {
    if (!ctx) return CD_E_INVALID;
    if (n && (!a9 || !b9 || !flags)) return fail(ctx, CD_E_INVALID, "null argument");
    if (mode != CD_MODE_PLAIN && mode != CD_MODE_BATCHED && mode != CD_MODE_REFERENCE)
        return fail(ctx, CD_E_INVALID, "unknown mode");
    if (mode == CD_MODE_REFERENCE) mode = CD_MODE_PLAIN;  // one pair per test: the reference test is the plain one
    hipStream_t st = ctx->stream;
    HCK(ctx, ctx->d_tt_a.ensure(n * 9 * sizeof(float) + 4));
    HCK(ctx, ctx->d_tt_b.ensure(n * 9 * sizeof(float) + 4));
    HCK(ctx, ctx->d_tt_flags.ensure(n * sizeof(int32_t) + 4));
    HCK(ctx, ctx->d_tt_seg.ensure(n * 6 * sizeof(float) + 4));
    const uint32_t tt_grid = uint32_t(ctx->num_cu) * 8;
    if (mode == CD_MODE_BATCHED) {
        if (n >= 0xFFFFFFFFull) return fail(ctx, CD_E_INVALID, "2^32 - 1 or more pairs in one exact batch");
        HCK(ctx, ctx->d_tt_list.ensure((64 + 2 * n) * sizeof(uint32_t)));
    }
    if (n) {
        HCK(ctx, hipMemcpyAsync(ctx->d_tt_a.p, a9, n * 9 * sizeof(float), hipMemcpyHostToDevice, st));
        HCK(ctx, hipMemcpyAsync(ctx->d_tt_b.p, b9, n * 9 * sizeof(float), hipMemcpyHostToDevice, st));
    }
    HCK(ctx, launch_zero(dev_counters(ctx), st));
    HCK(ctx, hipEventRecord(ctx->ev[0], st));
    if (mode == CD_MODE_PLAIN)
        HCK(ctx, launch_tritri_plain(ctx->d_tt_a.as<float>(), ctx->d_tt_b.as<float>(), n, ctx->d_tt_flags.as<int32_t>(),
                                     seg6 ? ctx->d_tt_seg.as<float>() : nullptr, st));
    else
        HCK(ctx, launch_tritri_exact(ctx->d_tt_a.as<float>(), ctx->d_tt_b.as<float>(), n, ctx->d_tt_flags.as<int32_t>(),
                                     seg6 ? ctx->d_tt_seg.as<float>() : nullptr, ctx->d_tt_list.as<uint32_t>(),
                                     tt_grid, dev_counters(ctx), st));
    HCK(ctx, hipEventRecord(ctx->ev[1], st));
    if (n) {
        HCK(ctx, hipMemcpyAsync(flags, ctx->d_tt_flags.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        if (seg6) HCK(ctx, hipMemcpyAsync(seg6, ctx->d_tt_seg.p, n * 6 * sizeof(float), hipMemcpyDeviceToHost, st));
    }
    unsigned long long lens[2] = {0, 0};
    if (mode == CD_MODE_BATCHED && n)
        HCK(ctx, hipMemcpyAsync(lens, ctx->d_tt_list.p, sizeof lens, hipMemcpyDeviceToHost, st));
    HCK(ctx, hipStreamSynchronize(st));
    if (std::max(lens[0], lens[1]) > n)  // a pass defers each pair at most once: unreachable
        return fail(ctx, CD_E_HIP, "exact stream: rerun list overflow");
    float ms = 0.f;
    HCK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    RunCounters h{};
    int rc = read_counters(ctx, h);
    if (rc) return rc;
    ctx->counters = cd_counters{};
    ctx->counters.tri_tests = n;
    ctx->counters.fp64_fallbacks = h.fp64;
    ctx->counters.exact_fallbacks = h.exact;
    ctx->timings = cd_timings{};
    ctx->timings.total_ms = ms;
    ctx->timings.narrow_ms = ms;
    ctx->timings.launches = mode == CD_MODE_PLAIN ? 1 : 4;
    return CD_OK;
}

int cd_tritri_batch_noalloc(cd_ctx* ctx, const float* a9, const float* b9, uint64_t n, uint32_t mode,
                            int32_t* flags, float* seg6) {
    if (!ctx) return CD_E_INVALID;
    if (mode != CD_MODE_PLAIN && mode != CD_MODE_BATCHED && mode != CD_MODE_REFERENCE)
        return fail(ctx, CD_E_INVALID, "unknown mode");
    if (mode == CD_MODE_REFERENCE) mode = CD_MODE_PLAIN;  // one pair per test: the reference test is the plain one
    if (n > CD_NOALLOC_MAX_PAIRS)
        return fail(ctx, CD_E_CAPACITY, "more than " + std::to_string(CD_NOALLOC_MAX_PAIRS) +
                                            " pairs in one allocation-free call");
    if (!ctx->scratch) return fail(ctx, CD_E_HIP, "allocation-free scratch unavailable: " + ctx->scratch_err);
    (void)hipSetDevice(ctx->device);
    if (n) {
        if (!a9 || !b9 || !flags) return fail(ctx, CD_E_INVALID, "null argument");
        if (!device_span(a9, n * 9 * sizeof(float)) || !device_span(b9, n * 9 * sizeof(float)))
            return fail(ctx, CD_E_INVALID, "a9 / b9 must be device memory holding n x 9 floats");
        if (!device_span(flags, n * sizeof(int32_t)) || (seg6 && !device_span(seg6, n * 6 * sizeof(float))))
            return fail(ctx, CD_E_INVALID, "flags / seg6 must be device memory holding n results");
    }
    hipStream_t st = ctx->stream;
    RunCounters* ctr = static_cast<RunCounters*>(ctx->scratch);
    uint32_t* lists = reinterpret_cast<uint32_t*>(static_cast<unsigned char*>(ctx->scratch) + NOALLOC_COUNTER_BYTES);
    // the array keeps the last call's contents: the counters start at zero,
    // and the exact launcher zeroes its list lengths itself (fill kernels)
    HCK(ctx, launch_zero(ctr, st));
    HCK(ctx, hipEventRecord(ctx->ev[0], st));
    if (mode == CD_MODE_PLAIN)
        HCK(ctx, launch_tritri_plain(a9, b9, n, flags, seg6, st));
    else
        HCK(ctx, launch_tritri_exact(a9, b9, n, flags, seg6, lists, uint32_t(ctx->num_cu) * 8, ctr, st));
    HCK(ctx, hipEventRecord(ctx->ev[1], st));
    RunCounters h{};
    unsigned long long lens[2] = {0, 0};
    HCK(ctx, hipMemcpyAsync(&h, ctr, sizeof h, hipMemcpyDeviceToHost, st));
    if (mode == CD_MODE_BATCHED && n) HCK(ctx, hipMemcpyAsync(lens, lists, sizeof lens, hipMemcpyDeviceToHost, st));
    HCK(ctx, hipStreamSynchronize(st));
    if (std::max(lens[0], lens[1]) > n)  // a pass defers each pair at most once: unreachable
        return fail(ctx, CD_E_HIP, "exact stream: rerun list overflow");
    float ms = 0.f;
    HCK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    ctx->counters = cd_counters{};
    ctx->counters.tri_tests = n;
    ctx->counters.fp64_fallbacks = h.fp64;
    ctx->counters.exact_fallbacks = h.exact;
    ctx->timings = cd_timings{};
    ctx->timings.total_ms = ms;
    ctx->timings.narrow_ms = ms;
    ctx->timings.launches = mode == CD_MODE_PLAIN ? 1 : 4;
    return CD_OK;
}

int cd_query_batch(cd_ctx* ctx, const cd_query* q, uint32_t nq, uint32_t mode, cd_contact* out, uint64_t cap,
                   uint64_t* n_out) {
    if (!ctx || !n_out) return CD_E_INVALID;
    (void)hipSetDevice(ctx->device);
    int rc = stage(ctx, q, nq);
    if (rc) return rc;
    uint64_t n = 0;
    rc = run_queries(ctx, nq, mode, &n);
    if (rc) return rc;
    ctx->staged_contacts = n;
    *n_out = n;
    if (n > cap || (n && !out)) return fail(ctx, CD_E_CAPACITY, "contact buffer too small");
    if (n) HCK(ctx, hipMemcpyAsync(out, ctx->d_contacts.p, n * sizeof(cd_contact), hipMemcpyDeviceToHost, ctx->stream));
    HCK(ctx, hipStreamSynchronize(ctx->stream));
    return CD_OK;
}

int cd_query_leafpairs(cd_ctx* ctx, const cd_query* q, uint32_t nq, uint32_t* out4, uint64_t cap, uint64_t* n_out) {
    if (!ctx || !n_out) return CD_E_INVALID;
    int rc = stage(ctx, q, nq);
    if (rc) return rc;
    uint64_t n = 0;
    rc = run_plain(ctx, nq, &n);
    if (rc) return rc;
    const uint64_t nlp = ctx->counters.leaf_pairs;
    *n_out = nlp;
    if (nlp > cap || (nlp && !out4)) return fail(ctx, CD_E_CAPACITY, "leaf-pair buffer too small");
    if (nlp) HCK(ctx, hipMemcpyAsync(out4, ctx->d_leafpairs.p, nlp * sizeof(uint4), hipMemcpyDeviceToHost, ctx->stream));
    HCK(ctx, hipStreamSynchronize(ctx->stream));
    return CD_OK;
}

int cd_stage_queries(cd_ctx* ctx, const cd_query* q, uint32_t nq) {
    if (!ctx) return CD_E_INVALID;
    return stage(ctx, q, nq);
}

int cd_run_staged(cd_ctx* ctx, uint32_t mode, uint64_t* n_contacts) {
    if (!ctx || !n_contacts) return CD_E_INVALID;
    uint64_t n = 0;
    int rc = run_queries(ctx, ctx->staged_nq, mode, &n);
    if (rc) return rc;
    ctx->staged_contacts = n;
    *n_contacts = n;
    return CD_OK;
}

int cd_fetch_contacts(cd_ctx* ctx, cd_contact* out, uint64_t cap, uint64_t* n_out) {
    if (!ctx || !n_out) return CD_E_INVALID;
    const uint64_t n = ctx->staged_contacts;
    *n_out = n;
    if (n > cap || (n && !out)) return fail(ctx, CD_E_CAPACITY, "contact buffer too small");
    if (n) HCK(ctx, hipMemcpyAsync(out, ctx->d_contacts.p, n * sizeof(cd_contact), hipMemcpyDeviceToHost, ctx->stream));
    HCK(ctx, hipStreamSynchronize(ctx->stream));
    return CD_OK;
}

}  // extern "C"

namespace {
// cd_frame's results from the nc sorted contacts hc (in host memory) and the
// frame's pairs (ctx->staged_ids): the pairs that produced contacts
// (CollisionDetection.cpp:52, :63 keep only those) with their ranges.
int frame_output(cd_ctx* ctx, const cd_contact* hc, uint64_t nc, cd_pair* pairs, uint32_t pair_cap,
                 uint32_t* n_pairs, cd_contact* contacts, uint64_t contact_cap, uint64_t* n_contacts) {
    const std::vector<uint2>& hp = ctx->staged_ids;
    std::vector<cd_pair> out_pairs;
    for (uint64_t i = 0; i < nc;) {
        uint64_t j = i;
        while (j < nc && hc[j].query == hc[i].query) ++j;
        const uint2 pr = hp[hc[i].query];
        out_pairs.push_back(cd_pair{pr.x, pr.y, uint32_t(i), uint32_t(j - i)});
        i = j;
    }
    *n_pairs = uint32_t(out_pairs.size());
    *n_contacts = nc;
    if (out_pairs.size() > pair_cap || nc > contact_cap || (nc && (!pairs || !contacts)))
        return fail(ctx, CD_E_CAPACITY, "pair or contact buffer too small");
    if (!out_pairs.empty()) std::memcpy(pairs, out_pairs.data(), out_pairs.size() * sizeof(cd_pair));
    if (nc) std::memcpy(contacts, hc, nc * sizeof(cd_contact));
    return CD_OK;
}

// cd_frame; to_host false (cd_step): the sorted contacts stay on the device,
// pairs and contacts are not returned.
int frame_impl(cd_ctx* ctx, const cd_entity* entities, uint32_t n, uint32_t mode, cd_pair* pairs, uint32_t pair_cap,
               uint32_t* n_pairs, cd_contact* contacts, uint64_t contact_cap, uint64_t* n_contacts, bool to_host) {
    if (!ctx || !n_pairs || !n_contacts) return CD_E_INVALID;
    (void)hipSetDevice(ctx->device);
    *n_pairs = 0;
    *n_contacts = 0;
    if (n < 2) {  // CollisionDetection.cpp:40
        ctx->counters = cd_counters{};
        return CD_OK;
    }
    uint32_t np = 0;
    std::vector<uint2> hp;
    ctx->last_run = 0;
    ctx->rays_ready = false;
    // The plain mode checks the pairs' tree depths first; the batched mode
    // takes the pairs after its run, whose count read-back waits for them.
    const bool defer = mode != CD_MODE_PLAIN;
    int rc = broad_phase(ctx, entities, n, &np, &hp, !defer);
    if (rc) return rc;
    ctx->frame_mesh.resize(n);
    for (uint32_t i = 0; i < n; ++i) ctx->frame_mesh[i] = entities[i].mesh_id;
    float broad_ms = ctx->timings.broad_ms;
    if (np == 0 && defer) {
        rc = finish_broad(ctx, np, &hp, &broad_ms);
        if (rc) return rc;
    }
    if (np == 0) {
        ctx->counters = cd_counters{};
        ctx->timings = cd_timings{};
        ctx->timings.broad_ms = broad_ms;
        ctx->timings.total_ms = broad_ms;
        ctx->staged_nq = 0;
        ctx->staged_contacts = 0;
        ctx->staged_ids.clear();
        ctx->last_run = 2;
        return CD_OK;
    }
    if (mode == CD_MODE_PLAIN) {  // the stack the plain walk of the deepest pair needs
        uint32_t need = 0;
        for (uint32_t i = 0; i < np; ++i)
            need = std::max(need, ctx->meshes[entities[hp[i].x].mesh_id].depth +
                                      ctx->meshes[entities[hp[i].y].mesh_id].depth + 2);
        ctx->plain_need = need;
    }
    ctx->staged_nq = np;
    uint64_t nc = 0;
    rc = run_queries(ctx, np, mode, &nc);
    if (rc) return rc;
    if (defer) {
        rc = finish_broad(ctx, np, &hp, &broad_ms);
        if (rc) return rc;
    }
    ctx->counters.broad_pairs = np;
    ctx->timings.broad_ms = broad_ms;
    ctx->timings.total_ms += broad_ms;
    ctx->staged_contacts = nc;
    // contacts grouped by pair, ordered by (pair, tri_a, tri_b), staged in
    // pinned host memory
    const cd_contact* hc = nullptr;
    if (nc) {
        if (nc > 0xFFFFFFF0ull) return fail(ctx, CD_E_CAPACITY, "too many contacts in one frame");
        hipStream_t st = ctx->stream;
        HCK(ctx, ctx->d_work[W_CSORT].ensure(nc * sizeof(cd_contact)));
        for (int w : {W_K0, W_V0, W_K1, W_V1}) HCK(ctx, ctx->d_work[w].ensure(nc * sizeof(uint32_t)));
        uint32_t* table = nullptr;
        HCK(ctx, sort_table(ctx, uint32_t(nc), &table));
        uint32_t bits[3];
        contact_field_bits(ctx, np, bits);
        HCK(ctx, launch_sort_contacts(ctx->d_contacts.as<cd_contact>(), uint32_t(nc), ctx->d_work[W_CSORT].as<cd_contact>(),
                                      bits, ctx->d_work[W_K0].as<uint32_t>(), ctx->d_work[W_V0].as<uint32_t>(),
                                      ctx->d_work[W_K1].as<uint32_t>(), ctx->d_work[W_V1].as<uint32_t>(), table, st));
        if (to_host) {
            HCK(ctx, ctx->h_stage.ensure(nc * sizeof(cd_contact)));
            HCK(ctx, hipMemcpyAsync(ctx->h_stage.p, ctx->d_work[W_CSORT].p, nc * sizeof(cd_contact),
                                    hipMemcpyDeviceToHost, st));
            HCK(ctx, hipStreamSynchronize(st));
            hc = static_cast<const cd_contact*>(ctx->h_stage.p);
        }
    }
    ctx->contacts_sorted = true;
    ctx->last_run = 2;
    ctx->staged_ids = hp;
    *n_contacts = nc;
    if (!to_host) return CD_OK;
    return frame_output(ctx, hc, nc, pairs, pair_cap, n_pairs, contacts, contact_cap, n_contacts);
}

// The slab records of the ray shooting's box tests for the given meshes,
// recomputed for a mesh whose boxes changed since (a build or a refit).
int ensure_slabs(cd_ctx* ctx, const uint32_t* mesh_ids, uint32_t n) {
    hipStream_t st = ctx->stream;
    ctx->slab_ok.resize(ctx->meshes.size(), 0);
    HCK(ctx, grow_keep(ctx->d_rayslab, (2 * ctx->nodes_used + 2) * sizeof(RaySlab), st));
    HCK(ctx, grow_keep(ctx->d_rootslab, (ctx->meshes.size() + 1) * sizeof(RaySlab), st));
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t m = mesh_ids[i];
        if (ctx->slab_ok[m]) continue;
        const MeshRec& r = ctx->meshes[m];
        HCK(ctx, launch_ray_slabs(ctx->d_nodes.as<NodeRec>() + r.node_base, r.nnodes,
                                  ctx->d_rayslab.as<RaySlab>() + 2 * size_t(r.node_base), ctx->d_meshes.as<MeshRec>(),
                                  m, ctx->d_rootslab.as<RaySlab>() + m, st));
        ctx->slab_ok[m] = 1;
    }
    return CD_OK;
}

// ---- rays (CreateUncollideRays.cpp:60-200) -----------------------------------
// Slots of the ray generation's vertex hash set for nc contacts: more than
// the vertices offered (3 per contact), so that every probe ends.
uint64_t ray_hash_slots(uint32_t nc) {
    uint64_t H = 64;
    while (H < 6ull * nc) H <<= 1;
    return H;
}
// Block sums of the ray generation's scans per side (one per 1,024 contacts).
uint32_t ray_sums(uint32_t nc) { return std::max(1u, (nc + 1023) / 1024); }

// Buffers of the ray generation for nc contacts (sorted in W_CSORT) and the
// pairs' counts, begins and averages at stride spacing per side; the
// per-candidate buffers and hash sets once per side.  W_RAYS holds 6 nc
// rays: every contact starts at most one candidate per side and a candidate
// gives at most three rays.
int prepare_rays(cd_ctx* ctx, uint32_t nc, uint32_t stride) {
    HCK(ctx, ctx->d_work[W_CSORT].ensure(size_t(nc) * sizeof(cd_contact)));
    for (int w : {W_K0, W_V0, W_K1, W_V1}) HCK(ctx, ctx->d_work[w].ensure((size_t(nc) + 1) * sizeof(uint32_t)));
    for (int w : {W_RSTART, W_RPOS, W_RCNT, W_RPOS2})
        HCK(ctx, ctx->d_work[w].ensure(2 * (size_t(nc) + 1) * sizeof(uint32_t)));
    for (int w : {W_RCAND, W_RDENSE}) HCK(ctx, ctx->d_work[w].ensure(2 * size_t(nc) * sizeof(RayCand)));
    HCK(ctx, ctx->d_work[W_RSUMS].ensure(2 * ray_sums(nc) * sizeof(uint32_t)));
    HCK(ctx, ctx->d_work[W_RQ].ensure(size_t(stride) * 4 * sizeof(uint32_t)));
    HCK(ctx, ctx->d_work[W_RAVG].ensure(size_t(stride) * 6 * sizeof(float)));
    HCK(ctx, ctx->d_work[W_RTOT].ensure(64 * sizeof(uint32_t)));
    const uint64_t H = ray_hash_slots(nc);
    HCK(ctx, ctx->d_work[W_RHKEY].ensure(2 * H * sizeof(unsigned long long)));
    HCK(ctx, ctx->d_work[W_RHVAL].ensure(2 * H * sizeof(unsigned long long)));
    HCK(ctx, ctx->d_work[W_RAYS].ensure((6 * size_t(nc) + 1) * sizeof(cd_ray)));
    uint32_t* table = nullptr;
    HCK(ctx, sort_table(ctx, nc, &table));
    return CD_OK;
}

// The ray generation on the stream, with no wait: the second side's order of
// the sorted contacts (unless order_b: already in W_V0), then both sides'
// candidates, claimed vertices, rays (the second side's after the first's)
// and averages.  nc and nq are the counts, or with nc_dev / nq_dev the
// capacities the launches are shaped for; tot gets the totals DC_CAND0 ..
// DC_RAYS1.
int enqueue_rays(cd_ctx* ctx, uint32_t nc, uint32_t nq, uint32_t stride, const uint32_t* nc_dev,
                 const uint32_t* nq_dev, uint32_t* tot, bool order_b = false) {
    hipStream_t st = ctx->stream;
    const cd_contact* sorted = ctx->d_work[W_CSORT].as<cd_contact>();
    if (!order_b) {  // second side's order: (query, tri_b, tri_a)
        uint32_t bits[3];
        contact_field_bits(ctx, nq, bits);
        HCK(ctx, launch_sort_perm(sorted, nc, 0, 2, 1, bits, ctx->d_work[W_K0].as<uint32_t>(),
                                  ctx->d_work[W_V0].as<uint32_t>(), ctx->d_work[W_K1].as<uint32_t>(),
                                  ctx->d_work[W_V1].as<uint32_t>(), ctx->d_work[W_SORTT].as<uint32_t>(), st, nc_dev));
    }
    RaySides b{};
    b.sorted = sorted;
    b.order_b = ctx->d_work[W_V0].as<uint32_t>();
    b.n = nc;
    b.n_dev = nc_dev;
    b.nq = nq;
    b.nq_dev = nq_dev;
    b.qstride = stride;
    b.cand = ctx->d_work[W_RCAND].as<RayCand>();
    b.dense = ctx->d_work[W_RDENSE].as<RayCand>();
    b.start = ctx->d_work[W_RSTART].as<uint32_t>();
    b.pos = ctx->d_work[W_RPOS].as<uint32_t>();
    b.cnt = ctx->d_work[W_RCNT].as<uint32_t>();
    b.pos2 = ctx->d_work[W_RPOS2].as<uint32_t>();
    b.sums = ctx->d_work[W_RSUMS].as<uint32_t>();
    b.nsums = ray_sums(nc);
    b.hkey = ctx->d_work[W_RHKEY].as<unsigned long long>();
    b.hval = ctx->d_work[W_RHVAL].as<unsigned long long>();
    b.H = ray_hash_slots(nc);
    b.qw = ctx->d_work[W_RQ].as<uint32_t>();
    b.avg = ctx->d_work[W_RAVG].as<float>();
    b.tot = tot;
    b.rays = ctx->d_work[W_RAYS].as<cd_ray>();
    HCK(ctx, launch_rays(b, ctx->d_queries.as<QueryRec>(), ctx->d_meshes.as<MeshRec>(), ctx->d_leafof.as<uint32_t>(),
                         ctx->d_tris.as<float>(), ctx->d_attrs.as<uint2>(), ctx->d_vid.as<uint32_t>(),
                         ctx->d_nrm.as<float>(), ctx->d_perm.as<uint32_t>(), st));
    return CD_OK;
}

// cd_uncollide_rays; to_host false: the rays stay on the device for the
// shooting, and neither the pairs nor the rays are returned.
int rays_impl(cd_ctx* ctx, cd_pair_rays* pairs, uint32_t pair_cap, uint32_t* n_pairs, cd_ray* rays,
              uint64_t ray_cap, uint64_t* n_rays, bool to_host) {
    if (!ctx || !n_pairs || !n_rays) return CD_E_INVALID;
    *n_pairs = 0;
    *n_rays = 0;
    if (ctx->last_run == 0) return fail(ctx, CD_E_INVALID, "no run to take contacts from");
    (void)hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    const uint32_t nq = ctx->staged_nq;
    const uint64_t nc64 = ctx->staged_contacts;
    ctx->rays_ready = false;
    if (nc64 == 0 || nq == 0) {
        ctx->ray_total = 0;
        ctx->ray_stride = nq;
        ctx->ray_counts.assign(2 * size_t(nq), 0u);
        ctx->rays_ready = true;
        return CD_OK;
    }
    if (nc64 > 0x3FFFFFFFull) return fail(ctx, CD_E_CAPACITY, "too many contacts for ray generation");
    const uint32_t nc = uint32_t(nc64);
    int rc = prepare_rays(ctx, nc, nq);
    if (rc) return rc;
    if (!ctx->contacts_sorted) {
        uint32_t bits[3];
        contact_field_bits(ctx, nq, bits);
        HCK(ctx, launch_sort_contacts(ctx->d_contacts.as<cd_contact>(), nc, ctx->d_work[W_CSORT].as<cd_contact>(), bits,
                                      ctx->d_work[W_K0].as<uint32_t>(), ctx->d_work[W_V0].as<uint32_t>(),
                                      ctx->d_work[W_K1].as<uint32_t>(), ctx->d_work[W_V1].as<uint32_t>(),
                                      ctx->d_work[W_SORTT].as<uint32_t>(), st));
        ctx->contacts_sorted = true;
    }
    uint32_t* tot = ctx->d_work[W_RTOT].as<uint32_t>();
    rc = enqueue_rays(ctx, nc, nq, nq, nullptr, nullptr, tot);
    if (rc) return rc;
    // one wait, for the totals, the pairs' counts and begins and the averages
    uint32_t t4[4];
    std::vector<uint32_t> hq(size_t(nq) * 4);
    std::vector<float> ha(size_t(nq) * 6);
    HCK(ctx, hipMemcpyAsync(t4, tot, sizeof t4, hipMemcpyDeviceToHost, st));
    HCK(ctx, hipMemcpyAsync(hq.data(), ctx->d_work[W_RQ].p, hq.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HCK(ctx, hipMemcpyAsync(ha.data(), ctx->d_work[W_RAVG].p, ha.size() * sizeof(float), hipMemcpyDeviceToHost, st));
    HCK(ctx, hipStreamSynchronize(st));
    const uint32_t side_rays[2] = {t4[DC_RAYS0], t4[DC_RAYS1]};
    // pairs with rays (CollisionDetection.cpp:63); rays of the first sides, then of the second sides
    std::vector<cd_pair_rays> out;
    for (uint32_t q = 0; q < nq && to_host; ++q) {
        const uint32_t c0 = hq[q], c1 = hq[2 * size_t(nq) + q];
        if (c0 == 0 && c1 == 0) continue;
        cd_pair_rays r{};
        r.query = q;
        r.first = ctx->staged_ids[q].x;
        r.second = ctx->staged_ids[q].y;
        r.n_first = c0;
        r.n_second = c1;
        r.begin_first = hq[size_t(nq) + q];
        r.begin_second = uint64_t(side_rays[0]) + hq[3 * size_t(nq) + q];
        for (int k = 0; k < 3; ++k) {
            r.avg_first[k] = ha[3 * size_t(q) + k];
            r.avg_second[k] = ha[3 * (size_t(nq) + q) + k];
        }
        out.push_back(r);
    }
    const uint64_t total = uint64_t(side_rays[0]) + side_rays[1];
    ctx->ray_base1 = side_rays[0];
    ctx->ray_total = total;
    ctx->ray_stride = nq;
    ctx->ray_counts.resize(2 * size_t(nq));
    for (uint32_t q = 0; q < nq; ++q) {
        ctx->ray_counts[q] = hq[q];
        ctx->ray_counts[size_t(nq) + q] = hq[2 * size_t(nq) + q];
    }
    ctx->rays_ready = true;
    *n_pairs = uint32_t(out.size());
    *n_rays = total;
    if (!to_host) return CD_OK;
    if (out.size() > pair_cap || total > ray_cap || (total && (!pairs || !rays)))
        return fail(ctx, CD_E_CAPACITY, "pair or ray buffer too small");
    if (!out.empty()) std::memcpy(pairs, out.data(), out.size() * sizeof(cd_pair_rays));
    if (total) HCK(ctx, hipMemcpy(rays, ctx->d_work[W_RAYS].p, total * sizeof(cd_ray), hipMemcpyDeviceToHost));
    return CD_OK;
}

// The response's parameters: radians and cosines as the reference computes
// them (glm::radians, std::cos on float), defaults of
// CollisionDetection.cpp:70-72.
void shoot_params(const cd_shoot_params* params, float* cos_start, float* cos_finish, float* bias) {
    const float deg = 0.01745329251994329576923690768489f;
    const float a0 = params ? params->smoothstep_start_rads : 40.f * deg;
    const float a1 = params ? params->smoothstep_finish_rads : 65.f * deg;
    *bias = params ? params->distance_bias : 1.01f;
    *cos_start = std::cos(a0);
    *cos_finish = std::cos(a1);
}

// ---- the frame and the scene step on the device -----------------------------
// cd_step (and cd_frame) with every count kept on the device: the broad
// phase, the batched traversal and narrow phase, the contact order, and for
// the step the rays, the shots and the responses are enqueued at once (and
// replayed as one graph), shaped for capacities of pairs, contacts and shot
// rays.  The host reads one block of results after it and runs it again,
// with larger capacities, when a count exceeded one.
uint32_t pow2_at_least(uint64_t v) {
    uint64_t p = 64;
    while (p < v) p <<= 1;
    return uint32_t(std::min<uint64_t>(p, 1ull << 31));
}

struct StepArgs {
    uint32_t n;                // entities
    bool shoot;                // the whole step; else the frame (broad, narrow, contact order)
    uint32_t np, nc, rays;     // capacities: pairs, contacts, rays shot
    cd_tuning t;
    bool ref;
    bool timed;                // launched one by one: events around the traversal and the narrow phase
    bool radix;                // the contact order by the radix sorts (a pair beyond the per-pair order)
    uint32_t stack_entries;    // the ray traversal's stack: the entities' deepest tree + 2
    float cos_start, cos_finish, bias;
};

// Word offsets of the result block (W_STEPOUT on the device, h_step on the
// host): the run's counters, the shots' counters, the count block, per pair
// the rays of each side, the pairs, the responses (a frame's block: the
// counters, the count block and the pairs).
struct StepLayout {
    uint32_t fctr, sctr, dc, qc0, qc1, pairs, resp, words;
    StepLayout(uint32_t np, bool shoot) {
        const uint32_t rc = uint32_t(sizeof(RunCounters) / 4);
        const uint32_t q = shoot ? np : 0u;
        fctr = 0;
        sctr = rc;
        dc = 2 * rc;
        qc0 = dc + DC_WORDS;
        qc1 = qc0 + q;
        pairs = qc1 + q;
        resp = pairs + 2 * np;
        words = resp + q * uint32_t(sizeof(cd_pair_response) / 4);
    }
};
static_assert(sizeof(RunCounters) % 4 == 0 && sizeof(cd_pair_response) % 4 == 0, "results copied as words");

int ensure_step(cd_ctx* ctx, const StepArgs& a, uint64_t* fcap) {
    const uint32_t n = a.n, P = a.np, C = a.nc, R = a.rays;
    HCK(ctx, ctx->d_work[W_ENTS].ensure(size_t(n) * sizeof(cd_entity)));
    HCK(ctx, ctx->d_work[W_IV].ensure(size_t(n) * 6 * sizeof(float)));
    for (int w : {W_K0, W_V0, W_K1, W_V1}) HCK(ctx, ctx->d_work[w].ensure(size_t(std::max(n, P)) * sizeof(uint32_t)));
    HCK(ctx, ctx->d_work[W_PCOUNT].ensure(sizeof(unsigned int)));
    HCK(ctx, ctx->d_work[W_PAIRS].ensure(size_t(P) * sizeof(uint2)));
    HCK(ctx, ctx->d_work[W_SPAIRS].ensure(size_t(P) * sizeof(uint2)));
    HCK(ctx, ctx->d_queries.ensure(size_t(P) * sizeof(QueryRec)));
    if (ctx->contact_cap < C) ctx->contact_cap = C;
    int rc = ensure_batched(ctx, P, P, a.t, fcap);
    if (rc) return rc;
    if (a.shoot) {
        rc = prepare_rays(ctx, C, P);
        if (rc) return rc;
    } else {  // the contact order's buffers (the radix sorts' too)
        HCK(ctx, ctx->d_work[W_CSORT].ensure(size_t(C) * sizeof(cd_contact)));
        for (int w : {W_K0, W_V0, W_K1, W_V1}) HCK(ctx, ctx->d_work[w].ensure((size_t(C) + 1) * sizeof(uint32_t)));
    }
    uint32_t* table = nullptr;
    HCK(ctx, sort_table(ctx, std::max(std::max(n, P), C), &table));
    HCK(ctx, ctx->d_work[W_SCTR].ensure(sizeof(RunCounters)));
    HCK(ctx, ctx->d_work[W_STEPC].ensure(DC_WORDS * sizeof(uint32_t)));
    HCK(ctx, ctx->d_work[W_SEGQ].ensure(3 * size_t(P) * sizeof(uint32_t)));
    HCK(ctx, ctx->d_work[W_CSEG].ensure(size_t(C) * sizeof(cd_contact)));
    if (a.shoot) {
        HCK(ctx, ctx->d_work[W_PIDS].ensure(2 * size_t(P) * sizeof(uint32_t)));
        HCK(ctx, ctx->d_work[W_SHOTS].ensure(size_t(R) * sizeof(RayShot)));
        HCK(ctx, ctx->d_work[W_RESP].ensure(size_t(P) * sizeof(cd_pair_response)));
        if (a.stack_entries > uint32_t(RAY_STACK_ENTRIES)) {  // trees deeper than the LDS stack holds
            const size_t blocks = (2 * size_t(R) + 63) / 64;
            HCK(ctx, ctx->d_work[W_GSTACK].ensure(blocks * 2 * a.stack_entries * 64 * sizeof(uint32_t)));
        }
    }
    const StepLayout L(P, a.shoot);
    HCK(ctx, ctx->d_work[W_STEPOUT].ensure(size_t(L.words) * sizeof(uint32_t)));
    HCK(ctx, ctx->h_step.ensure(size_t(L.words) * sizeof(uint32_t)));
    HCK(ctx, ctx->h_ents.ensure(size_t(n) * sizeof(cd_entity)));
    if (!a.shoot) HCK(ctx, ctx->h_stage.ensure(size_t(C) * sizeof(cd_contact)));
    return CD_OK;
}

// The whole step (or the frame) on the stream, no host wait inside (the
// entities already in W_ENTS); its last command copies the result block to
// h_step.
int enqueue_step(cd_ctx* ctx, const StepArgs& a) {
    hipStream_t st = ctx->stream;
    const uint32_t n = a.n, P = a.np, C = a.nc, R = a.rays;
    const cd_entity* de = ctx->d_work[W_ENTS].as<cd_entity>();
    uint32_t* dc = ctx->d_work[W_STEPC].as<uint32_t>();
    uint32_t* k0 = ctx->d_work[W_K0].as<uint32_t>();
    uint32_t* v0 = ctx->d_work[W_V0].as<uint32_t>();
    uint32_t* k1 = ctx->d_work[W_K1].as<uint32_t>();
    uint32_t* v1 = ctx->d_work[W_V1].as<uint32_t>();
    uint32_t* table = ctx->d_work[W_SORTT].as<uint32_t>();
    unsigned int* cnt = ctx->d_work[W_PCOUNT].as<unsigned int>();
    uint2* pairs = ctx->d_work[W_PAIRS].as<uint2>();
    uint2* spairs = ctx->d_work[W_SPAIRS].as<uint2>();
    QueryRec* dq = ctx->d_queries.as<QueryRec>();
    const MeshRec* dm = ctx->d_meshes.as<MeshRec>();
    // the entity records, from pinned staging the host fills before each launch
    HCK(ctx, hipMemcpyAsync(ctx->d_work[W_ENTS].p, ctx->h_ents.p, size_t(n) * sizeof(cd_entity), hipMemcpyHostToDevice,
                            st));
    HCK(ctx, launch_fill_u32(dc, DC_WORDS, 0u, st));
    // broad phase: sweep and prune, the pairs in (first, second) order, their relative matrices
    HCK(ctx, launch_sap_project(de, n, dm, sap_axes(), ctx->d_work[W_IV].as<float>(), k0, v0, st));
    HCK(ctx, launch_radix_sort(k0, v0, k1, v1, n, 32, table, st));
    HCK(ctx, launch_fill_u32(cnt, 1, 0u, st));
    HCK(ctx, launch_sap_sweep(v0, ctx->d_work[W_IV].as<float>(), de, n, pairs, P, cnt, st));
    HCK(ctx, launch_pair_keys(pairs, P, k0, v0, st, cnt));
    HCK(ctx, launch_radix_sort(k0, v0, k1, v1, P, key_bits_below(uint64_t(n) << 16), table, st, cnt));
    HCK(ctx, launch_pair_queries(pairs, v0, P, de, dq, spairs, st, cnt));
    HCK(ctx, launch_step_pairs(cnt, P, dc, ctx->d_nq.as<unsigned int>(), st));
    // the pairs' contacts, ordered by (pair, tri_a, tri_b) and by (pair, tri_b, tri_a)
    int rc = enqueue_batched(ctx, P, a.t, a.timed, a.ref);
    if (rc) return rc;
    const uint64_t ccap = std::min<uint64_t>(C, ctx->contact_cap);
    if (a.radix) {
        HCK(ctx, launch_step_contacts(dev_counters(ctx), ccap, dc, st));
        uint32_t bits[3];
        contact_field_bits(ctx, P, bits);
        HCK(ctx, launch_sort_contacts(ctx->d_contacts.as<cd_contact>(), C, ctx->d_work[W_CSORT].as<cd_contact>(), bits,
                                      k0, v0, k1, v1, table, st, &dc[DC_NC]));
    } else {
        uint32_t* qw3 = ctx->d_work[W_SEGQ].as<uint32_t>();
        HCK(ctx, launch_step_contacts(dev_counters(ctx), ccap, dc, st, qw3, P));
        HCK(ctx, launch_seg_order(ctx->d_contacts.as<cd_contact>(), C, P, dc, qw3, ctx->d_work[W_CSEG].as<cd_contact>(),
                                  ctx->d_work[W_CSORT].as<cd_contact>(), v0, st));
    }
    const StepLayout L(P, a.shoot);
    PackSpans sp{};
    auto span = [&](const void* src, uint32_t words, uint32_t at) {
        sp.src[sp.n] = static_cast<const uint32_t*>(src);
        sp.words[sp.n] = words;
        sp.dst[sp.n] = at;
        ++sp.n;
    };
    if (!a.shoot) {  // the frame: its counters, counts and pairs, and the sorted contacts (the capacity)
        span(ctx->d_counters.p, L.sctr - L.fctr, L.fctr);
        span(dc, DC_WORDS, L.dc);
        span(spairs, 2 * P, L.pairs);
        HCK(ctx, launch_pack(sp, ctx->d_work[W_STEPOUT].as<uint32_t>(), st));
        HCK(ctx, hipMemcpyAsync(ctx->h_step.p, ctx->d_work[W_STEPOUT].p, size_t(L.words) * sizeof(uint32_t),
                                hipMemcpyDeviceToHost, st));
        HCK(ctx, hipMemcpyAsync(ctx->h_stage.p, ctx->d_work[W_CSORT].p, size_t(C) * sizeof(cd_contact),
                                hipMemcpyDeviceToHost, st));
        return CD_OK;
    }
    // rays, shots, responses
    rc = enqueue_rays(ctx, C, P, P, &dc[DC_NC], &dc[DC_NQ], dc, !a.radix);
    if (rc) return rc;
    uint32_t* qw = ctx->d_work[W_RQ].as<uint32_t>();
    uint32_t* pbegin = ctx->d_work[W_PIDS].as<uint32_t>();
    RunCounters* sctr = ctx->d_work[W_SCTR].as<RunCounters>();
    uint32_t* gstack = a.stack_entries > uint32_t(RAY_STACK_ENTRIES) ? ctx->d_work[W_GSTACK].as<uint32_t>() : nullptr;
    HCK(ctx, launch_pair_begins(qw, P, dc, pbegin, st));
    HCK(ctx, launch_zero(sctr, st));
    HCK(ctx, launch_shoot(ctx->d_work[W_RAYS].as<cd_ray>(), R, 0, pbegin, P, dq, dm, ctx->d_nodes.as<NodeRec>(),
                          ctx->d_rayslab.as<RaySlab>(), ctx->d_rootslab.as<RaySlab>(), ctx->d_tris.as<float>(), ctx->d_attrs.as<uint2>(), ctx->d_vid.as<uint32_t>(),
                          ctx->d_nrm.as<float>(), ctx->d_perm.as<uint32_t>(), ctx->d_work[W_SHOTS].as<RayShot>(),
                          a.stack_entries, sctr, st, gstack, P, dc));
    HCK(ctx, launch_respond(ctx->d_work[W_SHOTS].as<RayShot>(), 0, qw, P, ctx->d_work[W_RAVG].as<float>(), dq, spairs, de,
                            a.cos_start, a.cos_finish, a.bias, ctx->d_work[W_RESP].as<cd_pair_response>(), st, P, dc,
                            R));
    // the results in one block, one copy to the host
    span(ctx->d_counters.p, L.sctr - L.fctr, L.fctr);
    span(sctr, L.dc - L.sctr, L.sctr);
    span(dc, DC_WORDS, L.dc);
    span(qw, P, L.qc0);
    span(qw + 2 * size_t(P), P, L.qc1);
    span(spairs, 2 * P, L.pairs);
    span(ctx->d_work[W_RESP].p, L.words - L.resp, L.resp);
    HCK(ctx, launch_pack(sp, ctx->d_work[W_STEPOUT].as<uint32_t>(), st));
    HCK(ctx, hipMemcpyAsync(ctx->h_step.p, ctx->d_work[W_STEPOUT].p, size_t(L.words) * sizeof(uint32_t),
                            hipMemcpyDeviceToHost, st));
    return CD_OK;
}

// What a captured step depends on: sizes, parameters and every buffer.
void step_key(const cd_ctx* ctx, const StepArgs& a, uint64_t fcap, std::vector<uint64_t>& key) {
    uint32_t f[3];
    std::memcpy(&f[0], &a.cos_start, 4);
    std::memcpy(&f[1], &a.cos_finish, 4);
    std::memcpy(&f[2], &a.bias, 4);
    const uint64_t k[] = {a.n, a.np, a.nc, a.rays, a.stack_entries, f[0], f[1], f[2], ctx->meshes.size(),
                          uint64_t(a.radix), uint64_t(a.shoot)};
    key.assign(std::begin(k), std::end(k));
    batched_key(ctx, a.np, a.t, a.ref, fcap, key);
    for (const DevBuf& b : ctx->d_work) key.push_back(reinterpret_cast<uint64_t>(b.p));
    for (const DevBuf* b : {&ctx->d_leafof, &ctx->d_attrs, &ctx->d_vid, &ctx->d_nrm, &ctx->d_perm})
        key.push_back(reinterpret_cast<uint64_t>(b->p));
    for (const HostBuf* b : {&ctx->h_step, &ctx->h_ents, &ctx->h_stage}) key.push_back(reinterpret_cast<uint64_t>(b->p));
    for (const DevBuf* b : {&ctx->d_rayslab, &ctx->d_rootslab}) key.push_back(reinterpret_cast<uint64_t>(b->p));
}

// The frame (shoot false) or the whole step on the device, run until its
// counts fit the capacities.  On success the result block is in h_step (its
// layout *L) and the context holds what the separate calls leave: the pairs
// and sorted contacts of the frame, and after a step its rays.
int device_run(cd_ctx* ctx, const cd_entity* ents, uint32_t n, bool ref, bool shoot, const cd_shoot_params* params,
               uint32_t stack_entries, StepLayout* L) {
    hipStream_t st = ctx->stream;
    StepArgs a{};
    a.n = n;
    a.shoot = shoot;
    a.ref = ref;
    a.stack_entries = stack_entries;
    shoot_params(params, &a.cos_start, &a.cos_finish, &a.bias);
    const bool graph = ctx->tuning.step_graph != 2;
    a.timed = !graph;
    hipGraphExec_t* exec = shoot ? &ctx->step_graph : &ctx->frame_graph;
    std::vector<uint64_t>& exec_key = shoot ? ctx->step_key : ctx->frame_key;
    // first capacities; a run that outgrows one runs again with it larger,
    // and later runs start from there
    if (!ctx->step_np_cap) ctx->step_np_cap = pow2_at_least(n);
    if (!ctx->step_nc_cap) ctx->step_nc_cap = 4096;
    if (!ctx->step_ray_cap) ctx->step_ray_cap = 4096;
    ctx->last_run = 0;
    ctx->rays_ready = false;
    ctx->contacts_sorted = false;
    ctx->frame_mesh.resize(n);
    for (uint32_t i = 0; i < n; ++i) ctx->frame_mesh[i] = ents[i].mesh_id;
    if (shoot) {
        int rc = ensure_slabs(ctx, ctx->frame_mesh.data(), n);
        if (rc) return rc;
    }
    for (int attempt = 0; attempt < 8; ++attempt) {
        a.np = ctx->step_np_cap;
        a.nc = ctx->step_nc_cap;
        a.rays = ctx->step_ray_cap;
        a.radix = ctx->step_radix;
        a.t = effective_tuning(ctx, a.np);
        uint64_t fcap = 0;
        int rc = ensure_step(ctx, a, &fcap);
        if (rc) return rc;
        std::memcpy(ctx->h_ents.p, ents, size_t(n) * sizeof(cd_entity));
        HCK(ctx, hipEventRecord(ctx->ev[4], st));
        if (graph) {
            std::vector<uint64_t> key;
            step_key(ctx, a, fcap, key);
            if (!(*exec && key == exec_key)) {
                exec_key.clear();
                rc = capture_graph(ctx, exec, [&] { return enqueue_step(ctx, a); });
                if (rc) return rc;
                exec_key = key;
            }
            HCK(ctx, hipGraphLaunch(*exec, st));
        } else {
            rc = enqueue_step(ctx, a);
            if (rc) return rc;
        }
        HCK(ctx, hipEventRecord(ctx->ev[5], st));
        HCK(ctx, hipEventSynchronize(ctx->ev[5]));
        *L = StepLayout(a.np, shoot);
        const uint32_t* w = static_cast<const uint32_t*>(ctx->h_step.p);
        RunCounters fh{}, sh{};
        std::memcpy(&fh, w + L->fctr, sizeof fh);
        if (shoot) std::memcpy(&sh, w + L->sctr, sizeof sh);
        const uint32_t* dcw = w + L->dc;
        bool retry = false;
        rc = check_batched(ctx, fh, fcap, a.t, &retry);
        if (rc) return rc;
        if (dcw[DC_SWEEP] > a.np) {
            ctx->step_np_cap = pow2_at_least(uint64_t(dcw[DC_SWEEP]) + dcw[DC_SWEEP] / 4);
            retry = true;
        }
        if (fh.contacts > a.nc) {
            ctx->step_nc_cap = pow2_at_least(fh.contacts + fh.contacts / 4);
            retry = true;
        }
        if (dcw[DC_BIGSEG]) {
            ctx->step_radix = true;
            retry = true;
        }
        const uint64_t nrays = shoot ? uint64_t(dcw[DC_RAYS0]) + dcw[DC_RAYS1] : 0;
        if (nrays > a.rays) {
            ctx->step_ray_cap = pow2_at_least(nrays + nrays / 4);
            retry = true;
        }
        if (retry) continue;
        if (sh.err) return fail(ctx, CD_E_STACK, "ray traversal stack overflow");
        // what the separate calls would leave: the pairs, sorted contacts (and rays) of the frame
        const uint32_t nq = dcw[DC_NQ];
        const uint2* hp = reinterpret_cast<const uint2*>(w + L->pairs);
        ctx->staged_nq = nq;
        ctx->staged_ids.assign(hp, hp + nq);
        ctx->staged_contacts = fh.contacts;
        ctx->contacts_sorted = true;
        ctx->last_run = 2;
        if (shoot) {
            ctx->ray_base1 = dcw[DC_RAYS0];
            ctx->ray_total = nrays;
            ctx->ray_stride = a.np;
            ctx->ray_counts.resize(2 * size_t(nq));
            for (uint32_t q = 0; q < nq; ++q) {
                ctx->ray_counts[q] = w[L->qc0 + q];
                ctx->ray_counts[size_t(nq) + q] = w[L->qc1 + q];
            }
            ctx->rays_ready = true;
        }
        set_run_counters(ctx, fh);
        ctx->counters.broad_pairs = nq;
        ctx->timings = cd_timings{};
        HCK(ctx, hipEventElapsedTime(&ctx->timings.total_ms, ctx->ev[4], ctx->ev[5]));
        if (a.timed) {  // the batched run's kernels, as run_batched reports them
            float t_split = 0.f, t_dfs = 0.f, t_narrow = 0.f;
            HCK(ctx, hipEventElapsedTime(&t_split, ctx->ev[0], ctx->ev[1]));
            HCK(ctx, hipEventElapsedTime(&t_dfs, ctx->ev[1], ctx->ev[2]));
            HCK(ctx, hipEventElapsedTime(&t_narrow, ctx->ev[2], ctx->ev[3]));
            ctx->timings.traverse_ms = t_split + t_dfs;
            ctx->timings.narrow_ms = t_narrow;
        }
        ctx->timings.launches = 1;
        return CD_OK;
    }
    return fail(ctx, CD_E_NOMEM, "step sizing did not converge");
}

int step_device(cd_ctx* ctx, const cd_entity* ents, uint32_t n, bool ref, const cd_shoot_params* params,
                uint32_t stack_entries, cd_pair_response* out, uint32_t cap, uint32_t* n_out) {
    StepLayout L(0, true);
    int rc = device_run(ctx, ents, n, ref, true, params, stack_entries, &L);
    if (rc) return rc;
    const uint32_t* w = static_cast<const uint32_t*>(ctx->h_step.p);
    const uint32_t nq = ctx->staged_nq;
    uint32_t npairs = 0;
    for (uint32_t q = 0; q < nq; ++q) npairs += (w[L.qc0 + q] + w[L.qc1 + q]) ? 1u : 0u;
    *n_out = npairs;
    if (npairs == 0) return CD_OK;
    if (npairs > cap || !out) return fail(ctx, CD_E_CAPACITY, "response buffer too small");
    const cd_pair_response* resp = reinterpret_cast<const cd_pair_response*>(w + L.resp);
    uint32_t k = 0;
    for (uint32_t q = 0; q < nq; ++q)
        if (w[L.qc0 + q] + w[L.qc1 + q]) out[k++] = resp[q];
    return CD_OK;
}

// cd_frame on the device: one run, whose last copy brings the sorted contacts
// back too, then grouped by pair on the host.
int frame_device(cd_ctx* ctx, const cd_entity* ents, uint32_t n, bool ref, cd_pair* pairs, uint32_t pair_cap,
                 uint32_t* n_pairs, cd_contact* contacts, uint64_t contact_cap, uint64_t* n_contacts) {
    StepLayout L(0, false);
    int rc = device_run(ctx, ents, n, ref, false, nullptr, 2, &L);
    if (rc) return rc;
    // the run copied its sorted contacts (up to the capacity, which they fit) to h_stage
    const uint64_t nc = ctx->staged_contacts;
    const cd_contact* hc = static_cast<const cd_contact*>(ctx->h_stage.p);
    return frame_output(ctx, hc, nc, pairs, pair_cap, n_pairs, contacts, contact_cap, n_contacts);
}
}  // namespace

extern "C" {

int cd_frame(cd_ctx* ctx, const cd_entity* entities, uint32_t n, uint32_t mode, cd_pair* pairs, uint32_t pair_cap,
             uint32_t* n_pairs, cd_contact* contacts, uint64_t contact_cap, uint64_t* n_contacts) {
    if (ctx && n_pairs && n_contacts && (mode == CD_MODE_BATCHED || mode == CD_MODE_REFERENCE) && n >= 2) {
        (void)hipSetDevice(ctx->device);
        *n_pairs = 0;
        *n_contacts = 0;
        int rc = check_entities(ctx, entities, n);
        if (rc) return rc;
        return frame_device(ctx, entities, n, mode == CD_MODE_REFERENCE, pairs, pair_cap, n_pairs, contacts,
                            contact_cap, n_contacts);
    }
    return frame_impl(ctx, entities, n, mode, pairs, pair_cap, n_pairs, contacts, contact_cap, n_contacts, true);
}

int cd_uncollide_rays(cd_ctx* ctx, cd_pair_rays* pairs, uint32_t pair_cap, uint32_t* n_pairs, cd_ray* rays,
                      uint64_t ray_cap, uint64_t* n_rays) {
    return rays_impl(ctx, pairs, pair_cap, n_pairs, rays, ray_cap, n_rays, true);
}

int cd_step(cd_ctx* ctx, const cd_entity* entities, uint32_t n, uint32_t mode, const cd_shoot_params* params,
            cd_pair_response* out, uint32_t cap, uint32_t* n_out) {
    if (!ctx || !n_out) return CD_E_INVALID;
    *n_out = 0;
    if ((mode == CD_MODE_BATCHED || mode == CD_MODE_REFERENCE) && n >= 2) {
        (void)hipSetDevice(ctx->device);
        int rc = check_entities(ctx, entities, n);
        if (rc) return rc;
        // the ray traversal's stack: the deepest tree an entity has + 2; the
        // separate calls take a mesh beyond the ray traversal's limits
        uint32_t stack_entries = 2;
        bool fits = true;
        for (uint32_t i = 0; i < n; ++i) {
            const MeshRec& m = ctx->meshes[entities[i].mesh_id];
            stack_entries = std::max(stack_entries, m.depth + 2);
            fits = fits && m.ntri < (1u << RAY_LINK_BITS);
        }
        if (fits && ctx->leaf_max <= 64)
            return step_device(ctx, entities, n, mode == CD_MODE_REFERENCE, params, stack_entries, out, cap, n_out);
    }
    uint32_t np = 0;
    uint64_t nc = 0, nr = 0;
    int rc = frame_impl(ctx, entities, n, mode, nullptr, 0, &np, nullptr, 0, &nc, false);
    if (rc) return rc;
    if (n < 2 || ctx->staged_nq == 0) return CD_OK;  // no pairs: nothing to shoot
    rc = rays_impl(ctx, nullptr, 0, &np, nullptr, 0, &nr, false);
    if (rc) return rc;
    return cd_shoot_rays(ctx, params, out, cap, n_out);
}

// ---- response (ShootUncollideRays.cpp:14-179, CollisionDetection.cpp:70-99) ----
int cd_shoot_rays(cd_ctx* ctx, const cd_shoot_params* params, cd_pair_response* out, uint32_t cap, uint32_t* n_out) {
    if (!ctx || !n_out) return CD_E_INVALID;
    *n_out = 0;
    if (ctx->last_run != 2) return fail(ctx, CD_E_INVALID, "ray shooting needs the entity matrices of a cd_frame");
    if (!ctx->rays_ready) return fail(ctx, CD_E_INVALID, "no rays: call cd_uncollide_rays after cd_frame");
    (void)hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    const uint32_t nq = ctx->staged_nq;
    uint32_t npairs = 0;
    for (uint32_t q = 0; q < nq; ++q) npairs += (ctx->ray_counts[q] + ctx->ray_counts[size_t(nq) + q]) ? 1u : 0u;
    *n_out = npairs;
    if (npairs == 0) return CD_OK;
    if (npairs > cap || !out) return fail(ctx, CD_E_CAPACITY, "response buffer too small");
    uint32_t stack_entries = 2;  // the closest-hit stack holds a tree's depth + 2 entries
    for (uint32_t q = 0; q < nq; ++q)
        for (uint32_t e : {ctx->staged_ids[q].x, ctx->staged_ids[q].y}) {
            const MeshRec& m = ctx->meshes[ctx->frame_mesh[e]];
            stack_entries = std::max(stack_entries, m.depth + 2);
            if (m.ntri >= (1u << RAY_LINK_BITS) || ctx->leaf_max > 64)
                return fail(ctx, CD_E_CAPACITY, "mesh too large for the ray traversal stack (2^25 triangles)");
        }
    float cos_start = 0.f, cos_finish = 0.f, bias = 0.f;
    shoot_params(params, &cos_start, &cos_finish, &bias);
    HCK(ctx, ctx->d_work[W_SHOTS].ensure(std::max<uint64_t>(ctx->ray_total, 1) * sizeof(RayShot)));
    HCK(ctx, ctx->d_work[W_RESP].ensure(size_t(nq) * sizeof(cd_pair_response)));
    // pair ids, then per side the first ray of every pair (an exclusive scan
    // of the counts: the device keeps a begin only for pairs with rays)
    std::vector<uint32_t> up(4 * size_t(nq));
    for (uint32_t q = 0; q < nq; ++q) {
        up[2 * size_t(q)] = ctx->staged_ids[q].x;
        up[2 * size_t(q) + 1] = ctx->staged_ids[q].y;
    }
    for (int side = 0; side < 2; ++side) {
        uint32_t run = 0;
        for (uint32_t q = 0; q < nq; ++q) {
            up[2 * size_t(nq) + size_t(nq) * side + q] = run;
            run += ctx->ray_counts[size_t(nq) * side + q];
        }
    }
    HCK(ctx, ctx->d_work[W_PIDS].ensure(up.size() * sizeof(uint32_t)));
    HCK(ctx, hipMemcpyAsync(ctx->d_work[W_PIDS].p, up.data(), up.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    const uint32_t* qw = ctx->d_work[W_RQ].as<uint32_t>();
    const uint32_t* pbegin = ctx->d_work[W_PIDS].as<uint32_t>() + 2 * size_t(nq);
    uint32_t* gstack = nullptr;  // trees deeper than the LDS stack holds: the stacks in HBM
    if (stack_entries > uint32_t(RAY_STACK_ENTRIES)) {
        const size_t blocks = (2 * ctx->ray_total + 63) / 64;
        HCK(ctx, ctx->d_work[W_GSTACK].ensure(blocks * 2 * stack_entries * 64 * sizeof(uint32_t)));
        gstack = ctx->d_work[W_GSTACK].as<uint32_t>();
    }
    {
        std::vector<uint32_t> used;
        used.reserve(2 * size_t(nq));
        for (uint32_t q = 0; q < nq; ++q)
            for (uint32_t e : {ctx->staged_ids[q].x, ctx->staged_ids[q].y}) used.push_back(ctx->frame_mesh[e]);
        int rc = ensure_slabs(ctx, used.data(), uint32_t(used.size()));
        if (rc) return rc;
    }
    HCK(ctx, launch_zero(dev_counters(ctx), st));
    HCK(ctx, hipEventRecord(ctx->ev[0], st));
    HCK(ctx, launch_shoot(ctx->d_work[W_RAYS].as<cd_ray>(), uint32_t(ctx->ray_total), ctx->ray_base1, pbegin, nq,
                          ctx->d_queries.as<QueryRec>(), ctx->d_meshes.as<MeshRec>(), ctx->d_nodes.as<NodeRec>(),
                          ctx->d_rayslab.as<RaySlab>(), ctx->d_rootslab.as<RaySlab>(), ctx->d_tris.as<float>(), ctx->d_attrs.as<uint2>(), ctx->d_vid.as<uint32_t>(),
                          ctx->d_nrm.as<float>(), ctx->d_perm.as<uint32_t>(), ctx->d_work[W_SHOTS].as<RayShot>(),
                          stack_entries, dev_counters(ctx), st, gstack));
    HCK(ctx, hipEventRecord(ctx->ev[1], st));
    HCK(ctx, launch_respond(ctx->d_work[W_SHOTS].as<RayShot>(), ctx->ray_base1, qw, nq, ctx->d_work[W_RAVG].as<float>(),
                            ctx->d_queries.as<QueryRec>(), ctx->d_work[W_PIDS].as<uint2>(),
                            ctx->d_work[W_ENTS].as<cd_entity>(), cos_start, cos_finish, bias,
                            ctx->d_work[W_RESP].as<cd_pair_response>(), st, ctx->ray_stride));
    std::vector<cd_pair_response> all(nq);
    HCK(ctx, hipMemcpyAsync(all.data(), ctx->d_work[W_RESP].p, size_t(nq) * sizeof(cd_pair_response),
                            hipMemcpyDeviceToHost, st));
    RunCounters h{};
    HCK(ctx, hipEventRecord(ctx->ev[2], st));
    int rc = read_counters(ctx, h);  // synchronises the stream
    if (rc) return rc;
    if (h.err) return fail(ctx, CD_E_STACK, "ray traversal stack overflow");
    // the shot's work and time: boxes tested (sat_tests), ray-triangle tests
    // (tri_tests), the shooting kernel (traverse_ms) and with the response
    // (total_ms)
    float t_shoot = 0.f, t_all = 0.f;
    HCK(ctx, hipEventElapsedTime(&t_shoot, ctx->ev[0], ctx->ev[1]));
    HCK(ctx, hipEventElapsedTime(&t_all, ctx->ev[0], ctx->ev[2]));
    ctx->counters = cd_counters{};
    ctx->counters.sat_tests = h.sat;
    ctx->counters.tri_tests = h.tritests;
    ctx->timings = cd_timings{};
    ctx->timings.traverse_ms = t_shoot;
    ctx->timings.total_ms = t_all;
    ctx->timings.launches = 2;
    uint32_t k = 0;
    for (uint32_t q = 0; q < nq; ++q)
        if (ctx->ray_counts[q] + ctx->ray_counts[size_t(nq) + q]) out[k++] = all[q];
    return CD_OK;
}

int cd_broad_phase(cd_ctx* ctx, const cd_entity* entities, uint32_t n, uint32_t* pairs2, uint32_t cap, uint32_t* n_pairs) {
    if (!ctx || !n_pairs) return CD_E_INVALID;
    (void)hipSetDevice(ctx->device);
    *n_pairs = 0;
    if (n < 2) return CD_OK;
    // the broad phase overwrites the staged queries (relative matrices) and
    // the entity copy the last run's rays and response read: that run is gone
    ctx->last_run = 0;
    ctx->rays_ready = false;
    ctx->contacts_sorted = false;
    ctx->staged_nq = 0;
    ctx->staged_contacts = 0;
    ctx->staged_ids.clear();
    uint32_t np = 0;
    std::vector<uint2> hp;
    int rc = broad_phase(ctx, entities, n, &np, &hp);
    if (rc) return rc;
    *n_pairs = np;
    if (np > cap || (np && !pairs2)) return fail(ctx, CD_E_CAPACITY, "pair buffer too small");
    for (uint32_t i = 0; i < np; ++i) {
        pairs2[2 * i] = hp[i].x;
        pairs2[2 * i + 1] = hp[i].y;
    }
    return CD_OK;
}

int cd_export_triangles(cd_ctx* ctx, uint32_t mesh, float* tri9) {
    if (!ctx) return CD_E_INVALID;
    if (mesh >= ctx->meshes.size()) return fail(ctx, CD_E_INVALID, "unknown mesh");
    const MeshRec& m = ctx->meshes[mesh];
    if (m.ntri && !tri9) return fail(ctx, CD_E_INVALID, "null output");
    if (m.ntri)
        HCK(ctx, hipMemcpyAsync(tri9, ctx->d_tris.as<float>() + size_t(m.tri_base) * 9, size_t(m.ntri) * 9 * sizeof(float),
                                hipMemcpyDeviceToHost, ctx->stream));
    HCK(ctx, hipStreamSynchronize(ctx->stream));
    return CD_OK;
}

int cd_register_skinned(cd_ctx* ctx, const float* bind_positions, uint32_t nvert, const uint32_t* indices,
                        uint32_t ntri, const uint16_t* joints4, const float* weights4, uint32_t nbones,
                        uint32_t* mesh_id) {
    if (!ctx || !mesh_id) return CD_E_INVALID;
    if (nbones == 0 || nbones > 0xFFFFu) return fail(ctx, CD_E_INVALID, "bone count out of range");
    if (nvert && (!bind_positions || !joints4 || !weights4)) return fail(ctx, CD_E_INVALID, "null skin arrays");
    for (size_t i = 0; i < size_t(nvert) * 4; ++i)
        if (joints4[i] >= nbones) return fail(ctx, CD_E_INVALID, "joint index out of range");
    if (!all_finite(weights4, size_t(nvert) * 4)) return fail(ctx, CD_E_INVALID, "non-finite weight");
    uint32_t id = 0;
    int rc = cd_register_mesh(ctx, bind_positions, nvert, indices, ntri, &id);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    const size_t vbase = ctx->skin_verts, vidx_base = ctx->skin_vidx;
    if (vbase + nvert >= 0xFFFFFFF0ull || vidx_base + size_t(ntri) * 3 >= 0xFFFFFFF0ull)
        return fail(ctx, CD_E_INVALID, "skinned vertex arena exceeds 32-bit indexing");
    HCK(ctx, grow_keep(ctx->d_bind, (vbase + nvert + 1) * 3 * sizeof(float), st));
    HCK(ctx, grow_keep(ctx->d_joints, (vbase + nvert + 1) * sizeof(ushort4), st));
    HCK(ctx, grow_keep(ctx->d_weights, (vbase + nvert + 1) * sizeof(float4), st));
    HCK(ctx, grow_keep(ctx->d_vidx, (vidx_base + size_t(ntri) * 3 + 1) * sizeof(uint32_t), st));
    HCK(ctx, grow_keep(ctx->d_skinpos, (vbase + nvert + 1) * 3 * sizeof(float), st));
    if (nvert) {
        HCK(ctx, hipMemcpyAsync(ctx->d_bind.as<float>() + vbase * 3, bind_positions, size_t(nvert) * 3 * sizeof(float),
                                hipMemcpyHostToDevice, st));
        HCK(ctx, hipMemcpyAsync(ctx->d_skinpos.as<float>() + vbase * 3, bind_positions, size_t(nvert) * 3 * sizeof(float),
                                hipMemcpyHostToDevice, st));
        HCK(ctx, hipMemcpyAsync(ctx->d_joints.as<ushort4>() + vbase, joints4, size_t(nvert) * sizeof(ushort4),
                                hipMemcpyHostToDevice, st));
        HCK(ctx, hipMemcpyAsync(ctx->d_weights.as<float4>() + vbase, weights4, size_t(nvert) * sizeof(float4),
                                hipMemcpyHostToDevice, st));
    }
    if (ntri) {
        DevBuf idx;
        HCK(ctx, idx.ensure(size_t(ntri) * 3 * sizeof(uint32_t)));
        HCK(ctx, hipMemcpyAsync(idx.p, indices, size_t(ntri) * 3 * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        const MeshRec& m = ctx->meshes[id];
        HCK(ctx, launch_tree_vidx(ctx->d_perm.as<uint32_t>() + m.tri_base, ntri, idx.as<uint32_t>(), uint32_t(vbase),
                                  ctx->d_vidx.as<uint32_t>() + vidx_base, st));
        HCK(ctx, hipStreamSynchronize(st));
        idx.release();
    }
    ctx->skins.push_back({id, cd_ctx::SkinHost{uint32_t(vbase), nvert, nbones, uint32_t(vidx_base)}});
    ctx->skin_verts += nvert;
    ctx->skin_vidx += size_t(ntri) * 3;
    *mesh_id = id;
    return CD_OK;
}

int cd_skin_refit(cd_ctx* ctx, const uint32_t* mesh_ids, uint32_t n, const float* bone_matrices) {
    if (!ctx) return CD_E_INVALID;
    if (n == 0) return CD_OK;
    if (!mesh_ids || !bone_matrices) return fail(ctx, CD_E_INVALID, "null argument");
    for (uint32_t i = 0; i < n; ++i)  // their boxes change: the ray shooting's slab records too
        if (mesh_ids[i] < ctx->slab_ok.size()) ctx->slab_ok[mesh_ids[i]] = 0;
    (void)hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    std::vector<SkinRec> recs(n);
    std::vector<uint32_t> vstart(n), tstart(n);
    uint32_t vtot = 0, ttot = 0, btot = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const cd_ctx::SkinHost* sh = nullptr;
        for (const auto& kv : ctx->skins)
            if (kv.first == mesh_ids[i]) sh = &kv.second;
        if (!sh) return fail(ctx, CD_E_INVALID, "mesh is not a registered skinned mesh");
        const MeshRec& m = ctx->meshes[mesh_ids[i]];
        recs[i] = SkinRec{mesh_ids[i], sh->vbase, sh->nvert, sh->nbones, btot, m.tri_base, m.ntri, sh->vidx_base};
        vstart[i] = vtot;
        tstart[i] = ttot;
        vtot += sh->nvert;
        ttot += m.ntri;
        btot += sh->nbones;
    }
    if (!all_finite(bone_matrices, size_t(btot) * 16)) return fail(ctx, CD_E_INVALID, "non-finite bone matrix");
    const size_t rec_bytes = n * sizeof(SkinRec), pre_bytes = n * sizeof(uint32_t);
    HCK(ctx, ctx->d_skinrec.ensure(rec_bytes + 2 * pre_bytes));
    char* base = ctx->d_skinrec.as<char>();
    HCK(ctx, hipMemcpyAsync(base, recs.data(), rec_bytes, hipMemcpyHostToDevice, st));
    HCK(ctx, hipMemcpyAsync(base + rec_bytes, vstart.data(), pre_bytes, hipMemcpyHostToDevice, st));
    HCK(ctx, hipMemcpyAsync(base + rec_bytes + pre_bytes, tstart.data(), pre_bytes, hipMemcpyHostToDevice, st));
    HCK(ctx, ctx->d_bones.ensure(size_t(btot) * 16 * sizeof(float)));
    HCK(ctx, hipMemcpyAsync(ctx->d_bones.p, bone_matrices, size_t(btot) * 16 * sizeof(float), hipMemcpyHostToDevice, st));
    const SkinRec* dsk = reinterpret_cast<const SkinRec*>(base);
    const uint32_t* dvs = reinterpret_cast<const uint32_t*>(base + rec_bytes);
    const uint32_t* dts = reinterpret_cast<const uint32_t*>(base + rec_bytes + pre_bytes);
    // refit items: the root and both child boxes of every internal node (cached per mesh list)
    std::vector<uint32_t> key(mesh_ids, mesh_ids + n);
    if (key != ctx->refit_key) {
        std::vector<uint2> items;
        for (uint32_t i = 0; i < n; ++i) {
            const MeshRec& m = ctx->meshes[mesh_ids[i]];
            items.push_back(make_uint2(mesh_ids[i], ROOT_CODE));
            for (uint32_t nd = 0; nd < m.nnodes; ++nd) {
                items.push_back(make_uint2(mesh_ids[i], nd << 1));
                items.push_back(make_uint2(mesh_ids[i], (nd << 1) | 1u));
            }
        }
        HCK(ctx, ctx->d_refit_items.ensure(items.size() * sizeof(uint2) + n * sizeof(uint32_t)));
        HCK(ctx, hipMemcpyAsync(ctx->d_refit_items.p, items.data(), items.size() * sizeof(uint2), hipMemcpyHostToDevice, st));
        HCK(ctx, hipMemcpyAsync(ctx->d_refit_items.as<char>() + items.size() * sizeof(uint2), mesh_ids,
                                n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        // plan: each box's triangle range, then small boxes and chunks of the large ones
        const uint32_t ni = uint32_t(items.size());
        HCK(ctx, ctx->d_refit_ranges.ensure((size_t(ni) + 1) * sizeof(uint2)));
        HCK(ctx, launch_item_ranges(ctx->d_refit_items.as<uint2>(), ni, ctx->d_meshes.as<MeshRec>(),
                                    ctx->d_nodes.as<NodeRec>(), ctx->d_refit_ranges.as<uint2>(), st));
        std::vector<uint2> ranges(ni);
        HCK(ctx, hipMemcpyAsync(ranges.data(), ctx->d_refit_ranges.p, ni * sizeof(uint2), hipMemcpyDeviceToHost, st));
        HCK(ctx, hipStreamSynchronize(st));
        std::vector<uint32_t> small;
        std::vector<uint4> chunks, big;
        const uint32_t smax = refit_small_max(), ct = refit_chunk_tris();
        for (uint32_t it = 0; it < ni; ++it) {
            if (ranges[it].y <= smax) {
                small.push_back(it);
                continue;
            }
            big.push_back(make_uint4(it, uint32_t(chunks.size()), (ranges[it].y + ct - 1) / ct, 0u));
            for (uint32_t t = 0; t < ranges[it].y; t += ct)
                chunks.push_back(make_uint4(it, ranges[it].x + t, std::min(ct, ranges[it].y - t), 0u));
        }
        const size_t bs = small.size() * sizeof(uint32_t), bc = chunks.size() * sizeof(uint4),
                     bb = big.size() * sizeof(uint4);
        HCK(ctx, ctx->d_refit_plan.ensure(bs + bc + bb + 64));
        char* pb = ctx->d_refit_plan.as<char>();
        const size_t oc = (bs + 15) / 16 * 16, ob = oc + bc;
        if (bs) HCK(ctx, hipMemcpyAsync(pb, small.data(), bs, hipMemcpyHostToDevice, st));
        if (bc) HCK(ctx, hipMemcpyAsync(pb + oc, chunks.data(), bc, hipMemcpyHostToDevice, st));
        if (bb) HCK(ctx, hipMemcpyAsync(pb + ob, big.data(), bb, hipMemcpyHostToDevice, st));
        HCK(ctx, hipStreamSynchronize(st));
        ctx->refit_plan = RefitPlan{ctx->d_refit_items.as<uint2>(), ctx->d_refit_ranges.as<uint2>(),
                                    reinterpret_cast<const uint32_t*>(pb), uint32_t(small.size()),
                                    reinterpret_cast<const uint4*>(pb + oc), uint32_t(chunks.size()),
                                    reinterpret_cast<const uint4*>(pb + ob), uint32_t(big.size())};
        HCK(ctx, ctx->d_refit_partial.ensure((chunks.size() + 1) * 8 * sizeof(double)));
        ctx->refit_key = key;
        ctx->refit_nitems = ni;
    }
    HCK(ctx, ctx->d_refit_scratch.ensure((size_t(ctx->refit_nitems) + 1) * 24 * sizeof(double)));
    const uint32_t nm = uint32_t(ctx->meshes.size());
    HCK(ctx, ctx->d_refit_bits.ensure((SLACK_REPS + 1) * size_t(nm) * sizeof(unsigned int)));
    unsigned int* bits = ctx->d_refit_bits.as<unsigned int>();
    HCK(ctx, hipEventRecord(ctx->ev[0], st));
    HCK(ctx, launch_skin(dsk, n, dvs, ctx->d_bind.as<float>(), ctx->d_joints.as<ushort4>(), ctx->d_weights.as<float4>(),
                         ctx->d_bones.as<float>(), vtot, ctx->d_skinpos.as<float>(), st));
    HCK(ctx, launch_gather_tris(dsk, n, dts, ctx->d_vidx.as<uint32_t>(), ctx->d_skinpos.as<float>(), ctx->d_tris.as<float>(),
                                ttot, st));
    HCK(ctx, launch_fill_u32(bits, (SLACK_REPS + 1) * nm, 0u, st));
    const uint2* items = ctx->d_refit_items.as<uint2>();
    const uint32_t* ids = reinterpret_cast<const uint32_t*>(ctx->d_refit_items.as<char>() + ctx->refit_nitems * sizeof(uint2));
    (void)items;
    HCK(ctx, launch_refit(ctx->refit_plan, ctx->d_meshes.as<MeshRec>(), ctx->d_nodes.as<NodeRec>(),
                          ctx->d_tris.as<float>(), ctx->d_refit_scratch.as<double>(),
                          ctx->d_refit_partial.as<double>(), bits, bits + size_t(SLACK_REPS) * nm, ids, n, st));
    HCK(ctx, hipEventRecord(ctx->ev[1], st));
    // keep the host copy of the mesh table in step (root boxes, slack, radius): one copy of the id span
    const uint32_t lo_id = *std::min_element(mesh_ids, mesh_ids + n), hi_id = *std::max_element(mesh_ids, mesh_ids + n);
    HCK(ctx, hipMemcpyAsync(&ctx->meshes[lo_id], ctx->d_meshes.as<MeshRec>() + lo_id,
                            size_t(hi_id - lo_id + 1) * sizeof(MeshRec), hipMemcpyDeviceToHost, st));
    HCK(ctx, hipStreamSynchronize(st));
    float ms = 0.f;
    HCK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    ctx->timings = cd_timings{};
    ctx->timings.total_ms = ms;
    ctx->timings.broad_ms = ms;
    return CD_OK;
}

int cd_get_counters(cd_ctx* ctx, cd_counters* out) {
    if (!ctx || !out) return CD_E_INVALID;
    *out = ctx->counters;
    return CD_OK;
}

int cd_set_tuning(cd_ctx* ctx, const cd_tuning* t) {
    if (!ctx || !t) return CD_E_INVALID;
    if (t->stack_entries && t->stack_entries != 512 && t->stack_entries != 1024 && t->stack_entries != 2048)
        return fail(ctx, CD_E_INVALID, "stack_entries must be 512, 1024 or 2048");
    if (t->pop_width > 64 || t->split_levels > 64 || t->waves_per_cu > 64 || t->work_sharing > 3 ||
        t->handoff_pct > 10000 || t->resplit_levels > MAX_RESPLIT || t->pipeline > 2 ||
        t->split_depth > 2 || t->step_graph > 2)
        return fail(ctx, CD_E_INVALID, "tuning out of range");
    ctx->tuning = *t;
    return CD_OK;
}

int cd_get_tuning(cd_ctx* ctx, cd_tuning* t) {
    if (!ctx || !t) return CD_E_INVALID;
    *t = effective_tuning(ctx);
    t->use_graph = ctx->tuning.use_graph;
    t->step_graph = ctx->tuning.step_graph == 2 ? 2 : 1;
    return CD_OK;
}

int cd_get_timings(cd_ctx* ctx, cd_timings* out) {
    if (!ctx || !out) return CD_E_INVALID;
    *out = ctx->timings;
    return CD_OK;
}

}  // extern "C"
