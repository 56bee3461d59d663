// Host-side launchers of the device kernels (one declaration per kernel family).
#pragma once

#include "internal.hpp"
#include "sortbits.hpp"

namespace cdk {

// build.hip (tri9 in host memory, or already on the device)
BuildResult build_soups(cd_ctx* ctx, const float* tri9, const uint32_t* counts, uint32_t nmesh,
                        bool tri9_on_device = false);

// nodes.hip
hipError_t launch_node_globals(const cd_node* nodes, const uint32_t* depth, uint32_t n, float* out, hipStream_t st);

// ingest.hip
hipError_t launch_assemble(const PrimSpan* prims, uint32_t nprim, const uint32_t* idx, const float* pos, uint32_t total,
                           float* tri9, uint32_t* vid3, hipStream_t st);

// plain.hip
hipError_t launch_zero(RunCounters* c, hipStream_t st);
hipError_t launch_tritri_plain(const float* a9, const float* b9, uint64_t n, int32_t* flags, float* seg6,
                               hipStream_t st);
// need: stack entries the deepest query needs (its trees' depths + 2); above
// PLAIN_STACK the stacks live in gstack, ceil(nq / 64) * need * 64 entries
hipError_t launch_traverse_plain(const QueryRec* qs, uint32_t nq, const MeshRec* meshes, const NodeRec* nodes,
                                 uint4* lp, uint64_t lp_cap, RunCounters* ctr, hipStream_t st, uint32_t need = 0,
                                 uint2* gstack = nullptr);
hipError_t launch_narrow_plain(const uint4* lps, uint64_t nlp, uint32_t leaf_max, const QueryRec* qs,
                               const MeshRec* meshes, const float* tris, cd_contact* out, uint64_t cap,
                               RunCounters* ctr, hipStream_t st);

// batched.hip
// The run's first launch: query records, the first frontier (count in
// small[0]), small[1 .. nsmall) and the counters cleared.
hipError_t launch_start(const QueryRec* qs, uint32_t nq_cap, const unsigned int* nq_p, const MeshRec* meshes,
                        QueryPre* qp, uint3* frontier, unsigned int* small, uint32_t nsmall, RunCounters* ctr,
                        hipStream_t st);
// ref (these three launchers): the reference's box test and descent measure
// instead of the conservative ones (reference-semantics mode).
hipError_t launch_split_level(const uint3* in, const unsigned int* nin, uint3* out, unsigned int* nout, uint32_t out_cap,
                              const QueryPre* qp, const MeshRec* meshes, const NodeRec* nodes, uint4* lp,
                              uint64_t lp_cap, uint3* tasks, unsigned int* ntasks, uint32_t split_tris,
                              RunCounters* ctr, uint32_t grid, hipStream_t st, bool ref = false);
// Two split levels in one launch: split_tris rules the first, split_tris2 the
// second (0xFFFFFFFF: the second is the last split level).
hipError_t launch_split_level2(const uint3* in, const unsigned int* nin, uint3* out, unsigned int* nout,
                               uint32_t out_cap, const QueryPre* qp, const MeshRec* meshes, const NodeRec* nodes,
                               uint4* lp, uint64_t lp_cap, uint3* tasks, unsigned int* ntasks, uint32_t split_tris,
                               uint32_t split_tris2, RunCounters* ctr, uint32_t grid, hipStream_t st, bool ref = false);
hipError_t launch_dfs(int cap_entries, const uint3* tasks, const uint4* tasks4, const unsigned int* task_lo,
                      const unsigned int* ntasks, uint32_t task_cap, unsigned int* head, const QueryPre* qp,
                      const MeshRec* meshes, const NodeRec* nodes, uint4* lp, uint64_t lp_cap, uint3* spill,
                      uint32_t spill_cap, uint32_t pop_width, uint32_t nwaves, uint4* dq, unsigned int* dq_ctl,
                      uint32_t dq_cap, uint32_t share, uint32_t spin_polls, const unsigned long long* budget,
                      uint32_t budget_pct, uint3* dump, unsigned int* ndump, uint32_t dump_cap, RunCounters* ctr,
                      hipStream_t st, bool pipelined = false, bool ref = false);
// Keeps this run's node-pair test count for the next run's budgeted pass.
hipError_t launch_keep_budget(const RunCounters* ctr, unsigned long long* budget, hipStream_t st);
// defer: 2 * defer_cap entries, the float64 and the exact rerun lists;
// hits: cap hit records (query, A slot, B slot, sides), made into cap contacts;
// cus: the device's compute units (the grids are sized from it)
hipError_t launch_narrow_exact(const uint4* lps, const unsigned long long* nlp, uint64_t lp_cap, const QueryPre* qp,
                               const float* tris, uint4* hits, cd_contact* out, uint64_t cap, uint4* defer,
                               uint64_t defer_cap, RunCounters* ctr, uint32_t cus, hipStream_t st);
// The reference's narrow phase (plain path's test and contact record) over
// the batched traversal's leaf pairs, contacts straight to `out`.
hipError_t launch_narrow_ref(const uint4* lps, const unsigned long long* nlp, uint64_t lp_cap, const QueryPre* qp,
                             const float* tris, cd_contact* out, uint64_t cap, RunCounters* ctr, uint32_t cus,
                             hipStream_t st);
// Queries the narrow launches' occupancy once, outside any graph capture.
void warm_launch_shapes();
// Workgroups per CU a split-level launch keeps resident.
uint32_t split_blocks_per_cu(bool ref, bool two_levels);
// lists: 64 + 2n words of device scratch (two rerun lists of n entries)
hipError_t launch_tritri_exact(const float* a9, const float* b9, uint64_t n, int32_t* flags, float* seg6,
                               uint32_t* lists, uint32_t grid, RunCounters* ctr, hipStream_t st);
hipError_t launch_fill_u32(unsigned int* p, uint32_t n, unsigned int v, hipStream_t st);

// broad.hip
hipError_t launch_sap_project(const cd_entity* ents, uint32_t n, const MeshRec* meshes, const SapAxes& ax, float* iv,
                              uint32_t* keys, uint32_t* vals, hipStream_t st);
// Stable sort of (key, value) pairs by keys below 2^key_bits; the result is
// in keys/vals.  Longer arrays run over many workgroups and need `table`,
// radix_table_words(n) words of device scratch (without it they run in one).
// n_dev (here, in the pair kernels and in the contact sorts): a device count,
// min(*n_dev, n) used, n then being the capacity the launches are shaped for.
hipError_t launch_radix_sort(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t n,
                             uint32_t key_bits, uint32_t* table, hipStream_t st, const uint32_t* n_dev = nullptr);
uint64_t radix_table_words(uint32_t n);
hipError_t launch_sap_sweep(const uint32_t* order, const float* iv, const cd_entity* ents, uint32_t n, uint2* pairs,
                            uint32_t cap, unsigned int* count, hipStream_t st);
hipError_t launch_pair_keys(const uint2* pairs, uint32_t n, uint32_t* keys, uint32_t* vals, hipStream_t st,
                            const uint32_t* n_dev = nullptr);
hipError_t launch_pair_queries(const uint2* pairs, const uint32_t* order, uint32_t n, const cd_entity* ents,
                               QueryRec* qs, uint2* sorted_pairs, hipStream_t st, const uint32_t* n_dev = nullptr);
// Contacts ordered by (query, tri_a, tri_b) into out.  field_bits: key bits
// of query, tri_a and tri_b (values below 2^bits).
hipError_t launch_sort_contacts(const cd_contact* in, uint32_t n, cd_contact* out, const uint32_t field_bits[3],
                                uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, uint32_t* table,
                                hipStream_t st, const uint32_t* n_dev = nullptr);
// Stable order of contacts by three fields (0 query, 1 tri_a, 2 tri_b), most
// significant first; the permutation ends in v0.
hipError_t launch_sort_perm(const cd_contact* in, uint32_t n, int f_hi, int f_mid, int f_lo,
                            const uint32_t field_bits[3], uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1,
                            uint32_t* table, hipStream_t st, const uint32_t* n_dev = nullptr);

// response.hip
// pbegin: per side (nq each) the first ray of every pair, non-decreasing;
// stack_entries: the deepest tree the rays meet + 2; above RAY_STACK_ENTRIES
// the stacks live in gstack, ceil(2 nrays / 64) * 2 * stack_entries * 64
// words; ctr->err gets bit 0 if a traversal stack would have overflowed,
// ctr->sat the boxes tested and ctr->tritests the ray-triangle tests.
// stride: the per-side spacing of pbegin (0: nq); dc: the step's count block
// (DC_*), from which the ray and pair counts are read when given, nrays and
// nq being the capacities.
// slabs / roots: the slab records of every node's children and every mesh's
// root (launch_ray_slabs), current for the meshes the rays meet.
hipError_t launch_shoot(const cd_ray* rays, uint32_t nrays, uint32_t base1, const uint32_t* pbegin, uint32_t nq,
                        const QueryRec* qs, const MeshRec* meshes, const NodeRec* nodes, const RaySlab* slabs,
                        const RaySlab* roots, const float* tris,
                        const uint2* attrs, const uint32_t* vid, const float* nrm, const uint32_t* perm, RayShot* shots,
                        uint32_t stack_entries, RunCounters* ctr, hipStream_t st, uint32_t* gstack = nullptr,
                        uint32_t stride = 0, const uint32_t* dc = nullptr);
// Slab records of count internal nodes' children (out: 2 per node) and of
// mesh m's root (root_out).
hipError_t launch_ray_slabs(const NodeRec* nodes, uint32_t count, RaySlab* out, const MeshRec* meshes, uint32_t m,
                            RaySlab* root_out, hipStream_t st);
// qw / avg per side at stride spacing (0: nq), dc as for launch_shoot, nshots
// the shots written.
hipError_t launch_respond(const RayShot* shots, uint32_t base1, const uint32_t* qw, uint32_t nq, const float* avg,
                          const QueryRec* qs, const uint2* pair_ids, const cd_entity* ents, float cos_start,
                          float cos_finish, float bias, cd_pair_response* out, hipStream_t st, uint32_t stride = 0,
                          const uint32_t* dc = nullptr, uint64_t nshots = ~0ull);

// rays.hip
// The ray generation's buffers, both sides: side s's part of a per-side
// buffer starts s times its size in.  n, nq: contacts and pairs (the
// capacities the launches are shaped for when n_dev / nq_dev give the counts
// on the device).
struct RaySides {
    const cd_contact* sorted;   // contacts by (pair, tri_a, tri_b)
    const uint32_t* order_b;    // their order by (pair, tri_b, tri_a)
    uint32_t n;
    const uint32_t* n_dev;
    uint32_t nq;
    const uint32_t* nq_dev;
    uint32_t qstride;           // per-side spacing of qw (counts, then begins) and avg (x 3)
    RayCand* cand;              // n per side
    RayCand* dense;             // n per side
    uint32_t* start;            // n + 1 per side (likewise pos, cnt, pos2)
    uint32_t* pos;
    uint32_t* cnt;
    uint32_t* pos2;
    uint32_t* sums;             // nsums per side: one per 1,024 contacts
    uint32_t nsums;
    unsigned long long* hkey;   // H per side (a power of two, at least 6 n)
    unsigned long long* hval;
    uint64_t H;
    uint32_t* qw;               // per side: the pairs' ray counts, then their first rays
    float* avg;                 // per side: the pairs' average ray origins
    uint32_t* tot;              // DC_CAND0, DC_RAYS0, DC_CAND1, DC_RAYS1
    cd_ray* rays;               // the first side's rays, then the second side's: 6 n at most
};
// The whole ray generation (twelve launches, both sides in each).
hipError_t launch_rays(const RaySides& b, const QueryRec* qs, const MeshRec* meshes, const uint32_t* leafof,
                       const float* tris, const uint2* attrs, const uint32_t* vid, const float* nrm,
                       const uint32_t* perm, hipStream_t st);

// step.hip (the one-call scene step; dc is its count block, DC_* words)
// Word spans gathered into one block: span s, words[s] words from src[s],
// lands at dst[s] words into the block.
struct PackSpans {
    static constexpr uint32_t MAX = 8;
    const uint32_t* src[MAX];
    uint32_t words[MAX], dst[MAX];
    uint32_t n;
};
// dc[DC_SWEEP] the pairs the sweep found, dc[DC_NQ] and *d_nq those of them kept
hipError_t launch_step_pairs(const unsigned int* sweep, uint32_t cap, uint32_t* dc, unsigned int* d_nq, hipStream_t st);
// dc[DC_NC] the contacts of the run, up to cap; qcnt (np words), when given,
// cleared for launch_seg_order
hipError_t launch_step_contacts(const RunCounters* ctr, uint64_t cap, uint32_t* dc, hipStream_t st,
                                uint32_t* qcnt = nullptr, uint32_t np = 0);
// The dc[DC_NC] contacts at c, of dc[DC_NQ] pairs, ordered by (pair, tri_a,
// tri_b) into sorted, and order_b their permutation by (pair, tri_b, tri_a);
// capacities nc_cap contacts and np_cap pairs; qw3 3 np_cap words (cleared
// counts first), seg nc_cap contacts of scratch.  dc[DC_BIGSEG] is set when
// a pair has more contacts than one workgroup ranks, and the order is then
// incomplete.
hipError_t launch_seg_order(const cd_contact* c, uint32_t nc_cap, uint32_t np_cap, uint32_t* dc, uint32_t* qw3,
                            cd_contact* seg, cd_contact* sorted, uint32_t* order_b, hipStream_t st);
// pbegin[stride side + q]: the first ray of pair q on each side, from the
// counts at qw + 2 stride side, for the dc[DC_NQ] pairs
hipError_t launch_pair_begins(const uint32_t* qw, uint32_t stride, const uint32_t* dc, uint32_t* pbegin, hipStream_t st);
hipError_t launch_pack(const PackSpans& sp, uint32_t* dst, hipStream_t st);

// refit.hip
hipError_t launch_tree_vidx(const uint32_t* perm, uint32_t ntri, const uint32_t* indices, uint32_t vbase,
                            uint32_t* out, hipStream_t st);
hipError_t launch_tree_vidx_many(const uint32_t* perm, const uint32_t* prefix, uint32_t nmesh, uint64_t ntri,
                                 const uint32_t* indices, uint32_t* out, hipStream_t st);
hipError_t launch_skin(const SkinRec* sk, uint32_t nsk, const uint32_t* vstart, const float* bind, const ushort4* joints,
                       const float4* weights, const float* bones, uint32_t total, float* out, hipStream_t st);
hipError_t launch_gather_tris(const SkinRec* sk, uint32_t nsk, const uint32_t* tstart, const uint32_t* vidx,
                              const float* pos, float* tris, uint32_t total, hipStream_t st);
hipError_t launch_item_ranges(const uint2* items, uint32_t nitems, const MeshRec* meshes, const NodeRec* nodes,
                              uint2* ranges, hipStream_t st);
hipError_t launch_refit(const RefitPlan& p, MeshRec* meshes, NodeRec* nodes, const float* tris, double* scratch,
                        double* partial, unsigned int* slack_bits, unsigned int* radius_bits, const uint32_t* ids,
                        uint32_t nids, hipStream_t st);
uint32_t refit_small_max();
uint32_t refit_chunk_tris();

}  // namespace cdk
