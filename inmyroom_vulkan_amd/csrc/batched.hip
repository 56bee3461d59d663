// Batched path: thousands of uneven mesh-pair queries in a few launches.
//
//  k_start        per query: relative matrix and the slack that makes the
//                 float32 separating-axis test conservative, and one root
//                 node pair in the first frontier; the run's counters cleared
//  k_split_level  breadth-first expansion of node pairs for the first
//                 levels (work splitting: a large query becomes many tasks)
//  k_dfs          persistent wavefronts; each pops up to 64 node pairs at a
//                 time from a stack in LDS (wave-private spill to HBM), tests
//                 them, pushes children with ballot/prefix compaction, and
//                 takes new tasks from a global counter when it runs low.
//                 With work sharing, wavefronts that run dry take stack
//                 bottoms donated by busy ones (bounded waits only), and a
//                 second launch takes whatever nobody claimed
//  k_narrow_exact candidate leaf pairs expanded densely to triangle tests
//                 (wave prefix scan), exact-sign triangle test, contacts
//                 compacted with one atomic per wavefront
//
// Node-pair semantics follow OBBtree::IntersectOBBtreesRecursive
// (inMyRoom_vulkan/src/Geometry/OBBtree.cpp:414-477) with the separating-axis
// test of Paralgram.cpp:17-173 made conservative: an axis separates only when
// the float gap exceeds a forward error bound plus the containment slack of
// both trees, so no pair of triangles that truly intersect is culled.
#include "exact.hpp"
#include "kernels.hpp"

namespace cdk {

namespace {

__device__ __forceinline__ unsigned long long lanes_below_b() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}
__device__ __forceinline__ unsigned long long reserve_b(unsigned long long mask, unsigned long long* counter) {
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(counter, static_cast<unsigned long long>(__popcll(mask)));
    base = __shfl(base, leader, 64);
    return base + static_cast<unsigned long long>(__popcll(mask & lanes_below_b()));
}
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Adds a per-lane count to a device counter with one atomic per block
// (256 threads; every thread of the block calls it).  Same-address atomics
// serialise in L2, so one per wavefront over a long stream costs more than
// the arithmetic.
__device__ __forceinline__ void block_add(unsigned long long v, unsigned long long* __restrict__ dst) {
    __shared__ unsigned long long part[4];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long t = part[0] + part[1] + part[2] + part[3];
        if (t) atomicAdd(dst, t);
    }
    __syncthreads();
}

constexpr float U32 = 5.9604644775390625e-08f;  // 2^-24

__device__ __forceinline__ float norm1(f3 v) { return (fabsf(v.x) + fabsf(v.y)) + fabsf(v.z); }

// Fused-multiply-add forms of the products the conservative test, the box
// transform and the descent rule of this path use.  A fused operation rounds
// once where a product and a sum round twice, so each of these stays within
// the bound its unfused form has (gamma_2 for a cross-product component,
// gamma_3 for a three-term dot product or matrix row): the rounding margins
// below and in the query set-up hold unchanged.  The plain path keeps the
// reference's unfused operation order; this one only has to be conservative.
__device__ __forceinline__ float dotf(f3 a, f3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
__device__ __forceinline__ f3 crossf(f3 a, f3 b) {
    return mk(fmaf(a.y, b.z, -(b.y * a.z)), fmaf(a.z, b.x, -(b.z * a.x)), fmaf(a.x, b.y, -(b.x * a.y)));
}
__device__ __forceinline__ f3 mul4f(const M4& M, f3 p) {  // M (p, 0)
    return mk(fmaf(M.m[8], p.z, fmaf(M.m[4], p.y, M.m[0] * p.x)), fmaf(M.m[9], p.z, fmaf(M.m[5], p.y, M.m[1] * p.x)),
              fmaf(M.m[10], p.z, fmaf(M.m[6], p.y, M.m[2] * p.x)));
}
__device__ __forceinline__ f3 xform_pointf(const M4& M, f3 p) {  // M (p, 1): every term rounded at most three times
    return mk(fmaf(M.m[8], p.z, fmaf(M.m[4], p.y, fmaf(M.m[0], p.x, M.m[12]))),
              fmaf(M.m[9], p.z, fmaf(M.m[5], p.y, fmaf(M.m[1], p.x, M.m[13]))),
              fmaf(M.m[10], p.z, fmaf(M.m[6], p.y, fmaf(M.m[2], p.x, M.m[14]))));
}
__device__ __forceinline__ Box xform_boxf(const M4& M, const Box& b) {
    Box r;
    r.c = xform_pointf(M, b.c);
    r.u = mul4f(M, b.u);
    r.v = mul4f(M, b.v);
    r.w = mul4f(M, b.w);
    return r;
}
// The descent rule's measure of a box: any consistent choice of which side to
// split keeps the traversal conservative, so the hardware square root (within
// one ulp) stands in for the correctly rounded one.
__device__ __forceinline__ float surfacef(const Box& b) {
    const f3 uv = crossf(b.u, b.v), uw = crossf(b.u, b.w), vw = crossf(b.v, b.w);
    return 2.f * ((__builtin_amdgcn_sqrtf(dotf(uv, uv)) + __builtin_amdgcn_sqrtf(dotf(uw, uw))) +
                  __builtin_amdgcn_sqrtf(dotf(vw, vw)));
}

// Conservative separating-axis test of two boxes (centre + three half-axis
// vectors each) on the 15 axes of Paralgram.cpp:17-173.  On an axis L the
// boxes are disjoint when |L.T| > sum_k |L.a_k| + sum_l |L.b_l|, T = r.c - l.c,
// a_k and b_l the two boxes' half-axes.  Every one of these dot products is a
// scalar triple product of box vectors and T, so the 15 axes share them:
//   L = a_p x a_q (face of l):  |L.a_m| = |det a|, |L.b_j| = D1[m][j] = |(a_p x a_q).b_j|
//   L = b_p x b_q (face of r):  |L.b_n| = |det b|, |L.a_i| = D2[i][n] = |a_i.(b_p x b_q)|
//   L = a_i x b_j (edges):      |L.a_k| = D1[m][j] for the m != i,
//                               |L.b_l| = D2[i][n] for the n != j,
//                               |L.T| = |a_i.(b_j x T)|
// (the remaining terms, L.a_p for an axis built from a_p, are exactly zero).
// Rounding: a float triple product x.(y z) is within 5u |x|_1 |y|_1 |z|_1 of
// the exact one and |x cross y|_1 <= |x|_1 |y|_1, so 16u S |x|_1 |y|_1, S the
// sum of the 1-norms of T and the six half-axes, bounds every error of an
// axis built from x and y, the sums included; sigma |x|_1 |y|_1 bounds how far
// the triangles stick out of their boxes along it.  An axis rejects only
// beyond both: no intersecting pair is culled.
__device__ inline bool sat_conservative(const Box& l, const Box& r, float sigma) {
    const f3 a[3] = {l.u, l.v, l.w};
    const f3 b[3] = {r.u, r.v, r.w};
    const f3 T = sub(r.c, l.c);
    float na[3], nb[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        na[i] = norm1(a[i]);
        nb[i] = norm1(b[i]);
    }
    const float S = norm1(T) + ((na[0] + na[1]) + na[2]) + ((nb[0] + nb[1]) + nb[2]);
    const float K = 16.f * U32 * S + sigma;
    // face normals in the reference's order: v x w, u x w, u x v
    const f3 NA[3] = {crossf(a[1], a[2]), crossf(a[0], a[2]), crossf(a[0], a[1])};
    const f3 NB[3] = {crossf(b[1], b[2]), crossf(b[0], b[2]), crossf(b[0], b[1])};
    const float fa[3] = {na[1] * na[2], na[0] * na[2], na[0] * na[1]};
    const float fb[3] = {nb[1] * nb[2], nb[0] * nb[2], nb[0] * nb[1]};
    const float detA = fabsf(dotf(NA[0], a[0])), detB = fabsf(dotf(NB[0], b[0]));
    float D1[3][3], D2[3][3];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            D1[m][j] = fabsf(dotf(NA[m], b[j]));
            D2[m][j] = fabsf(dotf(a[m], NB[j]));
        }
    bool sep = false;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        sep |= fabsf(dotf(NA[m], T)) - (detA + ((D1[m][0] + D1[m][1]) + D1[m][2])) > fa[m] * K;
        sep |= fabsf(dotf(NB[m], T)) - (((D2[0][m] + D2[1][m]) + D2[2][m]) + detB) > fb[m] * K;
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const f3 X = crossf(b[j], T);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float rA = D1[i == 0 ? 1 : 0][j] + D1[i == 2 ? 1 : 2][j];
            const float rB = D2[i][j == 0 ? 1 : 0] + D2[i][j == 2 ? 1 : 2];
            sep |= fabsf(dotf(a[i], X)) - (rA + rB) > (na[i] * nb[j]) * K;
        }
    }
    return !sep;
}

// Traversal codes of the batched path are absolute, so a node pair's two
// records can be fetched straight from the codes, together with the query's
// record, in one round trip to memory: ROOT_FLAG | mesh for a tree's root
// (its box lives in the mesh record), else (node << 1) | side with node an
// index into the whole node arena.
constexpr uint32_t ROOT_FLAG = 0x80000000u;

struct Code {  // resolved traversal code
    Box box;
    float surf;
    bool leaf;
    uint32_t link, cnt;  // link: mesh-local child node index, or triangle offset of a leaf
    uint32_t tris;       // triangles under the box
};
// A traversal code's child record as loaded, before it is decoded: the loads
// of a node pair are issued in one phase (load_code) and used in another
// (finish_code), so a wavefront can have one batch's loads in flight while it
// tests the previous batch.  A root's record sits in its mesh record and has
// the layout of every other child record, so the only lane-dependent choice is
// the record's address; the four 16-byte loads are the same for every lane.
struct RawCode {
    uint4 w[4];  // box (w[0..2]); surf, link, tris, toff (w[3])
    uint32_t code;
};
__device__ __forceinline__ void load_code(const MeshRec* __restrict__ meshes, const NodeRec* __restrict__ nodes,
                                          uint32_t code, RawCode& r) {
    const bool root = (code & ROOT_FLAG) != 0u;
    const MeshRec* mr = meshes + (root ? (code & ~ROOT_FLAG) : 0u);
    const NodeRec* nr = nodes + (root ? 0u : (code >> 1));
    const uint4* p = reinterpret_cast<const uint4*>(root ? &mr->root : &nr->c[code & 1u]);
#pragma unroll
    for (int k = 0; k < 4; ++k) r.w[k] = p[k];
    r.code = code;
}
__device__ __forceinline__ Code finish_code(const RawCode& w) {
    float bx[12];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        bx[4 * k] = __uint_as_float(w.w[k].x);
        bx[4 * k + 1] = __uint_as_float(w.w[k].y);
        bx[4 * k + 2] = __uint_as_float(w.w[k].z);
        bx[4 * k + 3] = __uint_as_float(w.w[k].w);
    }
    Code r;
    r.box = box_from12(bx);
    r.surf = __uint_as_float(w.w[3].x);
    r.link = w.w[3].y;
    r.leaf = (w.w[3].z & LEAF_BIT) != 0u;
    r.tris = w.w[3].z & ~LEAF_BIT;
    r.cnt = r.leaf ? r.tris : 0u;
    return r;
}

enum : int { PAIR_NONE = 0, PAIR_LEAF = 1, PAIR_SPLIT = 2 };

constexpr uint32_t LP_CHUNK = 256;  // leaf-pair slots a persistent wavefront reserves at a time (>= 64)

// Unused slots [next, end) of a wavefront's chunk become empty leaf pairs.
__device__ __forceinline__ void close_chunk(uint4* __restrict__ lp_out, uint64_t lp_cap, uint64_t next, uint64_t end,
                                            int lane) {
    for (uint64_t i = next + uint64_t(lane); i < end; i += 64)
        if (i < lp_cap) lp_out[i] = make_uint4(0u, 0u, 0u, 0u);
}

// The loads of one node pair: the query's record and both codes' records.
struct PairLoad {
    float m[16];
    float sigma;
    uint32_t base_a, base_b, tbase_a, tbase_b;
    RawCode A, B;
};
__device__ __forceinline__ void load_pair(const QueryPre* __restrict__ qp, const MeshRec* __restrict__ meshes,
                                          const NodeRec* __restrict__ nodes, uint32_t qi, uint32_t ca, uint32_t cb,
                                          PairLoad& P) {
    const QueryPre& Q = qp[qi];
#pragma unroll
    for (int k = 0; k < 16; ++k) P.m[k] = Q.m[k];
    P.sigma = Q.sigma;
    P.base_a = Q.node_base_a;
    P.base_b = Q.node_base_b;
    P.tbase_a = Q.tri_base_a;
    P.tbase_b = Q.tri_base_b;
    load_code(meshes, nodes, ca, P.A);
    load_code(meshes, nodes, cb, P.B);
}

// One node pair from its loads: the box test, then leaf pair or the two
// child pairs (first child on top), with the reference's descent rule.
// REF: the reference's float test, box transform and surface (the plain
// path's, operation for operation), so the node pairs visited and the leaf
// pairs found are the reference's; else the conservative test of the
// exact-predicate path.
template <bool REF>
__device__ __forceinline__ int eval_loaded(const PairLoad& P, uint32_t qi, uint32_t& n0a, uint32_t& n0b,
                                           uint32_t& n1a, uint32_t& n1b, uint4& leaf, uint32_t& cost) {
    M4 M;
#pragma unroll
    for (int k = 0; k < 16; ++k) M.m[k] = P.m[k];
    const Code A = finish_code(P.A);
    const Code B0 = finish_code(P.B);
    const uint32_t ca = P.A.code, cb = P.B.code;
    Box sb;
    bool hit;
    if constexpr (REF) {
        sb = xform_box(M, B0.box);
        hit = sat(A.box, sb);
    } else {
        sb = xform_boxf(M, B0.box);
        hit = sat_conservative(A.box, sb, P.sigma);
    }
    // Outputs are formed whatever the test says (selects, no branches), so
    // the compiler keeps every record load ahead of the test instead of
    // sinking some of them behind it into a second round trip.
    const bool both_leaf = A.leaf && B0.leaf;
    // leaf pair with arena-global triangle offsets: the narrow phase can fetch
    // triangles without first reading the query's record
    leaf = make_uint4(qi, P.tbase_a + A.link, P.tbase_b + B0.link, A.cnt | (B0.cnt << 16));
    const bool split_a = !A.leaf && (B0.leaf || A.surf >= (REF ? surface(sb) : surfacef(sb)));
    const uint32_t ka = (P.base_a + A.link) << 1, kb = (P.base_b + B0.link) << 1;
    n0a = split_a ? ka : ca;
    n1a = split_a ? (ka | 1u) : ca;
    n0b = split_a ? cb : kb;
    n1b = split_a ? cb : (kb | 1u);
    cost = A.tris + B0.tris;  // size of the pair's subtrees: how much work splitting it further would share out
    return hit ? (both_leaf ? PAIR_LEAF : PAIR_SPLIT) : PAIR_NONE;
}

template <bool REF>
__device__ __forceinline__ int eval_pair(const QueryPre* __restrict__ qp, const MeshRec* __restrict__ meshes,
                                         const NodeRec* __restrict__ nodes, uint32_t qi, uint32_t ca, uint32_t cb,
                                         uint32_t& n0a, uint32_t& n0b, uint32_t& n1a, uint32_t& n1b, uint4& leaf,
                                         uint32_t& cost) {
    PairLoad P;
    load_pair(qp, meshes, nodes, qi, ca, cb, P);
    return eval_loaded<REF>(P, qi, n0a, n0b, n1a, n1b, leaf, cost);
}

}  // namespace

// ---------------------------------------------------------------------------
// A query's record for the traversal: its matrix, the conservative test's
// bound and the two meshes' arena offsets.
__device__ __forceinline__ QueryPre query_pre(const QueryRec& q, const MeshRec* __restrict__ meshes) {
    const MeshRec& A = meshes[q.mesh_a];
    const MeshRec& B = meshes[q.mesh_b];
    QueryPre p;
    float kinf = 0.f, fro = 0.f, tmax = 0.f;
    for (int r = 0; r < 3; ++r) {
        float row = 0.f;
        for (int c = 0; c < 3; ++c) {
            row += fabsf(q.rel[c * 4 + r]);
            fro += q.rel[c * 4 + r] * q.rel[c * 4 + r];
        }
        kinf = fmaxf(kinf, row);
        tmax = fmaxf(tmax, fabsf(q.rel[12 + r]));
    }
    // Triangles of B sit within slack_B of B's boxes; M stretches that by at
    // most its Frobenius norm; rounding of M * p (vertices, box centre and
    // sides, all |.| <= 2 radius) is bounded with room to spare by
    // 16u (10 |M|_inf R_B + 2 |t|_max).  Twice the A-side slack, plus a
    // little for the float rounding of the slack values themselves.
    const float rb = A.slack + sqrtf(fro) * 1.0001f * B.slack + 16.f * U32 * (10.f * kinf * B.radius + 2.f * tmax) +
                     16.f * U32 * 2.f * A.radius;
    for (int k = 0; k < 16; ++k) p.m[k] = q.rel[k];
    p.sigma = rb * 1.0001f + 1e-30f;
    p.mesh_a = q.mesh_a;
    p.mesh_b = q.mesh_b;
    p.node_base_a = A.node_base;
    p.node_base_b = B.node_base;
    p.tri_base_a = A.tri_base;
    p.tri_base_b = B.tri_base;
    p.pad = 0;
    return p;
}

// A batched run's first launch: every query's record and its pair of roots
// in the first frontier (counted in small[0]), the rest of the run's small
// words and its counters cleared.
__global__ void k_start(const QueryRec* __restrict__ qs, const unsigned int* __restrict__ nq_p,
                        const MeshRec* __restrict__ meshes, QueryPre* __restrict__ qp, uint3* __restrict__ frontier,
                        unsigned int* __restrict__ small, uint32_t nsmall, RunCounters* __restrict__ ctr) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nq = *nq_p;
    if (i < nsmall) small[i] = i == 0 ? nq : 0u;
    if (i == 0) *ctr = RunCounters{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (i >= nq) return;
    const QueryPre p = query_pre(qs[i], meshes);
    qp[i] = p;
    frontier[i] = make_uint3(i, ROOT_FLAG | p.mesh_a, ROOT_FLAG | p.mesh_b);
}

template <bool REF>
__global__ __launch_bounds__(256) void k_split_level(const uint3* __restrict__ in, const unsigned int* __restrict__ nin,
                                                     uint3* __restrict__ out, unsigned int* __restrict__ nout,
                                                     uint32_t out_cap, const QueryPre* __restrict__ qp,
                                                     const MeshRec* __restrict__ meshes,
                                                     const NodeRec* __restrict__ nodes, uint4* __restrict__ lp_out,
                                                     uint64_t lp_cap, uint3* __restrict__ tasks,
                                                     unsigned int* __restrict__ ntasks, uint32_t split_tris,
                                                     RunCounters* __restrict__ ctr) {
    // A pair that passes is split; its two children go on to the next level
    // when the pair's subtrees hold more than split_tris triangles, and to the
    // persistent kernel's task list otherwise.
    const uint32_t n = min(*nin, out_cap);  // frontier buffers and the task list hold out_cap entries each
    const uint32_t stride = gridDim.x * blockDim.x;
    // slots are reserved per block: one atomic per output list and iteration
    __shared__ unsigned int s_n[4][3];  // per wavefront: leaf pairs, next-level entries, tasks
    __shared__ unsigned long long s_base[3];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long below = lanes_below_b();
    unsigned long long nsat = 0, nleaf = 0;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        int kind = PAIR_NONE;
        uint32_t n0a = 0, n0b = 0, n1a = 0, n1b = 0;
        uint4 leaf = make_uint4(0, 0, 0, 0);
        uint32_t qi = 0, cost = 0;
        if (i < n) {
            const uint3 e = in[i];
            qi = e.x;
            ++nsat;
            kind = eval_pair<REF>(qp, meshes, nodes, qi, e.y, e.z, n0a, n0b, n1a, n1b, leaf, cost);
        }
        const bool go_on = cost > split_tris;
        const bool is_leaf = kind == PAIR_LEAF;
        const bool to_next = kind == PAIR_SPLIT && go_on, to_task = kind == PAIR_SPLIT && !go_on;
        const unsigned long long mL = __ballot(is_leaf), mN = __ballot(to_next), mT = __ballot(to_task);
        nleaf += is_leaf ? 1u : 0u;
        if (lane == 0) {
            s_n[wave][0] = unsigned(__popcll(mL));
            s_n[wave][1] = 2u * unsigned(__popcll(mN));
            s_n[wave][2] = 2u * unsigned(__popcll(mT));
        }
        __syncthreads();
        if (threadIdx.x < 3) {
            const int c = int(threadIdx.x);
            const unsigned int tot = (s_n[0][c] + s_n[1][c]) + (s_n[2][c] + s_n[3][c]);
            unsigned long long b0 = 0;
            if (tot) {
                if (c == 0) b0 = atomicAdd(&ctr->lpslots, static_cast<unsigned long long>(tot));
                else b0 = atomicAdd(c == 1 ? nout : ntasks, tot);
            }
            s_base[c] = b0;
        }
        __syncthreads();
        unsigned long long first[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            unsigned long long f = s_base[c];
            for (int w = 0; w < wave; ++w) f += s_n[w][c];
            first[c] = f;
        }
        if (is_leaf) {
            const unsigned long long sl = first[0] + unsigned(__popcll(mL & below));
            if (sl < lp_cap) lp_out[sl] = leaf;
        }
        if (to_next || to_task) {
            const unsigned long long m = to_next ? mN : mT;
            const unsigned long long slot = first[to_next ? 1 : 2] + 2u * unsigned(__popcll(m & below));
            uint3* dst = to_next ? out : tasks;
            if (slot + 1 < out_cap) {
                dst[slot] = make_uint3(qi, n0a, n0b);
                dst[slot + 1] = make_uint3(qi, n1a, n1b);
            } else {
                atomicOr(reinterpret_cast<unsigned int*>(&ctr->err), 2u);
            }
        }
        __syncthreads();  // s_n and s_base are rewritten by the next iteration
    }
    block_add(nsat, &ctr->sat);
    block_add(nleaf, &ctr->leafpairs);
}

// Two split levels per launch.  A thread tests its pair and, when the pair
// splits with more than split_tris triangles under it, tests both children
// itself (their four records loaded together) instead of leaving them to the
// next launch; a child that splits sends its two children on to the next
// launch when it holds more than split_tris2 triangles, else to the task
// list.  The node pairs tested, and what becomes of each, are those of two
// k_split_level launches with split_tris and split_tris2; only the order of
// the outputs differs.  Half the launches, and half the count loads, entry
// loads and slot reservations the levels wait on in turn.
template <bool REF>
__global__ __launch_bounds__(256) void k_split_level2(const uint3* __restrict__ in, const unsigned int* __restrict__ nin,
                                                      uint3* __restrict__ out, unsigned int* __restrict__ nout,
                                                      uint32_t out_cap, const QueryPre* __restrict__ qp,
                                                      const MeshRec* __restrict__ meshes,
                                                      const NodeRec* __restrict__ nodes, uint4* __restrict__ lp_out,
                                                      uint64_t lp_cap, uint3* __restrict__ tasks,
                                                      unsigned int* __restrict__ ntasks, uint32_t split_tris,
                                                      uint32_t split_tris2, RunCounters* __restrict__ ctr) {
    const uint32_t n = min(*nin, out_cap);
    const uint32_t stride = gridDim.x * blockDim.x;
    __shared__ unsigned int s_n[4][3];  // per wavefront: leaf pairs, next-launch entries, tasks
    __shared__ unsigned long long s_base[3];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long below = lanes_below_b();
    unsigned long long nsat = 0, nleaf = 0;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        uint32_t qi = 0, n0a = 0, n0b = 0, n1a = 0, n1b = 0, cost = 0;
        int kind = PAIR_NONE;
        uint4 leaf0 = make_uint4(0, 0, 0, 0);
        PairLoad P;
        if (i < n) {
            const uint3 e = in[i];
            qi = e.x;
            ++nsat;
            load_pair(qp, meshes, nodes, qi, e.y, e.z, P);
            kind = eval_loaded<REF>(P, qi, n0a, n0b, n1a, n1b, leaf0, cost);
        }
        const bool deeper = kind == PAIR_SPLIT && cost > split_tris;
        // Two output groups of two entries each (g0, g1) and two leaf slots:
        // a pair that splits without going deeper fills g0 with its children
        // for the task list; one that goes deeper fills group c (and leaf
        // slot c) from its child c.
        uint32_t g[2][4] = {{n0a, n0b, n1a, n1b}, {0, 0, 0, 0}};
        int dest[2] = {kind == PAIR_SPLIT && !deeper ? 2 : 0, 0};  // 0 none, 1 next launch, 2 task list
        uint4 lf[2] = {leaf0, make_uint4(0, 0, 0, 0)};
        bool has_leaf[2] = {kind == PAIR_LEAF, false};
        if (deeper) {
            PairLoad C[2];
            const uint32_t codes[2][2] = {{n0a, n0b}, {n1a, n1b}};
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                C[c] = P;  // the query's record: the same for the children
                load_code(meshes, nodes, codes[c][0], C[c].A);
                load_code(meshes, nodes, codes[c][1], C[c].B);
            }
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                uint32_t ccost = 0;
                const int ck = eval_loaded<REF>(C[c], qi, g[c][0], g[c][1], g[c][2], g[c][3], lf[c], ccost);
                has_leaf[c] = ck == PAIR_LEAF;
                dest[c] = ck == PAIR_SPLIT ? (ccost > split_tris2 ? 1 : 2) : 0;
            }
            nsat += 2;
        }
        const unsigned long long mL0 = __ballot(has_leaf[0]), mL1 = __ballot(has_leaf[1]);
        const unsigned long long mN0 = __ballot(dest[0] == 1), mN1 = __ballot(dest[1] == 1);
        const unsigned long long mT0 = __ballot(dest[0] == 2), mT1 = __ballot(dest[1] == 2);
        nleaf += uint32_t(has_leaf[0]) + uint32_t(has_leaf[1]);
        if (lane == 0) {
            s_n[wave][0] = unsigned(__popcll(mL0) + __popcll(mL1));
            s_n[wave][1] = 2u * unsigned(__popcll(mN0) + __popcll(mN1));
            s_n[wave][2] = 2u * unsigned(__popcll(mT0) + __popcll(mT1));
        }
        __syncthreads();
        if (threadIdx.x < 3) {
            const int c = int(threadIdx.x);
            const unsigned int tot = (s_n[0][c] + s_n[1][c]) + (s_n[2][c] + s_n[3][c]);
            unsigned long long b0 = 0;
            if (tot) {
                if (c == 0) b0 = atomicAdd(&ctr->lpslots, static_cast<unsigned long long>(tot));
                else b0 = atomicAdd(c == 1 ? nout : ntasks, tot);
            }
            s_base[c] = b0;
        }
        __syncthreads();
        unsigned long long first[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            unsigned long long f = s_base[c];
            for (int w = 0; w < wave; ++w) f += s_n[w][c];
            first[c] = f;
        }
        unsigned long long sl = first[0] + unsigned(__popcll(mL0 & below) + __popcll(mL1 & below));
        unsigned long long sn = first[1] + 2u * unsigned(__popcll(mN0 & below) + __popcll(mN1 & below));
        unsigned long long st = first[2] + 2u * unsigned(__popcll(mT0 & below) + __popcll(mT1 & below));
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            if (has_leaf[c]) {
                if (sl < lp_cap) lp_out[sl] = lf[c];
                ++sl;
            }
            if (dest[c] != 0) {
                unsigned long long& slot = dest[c] == 1 ? sn : st;
                uint3* dst = dest[c] == 1 ? out : tasks;
                if (slot + 1 < out_cap) {
                    dst[slot] = make_uint3(qi, g[c][0], g[c][1]);
                    dst[slot + 1] = make_uint3(qi, g[c][2], g[c][3]);
                } else {
                    atomicOr(reinterpret_cast<unsigned int*>(&ctr->err), 2u);
                }
                slot += 2;
            }
        }
        __syncthreads();  // s_n and s_base are rewritten by the next iteration
    }
    block_add(nsat, &ctr->sat);
    block_add(nleaf, &ctr->leafpairs);
}

// Work sharing (share != 0).  A wavefront that runs out of tasks registers
// as idle and polls a donation queue (the first POLLERS of them; later ones
// leave); busy wavefronts that see idle ones push the bottom half of their
// LDS stack (the shallowest,
// hence largest, node pairs) into it.  Entries are written and read with
// atomics only (no cache-wide fences, which write back or drop a whole L2),
// the ready word last.  Every wait is bounded: an idle
// wavefront leaves after spin_polls empty polls (or as soon as every
// wavefront is idle), and a claimed entry whose ready word has not appeared
// after READY_POLLS polls is reported (err bit 8) instead of waited for.  A
// donor writes its entries right after reserving them and waits on nothing,
// so neither bound is reached in practice.  Entries nobody claimed are taken
// by a second launch with sharing off.
// dq_ctl: [0] entries reserved by donors, [1] entries claimed, [2] wavefronts
// idle or gone, [3] pollers waiting.
constexpr uint32_t READY_POLLS = 1u << 16;
constexpr uint32_t DONATE_MAX = 256;
constexpr uint32_t POLLERS = 128;  // idle wavefronts that wait for donations; later ones leave

// Read a word other wavefronts update, through the device's coherence point.
// The L2 caches of the dies are not coherent with each other for ordinary
// device memory, so a plain (or relaxed atomic) load can keep returning a
// line this die cached earlier; a read-modify-write cannot.
__device__ __forceinline__ unsigned int coherent_read(unsigned int* p) {
    unsigned int zero = 0u;
    asm volatile("" : "+v"(zero));  // opaque to the optimiser, which would turn "add 0" into a plain load
    return atomicAdd(p, zero);
}

template <int CAP, bool REF>
__global__ __launch_bounds__(64) void k_dfs(const uint3* __restrict__ tasks, const uint4* __restrict__ tasks4,
                                            const unsigned int* __restrict__ task_lo_p,
                                            const unsigned int* __restrict__ ntasks_p, uint32_t task_cap,
                                            unsigned int* __restrict__ head, const QueryPre* __restrict__ qp,
                                            const MeshRec* __restrict__ meshes, const NodeRec* __restrict__ nodes,
                                            uint4* __restrict__ lp_out, uint64_t lp_cap, uint3* __restrict__ spill_base,
                                            uint32_t spill_cap, uint32_t pop_width, uint4* __restrict__ dq,
                                            unsigned int* __restrict__ dq_ctl, uint32_t dq_cap, uint32_t share,
                                            uint32_t spin_polls, const unsigned long long* __restrict__ budget_p,
                                            uint32_t budget_pct, uint3* __restrict__ dump,
                                            unsigned int* __restrict__ ndump, uint32_t dump_cap,
                                            RunCounters* __restrict__ ctr) {
    __shared__ uint32_t sq[CAP], sa[CAP], sb[CAP];
    const int lane = threadIdx.x;
    uint3* spill = spill_base + size_t(blockIdx.x) * spill_cap;
    // tasks [lo, hi): the split levels' list, or (second launch) what nobody claimed from the queue
    const uint32_t lo = task_lo_p ? *task_lo_p : 0u;
    const uint32_t hi = min(*ntasks_p, task_cap);
    const uint32_t ntasks = hi > lo ? hi - lo : 0u;
    const uint32_t grab = max(1u, min(64u, ntasks / (2u * gridDim.x)));
    uint32_t size = 0, nspill = 0;
    bool tasks_done = false, overflow = false, lost = false, poller = false;
    // Budgeted pass (share == 3): after `budget` node-pair tests a wavefront
    // hands what is left on its stack to `dump` and leaves; the budget is
    // budget_pct percent of the previous run's tests per wavefront (none on
    // a first run).
    uint64_t budget = 0, wave_tests = 0;
    if (share == 3u && budget_p) budget = (*budget_p / gridDim.x) * budget_pct / 100u;
    unsigned int idle_read = 0, idle_seen = 0, step = 0;  // lane 0's last read of the idle count, and its broadcast
    unsigned long long nsat = 0, nleaf = 0;
    // Leaf-pair slots are taken LP_CHUNK at a time, so the shared counter
    // sees one atomic per chunk rather than one per step.
    uint64_t lp_next = 0, lp_end = 0;
    const unsigned long long below = lanes_below_b();
    for (;;) {
        // The idle count is read every 8th step, one step ahead of its use,
        // so the (device-coherent, hence slow) load hides behind the step's
        // own loads and a thousand wavefronts do not queue on one address.
        if (share == 2u) {
            // Hand-off: once half the grid has finished, the wavefronts still
            // busy write their whole stack (LDS and spill) to the queue and
            // leave; the second launch shares the queue out over the grid.
            if ((step & 7u) == 0u) {
                idle_seen = __shfl(idle_read, 0, 64);
                if (lane == 0) idle_read = coherent_read(&dq_ctl[2]);  // wavefronts finished
            }
            ++step;
            if (idle_seen >= gridDim.x / 2u && size + nspill > 0u) {
                const uint32_t n = size + nspill;
                unsigned int base = 0xFFFFFFFFu;
                if (lane == 0) {
                    unsigned int cur = coherent_read(&dq_ctl[0]);
                    while (cur + n <= dq_cap) {
                        const unsigned int prev = atomicCAS(&dq_ctl[0], cur, cur + n);
                        if (prev == cur) {
                            base = cur;
                            break;
                        }
                        cur = prev;
                    }
                }
                base = __shfl(base, 0, 64);
                if (base != 0xFFFFFFFFu) {  // read by the next launch only: plain stores suffice
                    for (uint32_t i = lane; i < size; i += 64) dq[base + i] = make_uint4(sq[i], sa[i], sb[i], 1u);
                    for (uint32_t i = lane; i < nspill; i += 64) {
                        const uint3 e = spill[i];
                        dq[base + size + i] = make_uint4(e.x, e.y, e.z, 1u);
                    }
                    size = 0;
                    nspill = 0;
                    tasks_done = true;  // the task list is empty by now: half the grid ran out of it
                    break;
                }
            }
        }
        if (share == 1u) {
            if ((step & 7u) == 0u) {
                idle_seen = __shfl(idle_read, 0, 64);
                if (lane == 0) idle_read = coherent_read(&dq_ctl[3]);  // pollers waiting
            }
            ++step;
            if (idle_seen > 0 && size >= 128u) {  // donate the bottom half of the stack
                const uint32_t k = min(size / 2u, DONATE_MAX);
                unsigned int base = 0xFFFFFFFFu;
                if (lane == 0) {  // reserve without overshooting the capacity (each failed CAS means another wave won)
                    unsigned int cur = coherent_read(&dq_ctl[0]);
                    while (cur + k <= dq_cap) {
                        const unsigned int prev = atomicCAS(&dq_ctl[0], cur, cur + k);
                        if (prev == cur) {
                            base = cur;
                            break;
                        }
                        cur = prev;
                    }
                }
                base = __shfl(base, 0, 64);
                if (base != 0xFFFFFFFFu) {
                    for (uint32_t i = lane; i < k; i += 64) {
                        // every word goes through the coherence point (atomics),
                        // the ready word only after the others have been performed
                        unsigned int* e = reinterpret_cast<unsigned int*>(dq + base + i);
                        const unsigned int r0 = atomicExch(&e[0], sq[i]);
                        const unsigned int r1 = atomicExch(&e[1], sa[i]);
                        const unsigned int r2 = atomicExch(&e[2], sb[i]);
                        asm volatile("" ::"v"(r0), "v"(r1), "v"(r2));  // wait for their completion
                        atomicExch(&e[3], 1u);
                    }
                    __syncthreads();
                    for (uint32_t i = 0; i + k < size; i += 64) {  // shift the rest down
                        const uint32_t j = i + lane;
                        uint32_t tq = 0, ta = 0, tb = 0;
                        if (j + k < size) { tq = sq[j + k]; ta = sa[j + k]; tb = sb[j + k]; }
                        __syncthreads();
                        if (j + k < size) { sq[j] = tq; sa[j] = ta; sb[j] = tb; }
                        __syncthreads();
                    }
                    size -= k;
                    idle_seen = 0;  // once per read of the idle count
                }
            }
        }
        if (size < 64 && nspill > 0) {  // bring back the spilled bottom block underneath
            const uint32_t k = min(uint32_t(CAP / 2), nspill);
            uint32_t tq = 0, ta = 0, tb = 0;
            if (uint32_t(lane) < size) { tq = sq[lane]; ta = sa[lane]; tb = sb[lane]; }
            __syncthreads();
            if (uint32_t(lane) < size) { sq[k + lane] = tq; sa[k + lane] = ta; sb[k + lane] = tb; }
            for (uint32_t i = lane; i < k; i += 64) {
                const uint3 e = spill[nspill - k + i];
                sq[i] = e.x; sa[i] = e.y; sb[i] = e.z;
            }
            nspill -= k;
            size += k;
            __syncthreads();
        }
        if (size < 64 && nspill == 0 && !tasks_done) {
            // take a fair share, not a full stack: a wave that grabbed 64 tasks
            // at once would leave most of the grid idle
            const uint32_t want = min(64 - size, grab);
            uint32_t t0 = 0;
            if (lane == 0) t0 = atomicAdd(head, want);
            t0 = __shfl(t0, 0, 64);
            const uint32_t got = t0 < ntasks ? min(want, ntasks - t0) : 0;
            if (got < want) tasks_done = true;
            if (uint32_t(lane) < got) {
                // Tasks are taken in a fixed pseudo-random order, k -> k P mod n
                // (P = 2^31 - 1 is prime, so for n < P this is a permutation):
                // neighbouring tasks come from the same query and cost alike, and
                // runs of them would pile one query's heavy subtrees onto a few
                // wavefronts.
                const uint32_t k = t0 + lane;
                const uint32_t idx =
                    lo + (ntasks < 2147483647u ? uint32_t((uint64_t(k) * 2147483647ull) % ntasks) : k);
                uint3 e;
                if (tasks4) {
                    const uint4 f = tasks4[idx];
                    e = make_uint3(f.x, f.y, f.z);
                } else {
                    e = tasks[idx];
                }
                sq[size + lane] = e.x; sa[size + lane] = e.y; sb[size + lane] = e.z;
            }
            size += got;
            __syncthreads();
        }
        if (size == 0) {
            if (!(tasks_done && nspill == 0)) continue;
            if (share != 1u) break;
            // Out of work.  Only the first POLLERS idle wavefronts wait
            // (boundedly) for donations; the others leave at once: a thousand
            // wavefronts polling one line would slow every atomic the busy
            // ones issue.
            unsigned int idle_before = 0;
            if (lane == 0) idle_before = atomicAdd(&dq_ctl[2], 1u);
            idle_before = __shfl(idle_before, 0, 64);
            if (!poller && idle_before >= POLLERS) break;
            poller = true;  // a poller keeps coming back here after each donation it works off
            if (lane == 0) atomicAdd(&dq_ctl[3], 1u);  // a poller is waiting
            bool leave = false;
            for (uint32_t polls = 0;; ++polls) {
                unsigned int base = 0, got = 0;
                if (lane == 0) {
                    const unsigned int t = min(coherent_read(&dq_ctl[0]), dq_cap);
                    unsigned int h = t ? coherent_read(&dq_ctl[1]) : 0u;
                    while (h < t) {  // each failed CAS means another wave claimed
                        const unsigned int k = min(t - h, 64u);
                        const unsigned int prev = atomicCAS(&dq_ctl[1], h, h + k);
                        if (prev == h) {
                            base = h;
                            got = k;
                            break;
                        }
                        h = prev;
                    }
                }
                got = __shfl(got, 0, 64);
                base = __shfl(base, 0, 64);
                if (got) {
                    bool ok = false;
                    uint4 e = make_uint4(0, 0, 0, 0);
                    if (uint32_t(lane) < got) {
                        unsigned int* w = reinterpret_cast<unsigned int*>(dq + base + lane);
                        for (uint32_t r = 0; r < READY_POLLS; ++r) {
                            if (coherent_read(&w[3])) {
                                ok = true;
                                break;
                            }
                            __builtin_amdgcn_s_sleep(1);
                        }
                        if (ok) e = make_uint4(coherent_read(&w[0]), coherent_read(&w[1]), coherent_read(&w[2]), 1u);
                    }
                    const unsigned long long mok = __ballot(ok);
                    if (__ballot(uint32_t(lane) < got && !ok)) lost = true;
                    if (ok) {
                        const uint32_t pos = uint32_t(__popcll(mok & below));
                        sq[pos] = e.x; sa[pos] = e.y; sb[pos] = e.z;
                    }
                    size = uint32_t(__popcll(mok));
                    __syncthreads();
                    if (lane == 0) {
                        atomicSub(&dq_ctl[3], 1u);
                        atomicSub(&dq_ctl[2], 1u);
                    }
                    break;
                }
                unsigned int idle_now = 0;
                if (lane == 0) idle_now = coherent_read(&dq_ctl[2]);
                idle_now = __shfl(idle_now, 0, 64);
                if (idle_now >= gridDim.x || polls + 1 >= spin_polls) {  // everyone idle or gone, or waited long enough
                    if (lane == 0) atomicSub(&dq_ctl[3], 1u);
                    leave = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(64);
            }
            if (leave) break;
            continue;
        }
        if (budget && wave_tests >= budget) {  // hand the rest of this stack on, then leave
            const uint32_t n = size + nspill;
            unsigned int base = 0;
            if (lane == 0) base = atomicAdd(ndump, n);
            base = __shfl(base, 0, 64);
            if (base + n <= dump_cap && base + n >= base) {
                for (uint32_t i = lane; i < size; i += 64) dump[base + i] = make_uint3(sq[i], sa[i], sb[i]);
                for (uint32_t i = lane; i < nspill; i += 64) dump[base + size + i] = spill[i];
            } else if (lane == 0) {
                atomicOr(reinterpret_cast<unsigned int*>(&ctr->err), 2u);  // out of frontier room: rerun larger
            }
            size = 0;
            nspill = 0;
            break;
        }
        const uint32_t p = min(size, pop_width);
        wave_tests += p;
        const bool have = uint32_t(lane) < p;
        uint32_t qi = 0, ca = 0, cb = 0;
        if (have) {
            const uint32_t idx = size - p + lane;
            qi = sq[idx]; ca = sa[idx]; cb = sb[idx];
        }
        size -= p;
        __syncthreads();
        int kind = PAIR_NONE;
        uint32_t n0a = 0, n0b = 0, n1a = 0, n1b = 0;
        uint4 leaf = make_uint4(0, 0, 0, 0);
        if (have) {
            ++nsat;
            uint32_t cost = 0;
            kind = eval_pair<REF>(qp, meshes, nodes, qi, ca, cb, n0a, n0b, n1a, n1b, leaf, cost);
        }
        const unsigned long long mL = __ballot(kind == PAIR_LEAF);
        if (mL) {
            const uint32_t nl = uint32_t(__popcll(mL));
            if (lp_next + nl > lp_end) {
                close_chunk(lp_out, lp_cap, lp_next, lp_end, lane);
                unsigned long long c0 = 0;
                if (lane == 0) c0 = atomicAdd(&ctr->lpslots, static_cast<unsigned long long>(LP_CHUNK));
                lp_next = __shfl(c0, 0, 64);
                lp_end = lp_next + LP_CHUNK;
            }
            const uint64_t s = lp_next + uint64_t(__popcll(mL & below));
            if (kind == PAIR_LEAF && s < lp_cap) lp_out[s] = leaf;
            lp_next += nl;
            nleaf += nl;
        }
        const unsigned long long mS = __ballot(kind == PAIR_SPLIT);
        const uint32_t ns = uint32_t(__popcll(mS));
        if (size + 2 * ns > uint32_t(CAP)) {  // spill the bottom half of the stack to this wave's HBM area
            const uint32_t k = CAP / 2;
            if (nspill + k <= spill_cap) {
                for (uint32_t i = lane; i < k; i += 64) spill[nspill + i] = make_uint3(sq[i], sa[i], sb[i]);
                __syncthreads();
                for (uint32_t i = 0; i + k < size; i += 64) {
                    const uint32_t j = i + lane;
                    uint32_t tq = 0, ta = 0, tb = 0;
                    if (j + k < size) { tq = sq[j + k]; ta = sa[j + k]; tb = sb[j + k]; }
                    __syncthreads();
                    if (j + k < size) { sq[j] = tq; sa[j] = ta; sb[j] = tb; }
                    __syncthreads();
                }
                nspill += k;
                size -= k;
            } else {
                overflow = true;  // cannot happen below ~spill_cap/64 levels; reported, never silently dropped
                break;
            }
        }
        if (kind == PAIR_SPLIT) {
            const uint32_t pos = size + 2u * uint32_t(__popcll(mS & below));
            sq[pos] = qi; sa[pos] = n1a; sb[pos] = n1b;
            sq[pos + 1] = qi; sa[pos + 1] = n0a; sb[pos + 1] = n0b;
        }
        size += 2 * ns;
        __syncthreads();
    }
    close_chunk(lp_out, lp_cap, lp_next, lp_end, lane);
    nsat = wave_sum(nsat);
    if (lane == 0) {
        atomicAdd(&ctr->sat, nsat);
        if (nleaf) atomicAdd(&ctr->leafpairs, nleaf);
        if (overflow) atomicOr(reinterpret_cast<unsigned int*>(&ctr->err), 4u);
        if (lost) atomicOr(reinterpret_cast<unsigned int*>(&ctr->err), 8u);
        if (share == 2u) atomicAdd(&dq_ctl[2], 1u);
    }
}

// Pipelined persistent traversal (no work sharing).  A wavefront keeps two
// batches of node pairs in flight: it pops the next batch and issues its
// record loads before it tests the current one, so the loads of one batch
// overlap the other's arithmetic instead of following it.  The next batch
// is popped before the current one's children are pushed, which changes
// only the order of the depth-first walk: every node pair is still tested
// exactly once and the leaf pairs are the same set.  The loop body is
// written twice with the two register sets' roles swapped, so no copy makes
// the compiler wait for a batch's loads before they are needed.
template <int CAP, bool REF>
__global__ __launch_bounds__(64) void k_dfs_pipe(const uint3* __restrict__ tasks,
                                                 const unsigned int* __restrict__ ntasks_p, uint32_t task_cap,
                                                 unsigned int* __restrict__ head, const QueryPre* __restrict__ qp,
                                                 const MeshRec* __restrict__ meshes, const NodeRec* __restrict__ nodes,
                                                 uint4* __restrict__ lp_out, uint64_t lp_cap,
                                                 uint3* __restrict__ spill_base, uint32_t spill_cap,
                                                 uint32_t pop_width, RunCounters* __restrict__ ctr) {
    __shared__ uint32_t sq[CAP], sa[CAP], sb[CAP];
    const int lane = threadIdx.x;
    uint3* spill = spill_base + size_t(blockIdx.x) * spill_cap;
    const uint32_t ntasks = min(*ntasks_p, task_cap);
    // a fair share per grab, not a full stack (see k_dfs); smaller grabs late
    // in the run cost more than they balance: every refill's returning atomic
    // waits for the batch loads in flight
    const uint32_t grab = max(1u, min(64u, ntasks / (2u * gridDim.x)));
    uint32_t size = 0, nspill = 0;
    bool tasks_done = false, overflow = false;
    unsigned long long nsat = 0, nleaf = 0;
    uint64_t lp_next = 0, lp_end = 0;
    const unsigned long long below = lanes_below_b();

    // Refill a low stack: the spilled bottom block first, else new tasks.
    auto refill = [&]() {
        if (size < 64 && nspill > 0) {
            const uint32_t k = min(uint32_t(CAP / 2), nspill);
            uint32_t tq = 0, ta = 0, tb = 0;
            if (uint32_t(lane) < size) { tq = sq[lane]; ta = sa[lane]; tb = sb[lane]; }
            __syncthreads();
            if (uint32_t(lane) < size) { sq[k + lane] = tq; sa[k + lane] = ta; sb[k + lane] = tb; }
            for (uint32_t i = lane; i < k; i += 64) {
                const uint3 e = spill[nspill - k + i];
                sq[i] = e.x; sa[i] = e.y; sb[i] = e.z;
            }
            nspill -= k;
            size += k;
            __syncthreads();
        }
        if (size < 64 && nspill == 0 && !tasks_done) {
            const uint32_t want = min(64 - size, grab);
            uint32_t t0 = 0;
            if (lane == 0) t0 = atomicAdd(head, want);
            t0 = __shfl(t0, 0, 64);
            const uint32_t got = t0 < ntasks ? min(want, ntasks - t0) : 0;
            if (got < want) tasks_done = true;
            if (uint32_t(lane) < got) {
                const uint32_t k = t0 + lane;  // the pseudo-random task order of k_dfs
                const uint32_t idx = ntasks < 2147483647u ? uint32_t((uint64_t(k) * 2147483647ull) % ntasks) : k;
                const uint3 e = tasks[idx];
                sq[size + lane] = e.x; sa[size + lane] = e.y; sb[size + lane] = e.z;
            }
            size += got;
            __syncthreads();
        }
    };
    // Pop up to pop_width node pairs and issue their loads.  Every lane
    // loads, a lane without an entry the root pair of query 0 (a valid
    // record: the kernel runs only for at least one query), and every lane
    // tests: with no lane-dependent branch around them, the compiler sees
    // that each batch's loads are consumed by the next test and needs no
    // early wait for them before issuing the following batch's.
    auto pop = [&](PairLoad& L, uint32_t& q, bool& h) {
        const uint32_t p = min(size, pop_width);
        h = uint32_t(lane) < p;
        const uint32_t idx = h ? size - p + lane : 0u;
        const uint32_t eq = sq[idx], ea = sa[idx], eb = sb[idx];
        q = h ? eq : 0u;
        const uint32_t ca = h ? ea : ROOT_FLAG, cb = h ? eb : ROOT_FLAG;
        size -= p;
        __syncthreads();
        load_pair(qp, meshes, nodes, q, ca, cb, L);
        // keep the previous batch's arithmetic below these loads: hoisted
        // above them, it would wait for its own loads before they are issued
        __builtin_amdgcn_sched_barrier(0);
    };
    // Test a loaded batch: leaf pairs out, children onto the stack.
    auto process = [&](const PairLoad& L, uint32_t q, bool h) {
        uint32_t n0a = 0, n0b = 0, n1a = 0, n1b = 0, cost = 0;
        uint4 leaf = make_uint4(0, 0, 0, 0);
        const int k_all = eval_loaded<REF>(L, q, n0a, n0b, n1a, n1b, leaf, cost);
        const int kind = h ? k_all : int(PAIR_NONE);
        nsat += h ? 1u : 0u;
        const unsigned long long mL = __ballot(kind == PAIR_LEAF);
        if (mL) {
            const uint32_t nl = uint32_t(__popcll(mL));
            if (lp_next + nl > lp_end) {
                close_chunk(lp_out, lp_cap, lp_next, lp_end, lane);
                unsigned long long c0 = 0;
                if (lane == 0) c0 = atomicAdd(&ctr->lpslots, static_cast<unsigned long long>(LP_CHUNK));
                lp_next = __shfl(c0, 0, 64);
                lp_end = lp_next + LP_CHUNK;
            }
            const uint64_t sl = lp_next + uint64_t(__popcll(mL & below));
            if (kind == PAIR_LEAF && sl < lp_cap) lp_out[sl] = leaf;
            lp_next += nl;
            nleaf += nl;
        }
        const unsigned long long mS = __ballot(kind == PAIR_SPLIT);
        const uint32_t ns = uint32_t(__popcll(mS));
        if (size + 2 * ns > uint32_t(CAP)) {  // spill the bottom half of the stack to this wave's HBM area
            const uint32_t k = CAP / 2;
            if (nspill + k <= spill_cap) {
                for (uint32_t i = lane; i < k; i += 64) spill[nspill + i] = make_uint3(sq[i], sa[i], sb[i]);
                __syncthreads();
                for (uint32_t i = 0; i + k < size; i += 64) {
                    const uint32_t j = i + lane;
                    uint32_t tq = 0, ta = 0, tb = 0;
                    if (j + k < size) { tq = sq[j + k]; ta = sa[j + k]; tb = sb[j + k]; }
                    __syncthreads();
                    if (j + k < size) { sq[j] = tq; sa[j] = ta; sb[j] = tb; }
                    __syncthreads();
                }
                nspill += k;
                size -= k;
            } else {
                overflow = true;  // reported, never silently dropped
                return;
            }
        }
        if (kind == PAIR_SPLIT) {
            const uint32_t pos = size + 2u * uint32_t(__popcll(mS & below));
            sq[pos] = q; sa[pos] = n1a; sb[pos] = n1b;
            sq[pos + 1] = q; sa[pos + 1] = n0a; sb[pos + 1] = n0b;
        }
        size += 2 * ns;
        __syncthreads();
    };
    PairLoad L0, L1;
    uint32_t q0 = 0, q1 = 0;
    bool h0 = false, h1 = false;
    refill();
    pop(L0, q0, h0);
    for (;;) {
        refill();
        pop(L1, q1, h1);
        process(L0, q0, h0);
        if (overflow || (size == 0 && nspill == 0 && tasks_done && __ballot(h1) == 0ull)) break;
        refill();
        pop(L0, q0, h0);
        process(L1, q1, h1);
        if (overflow || (size == 0 && nspill == 0 && tasks_done && __ballot(h0) == 0ull)) break;
    }
    close_chunk(lp_out, lp_cap, lp_next, lp_end, lane);
    nsat = wave_sum(nsat);
    if (lane == 0) {
        atomicAdd(&ctr->sat, nsat);
        if (nleaf) atomicAdd(&ctr->leafpairs, nleaf);
        if (overflow) atomicOr(reinterpret_cast<unsigned int*>(&ctr->err), 4u);
    }
}

// Hits of the batched narrow phase are recorded as {query, global A slot,
// global B slot, exact sides}; k_contacts turns record i into contact i.
// Building the contact record inside the test kernels would set their
// register peak (and so their occupancy) for the 1-2% of tests that hit.
constexpr uint32_t HSTAGE = 64;  // hit records a wavefront stages in LDS before one atomic reserves their slots

__device__ __forceinline__ void flush_hits(uint4* __restrict__ hits, uint64_t cap, const uint4* staged, uint32_t n,
                                           RunCounters* __restrict__ ctr, int lane) {
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(&ctr->contacts, static_cast<unsigned long long>(n));
    base = __shfl(base, 0, 64);
    if (uint32_t(lane) < n && base + lane < cap) hits[base + lane] = staged[lane];
}

// Stages the records of the lanes with `hit` set, flushing first when they
// do not fit.  Every lane calls it.
__device__ __forceinline__ void stage_hits(bool hit, uint4 rec, uint4* stage, uint32_t& staged,
                                           uint4* __restrict__ hits, uint64_t cap, RunCounters* __restrict__ ctr,
                                           int lane, unsigned long long below) {
    const unsigned long long m = __ballot(hit);
    if (!m) return;
    const uint32_t nh = uint32_t(__popcll(m));
    if (staged + nh > HSTAGE) {
        __threadfence_block();  // this wavefront's LDS writes are complete before other lanes copy them
        flush_hits(hits, cap, stage, staged, ctr, lane);
        staged = 0;
        __threadfence_block();
    }
    if (hit) stage[staged + uint32_t(__popcll(m & below))] = rec;
    staged += nh;
}

// Triangle pair of a test: A in its mesh's space, B moved into it by the
// query's matrix; ga, gb are arena-global triangle slots.
__device__ __forceinline__ void narrow_load(const QueryPre& Q, const float* __restrict__ tris, uint32_t ga, uint32_t gb,
                                            f3& A0, f3& A1, f3& A2, f3& B0, f3& B1, f3& B2) {
    M4 M;
#pragma unroll
    for (int e = 0; e < 16; ++e) M.m[e] = Q.m[e];
    const float* a = tris + size_t(ga) * 9;
    const float* b = tris + size_t(gb) * 9;
    A0 = mk(a[0], a[1], a[2]);
    A1 = mk(a[3], a[4], a[5]);
    A2 = mk(a[6], a[7], a[8]);
    B0 = xform_point(M, mk(b[0], b[1], b[2]));
    B1 = xform_point(M, mk(b[3], b[4], b[5]));
    B2 = xform_point(M, mk(b[6], b[7], b[8]));
}

// Appends the tests of the lanes with `defer` set to a rerun list (one
// atomic per wavefront).  The count keeps growing past the capacity, so the
// host sees how much room the run needed.  Every lane calls it.
__device__ __forceinline__ void defer_tests(bool defer, uint4 e, uint4* __restrict__ list, uint64_t list_cap,
                                            unsigned long long* __restrict__ count, int lane,
                                            unsigned long long below) {
    const unsigned long long m = __ballot(defer);
    if (!m) return;
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(count, static_cast<unsigned long long>(__popcll(m)));
    base = __shfl(base, 0, 64);
    if (defer) {
        const uint64_t s = base + uint64_t(__popcll(m & below));
        if (s < list_cap) list[s] = e;
    }
}

// Triangle tests of the leaf pairs, float32 predicates only: a test that a
// float32 predicate cannot decide goes to the rerun list (entries {query,
// global A slot, global B slot}), which k_narrow_rerun<1> and <2> finish in
// float64 and exactly.  Keeping the slow stages out of this kernel keeps its
// registers and scratch small.
//
// The test's three parts run as separate full-width batches: part 1 over
// every test, part 2 over the tests part 1 leaves open (~30%), part 3 over
// those part 2 leaves open (~6%).  Open tests wait in two queues per
// wavefront in LDS ({query, A slot, B slot, signs so far}); a queue is
// worked off 64 at a time once it holds 64, and drained at the end.  Run as
// one sequence per lane, nearly every wavefront would carry a lane through
// all three parts, and every test would pay for them.
constexpr uint32_t NQCAP = 128;  // entries of each queue: under 64 before a batch adds at most 64

// Appends e for the lanes with `on` set to a wavefront queue of n entries.
__device__ __forceinline__ void queue_push(bool on, uint4 e, uint4* q, uint32_t& n, unsigned long long below) {
    const unsigned long long m = __ballot(on);
    if (on) q[n + uint32_t(__popcll(m & below))] = e;
    n += uint32_t(__popcll(m));
    __threadfence_block();  // the entries are in LDS before any lane reads them
}

// One batch of part 2 (W == 2) or part 3 (W == 3) over the top (up to) 64
// entries of queue q; part 2 moves what stays open to queue q3.
template <int W>
__device__ __forceinline__ void narrow_batch(uint4* q, uint32_t& n, uint4* q3, uint32_t& n3,
                                             const QueryPre* __restrict__ qp, const float* __restrict__ tris,
                                             uint4* __restrict__ hits, uint64_t cap,
                                             uint4* __restrict__ defer_list, uint64_t defer_cap,
                                             RunCounters* __restrict__ ctr, uint4* stage, uint32_t& staged,
                                             int lane, unsigned long long below) {
    const uint32_t m = min(n, 64u);
    const bool have = uint32_t(lane) < m;
    uint4 e = make_uint4(0, 0, 0, 0);
    if (have) e = q[n - m + uint32_t(lane)];
    n -= m;
    __threadfence_block();  // read before the slots are written again
    bool hit = false, defer = false, more = false;
    uint32_t ds = e.w;
    if (have) {
        f3 A0, A1, A2, B0, B1, B2;
        narrow_load(qp[e.x], tris, e.y, e.z, A0, A1, A2, B0, B1, B2);
        PredStats st;
        const int res = W == 2 ? tt_part2<0>(A0, A1, A2, B0, B1, B2, ds, st)
                               : tt_part3<0>(A0, A1, A2, B0, B1, B2, ds, st);
        defer = res == DEFERRED;
        more = res == TT_MORE;
        hit = res == 1;
    }
    defer_tests(defer, make_uint4(e.x, e.y, e.z, 0), defer_list, defer_cap, &ctr->deferred, lane, below);
    if (W == 2) queue_push(more, make_uint4(e.x, e.y, e.z, ds), q3, n3, below);
    if (W == 3) stage_hits(hit, make_uint4(e.x, e.y, e.z, ds), stage, staged, hits, cap, ctr, lane, below);
}

__global__ __launch_bounds__(256) void k_narrow_exact(const uint4* __restrict__ lps,
                                                      const unsigned long long* __restrict__ nlp_p, uint64_t lp_cap,
                                                      const QueryPre* __restrict__ qp, const float* __restrict__ tris,
                                                      uint4* __restrict__ hits, uint64_t cap,
                                                      uint4* __restrict__ defer_list, uint64_t defer_cap,
                                                      RunCounters* __restrict__ ctr) {
    __shared__ uint4 s_hit[4][HSTAGE];
    __shared__ uint4 s_q2[4][NQCAP], s_q3[4][NQCAP];
    const uint64_t nlp = min(*nlp_p, lp_cap);
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    uint4* stage = s_hit[w];
    uint4* q2 = s_q2[w];
    uint4* q3 = s_q3[w];
    uint32_t staged = 0, n2 = 0, n3 = 0;
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint64_t wave = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = (uint64_t(gridDim.x) * blockDim.x) >> 6;
    unsigned long long ntt = 0;
    for (uint64_t chunk = wave; chunk * 64 < nlp; chunk += nwaves) {
        const uint64_t li = chunk * 64 + lane;
        uint4 lp = make_uint4(0, 0, 0, 0);
        if (li < nlp) lp = lps[li];
        const uint32_t cA = lp.w & 0xFFFFu, cB = lp.w >> 16;
        const uint32_t cnt = cA * cB;
        uint32_t incl = cnt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t v = __shfl_up(incl, off, 64);
            if (lane >= off) incl += v;
        }
        const uint32_t excl = incl - cnt;
        const uint32_t total = __shfl(incl, 63, 64);
        for (uint32_t t0 = 0; t0 < total; t0 += 64) {
            const uint32_t t = t0 + lane;
            const bool active = t < total;
            int k = 0;
#pragma unroll
            for (int step = 32; step > 0; step >>= 1) {
                const int cand = k + step;
                const uint32_t oc = __shfl(excl, cand & 63, 64);
                if (cand < 64 && oc <= t) k = cand;
            }
            const uint32_t kq = __shfl(lp.x, k, 64), ka = __shfl(lp.y, k, 64), kb = __shfl(lp.z, k, 64);
            const uint32_t kw = __shfl(lp.w, k, 64), ko = __shfl(excl, k, 64);
            bool defer = false, more = false;
            uint32_t ds = 0, ga = 0, gb = 0;
            if (active) {
                ++ntt;
                const uint32_t kcB = kw >> 16;
                const uint32_t r = t - ko;
                ga = ka + r / kcB;  // ka, kb: arena-global triangle offsets
                gb = kb + r % kcB;
                f3 A0, A1, A2, B0, B1, B2;
                narrow_load(qp[kq], tris, ga, gb, A0, A1, A2, B0, B1, B2);
                PredStats st;
                const int res = tt_part1<0>(A0, A1, A2, B0, B1, B2, ds, st);
                defer = res == DEFERRED;
                more = res == TT_MORE;
            }
            defer_tests(defer, make_uint4(kq, ga, gb, 0), defer_list, defer_cap, &ctr->deferred, lane, below);
            queue_push(more, make_uint4(kq, ga, gb, ds), q2, n2, below);
            if (n2 >= 64)
                narrow_batch<2>(q2, n2, q3, n3, qp, tris, hits, cap, defer_list, defer_cap, ctr, stage, staged, lane,
                                below);
            if (n3 >= 64)
                narrow_batch<3>(q3, n3, q3, n3, qp, tris, hits, cap, defer_list, defer_cap, ctr, stage, staged, lane,
                                below);
        }
    }
    while (n2 > 0) {
        narrow_batch<2>(q2, n2, q3, n3, qp, tris, hits, cap, defer_list, defer_cap, ctr, stage, staged, lane, below);
        if (n3 >= 64)
            narrow_batch<3>(q3, n3, q3, n3, qp, tris, hits, cap, defer_list, defer_cap, ctr, stage, staged, lane,
                            below);
    }
    while (n3 > 0)
        narrow_batch<3>(q3, n3, q3, n3, qp, tris, hits, cap, defer_list, defer_cap, ctr, stage, staged, lane, below);
    if (staged) {
        __threadfence_block();
        flush_hits(hits, cap, stage, staged, ctr, lane);
    }
    block_add(ntt, &ctr->tritests);
}

// Narrow phase with the reference's semantics over the leaf pairs of the
// batched traversal (CreateUncollideRays.cpp:74-115).  The leaf pairs of a
// wavefront's 64 are expanded densely to triangle tests over its lanes (as
// k_narrow_exact does).  Each lane runs the plane stage of the reference's
// Möller test (tri_tri_planes, the very code tri_tri_isect starts with) on the
// B triangle moved by the query's matrix in the reference's operation order;
// the few pairs it leaves open queue in LDS and are run through the whole
// test 64 at a time (ref_batch), which recomputes the plane stage with the
// same result and then writes a contact record (mesh-local triangle slots,
// planes, outside flags, segment) for each non-coplanar hit.  The tests, and
// so the contacts and counts, are the plain path's; only their order differs.
__device__ __forceinline__ void ref_batch(uint4* q, uint32_t& n, const QueryPre* __restrict__ qp,
                                          const float* __restrict__ tris, cd_contact* __restrict__ out, uint64_t cap,
                                          RunCounters* __restrict__ ctr, unsigned long long& nub, int lane,
                                          unsigned long long below) {
    const uint32_t m = min(n, 64u);
    const bool have = uint32_t(lane) < m;
    uint4 e = make_uint4(0, 0, 0, 0);
    if (have) e = q[n - m + uint32_t(lane)];
    n -= m;
    __threadfence_block();  // read before the slots are written again
    bool hit = false;
    cd_contact c;
    if (have) {
        const QueryPre& Q = qp[e.x];
        f3 A0, A1, A2, B0, B1, B2;
        narrow_load(Q, tris, e.y, e.z, A0, A1, A2, B0, B1, B2);
        f3 sp, tp;
        const uint32_t res = tri_tri_isect(A0, A1, A2, B0, B1, B2, sp, tp);
        nub += (res & TT_UB) ? 1u : 0u;
        hit = (res & TT_HIT) && !(res & TT_COPLANAR);
        if (hit) {
            const PlaneF pa = plane_of(A0, A1, A2), pb = plane_of(B0, B1, B2);
            c.query = e.x;
            c.tri_a = e.y - Q.tri_base_a;
            c.tri_b = e.z - Q.tri_base_b;
            c.flags = outside_flags(pa, pb, A0, A1, A2, B0, B1, B2);
            c.s[0] = sp.x; c.s[1] = sp.y; c.s[2] = sp.z;
            c.t[0] = tp.x; c.t[1] = tp.y; c.t[2] = tp.z;
            c.na[0] = pa.n.x; c.na[1] = pa.n.y; c.na[2] = pa.n.z;
            c.nb[0] = pb.n.x; c.nb[1] = pb.n.y; c.nb[2] = pb.n.z;
        }
    }
    const unsigned long long mh = __ballot(hit);
    if (mh) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(&ctr->contacts, static_cast<unsigned long long>(__popcll(mh)));
        const unsigned long long slot = __shfl(base, 0, 64) + static_cast<unsigned long long>(__popcll(mh & below));
        if (hit && slot < cap) out[slot] = c;
    }
}
__global__ __launch_bounds__(256) void k_narrow_ref(const uint4* __restrict__ lps,
                                                    const unsigned long long* __restrict__ nlp_p, uint64_t lp_cap,
                                                    const QueryPre* __restrict__ qp, const float* __restrict__ tris,
                                                    cd_contact* __restrict__ out, uint64_t cap,
                                                    RunCounters* __restrict__ ctr) {
    __shared__ uint4 s_q[4][NQCAP];
    const uint64_t nlp = min(*nlp_p, lp_cap);
    const int lane = threadIdx.x & 63;
    uint4* q = s_q[threadIdx.x >> 6];
    uint32_t nqd = 0;
    const unsigned long long below = lanes_below_b();
    const uint64_t wave = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = (uint64_t(gridDim.x) * blockDim.x) >> 6;
    unsigned long long ntt = 0, nub = 0;
    for (uint64_t chunk = wave; chunk * 64 < nlp; chunk += nwaves) {
        const uint64_t li = chunk * 64 + lane;
        uint4 lp = make_uint4(0, 0, 0, 0);
        if (li < nlp) lp = lps[li];
        const uint32_t cnt = (lp.w & 0xFFFFu) * (lp.w >> 16);
        uint32_t incl = cnt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t v = __shfl_up(incl, off, 64);
            if (lane >= off) incl += v;
        }
        const uint32_t excl = incl - cnt;
        const uint32_t total = __shfl(incl, 63, 64);
        for (uint32_t t0 = 0; t0 < total; t0 += 64) {
            const uint32_t t = t0 + lane;
            int k = 0;  // the leaf pair holding test t: the last with excl <= t
#pragma unroll
            for (int step = 32; step > 0; step >>= 1) {
                const int cand = k + step;
                const uint32_t oc = __shfl(excl, cand & 63, 64);
                if (cand < 64 && oc <= t) k = cand;
            }
            const uint32_t kq = __shfl(lp.x, k, 64), ka = __shfl(lp.y, k, 64), kb = __shfl(lp.z, k, 64);
            const uint32_t kw = __shfl(lp.w, k, 64), ko = __shfl(excl, k, 64);
            bool open = false;
            uint32_t ga = 0, gb = 0;
            if (t < total) {
                ++ntt;
                const uint32_t kcB = kw >> 16;
                const uint32_t r = t - ko;
                ga = ka + r / kcB;  // arena-global triangle slots
                gb = kb + r % kcB;
                f3 A0, A1, A2, B0, B1, B2;
                narrow_load(qp[kq], tris, ga, gb, A0, A1, A2, B0, B1, B2);
                TTPlanes pl;
                open = tri_tri_planes(A0, A1, A2, B0, B1, B2, pl);
            }
            queue_push(open, make_uint4(kq, ga, gb, 0u), q, nqd, below);
            if (nqd >= 64) ref_batch(q, nqd, qp, tris, out, cap, ctr, nub, lane, below);
        }
    }
    while (nqd > 0) ref_batch(q, nqd, qp, tris, out, cap, ctr, nub, lane, below);
    block_add(ntt, &ctr->tritests);
    block_add(nub, &ctr->ub);
}

// Rerun of deferred triangle tests at precision level L (1: float64, whose
// undecided tests go on to `next`; 2: exact).  Wavefront-stride over the
// list, whose length is read on the device.
template <int L>
__global__ __launch_bounds__(256) void k_narrow_rerun(const uint4* __restrict__ list,
                                                      const unsigned long long* __restrict__ nlist_p,
                                                      uint64_t list_cap, const QueryPre* __restrict__ qp,
                                                      const float* __restrict__ tris, uint4* __restrict__ hits,
                                                      uint64_t cap, uint4* __restrict__ next,
                                                      unsigned long long* __restrict__ nnext,
                                                      RunCounters* __restrict__ ctr) {
    __shared__ uint4 s_hit[4][HSTAGE];
    const uint64_t n = min(*nlist_p, list_cap);
    const int lane = threadIdx.x & 63;
    uint4* stage = s_hit[threadIdx.x >> 6];
    uint32_t staged = 0;
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint64_t wave = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = (uint64_t(gridDim.x) * blockDim.x) >> 6;
    unsigned long long nfp64 = 0, nexact = 0;
    for (uint64_t chunk = wave; chunk * 64 < n; chunk += nwaves) {
        const uint64_t li = chunk * 64 + lane;
        bool hit = false, defer = false;
        uint32_t ds = 0;
        uint4 e = make_uint4(0, 0, 0, 0);
        if (li < n) {
            e = list[li];
            f3 A0, A1, A2, B0, B1, B2;
            narrow_load(qp[e.x], tris, e.y, e.z, A0, A1, A2, B0, B1, B2);
            PredStats st;
            const int res = tri_tri_exact<L>(A0, A1, A2, B0, B1, B2, ds, st);
            hit = res == 1;
            defer = res == DEFERRED;
            if (!defer) {
                nfp64 += st.fp64;
                nexact += st.exact;
            }
        }
        if (L < 2) defer_tests(defer, e, next, list_cap, nnext, lane, below);
        stage_hits(hit, make_uint4(e.x, e.y, e.z, ds), stage, staged, hits, cap, ctr, lane, below);
    }
    if (staged) {
        __threadfence_block();
        flush_hits(hits, cap, stage, staged, ctr, lane);
    }
    block_add(nfp64, &ctr->fp64);
    block_add(nexact, &ctr->exact);
}

// Contact record of every hit record (record i -> contact i).
__global__ __launch_bounds__(256) void k_contacts(const uint4* __restrict__ hits,
                                                  const unsigned long long* __restrict__ nhits_p, uint64_t cap,
                                                  const QueryPre* __restrict__ qp, const float* __restrict__ tris,
                                                  cd_contact* __restrict__ out) {
    const uint64_t n = min(*nhits_p, cap);
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
        const uint4 e = hits[i];
        const QueryPre& Q = qp[e.x];
        f3 A0, A1, A2, B0, B1, B2;
        narrow_load(Q, tris, e.y, e.z, A0, A1, A2, B0, B1, B2);
        write_contact(out + i, A0, A1, A2, B0, B1, B2, e.w, e.x, e.y - Q.tri_base_a, e.z - Q.tri_base_b);
    }
}

// Exact-semantics stream of triangle pairs: flag 1 intersect (non-coplanar),
// 2 coplanar, 0 disjoint.
//
// A cascade of three launches, one precision level each, so that every
// wavefront runs one kind of arithmetic: the float32 pass decides almost
// every pair and appends the few it cannot decide to a list; the float64
// pass reruns those, appending what it cannot decide to a second list; the
// last pass reruns that remainder with the exact expansion.  A pair that
// needs float64 in one predicate would otherwise hold its whole wavefront on
// the slow path.  Each rerun repeats the pair from the start, so a pair's
// counters come from the pass that decides it and total the same as a
// single full-precision pass.
//
// The passes are persistent (grid-stride, block-uniform loops).  A block
// stages the pairs it defers in LDS and moves them to the list BSTAGE at a
// time with one atomic, so lists are dense and a pass makes few atomics.
constexpr uint32_t BSTAGE = 2048;

struct BlockStage {
    uint32_t buf[BSTAGE];
    uint32_t n;
    unsigned long long base;
};

__device__ __forceinline__ void stage_flush(BlockStage& bs, uint32_t* __restrict__ list, uint64_t cap,
                                            unsigned long long* __restrict__ count) {
    const uint32_t c = bs.n;  // every thread reads it after a barrier
    if (c == 0) return;
    if (threadIdx.x == 0) bs.base = atomicAdd(count, static_cast<unsigned long long>(c));
    __syncthreads();
    const unsigned long long b = bs.base;
    for (uint32_t i = threadIdx.x; i < c; i += blockDim.x)
        if (b + i < cap) list[b + i] = bs.buf[i];
    __syncthreads();
    if (threadIdx.x == 0) bs.n = 0;
    __syncthreads();
}

// Stages v for the threads with `on` set; every thread of the block calls
// it once per loop iteration (at most blockDim.x entries per call).
__device__ __forceinline__ void stage_push(bool on, uint32_t v, BlockStage& bs, uint32_t* __restrict__ list,
                                           uint64_t cap, unsigned long long* __restrict__ count, int lane,
                                           unsigned long long below) {
    const unsigned long long m = __ballot(on);
    uint32_t off = 0;
    if (lane == 0 && m) off = atomicAdd(&bs.n, uint32_t(__popcll(m)));
    off = __shfl(off, 0, 64);
    if (on) bs.buf[off + uint32_t(__popcll(m & below))] = v;
    __syncthreads();
    const uint32_t c = bs.n;
    __syncthreads();  // every thread holds c before any wavefront adds to bs.n again
    if (c > BSTAGE - blockDim.x) stage_flush(bs, list, cap, count);
}

__device__ __forceinline__ void tt_load(const float* __restrict__ a9, const float* __restrict__ b9, uint64_t i, f3& A0,
                                        f3& A1, f3& A2, f3& B0, f3& B1, f3& B2) {
    const float* a = a9 + i * 9;
    const float* b = b9 + i * 9;
    A0 = mk(a[0], a[1], a[2]); A1 = mk(a[3], a[4], a[5]); A2 = mk(a[6], a[7], a[8]);
    B0 = mk(b[0], b[1], b[2]); B1 = mk(b[3], b[4], b[5]); B2 = mk(b[6], b[7], b[8]);
}

__device__ __forceinline__ void tt_store(uint64_t i, int res, uint32_t ds, f3 A0, f3 A1, f3 A2, f3 B0, f3 B1, f3 B2,
                                         int32_t* __restrict__ flags, float* __restrict__ seg6) {
    flags[i] = res == 1 ? 1 : (res == 2 ? 2 : 0);
    if (seg6) {
        f3 s = mk(0.f, 0.f, 0.f), t = mk(0.f, 0.f, 0.f);
        if (res == 1) exact_segment(A0, A1, A2, B0, B1, B2, ds, s, t);
        float* o = seg6 + i * 6;
        o[0] = s.x; o[1] = s.y; o[2] = s.z;
        o[3] = t.x; o[4] = t.y; o[5] = t.z;
    }
}

__global__ __launch_bounds__(256) void k_tritri_f32(const float* __restrict__ a9, const float* __restrict__ b9,
                                                    uint64_t n, int32_t* __restrict__ flags,
                                                    float* __restrict__ seg6, uint32_t* __restrict__ list,
                                                    unsigned long long* __restrict__ nlist) {
    __shared__ BlockStage bs;
    if (threadIdx.x == 0) bs.n = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    for (uint64_t base = uint64_t(blockIdx.x) * blockDim.x; base < n; base += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t i = base + threadIdx.x;
        bool defer = false;
        if (i < n) {
            f3 A0, A1, A2, B0, B1, B2;
            tt_load(a9, b9, i, A0, A1, A2, B0, B1, B2);
            PredStats st;
            uint32_t ds = 0;
            const int res = tri_tri_exact<0>(A0, A1, A2, B0, B1, B2, ds, st);
            defer = res == DEFERRED;
            if (!defer) tt_store(i, res, ds, A0, A1, A2, B0, B1, B2, flags, seg6);
        }
        stage_push(defer, uint32_t(i), bs, list, n, nlist, lane, below);
    }
    stage_flush(bs, list, n, nlist);
}

// Reruns of listed pairs at precision level L (grid-stride; the list length
// is read on the device).
template <int L>
__global__ __launch_bounds__(256) void k_tritri_rerun(const float* __restrict__ a9, const float* __restrict__ b9,
                                                      const uint32_t* __restrict__ list,
                                                      const unsigned long long* __restrict__ nlist,
                                                      int32_t* __restrict__ flags, float* __restrict__ seg6,
                                                      uint64_t n, uint32_t* __restrict__ next,
                                                      unsigned long long* __restrict__ nnext,
                                                      RunCounters* __restrict__ ctr) {
    __shared__ BlockStage bs;
    if (L < 2 && threadIdx.x == 0) bs.n = 0;
    __syncthreads();
    const uint64_t count = min(static_cast<uint64_t>(*nlist), n);
    const int lane = threadIdx.x & 63;
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    unsigned long long nfp64 = 0, nexact = 0;
    for (uint64_t base = uint64_t(blockIdx.x) * blockDim.x; base < count; base += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t k = base + threadIdx.x;
        bool defer = false;
        uint32_t i = 0;
        if (k < count) {
            i = list[k];
            f3 A0, A1, A2, B0, B1, B2;
            tt_load(a9, b9, i, A0, A1, A2, B0, B1, B2);
            PredStats st;
            uint32_t ds = 0;
            const int res = tri_tri_exact<L>(A0, A1, A2, B0, B1, B2, ds, st);
            defer = res == DEFERRED;
            if (!defer) {
                tt_store(i, res, ds, A0, A1, A2, B0, B1, B2, flags, seg6);
                nfp64 += st.fp64;
                nexact += st.exact;
            }
        }
        if (L < 2) stage_push(defer, i, bs, next, n, nnext, lane, below);
    }
    if (L < 2) stage_flush(bs, next, n, nnext);
    block_add(nfp64, &ctr->fp64);
    block_add(nexact, &ctr->exact);
}

// Node-pair tests of this run, for the next run's budgeted pass.
__global__ void k_keep_budget(const RunCounters* __restrict__ ctr, unsigned long long* __restrict__ budget) {
    if (threadIdx.x == 0) *budget = ctr->sat;
}

// ---- host launchers -------------------------------------------------------
// nq_cap sizes the grids; the kernels read the query count from nq_p (<= nq_cap),
// so one captured graph serves every count up to its capacity
hipError_t launch_start(const QueryRec* qs, uint32_t nq_cap, const unsigned int* nq_p, const MeshRec* meshes,
                        QueryPre* qp, uint3* frontier, unsigned int* small, uint32_t nsmall, RunCounters* ctr,
                        hipStream_t st) {
    const uint32_t n = std::max(std::max(nq_cap, nsmall), 1u);
    hipLaunchKernelGGL(k_start, dim3((n + 127) / 128), dim3(128), 0, st, qs, nq_p, meshes, qp, frontier, small, nsmall,
                       ctr);
    return hipGetLastError();
}
hipError_t launch_split_level(const uint3* in, const unsigned int* nin, uint3* out, unsigned int* nout, uint32_t out_cap,
                              const QueryPre* qp, const MeshRec* meshes, const NodeRec* nodes, uint4* lp,
                              uint64_t lp_cap, uint3* tasks, unsigned int* ntasks, uint32_t split_tris,
                              RunCounters* ctr, uint32_t grid, hipStream_t st, bool ref) {
    if (ref)
        hipLaunchKernelGGL(k_split_level<true>, dim3(grid), dim3(256), 0, st, in, nin, out, nout, out_cap, qp, meshes,
                           nodes, lp, lp_cap, tasks, ntasks, split_tris, ctr);
    else
        hipLaunchKernelGGL(k_split_level<false>, dim3(grid), dim3(256), 0, st, in, nin, out, nout, out_cap, qp, meshes,
                           nodes, lp, lp_cap, tasks, ntasks, split_tris, ctr);
    return hipGetLastError();
}
hipError_t launch_split_level2(const uint3* in, const unsigned int* nin, uint3* out, unsigned int* nout,
                               uint32_t out_cap, const QueryPre* qp, const MeshRec* meshes, const NodeRec* nodes,
                               uint4* lp, uint64_t lp_cap, uint3* tasks, unsigned int* ntasks, uint32_t split_tris,
                               uint32_t split_tris2, RunCounters* ctr, uint32_t grid, hipStream_t st, bool ref) {
    if (ref)
        hipLaunchKernelGGL(k_split_level2<true>, dim3(grid), dim3(256), 0, st, in, nin, out, nout, out_cap, qp, meshes,
                           nodes, lp, lp_cap, tasks, ntasks, split_tris, split_tris2, ctr);
    else
        hipLaunchKernelGGL(k_split_level2<false>, dim3(grid), dim3(256), 0, st, in, nin, out, nout, out_cap, qp, meshes,
                           nodes, lp, lp_cap, tasks, ntasks, split_tris, split_tris2, ctr);
    return hipGetLastError();
}
hipError_t launch_dfs(int cap_entries, const uint3* tasks, const uint4* tasks4, const unsigned int* task_lo,
                      const unsigned int* ntasks, uint32_t task_cap, unsigned int* head, const QueryPre* qp,
                      const MeshRec* meshes, const NodeRec* nodes, uint4* lp, uint64_t lp_cap, uint3* spill,
                      uint32_t spill_cap, uint32_t pop_width, uint32_t nwaves, uint4* dq, unsigned int* dq_ctl,
                      uint32_t dq_cap, uint32_t share, uint32_t spin_polls, const unsigned long long* budget,
                      uint32_t budget_pct, uint3* dump, unsigned int* ndump, uint32_t dump_cap, RunCounters* ctr,
                      hipStream_t st, bool pipelined, bool ref) {
    if (pipelined && share == 0u && tasks && !tasks4 && !task_lo) {  // the pipelined kernel: plain task list, no sharing
#define CD_PIPE_ARGS tasks, ntasks, task_cap, head, qp, meshes, nodes, lp, lp_cap, spill, spill_cap, pop_width, ctr
#define CD_PIPE(C, R) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_dfs_pipe<C, R>), dim3(nwaves), dim3(64), 0, st, CD_PIPE_ARGS)
        switch (cap_entries) {
            case 512: if (ref) CD_PIPE(512, true); else CD_PIPE(512, false); break;
            case 2048: if (ref) CD_PIPE(2048, true); else CD_PIPE(2048, false); break;
            default: if (ref) CD_PIPE(1024, true); else CD_PIPE(1024, false); break;
        }
#undef CD_PIPE
#undef CD_PIPE_ARGS
        return hipGetLastError();
    }
#define CD_DFS_ARGS tasks, tasks4, task_lo, ntasks, task_cap, head, qp, meshes, nodes, lp, lp_cap, spill, spill_cap, \
                    pop_width, dq, dq_ctl, dq_cap, share, spin_polls, budget, budget_pct, dump, ndump, dump_cap, ctr
#define CD_DFS(C, R) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_dfs<C, R>), dim3(nwaves), dim3(64), 0, st, CD_DFS_ARGS)
    switch (cap_entries) {
        case 512: if (ref) CD_DFS(512, true); else CD_DFS(512, false); break;
        case 2048: if (ref) CD_DFS(2048, true); else CD_DFS(2048, false); break;
        default: if (ref) CD_DFS(1024, true); else CD_DFS(1024, false); break;
    }
#undef CD_DFS
#undef CD_DFS_ARGS
    return hipGetLastError();
}
hipError_t launch_keep_budget(const RunCounters* ctr, unsigned long long* budget, hipStream_t st) {
    hipLaunchKernelGGL(k_keep_budget, dim3(1), dim3(64), 0, st, ctr, budget);
    return hipGetLastError();
}
// Workgroups of 256 threads of `kernel` that one CU keeps resident (its
// registers and LDS decide).  The narrow phase's grid-stride kernels are
// launched with exactly that many per CU: with more, the extra ones wait
// for a second round that starts as the first ones end, and the phase
// lasts about a third longer.
template <class K>
static uint32_t resident_per_cu(K kernel) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 256, 0) != hipSuccess || nb < 1) nb = 1;
    return uint32_t(nb);
}
struct NarrowShapes {
    uint32_t exact, contacts, ref;
};
static const NarrowShapes& narrow_shapes() {
    static const NarrowShapes s = {resident_per_cu(k_narrow_exact), resident_per_cu(k_contacts),
                                   resident_per_cu(k_narrow_ref)};
    return s;
}
// The same for the split levels (reference or conservative test, one or
// two levels per launch).
uint32_t split_blocks_per_cu(bool ref, bool two_levels) {
    static const uint32_t b[4] = {resident_per_cu(k_split_level<false>), resident_per_cu(k_split_level<true>),
                                  resident_per_cu(k_split_level2<false>), resident_per_cu(k_split_level2<true>)};
    return b[(two_levels ? 2 : 0) + (ref ? 1 : 0)];
}
void warm_launch_shapes() {
    (void)narrow_shapes();
    (void)split_blocks_per_cu(false, false);
}

hipError_t launch_narrow_exact(const uint4* lps, const unsigned long long* nlp, uint64_t lp_cap, const QueryPre* qp,
                               const float* tris, uint4* hits, cd_contact* out, uint64_t cap, uint4* defer,
                               uint64_t defer_cap, RunCounters* ctr, uint32_t cus, hipStream_t st) {
    const NarrowShapes& sh = narrow_shapes();
    hipLaunchKernelGGL(k_narrow_exact, dim3(cus * sh.exact), dim3(256), 0, st, lps, nlp, lp_cap, qp, tris, hits, cap,
                       defer, defer_cap, ctr);
    hipLaunchKernelGGL(k_narrow_rerun<1>, dim3(cus), dim3(256), 0, st, defer, &ctr->deferred, defer_cap, qp, tris, hits,
                       cap, defer + defer_cap, &ctr->deferred2, ctr);
    hipLaunchKernelGGL(k_narrow_rerun<2>, dim3(cus), dim3(256), 0, st, defer + defer_cap, &ctr->deferred2, defer_cap,
                       qp, tris, hits, cap, (uint4*)nullptr, (unsigned long long*)nullptr, ctr);
    hipLaunchKernelGGL(k_contacts, dim3(cus * sh.contacts), dim3(256), 0, st, hits, &ctr->contacts, cap, qp, tris, out);
    return hipGetLastError();
}
hipError_t launch_narrow_ref(const uint4* lps, const unsigned long long* nlp, uint64_t lp_cap, const QueryPre* qp,
                             const float* tris, cd_contact* out, uint64_t cap, RunCounters* ctr, uint32_t cus,
                             hipStream_t st) {
    hipLaunchKernelGGL(k_narrow_ref, dim3(cus * narrow_shapes().ref), dim3(256), 0, st, lps, nlp, lp_cap, qp, tris,
                       out, cap, ctr);
    return hipGetLastError();
}
hipError_t launch_tritri_exact(const float* a9, const float* b9, uint64_t n, int32_t* flags, float* seg6,
                               uint32_t* lists, uint32_t grid, RunCounters* ctr, hipStream_t st) {
    if (!n) return hipSuccess;
    // lists: words 0-3 the two list lengths (64-bit, zeroed here), the
    // float32 pass's list (n entries) at word 64, the float64 pass's after it
    unsigned long long* nl = reinterpret_cast<unsigned long long*>(lists);
    uint32_t* l1 = lists + 64;
    uint32_t* l2 = lists + 64 + n;
    hipError_t e = launch_fill_u32(lists, 4, 0, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_tritri_f32, dim3(grid), dim3(256), 0, st, a9, b9, n, flags, seg6, l1, nl);
    hipLaunchKernelGGL(k_tritri_rerun<1>, dim3(grid), dim3(256), 0, st, a9, b9, l1, nl, flags, seg6, n, l2, nl + 1,
                       ctr);
    hipLaunchKernelGGL(k_tritri_rerun<2>, dim3(grid), dim3(256), 0, st, a9, b9, l2, nl + 1, flags, seg6, n,
                       (uint32_t*)nullptr, (unsigned long long*)nullptr, ctr);
    return hipGetLastError();
}

}  // namespace cdk
