// Plain path: reference semantics bit for bit.
//
//  k_tritri_plain   Triangle::IntersectTriangles over a stream of pairs
//                   (inMyRoom_vulkan/src/Geometry/Triangle.cpp:85-110)
//  k_traverse_plain OBBtree::IntersectOBBtreesRecursive, one lane per query,
//                   explicit stack in LDS, same visiting order
//                   (inMyRoom_vulkan/src/Geometry/OBBtree.cpp:414-477)
//  k_narrow_plain   the i x j triangle loop of CreateUncollideRays
//                   (inMyRoom_vulkan/src/CollisionDetection/CreateUncollideRays.cpp:74-115),
//                   contacts compacted with one atomic per wavefront
#include "internal.hpp"
#include "kernels.hpp"

namespace cdk {

__device__ __forceinline__ unsigned long long lanes_below() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Reserve popcount(mask) slots with one atomic per wavefront; returns this
// lane's slot (valid only where its bit of mask is set).
__device__ __forceinline__ unsigned long long wave_reserve(unsigned long long mask, unsigned long long* counter) {
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(counter, static_cast<unsigned long long>(__popcll(mask)));
    base = __shfl(base, leader, 64);
    return base + static_cast<unsigned long long>(__popcll(mask & lanes_below()));
}

__device__ __forceinline__ f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }

__global__ __launch_bounds__(256) void k_tritri_plain(const float* __restrict__ a9, const float* __restrict__ b9,
                                                      uint64_t n, int32_t* __restrict__ flags,
                                                      float* __restrict__ seg6) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* a = a9 + i * 9;
    const float* b = b9 + i * 9;
    f3 s = mk(0.f, 0.f, 0.f), t = mk(0.f, 0.f, 0.f);
    const uint32_t r = tri_tri_isect(ld3(a), ld3(a + 3), ld3(a + 6), ld3(b), ld3(b + 3), ld3(b + 6), s, t);
    flags[i] = int32_t(r);
    if (seg6) {
        float* o = seg6 + i * 6;
        const bool keep = (r & TT_HIT) && !(r & TT_COPLANAR);
        o[0] = keep ? s.x : 0.f;
        o[1] = keep ? s.y : 0.f;
        o[2] = keep ? s.z : 0.f;
        o[3] = keep ? t.x : 0.f;
        o[4] = keep ? t.y : 0.f;
        o[5] = keep ? t.z : 0.f;
    }
}

// Traversal codes: ROOT_CODE for a tree's root, else (node << 1) | side for
// the child box stored on side `side` of internal node `node`.
struct TreeView {
    const MeshRec* m;
    const NodeRec* nodes;  // this mesh's nodes (arena + node_base)
};

__device__ __forceinline__ const ChildRec& code_rec(const TreeView& t, uint32_t code) {
    return code == ROOT_CODE ? t.m->root : t.nodes[code >> 1].c[code & 1];
}
__device__ __forceinline__ void code_box(const TreeView& t, uint32_t code, Box& b, float& surf) {
    const ChildRec& c = code_rec(t, code);
    b = box_from12(c.box);
    surf = c.surf;
}
// leaf?  If so, triangle offset/count; else the child node index (0 for the
// root, the tree's node 0).
__device__ __forceinline__ bool code_leaf(const TreeView& t, uint32_t code, uint32_t& link, uint32_t& cnt) {
    const ChildRec& c = code_rec(t, code);
    link = c.link;
    cnt = child_cnt(c);
    return child_leaf(c);
}

// DEEP: the lanes' stacks live in HBM, `cap` entries each (entry e of lane l
// of block b at gstack[(b cap + e) 64 + l]), for trees deeper than the LDS
// stack holds; the walk is the same.
template <bool DEEP>
__global__ __launch_bounds__(64) void k_traverse_plain(const QueryRec* __restrict__ qs, uint32_t nq,
                                                       const MeshRec* __restrict__ meshes,
                                                       const NodeRec* __restrict__ nodes, uint4* __restrict__ lp_out,
                                                       uint64_t lp_cap, RunCounters* __restrict__ ctr,
                                                       uint2* __restrict__ gstack, uint32_t cap) {
    __shared__ uint2 lstack[DEEP ? 1 : PLAIN_STACK][64];
    uint2* stack = DEEP ? gstack + size_t(blockIdx.x) * cap * 64 : &lstack[0][0];
    const int scap = DEEP ? int(cap) : PLAIN_STACK;
    const int lane = threadIdx.x;
    const uint32_t qi = blockIdx.x * 64 + lane;
    if (qi >= nq) return;
    const QueryRec q = qs[qi];
    M4 M;
#pragma unroll
    for (int k = 0; k < 16; ++k) M.m[k] = q.rel[k];
    const TreeView A{&meshes[q.mesh_a], nodes + meshes[q.mesh_a].node_base};
    const TreeView B{&meshes[q.mesh_b], nodes + meshes[q.mesh_b].node_base};

    int sp = 0;
    stack[sp++ * 64 + lane] = make_uint2(ROOT_CODE, ROOT_CODE);
    unsigned long long nsat = 0, nlp = 0;
    bool overflow = false;
    while (sp > 0) {
        const uint2 top = stack[--sp * 64 + lane];
        const uint32_t ca = top.x, cb = top.y;
        Box fa, sb0;
        float sfa, unused;
        code_box(A, ca, fa, sfa);
        code_box(B, cb, sb0, unused);
        const Box sb = xform_box(M, sb0);
        ++nsat;
        if (!sat(fa, sb)) continue;
        uint32_t la_link, la_cnt, lb_link, lb_cnt;
        const bool la = code_leaf(A, ca, la_link, la_cnt);
        const bool lb = code_leaf(B, cb, lb_link, lb_cnt);
        if (la && lb) {
            const unsigned long long slot = atomicAdd(&ctr->leafpairs, 1ull);
            if (slot < lp_cap) lp_out[slot] = make_uint4(qi, la_link, lb_link, la_cnt | (lb_cnt << 16));
            ++nlp;
            continue;
        }
        if (sp + 2 > scap) {
            overflow = true;
            break;
        }
        // descend the first tree when it is the larger (OBBtree.cpp:426-447)
        const bool split_a = !la && (lb || sfa >= surface(sb));
        if (split_a) {
            stack[sp++ * 64 + lane] = make_uint2((la_link << 1) | 1u, cb);
            stack[sp++ * 64 + lane] = make_uint2((la_link << 1) | 0u, cb);
        } else {
            stack[sp++ * 64 + lane] = make_uint2(ca, (lb_link << 1) | 1u);
            stack[sp++ * 64 + lane] = make_uint2(ca, (lb_link << 1) | 0u);
        }
    }
    atomicAdd(&ctr->sat, nsat);
    if (overflow) atomicOr(reinterpret_cast<unsigned int*>(&ctr->err), 1u);
    (void)nlp;
}

__global__ __launch_bounds__(256) void k_narrow_plain(const uint4* __restrict__ lps, uint64_t nlp, uint32_t leaf_max,
                                                      const QueryRec* __restrict__ qs,
                                                      const MeshRec* __restrict__ meshes,
                                                      const float* __restrict__ tris, cd_contact* __restrict__ out,
                                                      uint64_t cap, RunCounters* __restrict__ ctr) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool valid = i < nlp;
    uint4 lp = make_uint4(0, 0, 0, 0);
    if (valid) lp = lps[i];
    const uint32_t qi = lp.x, cntA = lp.w & 0xFFFFu, cntB = lp.w >> 16;
    M4 M;
    const float* TA = tris;
    const float* TB = tris;
    if (valid) {
        const QueryRec& q = qs[qi];
#pragma unroll
        for (int k = 0; k < 16; ++k) M.m[k] = q.rel[k];
        TA = tris + (size_t(meshes[q.mesh_a].tri_base) + lp.y) * 9;
        TB = tris + (size_t(meshes[q.mesh_b].tri_base) + lp.z) * 9;
    }
    unsigned long long ntt = 0, nub = 0;
    const uint32_t L2 = leaf_max * leaf_max;
    for (uint32_t t = 0; t < L2; ++t) {
        const uint32_t ia = t / leaf_max, jb = t % leaf_max;
        const bool active = valid && ia < cntA && jb < cntB;
        bool hit = false;
        cd_contact c;
        if (active) {
            ++ntt;
            const float* a = TA + ia * 9;
            const float* b = TB + jb * 9;
            const f3 A0 = ld3(a), A1 = ld3(a + 3), A2 = ld3(a + 6);
            const f3 B0 = xform_point(M, ld3(b)), B1 = xform_point(M, ld3(b + 3)), B2 = xform_point(M, ld3(b + 6));
            f3 s, tt;
            const uint32_t r = tri_tri_isect(A0, A1, A2, B0, B1, B2, s, tt);
            nub += (r & TT_UB) ? 1 : 0;
            hit = (r & TT_HIT) && !(r & TT_COPLANAR);
            if (hit) {
                const PlaneF pa = plane_of(A0, A1, A2), pb = plane_of(B0, B1, B2);
                c.query = qi;
                c.tri_a = lp.y + ia;
                c.tri_b = lp.z + jb;
                c.flags = outside_flags(pa, pb, A0, A1, A2, B0, B1, B2);
                c.s[0] = s.x; c.s[1] = s.y; c.s[2] = s.z;
                c.t[0] = tt.x; c.t[1] = tt.y; c.t[2] = tt.z;
                c.na[0] = pa.n.x; c.na[1] = pa.n.y; c.na[2] = pa.n.z;
                c.nb[0] = pb.n.x; c.nb[1] = pb.n.y; c.nb[2] = pb.n.z;
            }
        }
        const unsigned long long m = __ballot(hit);
        if (m) {
            const unsigned long long slot = wave_reserve(m, &ctr->contacts);
            if (hit && slot < cap) out[slot] = c;
        }
    }
    // one atomic per wavefront for the counters
    for (int off = 32; off > 0; off >>= 1) {
        ntt += __shfl_xor(ntt, off, 64);
        nub += __shfl_xor(nub, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&ctr->tritests, ntt);
        if (nub) atomicAdd(&ctr->ub, nub);
    }
}

__global__ void k_zero_counters(RunCounters* c) {
    if (threadIdx.x == 0) *c = RunCounters{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
}

__global__ void k_fill_u32(unsigned int* p, uint32_t n, unsigned int v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---- host launchers -------------------------------------------------------
hipError_t launch_fill_u32(unsigned int* p, uint32_t n, unsigned int v, hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}
hipError_t launch_zero(RunCounters* c, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, st, c);
    return hipGetLastError();
}
hipError_t launch_tritri_plain(const float* a9, const float* b9, uint64_t n, int32_t* flags, float* seg6,
                               hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_tritri_plain, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, st, a9, b9, n, flags, seg6);
    return hipGetLastError();
}
hipError_t launch_traverse_plain(const QueryRec* qs, uint32_t nq, const MeshRec* meshes, const NodeRec* nodes,
                                 uint4* lp, uint64_t lp_cap, RunCounters* ctr, hipStream_t st, uint32_t need,
                                 uint2* gstack) {
    if (!nq) return hipSuccess;
    if (need <= uint32_t(PLAIN_STACK))
        hipLaunchKernelGGL(k_traverse_plain<false>, dim3((nq + 63) / 64), dim3(64), 0, st, qs, nq, meshes, nodes, lp,
                           lp_cap, ctr, (uint2*)nullptr, 0u);
    else if (!gstack)
        return hipErrorInvalidValue;
    else
        hipLaunchKernelGGL(k_traverse_plain<true>, dim3((nq + 63) / 64), dim3(64), 0, st, qs, nq, meshes, nodes, lp,
                           lp_cap, ctr, gstack, need);
    return hipGetLastError();
}
hipError_t launch_narrow_plain(const uint4* lps, uint64_t nlp, uint32_t leaf_max, const QueryRec* qs,
                               const MeshRec* meshes, const float* tris, cd_contact* out, uint64_t cap,
                               RunCounters* ctr, hipStream_t st) {
    if (!nlp) return hipSuccess;
    hipLaunchKernelGGL(k_narrow_plain, dim3(uint32_t((nlp + 255) / 256)), dim3(256), 0, st, lps, nlp, leaf_max, qs,
                       meshes, tris, out, cap, ctr);
    return hipGetLastError();
}

}  // namespace cdk
