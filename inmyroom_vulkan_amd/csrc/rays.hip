// Contact aggregation and uncollide-ray generation
// (CreateUncollideRays::ExecuteCreateUncollideRays,
//  inMyRoom_vulkan/src/CollisionDetection/CreateUncollideRays.cpp:60-200).
//
// Input: the contacts of a run sorted by (query, tri_a, tri_b), and a
// permutation of them sorted by (query, tri_b, tri_a).  Every launch runs
// both sides (grid y: the side, each side with its own part of the buffers,
// RaySides); for each side:
//  k_cand        one thread per run of contacts with the same (query,
//                triangle): weight |s - t|, weighted midpoint and the
//                vertex-not-outside mask, summed per leaf of the other tree
//                and merged leaf after leaf (:74-127).  The reference merges
//                in the order its DFS emits leaf pairs, and for a fixed leaf
//                of one tree that order meets the other tree's leaves in
//                tree order, which is the order of the sorted run: these
//                sums are the reference's, bit for bit.  A leaf whose
//                contacts all have zero weight is skipped, as IsNull() skips
//                it (:119, :125).
//  k_compact     dense candidate list (exclusive scan of the run starts)
//  k_vert_claim  the not-outside vertices go into a hash set keyed by
//                (query, side, vertex id); the smallest (triangle, corner)
//                claims each vertex (:146-152 keep the first one the
//                unordered_map walk meets, an order no GPU can reproduce)
//  k_ray_count / k_ray_emit  one ray per claimed vertex (origin the vertex,
//                direction minus its normal), or one at the weighted contact
//                point with the barycentric normal when no vertex is
//                inside (:155-167); the second entity's rays are in the
//                first's space (:135, :146, :160)
//  k_ray_avg     average origin per query and side (:179-197), one
//                wavefront per query
// k_cand also clears what the later kernels accumulate into (the pairs' ray
// counts, the hash sets, the per-candidate ray counts), so the chain is
// twelve launches with no fills.
#include "rayattr.hpp"

namespace cdk {

namespace {

__device__ __forceinline__ const cd_contact& at(const cd_contact* c, const uint32_t* order, uint32_t k) {
    return c[order ? order[k] : k];
}
__device__ __forceinline__ uint32_t own_tri(const cd_contact& c, int side) { return side ? c.tri_b : c.tri_a; }

__device__ __forceinline__ uint64_t mix64(uint64_t x) {  // splitmix64 finaliser
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}
constexpr uint64_t EMPTY = ~0ull;

struct Corner {  // one corner of a candidate triangle, in the space its rays use
    f3 p[3];     // triangle positions (the second's taken into the first's space)
    uint32_t mesh, tri;
};

}  // namespace

__global__ void k_cand(RaySides b, const QueryRec* __restrict__ qs, const MeshRec* __restrict__ meshes,
                       const uint32_t* __restrict__ leafof) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int side = int(blockIdx.y);
    const cd_contact* __restrict__ c = b.sorted;
    const uint32_t* __restrict__ order = side ? b.order_b : nullptr;
    RayCand* __restrict__ out = b.cand + size_t(b.n) * side;
    uint32_t* __restrict__ start = b.start + (size_t(b.n) + 1) * side;
    {  // clears: the hash sets and the pairs' counts (by side 0's threads), this side's ray counts
        const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
        if (side == 0) {
            for (uint64_t k = i; k < 2 * b.H; k += stride) {
                b.hkey[k] = ~0ull;
                b.hval[k] = ~0ull;
            }
            for (uint64_t k = i; k < 4 * uint64_t(b.qstride); k += stride) b.qw[k] = 0u;
        }
        uint32_t* cnt = b.cnt + (size_t(b.n) + 1) * side;
        for (uint64_t k = i; k <= b.n; k += stride) cnt[k] = 0u;
    }
    const uint32_t n = count_of(b.n, b.n_dev);
    if (i >= n) return;
    const cd_contact& ci = at(c, order, i);
    const uint32_t tri = own_tri(ci, side);
    bool first = i == 0;
    if (!first) {
        const cd_contact& cp = at(c, order, i - 1);
        first = cp.query != ci.query || own_tri(cp, side) != tri;
    }
    start[i] = 0u;
    if (!first) return;
    const QueryRec& q = qs[ci.query];
    const uint32_t other_base = meshes[side ? q.mesh_a : q.mesh_b].tri_base;
    const uint32_t shift = side ? 3u : 0u;
    // merged candidate (the unordered_map entry) and the current leaf's local one
    bool have = false;
    f3 twp = mk(0.f, 0.f, 0.f);
    float tw = 0.f;
    uint32_t tmask = 7u;
    f3 gwp = mk(0.f, 0.f, 0.f);
    float gw = 0.f;
    uint32_t gmask = 7u, leaf = 0u;
    auto flush = [&]() {
        if (gw == 0.f) return;  // IsNull (:19-22, :119)
        if (!have) {
            twp = gwp;
            tw = gw;
            tmask = gmask;
            have = true;
        } else {  // Merge (:52-57)
            tmask &= gmask;
            twp = add(twp, gwp);
            tw += gw;
        }
    };
    // The run is walked in order (the sums are order-dependent); its contacts
    // and their leaves are fetched RUN_FETCH at a time so that their loads
    // overlap instead of forming one round trip per contact.
    constexpr uint32_t RUN_FETCH = 4;
    bool more = true;
    for (uint32_t j0 = i; more && j0 < n; j0 += RUN_FETCH) {
        const uint32_t m = min(RUN_FETCH, n - j0);
        cd_contact cj[RUN_FETCH];
        uint32_t lf[RUN_FETCH];
#pragma unroll
        for (uint32_t u = 0; u < RUN_FETCH; ++u)
            if (u < m) cj[u] = at(c, order, j0 + u);
#pragma unroll
        for (uint32_t u = 0; u < RUN_FETCH; ++u) {
            const bool in_run = u < m && cj[u].query == ci.query && own_tri(cj[u], side) == tri;
            lf[u] = in_run ? leafof[size_t(other_base) + (side ? cj[u].tri_a : cj[u].tri_b)] : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < RUN_FETCH; ++u) {
            if (!more) break;
            if (u >= m || cj[u].query != ci.query || own_tri(cj[u], side) != tri) {
                more = false;
                break;
            }
            const uint32_t j = j0 + u;
            if (j > i && lf[u] != leaf) {
                flush();
                gwp = mk(0.f, 0.f, 0.f);
                gw = 0.f;
                gmask = 7u;
            }
            leaf = lf[u];
            const f3 s = mk(cj[u].s[0], cj[u].s[1], cj[u].s[2]), t = mk(cj[u].t[0], cj[u].t[1], cj[u].t[2]);
            const float weight = vlength(sub(s, t));                 // :93
            const f3 avg = divs(scale(add(s, t), weight), 2.f);      // :94
            gwp = add(gwp, avg);
            gw += weight;
            gmask &= ~((cj[u].flags >> shift) & 7u);                // :102-112
        }
    }
    flush();
    if (have) {
        RayCand r;
        r.query = ci.query;
        r.tri = tri;
        r.mask = tmask;
        r.pad = 0u;
        r.w = tw;
        r.wp[0] = twp.x;
        r.wp[1] = twp.y;
        r.wp[2] = twp.z;
        out[i] = r;
        start[i] = 1u;
    }
}

// Exclusive scan, three launches: per-block sums, one block over the sums
// (in chunks), then the per-element offsets.  1024 elements per block.
constexpr uint32_t SCAN_BLOCK = 1024;
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_w, uint32_t& total) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    uint32_t incl = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(incl, o, 64);
        if (lane >= uint32_t(o)) incl += u;
    }
    if (lane == 63) s_w[w] = incl;
    __syncthreads();
    uint32_t base = 0, all = 0;
    for (uint32_t k = 0; k < blockDim.x / 64; ++k) {
        if (k < w) base += s_w[k];
        all += s_w[k];
    }
    __syncthreads();
    total = all;
    return base + incl - v;
}
// Grid y: the side, whose input and output start y (n + 1) words in and
// whose block sums y nb words in (nb: the blocks per side); total per side
// at total + 2 y.
__global__ __launch_bounds__(256) void k_scan_sums(const uint32_t* __restrict__ in, uint32_t n,
                                                   const uint32_t* __restrict__ n_dev, uint32_t* __restrict__ sums) {
    __shared__ uint32_t s_w[4];
    const uint32_t b = blockIdx.x, tid = threadIdx.x;
    in += (size_t(n) + 1) * blockIdx.y;
    sums += size_t(gridDim.x) * blockIdx.y;
    n = count_of(n, n_dev);
    uint32_t v = 0;
    for (uint32_t k = 0; k < 4; ++k) {
        const uint32_t i = b * SCAN_BLOCK + tid * 4 + k;
        if (i < n) v += in[i];
    }
    uint32_t total = 0;
    (void)block_excl_scan(v, s_w, total);
    if (tid == 0) sums[b] = total;
}
__global__ __launch_bounds__(256) void k_scan_top(uint32_t* __restrict__ sums, uint32_t nb, uint32_t* __restrict__ total) {
    __shared__ uint32_t s_w[4];
    sums += size_t(nb) * blockIdx.y;
    total += 2 * blockIdx.y;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += 256) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nb ? sums[i] : 0u;
        uint32_t t = 0;
        const uint32_t ex = block_excl_scan(v, s_w, t);
        if (i < nb) sums[i] = carry + ex;
        carry += t;
    }
    if (threadIdx.x == 0) *total = carry;
}
__global__ __launch_bounds__(256) void k_scan_apply(const uint32_t* __restrict__ in, uint32_t n,
                                                    const uint32_t* __restrict__ n_dev,
                                                    const uint32_t* __restrict__ sums, uint32_t* __restrict__ out) {
    __shared__ uint32_t s_w[4];
    const uint32_t b = blockIdx.x, tid = threadIdx.x;
    in += (size_t(n) + 1) * blockIdx.y;
    out += (size_t(n) + 1) * blockIdx.y;
    sums += size_t(gridDim.x) * blockIdx.y;
    n = count_of(n, n_dev);
    uint32_t v[4], s = 0;
    for (uint32_t k = 0; k < 4; ++k) {
        const uint32_t i = b * SCAN_BLOCK + tid * 4 + k;
        v[k] = i < n ? in[i] : 0u;
        s += v[k];
    }
    uint32_t total = 0;
    uint32_t run = sums[b] + block_excl_scan(s, s_w, total);
    for (uint32_t k = 0; k < 4; ++k) {
        const uint32_t i = b * SCAN_BLOCK + tid * 4 + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
}

__global__ void k_compact(RaySides b) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t side = blockIdx.y;
    const size_t cs = size_t(b.n) * side, us = (size_t(b.n) + 1) * side;
    if (i < count_of(b.n, b.n_dev) && b.start[us + i]) b.dense[cs + b.pos[us + i]] = b.cand[cs + i];
}

// Hash set of claimed vertices: key (query, side, vertex id), value the
// smallest (triangle << 2 | corner) that wants the vertex.  Linear probing in
// a power-of-two table with more slots than keys, so every probe ends.
__device__ __forceinline__ uint64_t vkey(uint32_t query, int side, uint32_t v) {
    return (uint64_t(query) << 33) | (uint64_t(side) << 32) | uint64_t(v);
}
__global__ void k_vert_claim(RaySides b, const QueryRec* __restrict__ qs, const MeshRec* __restrict__ meshes,
                             const uint2* __restrict__ attrs, const uint32_t* __restrict__ vid,
                             const uint32_t* __restrict__ perm) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int side = int(blockIdx.y);
    if (i >= b.tot[2 * side]) return;
    const RayCand* __restrict__ cand = b.dense + size_t(b.n) * side;
    unsigned long long* __restrict__ hkey = b.hkey + b.H * side;
    unsigned long long* __restrict__ hval = b.hval + b.H * side;
    const uint64_t hmask = b.H - 1;
    const RayCand r = cand[i];
    if (r.mask == 0u) return;
    const uint32_t mesh = side ? qs[r.query].mesh_b : qs[r.query].mesh_a;
    const MeshRec& m = meshes[mesh];
    for (int k = 0; k < 3; ++k) {
        if (!(r.mask & (1u << k))) continue;
        const uint64_t key = vkey(r.query, side, vertex_id(m, attrs[mesh], vid, perm, r.tri, k));
        uint64_t h = mix64(key) & hmask;
        for (;;) {
            const unsigned long long prev = atomicCAS(&hkey[h], static_cast<unsigned long long>(EMPTY),
                                                      static_cast<unsigned long long>(key));
            if (prev == EMPTY || prev == key) break;
            h = (h + 1) & hmask;
        }
        atomicMin(&hval[h], (static_cast<unsigned long long>(r.tri) << 2) | static_cast<unsigned long long>(k));
    }
}
__device__ __forceinline__ bool claimed(const unsigned long long* hkey, const unsigned long long* hval, uint64_t hmask,
                                        uint64_t key, uint32_t tri, int k) {
    uint64_t h = mix64(key) & hmask;
    while (hkey[h] != key) h = (h + 1) & hmask;  // present: inserted by k_vert_claim
    return hval[h] == ((static_cast<unsigned long long>(tri) << 2) | static_cast<unsigned long long>(k));
}

__global__ void k_ray_count(RaySides b, const QueryRec* __restrict__ qs, const MeshRec* __restrict__ meshes,
                            const uint2* __restrict__ attrs, const uint32_t* __restrict__ vid,
                            const uint32_t* __restrict__ perm) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int side = int(blockIdx.y);
    if (i >= b.tot[2 * side]) return;
    const RayCand* __restrict__ cand = b.dense + size_t(b.n) * side;
    const unsigned long long* __restrict__ hkey = b.hkey + b.H * side;
    const unsigned long long* __restrict__ hval = b.hval + b.H * side;
    const uint64_t hmask = b.H - 1;
    uint32_t* __restrict__ cnt = b.cnt + (size_t(b.n) + 1) * side;
    uint32_t* __restrict__ qcount = b.qw + 2 * size_t(b.qstride) * side;
    const RayCand r = cand[i];
    uint32_t k_rays = 1u;  // fallback ray (:155-167)
    if (r.mask != 0u) {
        const uint32_t mesh = side ? qs[r.query].mesh_b : qs[r.query].mesh_a;
        const MeshRec& m = meshes[mesh];
        k_rays = 0u;
        for (int k = 0; k < 3; ++k)
            if ((r.mask & (1u << k)) &&
                claimed(hkey, hval, hmask, vkey(r.query, side, vertex_id(m, attrs[mesh], vid, perm, r.tri, k)), r.tri, k))
                ++k_rays;
    }
    cnt[i] = k_rays;
    atomicAdd(&qcount[r.query], k_rays);
}

__global__ void k_ray_emit(RaySides b, const QueryRec* __restrict__ qs, const MeshRec* __restrict__ meshes,
                           const float* __restrict__ tris, const uint2* __restrict__ attrs,
                           const uint32_t* __restrict__ vid, const float* __restrict__ nrm,
                           const uint32_t* __restrict__ perm) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int side = int(blockIdx.y);
    if (i >= b.tot[2 * side]) return;
    const RayCand* __restrict__ cand = b.dense + size_t(b.n) * side;
    const unsigned long long* __restrict__ hkey = b.hkey + b.H * side;
    const unsigned long long* __restrict__ hval = b.hval + b.H * side;
    const uint64_t hmask = b.H - 1;
    const uint32_t* __restrict__ pos = b.pos2 + (size_t(b.n) + 1) * side;
    const uint64_t ray_base = side ? b.tot[DC_RAYS0] : 0u;  // the second side's rays follow the first's
    cd_ray* __restrict__ rays = b.rays;
    uint32_t* __restrict__ qbegin = b.qw + 2 * size_t(b.qstride) * side + b.qstride;
    const RayCand r = cand[i];
    if (i == 0 || cand[i - 1].query != r.query) qbegin[r.query] = pos[i];  // candidates are in query order
    const QueryRec& q = qs[r.query];
    const uint32_t mesh = side ? q.mesh_b : q.mesh_a;
    const MeshRec& m = meshes[mesh];
    const uint2 attr = attrs[mesh];
    M4 M;
    for (int e = 0; e < 16; ++e) M.m[e] = q.rel[e];
    const float* t9 = tris + (size_t(m.tri_base) + r.tri) * 9;
    f3 P[3];
    for (int k = 0; k < 3; ++k) {
        const f3 p = mk(t9[3 * k], t9[3 * k + 1], t9[3 * k + 2]);
        P[k] = side ? mul4(M, p, 1.f) : p;  // :135-136
    }
    cd_ray* out = rays + ray_base + pos[i];
    uint32_t w = 0;
    if (r.mask != 0u) {
        const M3 NM = adjoint_transpose(q.rel);
        for (int k = 0; k < 3; ++k) {
            if (!(r.mask & (1u << k))) continue;
            const uint32_t v = vertex_id(m, attr, vid, perm, r.tri, k);
            if (!claimed(hkey, hval, hmask, vkey(r.query, side, v), r.tri, k)) continue;
            const f3 n = corner_normal(attr, nrm, v, t9);
            const f3 nn = side ? normalize(mul3(NM, n)) : normalize(n);  // Triangle.cpp:180-193
            cd_ray ray;
            ray.origin[0] = P[k].x; ray.origin[1] = P[k].y; ray.origin[2] = P[k].z;
            ray.direction[0] = -nn.x; ray.direction[1] = -nn.y; ray.direction[2] = -nn.z;
            out[w++] = ray;
        }
    } else {
        // weighted average point, barycentric normal (Triangle.cpp:148-163, 195-212)
        const f3 pt = divs(mk(r.wp[0], r.wp[1], r.wp[2]), r.w);
        const f3 v0 = sub(P[1], P[0]), v1 = sub(P[2], P[0]), v2 = sub(pt, P[0]);
        const float d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(v2, v0), d21 = dot(v2, v1);
        const float denom = d00 * d11 - d01 * d01;
        const float bx = (d11 * d20 - d01 * d21) / denom;
        const float by = (d00 * d21 - d01 * d20) / denom;
        f3 n[3];
        for (int k = 0; k < 3; ++k) n[k] = corner_normal(attr, nrm, vertex_id(m, attr, vid, perm, r.tri, k), t9);
        const f3 in = add(add(scale(n[0], (1.f - bx) - by), scale(n[1], bx)), scale(n[2], by));
        const f3 nn = side ? normalize(mul3(adjoint_transpose(q.rel), in)) : normalize(in);
        cd_ray ray;
        ray.origin[0] = pt.x; ray.origin[1] = pt.y; ray.origin[2] = pt.z;
        ray.direction[0] = -nn.x; ray.direction[1] = -nn.y; ray.direction[2] = -nn.z;
        out[0] = ray;
    }
}

// Average origin of each query's rays on one side (:179-197); the second's is
// taken back to its own model space by inverse(rel).
// One wavefront per query: lane-strided partial sums, then a fixed shuffle
// tree, so the sum is the same on every run (the reference sums in its
// unordered_map's order, which no GPU reproduces).
__global__ __launch_bounds__(64) void k_ray_avg(RaySides b, const QueryRec* __restrict__ qs) {
    const uint32_t qi = blockIdx.x;
    const int lane = threadIdx.x;
    const int side = int(blockIdx.y);
    if (qi >= count_of(b.nq, b.nq_dev)) return;
    const cd_ray* __restrict__ rays = b.rays;
    const uint64_t ray_base = side ? b.tot[DC_RAYS0] : 0u;
    const uint32_t* __restrict__ qcount = b.qw + 2 * size_t(b.qstride) * side;
    const uint32_t* __restrict__ qbegin = qcount + b.qstride;
    float* __restrict__ avg = b.avg + 3 * size_t(b.qstride) * side;
    const uint32_t n = qcount[qi];
    f3 acc = mk(0.f, 0.f, 0.f);
    if (n) {
        const cd_ray* r = rays + ray_base + qbegin[qi];
        for (uint32_t k = uint32_t(lane); k < n; k += 64) acc = add(acc, mk(r[k].origin[0], r[k].origin[1], r[k].origin[2]));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        acc = add(acc, mk(__shfl_down(acc.x, off, 64), __shfl_down(acc.y, off, 64), __shfl_down(acc.z, off, 64)));
    if (lane != 0) return;
    f3 a = divs(acc, float(n));
    if (side) {
        M4 M;
        for (int e = 0; e < 16; ++e) M.m[e] = qs[qi].rel[e];
        a = mul4(mat_inverse(M), a, 1.f);
    }
    avg[3 * qi] = a.x;
    avg[3 * qi + 1] = a.y;
    avg[3 * qi + 2] = a.z;
}

// ---- launchers ---------------------------------------------------------------
static inline dim3 gr2(uint64_t n) { return dim3(uint32_t((std::max<uint64_t>(n, 1) + 255) / 256), 2); }

hipError_t launch_rays(const RaySides& b, const QueryRec* qs, const MeshRec* meshes, const uint32_t* leafof,
                       const float* tris, const uint2* attrs, const uint32_t* vid, const float* nrm,
                       const uint32_t* perm, hipStream_t st) {
    if (b.H & (b.H - 1)) return hipErrorInvalidValue;
    const uint32_t nb = std::max(1u, (b.n + SCAN_BLOCK - 1) / SCAN_BLOCK);
    if (b.nsums < nb) return hipErrorInvalidValue;
    const dim3 g = gr2(b.n), gs(nb, 2);
    hipLaunchKernelGGL(k_cand, g, dim3(256), 0, st, b, qs, meshes, leafof);
    // candidates: dense list at the exclusive scan of the run starts
    hipLaunchKernelGGL(k_scan_sums, gs, dim3(256), 0, st, b.start, b.n, b.n_dev, b.sums);
    hipLaunchKernelGGL(k_scan_top, dim3(1, 2), dim3(256), 0, st, b.sums, nb, b.tot + DC_CAND0);
    hipLaunchKernelGGL(k_scan_apply, gs, dim3(256), 0, st, b.start, b.n, b.n_dev, b.sums, b.pos);
    hipLaunchKernelGGL(k_compact, g, dim3(256), 0, st, b);
    hipLaunchKernelGGL(k_vert_claim, g, dim3(256), 0, st, b, qs, meshes, attrs, vid, perm);
    hipLaunchKernelGGL(k_ray_count, g, dim3(256), 0, st, b, qs, meshes, attrs, vid, perm);
    // rays: each candidate's at the exclusive scan of the ray counts
    hipLaunchKernelGGL(k_scan_sums, gs, dim3(256), 0, st, b.cnt, b.n, b.n_dev, b.sums);
    hipLaunchKernelGGL(k_scan_top, dim3(1, 2), dim3(256), 0, st, b.sums, nb, b.tot + DC_RAYS0);
    hipLaunchKernelGGL(k_scan_apply, gs, dim3(256), 0, st, b.cnt, b.n, b.n_dev, b.sums, b.pos2);
    hipLaunchKernelGGL(k_ray_emit, g, dim3(256), 0, st, b, qs, meshes, tris, attrs, vid, nrm, perm);
    if (b.nq) hipLaunchKernelGGL(k_ray_avg, dim3(b.nq, 2), dim3(64), 0, st, b, qs);
    return hipGetLastError();
}

}  // namespace cdk
