// Broad phase and frame assembly.
//
//  k_sap_project  world root box = current * root OBB and its intervals on
//                 the fixed U/V/W frame (SweepAndPrune.cpp:20-38,
//                 CollisionDetection.cpp:9-13); the sort key is U.min
//  k_radix_sort   stable LSD radix sort of (key, value) pairs, 4-bit digits,
//                 one workgroup; ranks within a wavefront from ballots on
//                 the digit bits, across wavefronts from an LDS table
//  k_sap_sweep    one lane per entity in U order: every later entity whose
//                 U.min <= this U.max, kept when V and W overlap and either
//                 entity wants callbacks (SweepAndPrune.cpp:50-85); the pair
//                 is oriented (earlier, later) as the reference orients it
//  k_pair_queries relative matrix inverse(first) * second per pair
//                 (OBBtreesCollision.cpp:15) with glm's formulas
//
// Ties on U.min are ordered by entity index (the reference's std::sort
// leaves them unspecified); +0 and -0 are the same key.
#include "internal.hpp"
#include "kernels.hpp"

namespace cdk {

__device__ __forceinline__ uint32_t orderable(float f) {
    if (f == 0.0f) f = 0.0f;  // -0 sorts with +0
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ void k_sap_project(const cd_entity* __restrict__ ents, uint32_t n, const MeshRec* __restrict__ meshes,
                              SapAxes ax, float* __restrict__ iv, uint32_t* __restrict__ keys,
                              uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    M4 M;
    for (int k = 0; k < 16; ++k) M.m[k] = ents[i].current[k];
    const Box P = xform_box(M, box_from12(meshes[ents[i].mesh_id].root));
    const f3 axes[3] = {ax.u, ax.v, ax.w};
    for (int a = 0; a < 3; ++a) {
        float lo, hi;
        project(P, axes[a], lo, hi);
        iv[size_t(i) * 6 + 2 * a] = lo;
        iv[size_t(i) * 6 + 2 * a + 1] = hi;
    }
    keys[i] = orderable(iv[size_t(i) * 6]);
    vals[i] = i;
}

__global__ __launch_bounds__(1024) void k_radix_sort(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp,
                                                     uint32_t* vals_tmp, uint32_t n, uint32_t passes) {
    __shared__ uint32_t s_cnt[16][16];  // [wavefront][digit]
    __shared__ uint32_t s_base[16];
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint32_t *sk = keys, *sv = vals, *dk = keys_tmp, *dv = vals_tmp;
    for (uint32_t pass = 0; pass < passes; ++pass) {
        const uint32_t shift = 4 * pass;
        if (tid < 16) s_base[tid] = 0;
        __syncthreads();
        uint32_t local[16];
        for (int d = 0; d < 16; ++d) local[d] = 0;
        for (uint32_t i = tid; i < n; i += 1024) local[(sk[i] >> shift) & 15u]++;
        for (int d = 0; d < 16; ++d)
            if (local[d]) atomicAdd(&s_base[d], local[d]);
        __syncthreads();
        if (tid == 0) {
            uint32_t run = 0;
            for (int d = 0; d < 16; ++d) {
                const uint32_t c = s_base[d];
                s_base[d] = run;
                run += c;
            }
        }
        __syncthreads();
        for (uint32_t c0 = 0; c0 < n; c0 += 1024) {
            const uint32_t i = c0 + tid;
            const bool valid = i < n;
            const uint32_t k = valid ? sk[i] : 0u, v = valid ? sv[i] : 0u;
            const uint32_t d = (k >> shift) & 15u;
            unsigned long long peers = __ballot(valid);
            for (int b = 0; b < 4; ++b) {
                const unsigned long long bm = __ballot(valid && ((d >> b) & 1u));
                peers &= ((d >> b) & 1u) ? bm : ~bm;
            }
            const uint32_t rank = uint32_t(__popcll(peers & below));
            if (tid < 256) s_cnt[tid >> 4][tid & 15] = 0;
            __syncthreads();
            if (valid && rank == 0) s_cnt[wave][d] = uint32_t(__popcll(peers));
            __syncthreads();
            if (valid) {
                uint32_t off = s_base[d];
                for (uint32_t w = 0; w < wave; ++w) off += s_cnt[w][d];
                dk[off + rank] = k;
                dv[off + rank] = v;
            }
            __syncthreads();
            if (tid < 16) {
                uint32_t tot = 0;
                for (int w = 0; w < 16; ++w) tot += s_cnt[w][tid];
                s_base[tid] += tot;
            }
            __syncthreads();
        }
        uint32_t* t = sk; sk = dk; dk = t;
        t = sv; sv = dv; dv = t;
    }
}

__global__ void k_sap_sweep(const uint32_t* __restrict__ order, const float* __restrict__ iv,
                            const cd_entity* __restrict__ ents, uint32_t n, uint2* __restrict__ pairs, uint32_t cap,
                            unsigned int* __restrict__ count) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t e = order[p];
    const float* a = iv + size_t(e) * 6;
    const bool cbe = ents[e].should_callback != 0;
    for (uint32_t q = p + 1; q < n; ++q) {
        const uint32_t f = order[q];
        const float* b = iv + size_t(f) * 6;
        if (a[1] < b[0]) break;  // retired: this U.max is below every later U.min (SweepAndPrune.cpp:58-59)
        if (!(cbe || ents[f].should_callback)) continue;
        if ((a[3] >= b[2]) && (b[3] >= a[2]) && (a[5] >= b[4]) && (b[5] >= a[4])) {
            const unsigned int s = atomicAdd(count, 1u);
            if (s < cap) pairs[s] = make_uint2(e, f);
        }
    }
}

__global__ void k_pair_keys(const uint2* __restrict__ pairs, uint32_t n, uint32_t* __restrict__ keys,
                            uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = (pairs[i].x << 16) | pairs[i].y;
    vals[i] = i;
}

__global__ void k_pair_queries(const uint2* __restrict__ pairs, const uint32_t* __restrict__ order, uint32_t n,
                               const cd_entity* __restrict__ ents, QueryRec* __restrict__ qs,
                               uint2* __restrict__ sorted_pairs) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint2 pr = pairs[order[i]];
    M4 A, B;
    for (int k = 0; k < 16; ++k) {
        A.m[k] = ents[pr.x].current[k];
        B.m[k] = ents[pr.y].current[k];
    }
    const M4 R = mat_mul(mat_inverse(A), B);
    QueryRec q;
    q.mesh_a = ents[pr.x].mesh_id;
    q.mesh_b = ents[pr.y].mesh_id;
    for (int k = 0; k < 16; ++k) q.rel[k] = R.m[k];
    qs[i] = q;
    sorted_pairs[i] = pr;
}

// Contact grouping: keys of one field of the contact records.
__global__ void k_contact_keys(const cd_contact* __restrict__ c, uint32_t n, int field, const uint32_t* __restrict__ order,
                               uint32_t* __restrict__ keys) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const cd_contact& r = c[order[i]];
    keys[i] = field == 0 ? r.query : (field == 1 ? r.tri_a : r.tri_b);
}
__global__ void k_iota(uint32_t* v, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = i;
}
__global__ void k_gather_contacts(const cd_contact* __restrict__ in, const uint32_t* __restrict__ order, uint32_t n,
                                  cd_contact* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[order[i]];
}

// ---- launchers --------------------------------------------------------------
static inline dim3 g256(uint32_t n) { return dim3((std::max(n, 1u) + 255) / 256); }

hipError_t launch_sap_project(const cd_entity* ents, uint32_t n, const MeshRec* meshes, const SapAxes& ax, float* iv,
                              uint32_t* keys, uint32_t* vals, hipStream_t st) {
    hipLaunchKernelGGL(k_sap_project, g256(n), dim3(256), 0, st, ents, n, meshes, ax, iv, keys, vals);
    return hipGetLastError();
}
hipError_t launch_radix_sort(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t n,
                             uint32_t key_bits, hipStream_t st) {
    uint32_t passes = (key_bits + 3) / 4;
    passes += passes & 1u;  // even, so the result lands back in keys/vals
    hipLaunchKernelGGL(k_radix_sort, dim3(1), dim3(1024), 0, st, keys, vals, keys_tmp, vals_tmp, n, passes);
    return hipGetLastError();
}
hipError_t launch_sap_sweep(const uint32_t* order, const float* iv, const cd_entity* ents, uint32_t n, uint2* pairs,
                            uint32_t cap, unsigned int* count, hipStream_t st) {
    hipLaunchKernelGGL(k_sap_sweep, g256(n), dim3(256), 0, st, order, iv, ents, n, pairs, cap, count);
    return hipGetLastError();
}
hipError_t launch_pair_keys(const uint2* pairs, uint32_t n, uint32_t* keys, uint32_t* vals, hipStream_t st) {
    hipLaunchKernelGGL(k_pair_keys, g256(n), dim3(256), 0, st, pairs, n, keys, vals);
    return hipGetLastError();
}
hipError_t launch_pair_queries(const uint2* pairs, const uint32_t* order, uint32_t n, const cd_entity* ents,
                               QueryRec* qs, uint2* sorted_pairs, hipStream_t st) {
    hipLaunchKernelGGL(k_pair_queries, g256(n), dim3(256), 0, st, pairs, order, n, ents, qs, sorted_pairs);
    return hipGetLastError();
}
hipError_t launch_sort_contacts(const cd_contact* in, uint32_t n, cd_contact* out, uint32_t* k0, uint32_t* v0,
                                uint32_t* k1, uint32_t* v1, hipStream_t st) {
    // stable LSD over (query, tri_a, tri_b): least significant field first
    hipLaunchKernelGGL(k_iota, g256(n), dim3(256), 0, st, v0, n);
    for (int field = 2; field >= 0; --field) {
        hipLaunchKernelGGL(k_contact_keys, g256(n), dim3(256), 0, st, in, n, field, v0, k0);
        hipError_t e = launch_radix_sort(k0, v0, k1, v1, n, 32, st);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_gather_contacts, g256(n), dim3(256), 0, st, in, v0, n, out);
    return hipGetLastError();
}
hipError_t launch_sort_perm(const cd_contact* in, uint32_t n, int f_hi, int f_mid, int f_lo, uint32_t* k0,
                            uint32_t* v0, uint32_t* k1, uint32_t* v1, hipStream_t st) {
    hipLaunchKernelGGL(k_iota, g256(n), dim3(256), 0, st, v0, n);
    for (int field : {f_lo, f_mid, f_hi}) {
        hipLaunchKernelGGL(k_contact_keys, g256(n), dim3(256), 0, st, in, n, field, v0, k0);
        hipError_t e = launch_radix_sort(k0, v0, k1, v1, n, 32, st);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

}  // namespace cdk
