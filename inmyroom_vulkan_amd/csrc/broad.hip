// Broad phase and frame assembly.
//
//  k_sap_project  world root box = current * root OBB and its intervals on
//                 the fixed U/V/W frame (SweepAndPrune.cpp:20-38,
//                 CollisionDetection.cpp:9-13); the sort key is U.min
//  k_radix_sort   stable LSD radix sort of (key, value) pairs, 4-bit digits,
//                 one workgroup; ranks within a wavefront from ballots on
//                 the digit bits, across wavefronts from an LDS table
//                 (up to RS_SMALL keys)
//  k_rank_sort    the small sorts (up to RANK_MAX keys): every key ranked
//                 against all others in LDS, same stable order, 64 keys
//                 per workgroup
//  k_rs_*         the same sort over many workgroups for longer arrays
//                 (contacts): 8-bit digits, per-tile digit counts, one
//                 scan of the count table, stable scatter
//  k_sap_sweep    one wavefront per entity in U order, the later entities
//                 tested 64 at a time: every later entity whose
//                 U.min <= this U.max, kept when V and W overlap and either
//                 entity wants callbacks (SweepAndPrune.cpp:50-85); the pair
//                 is oriented (earlier, later) as the reference orients it
//  k_pair_queries relative matrix inverse(first) * second per pair
//                 (OBBtreesCollision.cpp:15) with glm's formulas
//
// Ties on U.min are ordered by entity index (the reference's std::sort
// leaves them unspecified); +0 and -0 are the same key.
#include <utility>

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
    const Box P = xform_box(M, box_from12(meshes[ents[i].mesh_id].root.box));
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

// n_dev (here and in the sort and pair kernels below): when given, the count
// is min(*n_dev, n) and n is only the capacity the launch was shaped for, so
// that a captured graph serves every count up to it.
__global__ __launch_bounds__(1024) void k_radix_sort(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp,
                                                     uint32_t* vals_tmp, uint32_t n, const uint32_t* __restrict__ n_dev,
                                                     uint32_t passes) {
    __shared__ uint32_t s_cnt[16][16];  // [wavefront][digit]
    __shared__ uint32_t s_base[16];
    n = count_of(n, n_dev);
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

// Small sorts (the broad phase's keys and pairs of a scene): every key is
// ranked against all others in LDS, rank = (keys below) + (equal keys before
// it), the stable order an LSD radix sort over the same bits gives, with no
// passes and no scans.  A workgroup ranks RANK_KEYS keys, one wavefront per
// quarter of the others (the keys it compares against are read by every
// lane at once), and scatters them into the second buffers; k_copy_back
// returns them.  One workgroup ranking everything took ~40 us for 512 keys.
constexpr uint32_t RANK_MAX = 2048;
constexpr uint32_t RANK_KEYS = 64;
constexpr uint32_t RANK_SPLIT = 4;

__global__ __launch_bounds__(RANK_KEYS * RANK_SPLIT) void k_rank_sort(const uint32_t* __restrict__ keys,
                                                                      const uint32_t* __restrict__ vals, uint32_t n,
                                                                      const uint32_t* __restrict__ n_dev, uint32_t mask,
                                                                      uint32_t* __restrict__ keys_out,
                                                                      uint32_t* __restrict__ vals_out) {
    __shared__ uint32_t s_k[RANK_MAX];
    __shared__ uint32_t s_part[RANK_SPLIT][RANK_KEYS];
    n = count_of(n, n_dev);
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s_k[i] = keys[i] & mask;
    __syncthreads();
    const uint32_t kl = threadIdx.x % RANK_KEYS, part = threadIdx.x / RANK_KEYS;
    const uint32_t i = blockIdx.x * RANK_KEYS + kl;
    const uint32_t chunk = (n + RANK_SPLIT - 1) / RANK_SPLIT;
    const uint32_t j0 = min(n, part * chunk), j1 = min(n, j0 + chunk);
    uint32_t r = 0;
    if (i < n) {
        const uint32_t k = s_k[i];
#pragma unroll 8
        for (uint32_t j = j0; j < j1; ++j) {
            const uint32_t kj = s_k[j];
            r += (kj < k || (kj == k && j < i)) ? 1u : 0u;
        }
    }
    s_part[part][kl] = r;
    __syncthreads();
    if (part == 0 && i < n) {
        uint32_t rank = 0;
#pragma unroll
        for (uint32_t s = 0; s < RANK_SPLIT; ++s) rank += s_part[s][kl];
        keys_out[rank] = keys[i];
        vals_out[rank] = vals[i];
    }
}

__global__ void k_copy_back(const uint32_t* __restrict__ ka, const uint32_t* __restrict__ va, uint32_t n,
                            const uint32_t* __restrict__ n_dev, uint32_t* __restrict__ kb, uint32_t* __restrict__ vb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count_of(n, n_dev)) {
        kb[i] = ka[i];
        vb[i] = va[i];
    }
}

// Multi-workgroup pass of the LSD sort: tiles of RS_TILE keys; the count
// table is digit-major ([digit][tile]) so that one exclusive scan gives
// every tile the start of each of its digit runs in tile order (stable).
constexpr uint32_t RS_TILE = 1024;  // keys per workgroup per pass: 4 rounds of 256
constexpr uint32_t RS_SMALL = 4096;

__global__ __launch_bounds__(256) void k_rs_count(const uint32_t* __restrict__ keys, uint32_t n,
                                                  const uint32_t* __restrict__ n_dev, uint32_t shift,
                                                  uint32_t* __restrict__ table) {
    __shared__ uint32_t h[256];
    n = count_of(n, n_dev);
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t lo = blockIdx.x * RS_TILE, hi = min(lo + RS_TILE, n);
    for (uint32_t i = lo + threadIdx.x; i < hi; i += 256) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    __syncthreads();
    table[size_t(threadIdx.x) * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of the m-entry count table in place, one workgroup: each
// thread sums a run of the table (up to RS_SCAN_REGS entries loaded at once
// and kept in registers), the run sums are scanned within each wavefront by
// shuffles and across the 16 wavefronts through LDS.
constexpr uint32_t RS_SCAN_REGS = 16;

__global__ __launch_bounds__(1024) void k_rs_scan(uint32_t* __restrict__ table, uint32_t m) {
    __shared__ uint32_t wsum[16];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t per = (m + 1023) / 1024;
    const uint32_t lo = min(m, tid * per), hi = min(lo + per, m);
    uint32_t v[RS_SCAN_REGS];
    uint32_t sum = 0;
    if (per <= RS_SCAN_REGS) {
#pragma unroll
        for (uint32_t k = 0; k < RS_SCAN_REGS; ++k) v[k] = lo + k < hi ? table[lo + k] : 0u;
#pragma unroll
        for (uint32_t k = 0; k < RS_SCAN_REGS; ++k) sum += v[k];
    } else {
        for (uint32_t i = lo; i < hi; ++i) sum += table[i];
    }
    uint32_t x = sum;  // inclusive scan of the run sums within the wavefront
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    if (wave == 0) {
        uint32_t w = lane < 16 ? wsum[lane] : 0u;
#pragma unroll
        for (uint32_t off = 1; off < 16; off <<= 1) {
            const uint32_t y = __shfl_up(w, off, 64);
            if (lane >= off) w += y;
        }
        if (lane < 16) wsum[lane] = w;
    }
    __syncthreads();
    uint32_t run = x - sum + (wave ? wsum[wave - 1] : 0u);
    if (per <= RS_SCAN_REGS) {
#pragma unroll
        for (uint32_t k = 0; k < RS_SCAN_REGS; ++k)
            if (lo + k < hi) {
                table[lo + k] = run;
                run += v[k];
            }
    } else {
        for (uint32_t i = lo; i < hi; ++i) {
            const uint32_t c = table[i];
            table[i] = run;
            run += c;
        }
    }
}

__global__ __launch_bounds__(256) void k_rs_scatter(const uint32_t* __restrict__ sk, const uint32_t* __restrict__ sv,
                                                    uint32_t n, const uint32_t* __restrict__ n_dev, uint32_t shift,
                                                    const uint32_t* __restrict__ table, uint32_t* __restrict__ dk,
                                                    uint32_t* __restrict__ dv) {
    __shared__ uint32_t start[256];    // next output slot of each digit for this tile
    __shared__ uint32_t wcnt[4][256];  // keys of each digit per wavefront in the current round
    n = count_of(n, n_dev);
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    start[tid] = table[size_t(tid) * gridDim.x + blockIdx.x];
    const uint32_t lo = blockIdx.x * RS_TILE, hi = min(lo + RS_TILE, n);
    for (uint32_t r0 = lo; r0 < hi; r0 += 256) {
        const uint32_t i = r0 + tid;
        const bool valid = i < hi;
        const uint32_t k = valid ? sk[i] : 0u, v = valid ? sv[i] : 0u;
        const uint32_t d = (k >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const unsigned long long bm = __ballot(valid && ((d >> b) & 1u));
            peers &= ((d >> b) & 1u) ? bm : ~bm;
        }
        const uint32_t rank = uint32_t(__popcll(peers & below));
        for (uint32_t e = tid; e < 4 * 256; e += 256) (&wcnt[0][0])[e] = 0;
        __syncthreads();
        if (valid && rank == 0) wcnt[wave][d] = uint32_t(__popcll(peers));
        __syncthreads();
        if (valid) {
            uint32_t off = start[d] + rank;
            for (uint32_t w = 0; w < wave; ++w) off += wcnt[w][d];
            dk[off] = k;
            dv[off] = v;
        }
        __syncthreads();
        start[tid] += wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
        __syncthreads();
    }
}

constexpr uint32_t SAP_WAVES = 4;  // wavefronts (entities) per workgroup

// One wavefront per entity in U order: its lanes test the later entities 64
// at a time, up to the first chunk holding one whose U.min is above this
// U.max.  The U order makes every entity after that one retired too, so the
// pairs are the ones the sequential sweep finds.  (One lane per entity
// stepping through the later ones serialised on the entities that overlap
// many others: 78-113 us for 512 entities.)
__global__ __launch_bounds__(64 * SAP_WAVES) void k_sap_sweep(const uint32_t* __restrict__ order,
                                                              const float* __restrict__ iv,
                                                              const cd_entity* __restrict__ ents, uint32_t n,
                                                              uint2* __restrict__ pairs, uint32_t cap,
                                                              unsigned int* __restrict__ count) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t p = blockIdx.x * SAP_WAVES + (threadIdx.x >> 6);  // the same in every lane of a wavefront
    if (p >= n) return;
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t e = order[p];
    float a[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) a[k] = iv[size_t(e) * 6 + k];
    const bool cbe = ents[e].should_callback != 0;
    for (uint32_t q0 = p + 1; q0 < n; q0 += 64) {
        const uint32_t q = q0 + lane;
        bool live = false, hit = false;
        uint32_t f = 0;
        if (q < n) {
            f = order[q];
            float b[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) b[k] = iv[size_t(f) * 6 + k];
            live = !(a[1] < b[0]);  // retired: this U.max is below that U.min (SweepAndPrune.cpp:58-59)
            hit = live && (a[3] >= b[2]) && (b[3] >= a[2]) && (a[5] >= b[4]) && (b[5] >= a[4]) &&
                  (cbe || ents[f].should_callback != 0);
        }
        const unsigned long long m = __ballot(hit);
        if (m) {  // one atomic per wavefront and chunk
            const int leader = __ffsll(static_cast<long long>(m)) - 1;
            unsigned int base = 0;
            if (int(lane) == leader) base = atomicAdd(count, unsigned(__popcll(m)));
            base = __shfl(base, leader, 64);
            if (hit) {
                const unsigned int sl = base + unsigned(__popcll(m & below));
                if (sl < cap) pairs[sl] = make_uint2(e, f);
            }
        }
        if (__ballot(!live) != 0ull) break;  // a retired lane (or the end): every later entity is retired
    }
}

__global__ void k_pair_keys(const uint2* __restrict__ pairs, uint32_t n, const uint32_t* __restrict__ n_dev,
                            uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count_of(n, n_dev)) return;
    keys[i] = (pairs[i].x << 16) | pairs[i].y;
    vals[i] = i;
}

__global__ void k_pair_queries(const uint2* __restrict__ pairs, const uint32_t* __restrict__ order, uint32_t n,
                               const uint32_t* __restrict__ n_dev, const cd_entity* __restrict__ ents,
                               QueryRec* __restrict__ qs, uint2* __restrict__ sorted_pairs) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count_of(n, n_dev)) return;
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
__global__ void k_contact_keys(const cd_contact* __restrict__ c, uint32_t n, const uint32_t* __restrict__ n_dev,
                               int field, const uint32_t* __restrict__ order, uint32_t* __restrict__ keys) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count_of(n, n_dev)) return;
    const cd_contact& r = c[order[i]];
    keys[i] = field == 0 ? r.query : (field == 1 ? r.tri_a : r.tri_b);
}
__global__ void k_iota(uint32_t* v, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = i;
}
__global__ void k_gather_contacts(const cd_contact* __restrict__ in, const uint32_t* __restrict__ order, uint32_t n,
                                  const uint32_t* __restrict__ n_dev, cd_contact* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count_of(n, n_dev)) out[i] = in[order[i]];
}

// ---- launchers --------------------------------------------------------------
static inline dim3 g256(uint32_t n) { return dim3((std::max(n, 1u) + 255) / 256); }

hipError_t launch_sap_project(const cd_entity* ents, uint32_t n, const MeshRec* meshes, const SapAxes& ax, float* iv,
                              uint32_t* keys, uint32_t* vals, hipStream_t st) {
    hipLaunchKernelGGL(k_sap_project, g256(n), dim3(256), 0, st, ents, n, meshes, ax, iv, keys, vals);
    return hipGetLastError();
}
uint64_t radix_table_words(uint32_t n) { return uint64_t(256) * ((uint64_t(n) + RS_TILE - 1) / RS_TILE); }


hipError_t launch_radix_sort(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp, uint32_t n,
                             uint32_t key_bits, uint32_t* table, hipStream_t st, const uint32_t* n_dev) {
    if (n < 2) return hipSuccess;
    if (n <= RANK_MAX) {
        const uint32_t mask = rank_sort_mask(key_bits);
        hipLaunchKernelGGL(k_rank_sort, dim3((n + RANK_KEYS - 1) / RANK_KEYS), dim3(RANK_KEYS * RANK_SPLIT), 0, st, keys,
                           vals, n, n_dev, mask, keys_tmp, vals_tmp);
        hipLaunchKernelGGL(k_copy_back, g256(n), dim3(256), 0, st, keys_tmp, vals_tmp, n, n_dev, keys, vals);
        return hipGetLastError();
    }
    if (n <= RS_SMALL || !table) {
        uint32_t passes = (key_bits + 3) / 4;
        passes += passes & 1u;  // even, so the result lands back in keys/vals
        hipLaunchKernelGGL(k_radix_sort, dim3(1), dim3(1024), 0, st, keys, vals, keys_tmp, vals_tmp, n, n_dev, passes);
        return hipGetLastError();
    }
    const uint32_t tiles = uint32_t((uint64_t(n) + RS_TILE - 1) / RS_TILE);
    const uint32_t m = 256 * tiles;
    uint32_t *sk = keys, *sv = vals, *dk = keys_tmp, *dv = vals_tmp;
    for (uint32_t shift = 0; shift < key_bits; shift += 8) {
        hipLaunchKernelGGL(k_rs_count, dim3(tiles), dim3(256), 0, st, sk, n, n_dev, shift, table);
        hipLaunchKernelGGL(k_rs_scan, dim3(1), dim3(1024), 0, st, table, m);
        hipLaunchKernelGGL(k_rs_scatter, dim3(tiles), dim3(256), 0, st, sk, sv, n, n_dev, shift, table, dk, dv);
        std::swap(sk, dk);
        std::swap(sv, dv);
    }
    if (sk != keys) {  // odd number of passes: the result is in the temporaries
        hipError_t e = hipMemcpyAsync(keys, sk, size_t(n) * sizeof(uint32_t), hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(vals, sv, size_t(n) * sizeof(uint32_t), hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}
hipError_t launch_sap_sweep(const uint32_t* order, const float* iv, const cd_entity* ents, uint32_t n, uint2* pairs,
                            uint32_t cap, unsigned int* count, hipStream_t st) {
    hipLaunchKernelGGL(k_sap_sweep, dim3((n + SAP_WAVES - 1) / SAP_WAVES), dim3(64 * SAP_WAVES), 0, st, order, iv, ents, n,
                       pairs, cap, count);
    return hipGetLastError();
}
hipError_t launch_pair_keys(const uint2* pairs, uint32_t n, uint32_t* keys, uint32_t* vals, hipStream_t st,
                            const uint32_t* n_dev) {
    hipLaunchKernelGGL(k_pair_keys, g256(n), dim3(256), 0, st, pairs, n, n_dev, keys, vals);
    return hipGetLastError();
}
hipError_t launch_pair_queries(const uint2* pairs, const uint32_t* order, uint32_t n, const cd_entity* ents,
                               QueryRec* qs, uint2* sorted_pairs, hipStream_t st, const uint32_t* n_dev) {
    hipLaunchKernelGGL(k_pair_queries, g256(n), dim3(256), 0, st, pairs, order, n, n_dev, ents, qs, sorted_pairs);
    return hipGetLastError();
}
hipError_t launch_sort_perm(const cd_contact* in, uint32_t n, int f_hi, int f_mid, int f_lo,
                            const uint32_t field_bits[3], uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1,
                            uint32_t* table, hipStream_t st, const uint32_t* n_dev) {
    hipLaunchKernelGGL(k_iota, g256(n), dim3(256), 0, st, v0, n);
    for (int field : {f_lo, f_mid, f_hi}) {  // stable LSD: least significant field first
        hipLaunchKernelGGL(k_contact_keys, g256(n), dim3(256), 0, st, in, n, n_dev, field, v0, k0);
        hipError_t e = launch_radix_sort(k0, v0, k1, v1, n, field_bits[field], table, st, n_dev);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}
hipError_t launch_sort_contacts(const cd_contact* in, uint32_t n, cd_contact* out, const uint32_t field_bits[3],
                                uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, uint32_t* table,
                                hipStream_t st, const uint32_t* n_dev) {
    hipError_t e = launch_sort_perm(in, n, 0, 1, 2, field_bits, k0, v0, k1, v1, table, st, n_dev);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_gather_contacts, g256(n), dim3(256), 0, st, in, v0, n, n_dev, out);
    return hipGetLastError();
}

}  // namespace cdk
