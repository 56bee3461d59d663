// Glue of the one-call scene step (cd_step): the counts that the separate
// calls read back between their stages stay on the device, in a count block
// (DC_* words), so that the whole step, from the entity matrices to the pair
// responses, is enqueued at once and can be replayed as one graph.  Every
// stage is shaped for a capacity and reads its count from the block; the
// host checks the counts once, after the step, and runs it again with larger
// capacities when one was exceeded.
//  k_step_pairs   pairs the sweep kept: the batched traversal's query count
//  k_step_contacts contacts the later stages take
//  k_seg_*        the contacts grouped by pair and ordered within each pair
//                 (below)
//  k_pair_begins  per side, the first ray of every pair (exclusive scan of
//                 the pairs' ray counts)
//  k_pack         the step's results gathered into one block for a single
//                 copy to the host
#include "kernels.hpp"

namespace cdk {

__global__ void k_step_pairs(const unsigned int* __restrict__ sweep, uint32_t cap, uint32_t* __restrict__ dc,
                             unsigned int* __restrict__ d_nq) {
    if (threadIdx.x != 0) return;
    const uint32_t found = *sweep;
    const uint32_t kept = min(found, cap);
    dc[DC_SWEEP] = found;
    dc[DC_NQ] = kept;
    *d_nq = kept;
}

// qcnt (np words, when given) cleared for k_seg_count
__global__ __launch_bounds__(1024) void k_step_contacts(const RunCounters* __restrict__ ctr, uint64_t cap,
                                                        uint32_t* __restrict__ dc, uint32_t* __restrict__ qcnt,
                                                        uint32_t np) {
    if (threadIdx.x == 0) dc[DC_NC] = uint32_t(min(ctr->contacts, cap));
    if (qcnt)
        for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) qcnt[i] = 0u;
}

// ---- contact order ------------------------------------------------------------
// The rays need a run's contacts ordered by (pair, tri_a, tri_b), and a
// permutation of them by (pair, tri_b, tri_a).  A run tests a triangle pair
// of an entity pair once, so these keys are distinct and every sort gives
// the order the separate calls' multi-pass radix sorts give.  Here:
//  k_seg_count    contacts per pair
//  k_scan_one     each pair's first slot (and a copy the scatter advances)
//  k_seg_scatter  every contact into its pair's slots, in any order
//  k_seg_sort     one workgroup per pair: its contacts' (tri_a, tri_b) keys
//                 sorted in LDS by a bitonic network (the slot of each in the
//                 sorted array), then their (tri_b, tri_a) keys with those
//                 slots (the second order).  A pair of more than SEG_MAX
//                 contacts sets dc[DC_BIGSEG]; the step then runs again with
//                 the radix sorts.
// The count and the scatter add to a pair's counter once per wavefront and
// pair (a run's contacts of one pair lie together), not once per contact.
constexpr uint32_t SEG_MAX = 2048;
constexpr uint32_t SEG_THREADS = 1024;  // one compare-exchange per thread and stage at SEG_MAX

// Slot of each lane's contact among its pair's (base + rank among the
// wavefront's lanes of that pair); valid lanes only.  One atomic add per
// pair the wavefront holds.
__device__ __forceinline__ uint32_t seg_slot(uint32_t* __restrict__ ctr, uint32_t q, bool valid) {
    const uint32_t lane = threadIdx.x & 63;
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    unsigned long long todo = __ballot(valid);
    uint32_t slot = 0;
    while (todo) {
        const int leader = __ffsll(static_cast<long long>(todo)) - 1;
        const uint32_t lq = __shfl(q, leader, 64);
        const unsigned long long peers = __ballot(valid && q == lq);
        uint32_t base = 0;
        if (int(lane) == leader) base = atomicAdd(&ctr[lq], uint32_t(__popcll(peers)));
        base = __shfl(base, leader, 64);
        if (valid && q == lq) slot = base + uint32_t(__popcll(peers & below));
        todo &= ~peers;
    }
    return slot;
}

__global__ __launch_bounds__(256) void k_seg_count(const cd_contact* __restrict__ c, uint32_t cap,
                                                   const uint32_t* __restrict__ dc, uint32_t* __restrict__ qcnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < min(dc[DC_NC], cap);
    (void)seg_slot(qcnt, valid ? c[i].query : 0u, valid);
}

// Exclusive scan of n_dev (min with cap) counts in one workgroup, into out
// and out2.
__global__ __launch_bounds__(1024) void k_scan_one(const uint32_t* __restrict__ in, uint32_t cap,
                                                   const uint32_t* __restrict__ n_dev, uint32_t* __restrict__ out,
                                                   uint32_t* __restrict__ out2) {
    __shared__ uint32_t s_w[16];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t n = min(*n_dev, cap);
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n; base += 1024) {
        const uint32_t i = base + tid;
        const uint32_t v = i < n ? in[i] : 0u;
        uint32_t x = v;
#pragma unroll
        for (uint32_t off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63) s_w[wave] = x;
        __syncthreads();
        uint32_t before = 0, all = 0;
        for (uint32_t w = 0; w < 16; ++w) {
            before += w < wave ? s_w[w] : 0u;
            all += s_w[w];
        }
        if (i < n) {
            out[i] = carry + before + x - v;
            out2[i] = carry + before + x - v;
        }
        carry += all;
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_seg_scatter(const cd_contact* __restrict__ c, uint32_t cap,
                                                     const uint32_t* __restrict__ dc, uint32_t* __restrict__ qfill,
                                                     cd_contact* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < min(dc[DC_NC], cap);
    cd_contact r = {};
    if (valid) r = c[i];
    const uint32_t slot = seg_slot(qfill, r.query, valid);
    if (valid) out[slot] = r;
}

// Ascending bitonic sort of m (a power of two, up to SEG_MAX) keys with
// their values in LDS, by the workgroup.
__device__ __forceinline__ void bitonic_lds(unsigned long long* key, uint32_t* val, uint32_t m) {
    for (uint32_t k = 2; k <= m; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t t = threadIdx.x; t < m / 2; t += blockDim.x) {
                const uint32_t i = 2 * t - (t & (j - 1)), l = i + j;
                const unsigned long long ki = key[i], kl = key[l];
                if ((ki > kl) == ((i & k) == 0)) {
                    key[i] = kl;
                    key[l] = ki;
                    const uint32_t v = val[i];
                    val[i] = val[l];
                    val[l] = v;
                }
            }
            __syncthreads();
        }
}

__global__ __launch_bounds__(SEG_THREADS) void k_seg_sort(const cd_contact* __restrict__ seg,
                                                          const uint32_t* __restrict__ qoff,
                                                          const uint32_t* __restrict__ qcnt, uint32_t* __restrict__ dc,
                                                          uint32_t cap, cd_contact* __restrict__ sorted,
                                                          uint32_t* __restrict__ order_b) {
    __shared__ unsigned long long s_key[SEG_MAX], s_kb[SEG_MAX];  // sort keys; (tri_b, tri_a) of each contact
    __shared__ uint32_t s_val[SEG_MAX];
    const uint32_t q = blockIdx.x;
    if (q >= min(dc[DC_NQ], cap)) return;
    const uint32_t n = qcnt[q], base = qoff[q];
    if (n == 0) return;
    if (n > SEG_MAX) {
        if (threadIdx.x == 0) dc[DC_BIGSEG] = 1u;
        return;
    }
    uint32_t m = 2;
    while (m < n) m <<= 1;
    // keys of the pair's contacts (a run's triangle indices are below 2^25,
    // so the padding key ~0 sorts after them all)
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
        unsigned long long ka = ~0ull, kb = ~0ull;
        if (i < n) {
            const cd_contact& r = seg[base + i];
            ka = (static_cast<unsigned long long>(r.tri_a) << 32) | r.tri_b;
            kb = (static_cast<unsigned long long>(r.tri_b) << 32) | r.tri_a;
        }
        s_key[i] = ka;
        s_kb[i] = kb;
        s_val[i] = i;
    }
    __syncthreads();
    bitonic_lds(s_key, s_val, m);
    // slot r of the sorted array holds contact s_val[r]; the second order
    // sorts the contacts' (tri_b, tri_a) keys with their slots
    for (uint32_t r = threadIdx.x; r < n; r += blockDim.x) sorted[base + r] = seg[base + s_val[r]];
    for (uint32_t r = threadIdx.x; r < m; r += blockDim.x) s_key[r] = r < n ? s_kb[s_val[r]] : ~0ull;
    __syncthreads();
    for (uint32_t r = threadIdx.x; r < m; r += blockDim.x) s_val[r] = r;
    __syncthreads();
    bitonic_lds(s_key, s_val, m);
    for (uint32_t r = threadIdx.x; r < n; r += blockDim.x) order_b[base + r] = base + s_val[r];
}

// One workgroup: each side's counts (qw + 2 stride side) scanned 1024 at a
// time, the wavefronts' sums combined through LDS.
__global__ __launch_bounds__(1024) void k_pair_begins(const uint32_t* __restrict__ qw, uint32_t stride,
                                                      const uint32_t* __restrict__ dc, uint32_t* __restrict__ pbegin) {
    __shared__ uint32_t s_w[16];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t nq = min(dc[DC_NQ], stride);
    for (int side = 0; side < 2; ++side) {
        const uint32_t* cnt = qw + 2 * size_t(stride) * side;
        uint32_t* out = pbegin + size_t(stride) * side;
        uint32_t carry = 0;
        for (uint32_t base = 0; base < nq; base += 1024) {
            const uint32_t i = base + tid;
            const uint32_t v = i < nq ? cnt[i] : 0u;
            uint32_t x = v;
#pragma unroll
            for (uint32_t off = 1; off < 64; off <<= 1) {
                const uint32_t y = __shfl_up(x, off, 64);
                if (lane >= off) x += y;
            }
            if (lane == 63) s_w[wave] = x;
            __syncthreads();
            uint32_t before = 0, all = 0;
            for (uint32_t w = 0; w < 16; ++w) {
                before += w < wave ? s_w[w] : 0u;
                all += s_w[w];
            }
            if (i < nq) out[i] = carry + before + x - v;
            carry += all;
            __syncthreads();
        }
    }
}

__global__ void k_pack(PackSpans sp, uint32_t* __restrict__ dst) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t s = 0; s < sp.n; ++s) {
        const uint32_t* src = sp.src[s];
        uint32_t* d = dst + sp.dst[s];
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < sp.words[s]; i += stride) d[i] = src[i];
    }
}

hipError_t launch_step_pairs(const unsigned int* sweep, uint32_t cap, uint32_t* dc, unsigned int* d_nq, hipStream_t st) {
    hipLaunchKernelGGL(k_step_pairs, dim3(1), dim3(64), 0, st, sweep, cap, dc, d_nq);
    return hipGetLastError();
}
hipError_t launch_step_contacts(const RunCounters* ctr, uint64_t cap, uint32_t* dc, hipStream_t st, uint32_t* qcnt,
                                uint32_t np) {
    hipLaunchKernelGGL(k_step_contacts, dim3(1), dim3(1024), 0, st, ctr, cap, dc, qcnt, np);
    return hipGetLastError();
}
hipError_t launch_seg_order(const cd_contact* c, uint32_t nc_cap, uint32_t np_cap, uint32_t* dc, uint32_t* qw3,
                            cd_contact* seg, cd_contact* sorted, uint32_t* order_b, hipStream_t st) {
    uint32_t* qcnt = qw3;
    uint32_t* qoff = qw3 + np_cap;
    uint32_t* qfill = qw3 + 2 * size_t(np_cap);
    const dim3 gc((std::max(nc_cap, 1u) + 255) / 256);
    hipLaunchKernelGGL(k_seg_count, gc, dim3(256), 0, st, c, nc_cap, dc, qcnt);
    hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(1024), 0, st, qcnt, np_cap, dc + DC_NQ, qoff, qfill);
    hipLaunchKernelGGL(k_seg_scatter, gc, dim3(256), 0, st, c, nc_cap, dc, qfill, seg);
    hipLaunchKernelGGL(k_seg_sort, dim3(std::max(np_cap, 1u)), dim3(SEG_THREADS), 0, st, seg, qoff, qcnt, dc, np_cap,
                       sorted, order_b);
    return hipGetLastError();
}
hipError_t launch_pair_begins(const uint32_t* qw, uint32_t stride, const uint32_t* dc, uint32_t* pbegin, hipStream_t st) {
    hipLaunchKernelGGL(k_pair_begins, dim3(1), dim3(1024), 0, st, qw, stride, dc, pbegin);
    return hipGetLastError();
}
hipError_t launch_pack(const PackSpans& sp, uint32_t* dst, hipStream_t st) {
    if (sp.n > PackSpans::MAX) return hipErrorInvalidValue;
    uint64_t most = 0;
    for (uint32_t s = 0; s < sp.n; ++s) most = std::max<uint64_t>(most, sp.words[s]);
    const uint32_t blocks = uint32_t(std::min<uint64_t>(256, std::max<uint64_t>(1, (most + 255) / 256)));
    hipLaunchKernelGGL(k_pack, dim3(blocks), dim3(256), 0, st, sp, dst);
    return hipGetLastError();
}

}  // namespace cdk
