// Glue of the one-call scene step (cd_step): the counts that the separate
// calls read back between their stages stay on the device, in a count block
// (DC_* words), so that the whole step, from the entity matrices to the pair
// responses, is enqueued at once and can be replayed as one graph.  Every
// stage is shaped for a capacity and reads its count from the block; the
// host checks the counts once, after the step, and runs it again with larger
// capacities when one was exceeded.
//  k_step_pairs   pairs the sweep kept: the batched traversal's query count
//  k_step_contacts contacts the later stages take
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

__global__ void k_step_contacts(const RunCounters* __restrict__ ctr, uint64_t cap, uint32_t* __restrict__ dc) {
    if (threadIdx.x == 0) dc[DC_NC] = uint32_t(min(ctr->contacts, cap));
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
hipError_t launch_step_contacts(const RunCounters* ctr, uint64_t cap, uint32_t* dc, hipStream_t st) {
    hipLaunchKernelGGL(k_step_contacts, dim3(1), dim3(64), 0, st, ctr, cap, dc);
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
