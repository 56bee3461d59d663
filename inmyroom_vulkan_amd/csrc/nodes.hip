// Global matrices of a node hierarchy (the transform producer in front of
// the collision frame): NodeDataCompEntity::GetGlobalMatrix
// (src/ECS/GeneralCompEntities/NodeDataCompEntity.cpp:63-72) applied from each
// root down, as Early/LateNodeGlobalMatrixCompEntity::Update do one entity at
// a time (EarlyNodeGlobalMatrixCompEntity.cpp:63-78).
//
// One lane per node: the lane rebuilds its own chain from the root, each
// ancestor's matrix by the same float operations, so every lane reproduces
// the matrices its ancestors' lanes write and no level-by-level launches or
// parent-first ordering are needed.  A chain of depth d costs d products of
// five 4x4 matrices and d(d+1)/2 parent reads; scene hierarchies are shallow.
#include "kernels.hpp"

namespace cdk {

__global__ void k_node_globals(const cd_node* __restrict__ nodes, const uint32_t* __restrict__ depth, uint32_t n,
                               float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t d = depth[i];
    M4 G = identity4();
    for (uint32_t level = 0; level <= d; ++level) {
        uint32_t a = i;  // the ancestor at this level: d - level steps up
        for (uint32_t s = level; s < d; ++s) a = nodes[a].parent;
        const cd_node& nd = nodes[a];
        const float rec[13] = {nd.local_scale[0], nd.local_scale[1], nd.local_scale[2], nd.local_rotation[0],
                               nd.local_rotation[1], nd.local_rotation[2], nd.local_rotation[3],
                               nd.local_translation[0], nd.local_translation[1], nd.local_translation[2],
                               nd.global_translation[0], nd.global_translation[1], nd.global_translation[2]};
        G = node_global(rec, G);
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) out[size_t(i) * 16 + k] = G.m[k];
}

hipError_t launch_node_globals(const cd_node* nodes, const uint32_t* depth, uint32_t n, float* out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_node_globals, dim3((n + 255) / 256), dim3(256), 0, st, nodes, depth, n, out);
    return hipGetLastError();
}

}  // namespace cdk
