// Mesh ingest on the GPU: the index lists of the glTF primitives of many
// meshes become their triangle soups and per-corner vertex ids in one launch,
// ready for the tree build.
//
// Triangle j of a primitive takes the index-list entries its draw mode names
// (Triangle::CreateIndicesTriplets, src/Geometry/Triangle.cpp:10-62):
//   points (j, j, j)            lines (2j, 2j, 2j+1)       line strip (j, j, j+1)
//   triangles (3j, 3j+1, 3j+2)  strip (j, j+1+j%2, j+2-j%2) fan (j+1, j+2, 0)
// Positions are gathered through those indices (TrianglePosition), vertex ids
// are the indices themselves (TriangleIndices, :236-257) offset by the
// primitive's first vertex within its mesh.  The work is a gather: 36 bytes
// of positions and 12 of indices read, 48 bytes written per triangle.
#include "kernels.hpp"

namespace cdk {

__global__ void k_assemble(const PrimSpan* __restrict__ prims, uint32_t nprim, const uint32_t* __restrict__ idx,
                           const float* __restrict__ pos, uint32_t total, float* __restrict__ tri9,
                           uint32_t* __restrict__ vid3) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    uint32_t lo = 0, hi = nprim;  // the last primitive whose first triangle is <= t
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (prims[mid].tri_off <= t) lo = mid;
        else hi = mid;
    }
    const PrimSpan p = prims[lo];
    const uint32_t j = t - p.tri_off;
    uint32_t e0, e1, e2;
    switch (p.mode) {
        case 0: e0 = e1 = e2 = j; break;
        case 1: e0 = e1 = 2u * j; e2 = 2u * j + 1u; break;
        case 3: e0 = e1 = j; e2 = j + 1u; break;
        case 5: e0 = j; e1 = j + 1u + (j & 1u); e2 = j + 2u - (j & 1u); break;
        case 6: e0 = j + 1u; e1 = j + 2u; e2 = 0u; break;
        default: e0 = 3u * j; e1 = 3u * j + 1u; e2 = 3u * j + 2u; break;
    }
    const uint32_t v[3] = {idx[p.idx_off + e0], idx[p.idx_off + e1], idx[p.idx_off + e2]};
    float* o = tri9 + size_t(t) * 9;
    for (int k = 0; k < 3; ++k) {
        const float* q = pos + (size_t(p.vbase) + v[k]) * 3;
        o[3 * k] = q[0];
        o[3 * k + 1] = q[1];
        o[3 * k + 2] = q[2];
        vid3[size_t(t) * 3 + k] = p.vid_base + v[k];
    }
}

hipError_t launch_assemble(const PrimSpan* prims, uint32_t nprim, const uint32_t* idx, const float* pos, uint32_t total,
                           float* tri9, uint32_t* vid3, hipStream_t st) {
    if (total == 0 || nprim == 0) return hipSuccess;
    hipLaunchKernelGGL(k_assemble, dim3((total + 255) / 256), dim3(256), 0, st, prims, nprim, idx, pos, total, tri9, vid3);
    return hipGetLastError();
}

}  // namespace cdk
