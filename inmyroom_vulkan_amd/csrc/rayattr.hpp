// Triangle attributes shared by ray generation and ray shooting: the
// normal matrix, corner normals and vertex ids of tree-order triangles.
#pragma once
#include "kernels.hpp"

namespace cdk {

// glm::adjointTranspose(mat3(M)) (Triangle.cpp:175-178): the cofactor matrix
// of the upper-left 3x3, column-major like glm.
struct M3 { float m[9]; };
__device__ __forceinline__ M3 adjoint_transpose(const float* M) {
    auto a = [&](int c, int r) { return M[c * 4 + r]; };
    M3 R;
    R.m[0] = +(a(1, 1) * a(2, 2) - a(2, 1) * a(1, 2));
    R.m[1] = -(a(1, 0) * a(2, 2) - a(2, 0) * a(1, 2));
    R.m[2] = +(a(1, 0) * a(2, 1) - a(2, 0) * a(1, 1));
    R.m[3] = -(a(0, 1) * a(2, 2) - a(2, 1) * a(0, 2));
    R.m[4] = +(a(0, 0) * a(2, 2) - a(2, 0) * a(0, 2));
    R.m[5] = -(a(0, 0) * a(2, 1) - a(2, 0) * a(0, 1));
    R.m[6] = +(a(0, 1) * a(1, 2) - a(1, 1) * a(0, 2));
    R.m[7] = -(a(0, 0) * a(1, 2) - a(1, 0) * a(0, 2));
    R.m[8] = +(a(0, 0) * a(1, 1) - a(1, 0) * a(0, 1));
    return R;
}
__device__ __forceinline__ f3 mul3(const M3& M, f3 v) {
    return mk((M.m[0] * v.x + M.m[3] * v.y) + M.m[6] * v.z, (M.m[1] * v.x + M.m[4] * v.y) + M.m[7] * v.z,
              (M.m[2] * v.x + M.m[5] * v.y) + M.m[8] * v.z);
}
__device__ __forceinline__ f3 scale(f3 v, float s) { return mk(v.x * s, v.y * s, v.z * s); }
__device__ __forceinline__ f3 divs(f3 v, float s) { return mk(v.x / s, v.y / s, v.z / s); }

__device__ __forceinline__ uint32_t vertex_id(const MeshRec& m, const uint2& attr, const uint32_t* vid,
                                              const uint32_t* perm, uint32_t tri, int k) {
    // index-buffer value of the corner, or 3 t + k for a triangle soup
    return attr.x == ~0u ? 3u * perm[m.tri_base + tri] + uint32_t(k) : vid[attr.x + 3u * tri + uint32_t(k)];
}

// Corner normal before the normal matrix: the mesh's vertex normal, or the
// face normal of the model-space triangle (TriangleNormal, Triangle.cpp:140-146, 214-234).
__device__ __forceinline__ f3 corner_normal(const uint2& attr, const float* nrm, uint32_t v, const float* tri9) {
    if (attr.y != ~0u) {
        const float* n = nrm + size_t(attr.y) + 3u * size_t(v);
        return mk(n[0], n[1], n[2]);
    }
    const f3 p0 = mk(tri9[0], tri9[1], tri9[2]), p1 = mk(tri9[3], tri9[4], tri9[5]), p2 = mk(tri9[6], tri9[7], tri9[8]);
    return normalize(cross(sub(p1, p0), sub(p2, p0)));
}

}  // namespace cdk
