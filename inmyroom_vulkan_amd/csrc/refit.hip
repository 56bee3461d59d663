// Deformable meshes: GPU skinning and OBB-tree refit.
//
// The reference never collides skinned meshes (only non-skin, non-morph
// primitives get trees: src/Graphics/Meshes/PrimitivesOfMeshes.cpp:642) and
// never refits a tree; this is new work on the reference's data model.
//  k_skin        linear blend skinning with the semantics of the reference's
//                skinning shader (shaders/dynamicMeshShader_glsl.comp:120-145):
//                p' = sum_k w_k (S_k * (p, 1)), S_k = joint matrix * inverse bind
//  k_gather_tris rebuilds the tree-order triangle arena from skinned vertices
//  k_refit_*     every box of the tree (root and both children of every
//                internal node) is recomputed over its own triangles with the
//                PCA construction of OBB.cpp:33-166 (mean, covariance, eig3,
//                axes = rows of V, extents + 2 FLT_EPSILON); the topology built
//                at registration is kept.  The FP64 sums are parallel
//                wavefront reductions, so the bits differ from a sequential
//                build; boxes still contain their triangles, and the
//                containment slack and radius of each mesh are recomputed.
#include <float.h>

#include "eig3.hpp"
#include "kernels.hpp"

namespace cdk {

__global__ void k_skin(const SkinRec* __restrict__ sk, uint32_t nsk, const uint32_t* __restrict__ vstart,
                       const float* __restrict__ bind, const ushort4* __restrict__ joints,
                       const float4* __restrict__ weights, const float* __restrict__ bones, uint32_t total,
                       float* __restrict__ out) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    // which listed mesh (vstart is the exclusive prefix of vertex counts)
    uint32_t lo = 0, hi = nsk;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (vstart[mid] <= g) lo = mid; else hi = mid;
    }
    const SkinRec& s = sk[lo];
    const uint32_t v = s.vbase + (g - vstart[lo]);
    const f3 p = mk(bind[3 * size_t(v)], bind[3 * size_t(v) + 1], bind[3 * size_t(v) + 2]);
    const ushort4 j = joints[v];
    const float4 w = weights[v];
    const float* B = bones + size_t(s.bone_base) * 16;
    const uint32_t jj[4] = {j.x, j.y, j.z, j.w};
    const float ww[4] = {w.x, w.y, w.z, w.w};
    f3 r = mk(0.f, 0.f, 0.f);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        M4 S;
        const float* src = B + size_t(min(jj[k], s.nbones - 1)) * 16;
#pragma unroll
        for (int e = 0; e < 16; ++e) S.m[e] = src[e];
        const f3 q = mul4(S, p, 1.f);
        r = add(r, mk(ww[k] * q.x, ww[k] * q.y, ww[k] * q.z));
    }
    out[3 * size_t(v)] = r.x;
    out[3 * size_t(v) + 1] = r.y;
    out[3 * size_t(v) + 2] = r.z;
}

__global__ void k_gather_tris(const SkinRec* __restrict__ sk, uint32_t nsk, const uint32_t* __restrict__ tstart,
                              const uint32_t* __restrict__ vidx, const float* __restrict__ pos, float* __restrict__ tris,
                              uint32_t total) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    uint32_t lo = 0, hi = nsk;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (tstart[mid] <= g) lo = mid; else hi = mid;
    }
    const SkinRec& s = sk[lo];
    const uint32_t local = g - tstart[lo];
    float* o = tris + (size_t(s.tri_base) + local) * 9;
    for (int k = 0; k < 3; ++k) {
        const uint32_t v = vidx[size_t(s.vidx_base) + size_t(local) * 3 + k];
        o[3 * k] = pos[3 * size_t(v)];
        o[3 * k + 1] = pos[3 * size_t(v) + 1];
        o[3 * k + 2] = pos[3 * size_t(v) + 2];
    }
}

// Triangle range (tree order, mesh-local) of the box a refit item names.
__device__ __forceinline__ void item_range(const MeshRec& m, const NodeRec* nodes, uint32_t code, uint32_t& off,
                                           uint32_t& cnt) {
    if (code == ROOT_CODE) {
        off = 0;
        cnt = m.ntri;
        return;
    }
    const NodeRec& n = nodes[m.node_base + (code >> 1)];
    const uint32_t side = code & 1;
    const uint32_t c = (n.cnt >> (16 * side)) & 0xFFFFu;
    if (c) {
        off = n.link[side];
        cnt = c;
    } else {
        const NodeRec& k = nodes[m.node_base + n.link[side]];
        off = k.pad[0];
        cnt = k.pad[1];
    }
}

// Wavefront reductions (butterfly: min and max are the same on every lane;
// sums are read from lane 0 only).
__device__ __forceinline__ double wave_sum_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_min_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_max_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Each box is refit in three launches, so the serial eigen-solve runs one
// box per lane instead of one lane per workgroup:
//   k_refit_moments  one wavefront per box: distinct-point check, mean and
//                    covariance (two passes, FP64)
//   k_refit_eig      one thread per box: eig3, axes = rows of V
//   k_refit_box      one wavefront per box: extents along the axes, the
//                    box and its surface, then the containment slack
// Scratch per box: RS_WORDS doubles.
constexpr int RS_MEAN = 0, RS_COV = 3, RS_AX = 9, RS_KIND = 18, RS_WORDS = 24;
constexpr int REFIT_WAVES = 4;  // boxes per 256-thread workgroup

__global__ __launch_bounds__(64 * REFIT_WAVES) void k_refit_moments(const uint2* __restrict__ items, uint32_t nitems,
                                                                    const MeshRec* __restrict__ meshes,
                                                                    const NodeRec* __restrict__ nodes,
                                                                    const float* __restrict__ tris,
                                                                    double* __restrict__ scratch) {
    const uint32_t it = blockIdx.x * REFIT_WAVES + (threadIdx.x >> 6);
    if (it >= nitems) return;  // whole wavefronts leave; no workgroup barrier follows
    const int lane = threadIdx.x & 63;
    const uint2 item = items[it];
    const MeshRec& m = meshes[item.x];
    uint32_t off, cnt;
    item_range(m, nodes, item.y, off, cnt);
    const float* P = tris + (size_t(m.tri_base) + off) * 9;
    const uint32_t npts = 3 * cnt;
    double* out = scratch + size_t(it) * RS_WORDS;
    // distinct-point check (OBB.cpp:35-50): more than one distinct point -> PCA
    bool differs = false;
    for (uint32_t k = 1 + lane; k < npts; k += 64)
        differs |= !(P[3 * k] == P[0] && P[3 * k + 1] == P[1] && P[3 * k + 2] == P[2]);
    const bool multi = __ballot(differs) != 0ull;
    if (npts == 0 || !multi) {
        if (lane == 0) out[RS_KIND] = npts == 0 ? 0.0 : 1.0;
        return;
    }
    const double N = double(npts);
    double sx = 0, sy = 0, sz = 0;
    for (uint32_t k = lane; k < npts; k += 64) {
        sx += double(P[3 * k]);
        sy += double(P[3 * k + 1]);
        sz += double(P[3 * k + 2]);
    }
    // every lane needs the mean: take lane 0's sums
    const double mx = __shfl(wave_sum_d(sx), 0, 64) / N;
    const double my = __shfl(wave_sum_d(sy), 0, 64) / N;
    const double mz = __shfl(wave_sum_d(sz), 0, 64) / N;
    double cxx = 0, cyy = 0, czz = 0, cxy = 0, cxz = 0, cyz = 0;
    for (uint32_t k = lane; k < npts; k += 64) {
        const double dx = double(P[3 * k]) - mx, dy = double(P[3 * k + 1]) - my, dz = double(P[3 * k + 2]) - mz;
        cxx += dx * dx; cyy += dy * dy; czz += dz * dz;
        cxy += dx * dy; cxz += dx * dz; cyz += dy * dz;
    }
    cxx = wave_sum_d(cxx); cyy = wave_sum_d(cyy); czz = wave_sum_d(czz);
    cxy = wave_sum_d(cxy); cxz = wave_sum_d(cxz); cyz = wave_sum_d(cyz);
    if (lane == 0) {
        out[RS_MEAN] = mx; out[RS_MEAN + 1] = my; out[RS_MEAN + 2] = mz;
        out[RS_COV] = cxx / N; out[RS_COV + 1] = cyy / N; out[RS_COV + 2] = czz / N;
        out[RS_COV + 3] = cxy / N; out[RS_COV + 4] = cxz / N; out[RS_COV + 5] = cyz / N;
        out[RS_KIND] = 2.0;
    }
}

__global__ __launch_bounds__(256) void k_refit_eig(uint32_t nitems, double* __restrict__ scratch) {
    const uint32_t it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= nitems) return;
    double* s = scratch + size_t(it) * RS_WORDS;
    if (s[RS_KIND] != 2.0) return;
    const double xx = s[RS_COV], yy = s[RS_COV + 1], zz = s[RS_COV + 2];
    const double xy = s[RS_COV + 3], xz = s[RS_COV + 4], yz = s[RS_COV + 5];
    double A[3][3] = {{xx, xy, xz}, {xy, yy, yz}, {xz, yz, zz}};
    double V[3][3], d[3];
    eig3(A, V, d);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) s[RS_AX + r * 3 + c] = V[r][c];  // rows of V (OBB.cpp:86-87)
}

__global__ __launch_bounds__(64 * REFIT_WAVES) void k_refit_box(const uint2* __restrict__ items, uint32_t nitems,
                                                                MeshRec* __restrict__ meshes,
                                                                NodeRec* __restrict__ nodes,
                                                                const float* __restrict__ tris,
                                                                const double* __restrict__ scratch,
                                                                unsigned int* __restrict__ slack_bits,
                                                                unsigned int* __restrict__ radius_bits) {
    const uint32_t it = blockIdx.x * REFIT_WAVES + (threadIdx.x >> 6);
    if (it >= nitems) return;
    const int lane = threadIdx.x & 63;
    const uint2 item = items[it];
    MeshRec& m = meshes[item.x];
    uint32_t off, cnt;
    item_range(m, nodes, item.y, off, cnt);
    const float* P = tris + (size_t(m.tri_base) + off) * 9;
    const uint32_t npts = 3 * cnt;
    const double* s = scratch + size_t(it) * RS_WORDS;
    Box box;
    if (s[RS_KIND] != 2.0) {  // OBB::EmptyOBB (OBB.cpp:168-178), centred on the single point if there is one
        const float e = FLT_EPSILON;
        box.c = npts ? mk(P[0], P[1], P[2]) : mk(0.f, 0.f, 0.f);
        box.u = mk(e, 0.f, 0.f);
        box.v = mk(0.f, e, 0.f);
        box.w = mk(0.f, 0.f, e);
    } else {
        double ax[9];
        for (int t = 0; t < 9; ++t) ax[t] = s[RS_AX + t];
        // OBB::CreateAABBfromPoints (OBB.cpp:105-166)
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t k = lane; k < npts; k += 64) {
            const double px = P[3 * k], py = P[3 * k + 1], pz = P[3 * k + 2];
            for (int q = 0; q < 3; ++q) {
                const double pr = (ax[3 * q] * px + ax[3 * q + 1] * py) + ax[3 * q + 2] * pz;
                lo[q] = fmin(lo[q], pr);
                hi[q] = fmax(hi[q], pr);
            }
        }
        for (int q = 0; q < 3; ++q) {
            lo[q] = wave_min_d(lo[q]);
            hi[q] = wave_max_d(hi[q]);
        }
        double cx = 0.0, cy = 0.0, cz = 0.0;
        float side[3][3];
        for (int q = 0; q < 3; ++q) {
            const double delta = (hi[q] - lo[q]) + 2.0 * double(FLT_EPSILON);
            const double mid = (hi[q] + lo[q]) / 2.0;
            cx = cx + mid * ax[3 * q];
            cy = cy + mid * ax[3 * q + 1];
            cz = cz + mid * ax[3 * q + 2];
            const double half = delta / 2.0;
            side[q][0] = float(half * ax[3 * q]);
            side[q][1] = float(half * ax[3 * q + 1]);
            side[q][2] = float(half * ax[3 * q + 2]);
        }
        box.c = mk(float(cx), float(cy), float(cz));
        box.u = mk(side[0][0], side[0][1], side[0][2]);
        box.v = mk(side[1][0], side[1][1], side[1][2]);
        box.w = mk(side[2][0], side[2][1], side[2][2]);
    }
    // every lane holds the same box (min and max reductions are exact)
    if (lane == 0) {
        const float sf = surface(box);
        if (item.y == ROOT_CODE) {
            box_to12(box, m.root);
            m.root_surf = sf;
        } else {
            NodeRec& n = nodes[m.node_base + (item.y >> 1)];
            box_to12(box, n.box[item.y & 1]);
            n.surf[item.y & 1] = sf;
        }
    }
    // containment slack: how far any of the box's points lies outside it
    // (the batched traversal widens its separating-axis test by this much)
    const double c[3] = {box.c.x, box.c.y, box.c.z};
    const double u[3] = {box.u.x, box.u.y, box.u.z}, v[3] = {box.v.x, box.v.y, box.v.z};
    const double w[3] = {box.w.x, box.w.y, box.w.z};
    auto det3 = [](const double* a, const double* b, const double* cc) {
        return a[0] * (b[1] * cc[2] - b[2] * cc[1]) - a[1] * (b[0] * cc[2] - b[2] * cc[0]) +
               a[2] * (b[0] * cc[1] - b[1] * cc[0]);
    };
    const double D = det3(u, v, w);
    const double lu = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double lv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double lw = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    double worst = 0.0, rad = 0.0;
    for (uint32_t k = lane; k < npts; k += 64) {
        const double p[3] = {P[3 * k], P[3 * k + 1], P[3 * k + 2]};
        const double r[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
        double ex = INFINITY;
        if (fabs(D) > 0.0) {
            const double al = det3(r, v, w) / D, be = det3(u, r, w) / D, ga = det3(u, v, r) / D;
            ex = fmax(0.0, fabs(al) - 1.0) * lu + fmax(0.0, fabs(be) - 1.0) * lv + fmax(0.0, fabs(ga) - 1.0) * lw;
        }
        worst = fmax(worst, ex);
        rad = fmax(rad, sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]));
    }
    worst = wave_max_d(worst);
    rad = wave_max_d(rad);
    if (lane == 0) {
        float wf = float(worst);
        if (double(wf) < worst) wf = nextafterf(wf, INFINITY);
        float rf = float(rad);
        if (double(rf) < rad) rf = nextafterf(rf, INFINITY);
        atomicMax(&slack_bits[item.x], __float_as_uint(wf));
        if (item.y == ROOT_CODE) atomicMax(&radius_bits[item.x], __float_as_uint(rf));
    }
}

__global__ void k_store_slack(MeshRec* meshes, const uint32_t* ids, uint32_t n, const unsigned int* slack_bits,
                              const unsigned int* radius_bits) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    MeshRec& m = meshes[ids[i]];
    m.slack = __uint_as_float(slack_bits[ids[i]]);
    m.radius = __uint_as_float(radius_bits[ids[i]]);
}

__global__ void k_tree_vidx(const uint32_t* __restrict__ perm, uint32_t ntri, const uint32_t* __restrict__ indices,
                            uint32_t vbase, uint32_t* __restrict__ out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ntri) return;
    const uint32_t t = perm[s];
    for (int k = 0; k < 3; ++k) out[size_t(s) * 3 + k] = vbase + indices[size_t(t) * 3 + k];
}

// ---- launchers ---------------------------------------------------------------
static inline dim3 g256r(uint64_t n) { return dim3(uint32_t((std::max<uint64_t>(n, 1) + 255) / 256)); }

hipError_t launch_tree_vidx(const uint32_t* perm, uint32_t ntri, const uint32_t* indices, uint32_t vbase,
                            uint32_t* out, hipStream_t st) {
    hipLaunchKernelGGL(k_tree_vidx, g256r(ntri), dim3(256), 0, st, perm, ntri, indices, vbase, out);
    return hipGetLastError();
}
hipError_t launch_skin(const SkinRec* sk, uint32_t nsk, const uint32_t* vstart, const float* bind, const ushort4* joints,
                       const float4* weights, const float* bones, uint32_t total, float* out, hipStream_t st) {
    hipLaunchKernelGGL(k_skin, g256r(total), dim3(256), 0, st, sk, nsk, vstart, bind, joints, weights, bones, total, out);
    return hipGetLastError();
}
hipError_t launch_gather_tris(const SkinRec* sk, uint32_t nsk, const uint32_t* tstart, const uint32_t* vidx,
                              const float* pos, float* tris, uint32_t total, hipStream_t st) {
    hipLaunchKernelGGL(k_gather_tris, g256r(total), dim3(256), 0, st, sk, nsk, tstart, vidx, pos, tris, total);
    return hipGetLastError();
}
hipError_t launch_refit(const uint2* items, uint32_t nitems, MeshRec* meshes, NodeRec* nodes, const float* tris,
                        double* scratch, unsigned int* slack_bits, unsigned int* radius_bits, const uint32_t* ids,
                        uint32_t nids, hipStream_t st) {
    if (!nitems) return hipSuccess;
    const dim3 gw((nitems + REFIT_WAVES - 1) / REFIT_WAVES);
    hipLaunchKernelGGL(k_refit_moments, gw, dim3(64 * REFIT_WAVES), 0, st, items, nitems, meshes, nodes, tris, scratch);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_refit_eig, g256r(nitems), dim3(256), 0, st, nitems, scratch);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_refit_box, gw, dim3(64 * REFIT_WAVES), 0, st, items, nitems, meshes, nodes, tris, scratch,
                       slack_bits, radius_bits);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_store_slack, g256r(nids), dim3(256), 0, st, meshes, ids, nids, slack_bits, radius_bits);
    return hipGetLastError();
}

}  // namespace cdk
