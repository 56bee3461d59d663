// Deformable meshes: GPU skinning and OBB-tree refit.
//
// The reference never collides skinned meshes (only non-skin, non-morph
// primitives get trees: src/Graphics/Meshes/PrimitivesOfMeshes.cpp:642) and
// never refits a tree; this is new work on the reference's data model.
//  k_skin        linear blend skinning with the semantics of the reference's
//                skinning shader (shaders/dynamicMeshShader_glsl.comp:120-145):
//                p' = sum_k w_k (S_k * (p, 1)), S_k = joint matrix * inverse bind
//  k_gather_tris rebuilds the tree-order triangle arena from skinned vertices
//  k_rf_*        every box of the tree (root and both children of every
//                internal node) is recomputed over its own triangles with the
//                PCA construction of OBB.cpp:33-166 (mean, covariance, eig3,
//                axes = rows of V, extents + 2 FLT_EPSILON); the topology built
//                at registration is kept.  The FP64 sums of a large box are
//                chunked parallel reductions, so the bits differ from a sequential
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
    const ChildRec& c = nodes[m.node_base + (code >> 1)].c[code & 1];
    off = c.toff;
    cnt = child_tris(c);
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

// Boxes are refit by size class, so that neither the serial eigen-solve nor
// the passes over a large box's points sit on one wavefront:
//   k_rf_small   one thread per small box (<= RF_SMALL triangles): the whole
//                construction in registers
//   large boxes  their triangles split into chunks of RF_CHUNK; per pass one
//                wavefront per chunk writes a partial result and one thread
//                per box combines its chunks in chunk order:
//                sums -> mean, centred products -> covariance and eig3,
//                projections -> extents and the box, then the slack
// Containment slack and radius end in per-mesh atomic maxima.
constexpr uint32_t RF_SMALL = 16, RF_CHUNK = 512;
constexpr int RS_MEAN = 0, RS_COV = 3, RS_AX = 9, RS_KIND = 18, RS_WORDS = 24;
constexpr int PW = 8;  // doubles per chunk partial

__device__ __forceinline__ const float* item_tris(const MeshRec& m, const float* tris, uint32_t off) {
    return tris + (size_t(m.tri_base) + off) * 9;
}

// OBB::CreateAABBfromPoints tail (OBB.cpp:120-166): box from axes and extents.
__device__ inline Box box_from_extents(const double* ax, const double* lo, const double* hi) {
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
    Box box;
    box.c = mk(float(cx), float(cy), float(cz));
    box.u = mk(side[0][0], side[0][1], side[0][2]);
    box.v = mk(side[1][0], side[1][1], side[1][2]);
    box.w = mk(side[2][0], side[2][1], side[2][2]);
    return box;
}
// OBB::EmptyOBB (OBB.cpp:168-178), centred on the single point if there is one.
__device__ inline Box empty_box(const float* P, uint32_t npts) {
    const float e = FLT_EPSILON;
    Box box;
    box.c = npts ? mk(P[0], P[1], P[2]) : mk(0.f, 0.f, 0.f);
    box.u = mk(e, 0.f, 0.f);
    box.v = mk(0.f, e, 0.f);
    box.w = mk(0.f, 0.f, e);
    return box;
}
__device__ inline void axes_of(double xx, double yy, double zz, double xy, double xz, double yz, double* ax) {
    double A[3][3] = {{xx, xy, xz}, {xy, yy, yz}, {xz, yz, zz}};
    double V[3][3], d[3];
    eig3(A, V, d);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) ax[r * 3 + c] = V[r][c];  // rows of V (OBB.cpp:86-87)
}
__device__ inline void store_box(MeshRec& m, NodeRec* nodes, uint32_t code, const Box& box) {
    const float sf = surface(box);
    ChildRec& c = code == ROOT_CODE ? m.root : nodes[m.node_base + (code >> 1)].c[code & 1];
    box_to12(box, c.box);
    c.surf = sf;
}
// How far a point lies outside a box (distance along each side beyond the
// box), in FP64 from the stored float box.
// A box's frame for the containment slack: a point's coordinates in the
// box's side vectors are its offset dotted with the dual basis (v x w, w x u,
// u x v) / det(u, v, w), computed once per box.
struct SlackFrame {
    double c[3], du[3], dv[3], dw[3], lu, lv, lw;
    bool flat;
};
__device__ inline void cross_d(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ inline SlackFrame slack_frame(const float* bx) {
    SlackFrame f;
    double u[3], v[3], w[3];
    for (int k = 0; k < 3; ++k) {
        f.c[k] = bx[k];
        u[k] = bx[3 + k];
        v[k] = bx[6 + k];
        w[k] = bx[9 + k];
    }
    double vw[3], wu[3], uv[3];
    cross_d(v, w, vw);
    cross_d(w, u, wu);
    cross_d(u, v, uv);
    const double D = u[0] * vw[0] + u[1] * vw[1] + u[2] * vw[2];
    f.flat = !(fabs(D) > 0.0);
    for (int k = 0; k < 3; ++k) {
        f.du[k] = vw[k] / D;
        f.dv[k] = wu[k] / D;
        f.dw[k] = uv[k] / D;
    }
    f.lu = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    f.lv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    f.lw = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    return f;
}
// How far p lies outside the box, measured along its side vectors.
__device__ inline double excess(const SlackFrame& f, const double* p) {
    if (f.flat) return INFINITY;
    const double r[3] = {p[0] - f.c[0], p[1] - f.c[1], p[2] - f.c[2]};
    const double al = r[0] * f.du[0] + r[1] * f.du[1] + r[2] * f.du[2];
    const double be = r[0] * f.dv[0] + r[1] * f.dv[1] + r[2] * f.dv[2];
    const double ga = r[0] * f.dw[0] + r[1] * f.dw[1] + r[2] * f.dw[2];
    return fmax(0.0, fabs(al) - 1.0) * f.lu + fmax(0.0, fabs(be) - 1.0) * f.lv + fmax(0.0, fabs(ga) - 1.0) * f.lw;
}
// Per-mesh maxima: the slack bound goes to one of SLACK_REPS copies per mesh
// (chosen by the caller, e.g. its chunk or wavefront index), so the
// thousands of boxes of a mesh do not all queue on one address;
// k_store_slack takes the largest copy.
__device__ inline void post_slack(unsigned int* slack_bits, unsigned int* radius_bits, uint32_t mesh, bool root,
                                  double worst, double rad, uint32_t rep) {
    float wf = float(worst);
    if (double(wf) < worst) wf = nextafterf(wf, INFINITY);
    float rf = float(rad);
    if (double(rf) < rad) rf = nextafterf(rf, INFINITY);
    atomicMax(&slack_bits[size_t(mesh) * SLACK_REPS + (rep % SLACK_REPS)], __float_as_uint(wf));
    if (root) atomicMax(&radius_bits[mesh], __float_as_uint(rf));
}
__device__ inline const float* box_ptr(const MeshRec& m, const NodeRec* nodes, uint32_t code) {
    return code == ROOT_CODE ? m.root.box : nodes[m.node_base + (code >> 1)].c[code & 1].box;
}

// Triangle range of every item (tree order, mesh-local).
__global__ void k_item_ranges(const uint2* __restrict__ items, uint32_t nitems, const MeshRec* __restrict__ meshes,
                              const NodeRec* __restrict__ nodes, uint2* __restrict__ ranges) {
    const uint32_t it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= nitems) return;
    const uint2 item = items[it];
    uint32_t off, cnt;
    item_range(meshes[item.x], nodes, item.y, off, cnt);
    ranges[it] = make_uint2(off, cnt);
}

// Visits points first, first + step, ... below npts in that order, with the
// loads of RF_GROUP of them issued before any is used: the refit's loops
// walk a box's points one per step, and one point at a time waits on memory
// once per point.
constexpr int RF_GROUP = 8;
template <class F>
__device__ __forceinline__ void for_points(const float* __restrict__ P, uint32_t first, uint32_t npts, uint32_t step,
                                           F f) {
    for (uint32_t k0 = first; k0 < npts; k0 += step * RF_GROUP) {
        float x[RF_GROUP], y[RF_GROUP], z[RF_GROUP];
#pragma unroll
        for (int j = 0; j < RF_GROUP; ++j) {
            // clamped rather than guarded: a guarded load is moved behind
            // the guard of its use, back to one wait per point
            const size_t k = min(k0 + step * uint32_t(j), npts - 1);
            x[j] = P[3 * k];
            y[j] = P[3 * k + 1];
            z[j] = P[3 * k + 2];
        }
#pragma unroll
        for (int j = 0; j < RF_GROUP; ++j)
            if (k0 + step * uint32_t(j) < npts) f(x[j], y[j], z[j]);
    }
}

// A small box, whole, in one thread.
__global__ __launch_bounds__(256) void k_rf_small(const uint2* __restrict__ items, const uint32_t* __restrict__ list,
                                                  uint32_t n, const uint2* __restrict__ ranges,
                                                  MeshRec* __restrict__ meshes, NodeRec* __restrict__ nodes,
                                                  const float* __restrict__ tris, unsigned int* __restrict__ slack_bits,
                                                  unsigned int* __restrict__ radius_bits) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    // every lane stays to the end: the slack bound is reduced over the
    // wavefront before it goes to the per-mesh maximum (one atomic per
    // wavefront and mesh instead of one per box)
    const bool valid = i < n;
    const uint32_t it = valid ? list[i] : 0u;
    const uint2 item = valid ? items[it] : make_uint2(0u, 0u);
    double worst = 0.0, rad = 0.0;
    if (valid) {
        MeshRec& m = meshes[item.x];
        const uint2 rg = ranges[it];
        const float* P = item_tris(m, tris, rg.x);
        const uint32_t npts = 3 * rg.y;
        bool multi = false;
        for (uint32_t k = 1; k < npts && !multi; ++k)
            multi = !(P[3 * k] == P[0] && P[3 * k + 1] == P[1] && P[3 * k + 2] == P[2]);
        Box box;
        if (!multi) {
            box = empty_box(P, npts);
        } else {
            const double N = double(npts);
            double sx = 0, sy = 0, sz = 0;
            for_points(P, 0, npts, 1, [&](float x, float y, float z) {
                sx += double(x);
                sy += double(y);
                sz += double(z);
            });
            const double mx = sx / N, my = sy / N, mz = sz / N;
            double cxx = 0, cyy = 0, czz = 0, cxy = 0, cxz = 0, cyz = 0;
            for_points(P, 0, npts, 1, [&](float x, float y, float z) {
                const double dx = double(x) - mx, dy = double(y) - my, dz = double(z) - mz;
                cxx += dx * dx; cyy += dy * dy; czz += dz * dz;
                cxy += dx * dy; cxz += dx * dz; cyz += dy * dz;
            });
            double ax[9];
            axes_of(cxx / N, cyy / N, czz / N, cxy / N, cxz / N, cyz / N, ax);
            double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for_points(P, 0, npts, 1, [&](float x, float y, float z) {
                const double px = x, py = y, pz = z;
                for (int q = 0; q < 3; ++q) {
                    const double pr = (ax[3 * q] * px + ax[3 * q + 1] * py) + ax[3 * q + 2] * pz;
                    lo[q] = fmin(lo[q], pr);
                    hi[q] = fmax(hi[q], pr);
                }
            });
            box = box_from_extents(ax, lo, hi);
        }
        store_box(m, nodes, item.y, box);
        float b12[12];
        box_to12(box, b12);
        const SlackFrame f = slack_frame(b12);
        double rad2 = 0.0;
        for_points(P, 0, npts, 1, [&](float x, float y, float z) {
            const double p[3] = {x, y, z};
            worst = fmax(worst, excess(f, p));
            rad2 = fmax(rad2, p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        });
        rad = sqrt(rad2);  // sqrt is monotone: the root of the largest square is the largest root
    }
    const uint32_t mesh0 = __shfl(item.x, 0, 64);
    const bool one_mesh = __ballot(valid && item.x != mesh0) == 0ull;
    const bool root = valid && item.y == ROOT_CODE;
    if (one_mesh && __ballot(root) == 0ull) {  // the usual case: one mesh, no root box in this wavefront
        worst = wave_max_d(worst);
        if ((threadIdx.x & 63u) == 0) post_slack(slack_bits, radius_bits, mesh0, false, worst, 0.0, i >> 6);
    } else if (valid) {
        post_slack(slack_bits, radius_bits, item.x, root, worst, rad, i >> 6);
    }
}

// ---- large boxes: chunk passes ----
// chunk = {item, first triangle (mesh-local), triangles}
enum : int { RF_SUM = 0, RF_COV = 1, RF_EXT = 2, RF_SLACK = 3 };

template <int PASS>
__global__ __launch_bounds__(256) void k_rf_chunk(const uint2* __restrict__ items, const uint4* __restrict__ chunks,
                                                  uint32_t nchunks, const uint2* __restrict__ ranges,
                                                  const MeshRec* __restrict__ meshes, const NodeRec* __restrict__ nodes,
                                                  const float* __restrict__ tris, const double* __restrict__ scratch,
                                                  double* __restrict__ partial, unsigned int* __restrict__ slack_bits,
                                                  unsigned int* __restrict__ radius_bits) {
    const uint32_t ci = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (ci >= nchunks) return;  // whole wavefronts leave; no workgroup barrier follows
    const int lane = threadIdx.x & 63;
    const uint4 ch = chunks[ci];
    const uint2 item = items[ch.x];
    const MeshRec& m = meshes[item.x];
    const float* P = item_tris(m, tris, ch.y);
    const uint32_t npts = 3 * ch.z;
    const double* s = scratch + size_t(ch.x) * RS_WORDS;
    double* out = partial + size_t(ci) * PW;
    if (PASS == RF_SUM) {
        const float* P0 = item_tris(m, tris, ranges[ch.x].x);  // the box's first point
        bool differs = false;
        double sx = 0, sy = 0, sz = 0;
        const float x0 = P0[0], y0 = P0[1], z0 = P0[2];
        for_points(P, uint32_t(lane), npts, 64, [&](float x, float y, float z) {
            differs |= !(x == x0 && y == y0 && z == z0);
            sx += double(x);
            sy += double(y);
            sz += double(z);
        });
        const bool any = __ballot(differs) != 0ull;
        sx = wave_sum_d(sx); sy = wave_sum_d(sy); sz = wave_sum_d(sz);
        if (lane == 0) {
            out[0] = sx; out[1] = sy; out[2] = sz;
            out[7] = any ? 1.0 : 0.0;
        }
    } else if (PASS == RF_COV) {
        if (s[RS_KIND] != 2.0) return;
        const double mx = s[RS_MEAN], my = s[RS_MEAN + 1], mz = s[RS_MEAN + 2];
        double cxx = 0, cyy = 0, czz = 0, cxy = 0, cxz = 0, cyz = 0;
        for_points(P, uint32_t(lane), npts, 64, [&](float x, float y, float z) {
            const double dx = double(x) - mx, dy = double(y) - my, dz = double(z) - mz;
            cxx += dx * dx; cyy += dy * dy; czz += dz * dz;
            cxy += dx * dy; cxz += dx * dz; cyz += dy * dz;
        });
        cxx = wave_sum_d(cxx); cyy = wave_sum_d(cyy); czz = wave_sum_d(czz);
        cxy = wave_sum_d(cxy); cxz = wave_sum_d(cxz); cyz = wave_sum_d(cyz);
        if (lane == 0) {
            out[0] = cxx; out[1] = cyy; out[2] = czz; out[3] = cxy; out[4] = cxz; out[5] = cyz;
        }
    } else if (PASS == RF_EXT) {
        if (s[RS_KIND] != 2.0) return;
        double ax[9];
        for (int t = 0; t < 9; ++t) ax[t] = s[RS_AX + t];
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for_points(P, uint32_t(lane), npts, 64, [&](float x, float y, float z) {
            const double px = x, py = y, pz = z;
            for (int q = 0; q < 3; ++q) {
                const double pr = (ax[3 * q] * px + ax[3 * q + 1] * py) + ax[3 * q + 2] * pz;
                lo[q] = fmin(lo[q], pr);
                hi[q] = fmax(hi[q], pr);
            }
        });
        for (int q = 0; q < 3; ++q) {
            lo[q] = wave_min_d(lo[q]);
            hi[q] = wave_max_d(hi[q]);
        }
        if (lane == 0)
            for (int q = 0; q < 3; ++q) {
                out[q] = lo[q];
                out[3 + q] = hi[q];
            }
    } else {
        const SlackFrame f = slack_frame(box_ptr(m, nodes, item.y));
        const bool root = item.y == ROOT_CODE;  // only the root box's radius is kept
        double worst = 0.0, rad2 = 0.0;
        for_points(P, uint32_t(lane), npts, 64, [&](float x, float y, float z) {
            const double p[3] = {x, y, z};
            worst = fmax(worst, excess(f, p));
            if (root) rad2 = fmax(rad2, p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        });
        worst = wave_max_d(worst);
        rad2 = wave_max_d(rad2);  // sqrt is monotone: the root of the largest square is the largest root
        if (lane == 0) post_slack(slack_bits, radius_bits, item.x, root, worst, sqrt(rad2), ci);
    }
}

// One thread per large box: combine its chunks (in chunk order) after a pass.
// big = {item, first chunk, chunks}
template <int PASS>
__global__ __launch_bounds__(256) void k_rf_big(const uint2* __restrict__ items, const uint4* __restrict__ big,
                                                uint32_t nbig, const uint2* __restrict__ ranges,
                                                MeshRec* __restrict__ meshes, NodeRec* __restrict__ nodes,
                                                const float* __restrict__ tris, const double* __restrict__ partial,
                                                double* __restrict__ scratch) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nbig) return;
    const uint4 b = big[i];
    double* s = scratch + size_t(b.x) * RS_WORDS;
    const double* pp = partial + size_t(b.y) * PW;
    const uint2 rg = ranges[b.x];
    const double N = 3.0 * double(rg.y);
    if (PASS == RF_SUM) {
        double sx = 0, sy = 0, sz = 0;
        bool any = false;
        for (uint32_t c = 0; c < b.z; ++c) {
            sx += pp[c * PW]; sy += pp[c * PW + 1]; sz += pp[c * PW + 2];
            any |= pp[c * PW + 7] != 0.0;
        }
        s[RS_MEAN] = sx / N; s[RS_MEAN + 1] = sy / N; s[RS_MEAN + 2] = sz / N;
        s[RS_KIND] = any ? 2.0 : 1.0;
    } else if (PASS == RF_COV) {
        if (s[RS_KIND] != 2.0) return;
        double v[6] = {0, 0, 0, 0, 0, 0};
        for (uint32_t c = 0; c < b.z; ++c)
            for (int k = 0; k < 6; ++k) v[k] += pp[c * PW + k];
        axes_of(v[0] / N, v[1] / N, v[2] / N, v[3] / N, v[4] / N, v[5] / N, s + RS_AX);
    } else {  // RF_EXT: the box
        const uint2 item = items[b.x];
        MeshRec& m = meshes[item.x];
        Box box;
        if (s[RS_KIND] != 2.0) {
            box = empty_box(item_tris(m, tris, rg.x), 3 * rg.y);
        } else {
            double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (uint32_t c = 0; c < b.z; ++c)
                for (int q = 0; q < 3; ++q) {
                    lo[q] = fmin(lo[q], pp[c * PW + q]);
                    hi[q] = fmax(hi[q], pp[c * PW + 3 + q]);
                }
            box = box_from_extents(s + RS_AX, lo, hi);
        }
        store_box(m, nodes, item.y, box);
    }
}

__global__ void k_store_slack(MeshRec* meshes, const uint32_t* ids, uint32_t n, const unsigned int* slack_bits,
                              const unsigned int* radius_bits) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    MeshRec& m = meshes[ids[i]];
    unsigned int sb = 0;
    for (uint32_t r = 0; r < SLACK_REPS; ++r) sb = max(sb, slack_bits[size_t(ids[i]) * SLACK_REPS + r]);
    m.slack = __uint_as_float(sb);
    m.radius = __uint_as_float(radius_bits[ids[i]]);
}

__global__ void k_tree_vidx(const uint32_t* __restrict__ perm, uint32_t ntri, const uint32_t* __restrict__ indices,
                            uint32_t vbase, uint32_t* __restrict__ out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ntri) return;
    const uint32_t t = perm[s];
    for (int k = 0; k < 3; ++k) out[size_t(s) * 3 + k] = vbase + indices[size_t(t) * 3 + k];
}

// k_tree_vidx for every mesh of one build at once: the meshes' slots are
// back to back, prefix[m] is mesh m's first slot (prefix[nmesh] = all), and
// perm and the per-mesh index lists are mesh-local.
__global__ void k_tree_vidx_many(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ prefix,
                                 uint32_t nmesh, const uint32_t* __restrict__ indices, uint32_t* __restrict__ out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= prefix[nmesh]) return;
    uint32_t lo = 0, hi = nmesh;  // prefix[lo] <= s < prefix[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (prefix[mid] <= s) lo = mid;
        else hi = mid;
    }
    const uint32_t t = prefix[lo] + perm[s];
    for (int k = 0; k < 3; ++k) out[size_t(s) * 3 + k] = indices[size_t(t) * 3 + k];
}

// ---- launchers ---------------------------------------------------------------
static inline dim3 g256r(uint64_t n) { return dim3(uint32_t((std::max<uint64_t>(n, 1) + 255) / 256)); }

hipError_t launch_tree_vidx(const uint32_t* perm, uint32_t ntri, const uint32_t* indices, uint32_t vbase,
                            uint32_t* out, hipStream_t st) {
    hipLaunchKernelGGL(k_tree_vidx, g256r(ntri), dim3(256), 0, st, perm, ntri, indices, vbase, out);
    return hipGetLastError();
}
hipError_t launch_tree_vidx_many(const uint32_t* perm, const uint32_t* prefix, uint32_t nmesh, uint64_t ntri,
                                 const uint32_t* indices, uint32_t* out, hipStream_t st) {
    hipLaunchKernelGGL(k_tree_vidx_many, g256r(ntri), dim3(256), 0, st, perm, prefix, nmesh, indices, out);
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
hipError_t launch_item_ranges(const uint2* items, uint32_t nitems, const MeshRec* meshes, const NodeRec* nodes,
                              uint2* ranges, hipStream_t st) {
    hipLaunchKernelGGL(k_item_ranges, g256r(nitems), dim3(256), 0, st, items, nitems, meshes, nodes, ranges);
    return hipGetLastError();
}
hipError_t launch_refit(const RefitPlan& p, MeshRec* meshes, NodeRec* nodes, const float* tris, double* scratch,
                        double* partial, unsigned int* slack_bits, unsigned int* radius_bits, const uint32_t* ids,
                        uint32_t nids, hipStream_t st) {
    if (p.nsmall)
        hipLaunchKernelGGL(k_rf_small, g256r(p.nsmall), dim3(256), 0, st, p.items, p.small, p.nsmall, p.ranges, meshes,
                           nodes, tris, slack_bits, radius_bits);
    if (p.nbig) {
        const dim3 gc((p.nchunks + 3) / 4), gb = g256r(p.nbig);
        hipLaunchKernelGGL(k_rf_chunk<RF_SUM>, gc, dim3(256), 0, st, p.items, p.chunks, p.nchunks, p.ranges, meshes,
                           nodes, tris, scratch, partial, slack_bits, radius_bits);
        hipLaunchKernelGGL(k_rf_big<RF_SUM>, gb, dim3(256), 0, st, p.items, p.big, p.nbig, p.ranges, meshes, nodes,
                           tris, partial, scratch);
        hipLaunchKernelGGL(k_rf_chunk<RF_COV>, gc, dim3(256), 0, st, p.items, p.chunks, p.nchunks, p.ranges, meshes,
                           nodes, tris, scratch, partial, slack_bits, radius_bits);
        hipLaunchKernelGGL(k_rf_big<RF_COV>, gb, dim3(256), 0, st, p.items, p.big, p.nbig, p.ranges, meshes, nodes,
                           tris, partial, scratch);
        hipLaunchKernelGGL(k_rf_chunk<RF_EXT>, gc, dim3(256), 0, st, p.items, p.chunks, p.nchunks, p.ranges, meshes,
                           nodes, tris, scratch, partial, slack_bits, radius_bits);
        hipLaunchKernelGGL(k_rf_big<RF_EXT>, gb, dim3(256), 0, st, p.items, p.big, p.nbig, p.ranges, meshes, nodes,
                           tris, partial, scratch);
        hipLaunchKernelGGL(k_rf_chunk<RF_SLACK>, gc, dim3(256), 0, st, p.items, p.chunks, p.nchunks, p.ranges, meshes,
                           nodes, tris, scratch, partial, slack_bits, radius_bits);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_store_slack, g256r(nids), dim3(256), 0, st, meshes, ids, nids, slack_bits, radius_bits);
    return hipGetLastError();
}
uint32_t refit_small_max() { return RF_SMALL; }
uint32_t refit_chunk_tris() { return RF_CHUNK; }

}  // namespace cdk
