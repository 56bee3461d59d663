// OBB-tree build on the GPU, bit-identical to the reference builder.
//
// Reference: inMyRoom_vulkan/src/Geometry/OBB.cpp:33-178 (PCA box),
// inMyRoom_vulkan/src/Geometry/OBBtree.cpp:8-108 (recursive split) and
// :179-233, :321-394 (DFS flatten).
//
// Level-synchronous: the nodes of one depth, across every mesh of the
// batch, are one launch, one 64-lane workgroup per node at a time (levels
// are launched in batches; each launch reads its node range from the
// previous one).  Within a node:
//   * the distinct-point check is a wave ballot;
//   * the mean and the six covariance sums are SEQUENTIAL double chains in
//     point order (the reference's std::accumulate / std::inner_product),
//     one lane per chain adding terms the whole wave precomputes per LDS chunk;
//   * eig3 runs in FP64 on one lane;
//   * the three min/max projections are wave reductions whose tie rule
//     reproduces std::min/std::max visiting the points in order;
//   * the split is a stable partition built from ballots, scattering the
//     node's triangles into the next level's buffer.
// Afterwards the tree is flattened to the reference's DFS layout with
// per-level prefix passes (subtree counts bottom-up, preorder top-down).
#include <float.h>
#include <math.h>

#include <algorithm>
#include <cstring>

#include "eig3.hpp"
#include "internal.hpp"

namespace cdk {

struct BNode {
    uint32_t mesh, beg, end, level;
    uint32_t kid[2];
    uint32_t flags;  // bit0 = leaf
    uint32_t icount, tcount, pre, toff;
    float surf;
    float box[12];
};

struct LevelArgs {
    BNode* nodes;
    const float* pos_in;
    const uint32_t* id_in;
    float* pos_out;
    uint32_t* id_out;
    float* pos_fin;
    uint32_t* id_fin;
    unsigned int* mesh_depth;
    uint32_t leaf_max;
};

constexpr int CHUNK = 256;                 // points per step of the sequential chains
constexpr int CHUNK_REGS = CHUNK * 3 / 64;  // floats each lane fetches per chunk
static_assert(CHUNK * 3 % 64 == 0 && CHUNK % 64 == 0, "chunk split over a wave");

// The reference's sequential double sums over a node's points (OBB.cpp:
// 41-73: std::accumulate, then the six inner products), one chain per lane,
// in point order.  Only the final additions form the chain; the terms do
// not depend on it.  So the wave streams the points through LDS a chunk at a
// time: every lane fetches its share of the next chunk, all lanes turn the
// current chunk into its K terms per point (term(p, c, out) for the points
// lane, lane + 64, ...), and the first K lanes add their term of each point
// in order while the fetch is in flight.  The block is one wave.
// k_level's workgroup is one wavefront: its lanes run in lockstep and see
// each other's LDS accesses in program order, so a step between the phases
// of a node only has to keep the compiler from moving LDS accesses across
// it.  (__syncthreads would also wait for the global loads in flight, among
// them the prefetch of the next chunk of points.)
__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); }

template <int K, class Term>
__device__ __forceinline__ double stream_chain(const float* __restrict__ P, uint32_t npts, Term term, float* s_raw,
                                               double* s_val) {
    const int lane = threadIdx.x;
    float r[CHUNK_REGS];
    auto fetch = [&](uint32_t base) {
        const uint32_t m3 = 3u * min(uint32_t(CHUNK), npts - base);
#pragma unroll
        for (int j = 0; j < CHUNK_REGS; ++j) {
            const uint32_t t = uint32_t(lane) + 64u * j;
            r[j] = t < m3 ? P[3 * size_t(base) + t] : 0.f;
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int j = 0; j < CHUNK_REGS; ++j) s_raw[lane + 64 * j] = r[j];
    };
    fetch(0);
    store();
    wave_sync();
    double acc = 0.0;
    for (uint32_t base = 0; base < npts; base += CHUNK) {
        const uint32_t next = base + CHUNK;
        const uint32_t m = min(uint32_t(CHUNK), npts - base);
        if (next < npts) fetch(next);
#pragma unroll
        for (int j = 0; j < CHUNK / 64; ++j) {
            const uint32_t k = uint32_t(lane) + 64u * j;
            if (k < m) term(s_raw + 3 * k, s_val + K * k);
        }
        wave_sync();
        if (lane < K) {
            // 16 terms are read from LDS while the previous 16 are added,
            // so the chain waits on its additions only
            const double* q = s_val + lane;
            uint32_t k = 0;
            if (m >= 16) {
                double cur[16];
#pragma unroll
                for (int i = 0; i < 16; ++i) cur[i] = q[K * i];
                for (k = 16; k + 16 <= m; k += 16) {
                    double nxt[16];
#pragma unroll
                    for (int i = 0; i < 16; ++i) nxt[i] = q[K * (k + uint32_t(i))];
#pragma unroll
                    for (int i = 0; i < 16; ++i) acc = acc + cur[i];
#pragma unroll
                    for (int i = 0; i < 16; ++i) cur[i] = nxt[i];
                }
#pragma unroll
                for (int i = 0; i < 16; ++i) acc = acc + cur[i];
            }
            for (; k < m; ++k) acc = acc + q[K * k];
        }
        wave_sync();
        if (next < npts) store();
        wave_sync();
    }
    return acc;
}

__device__ inline void wave_min_last(double& v, int& i) {
    // smaller value wins; equal values (== , so -0 and +0 too): larger index wins
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double ov = __shfl_xor(v, off, 64);
        int oi = __shfl_xor(i, off, 64);
        if (ov < v || (ov == v && oi > i)) {
            v = ov;
            i = oi;
        }
    }
}
__device__ inline void wave_max_first(double& v, int& i) {
    // larger value wins; equal values: smaller index wins
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double ov = __shfl_xor(v, off, 64);
        int oi = __shfl_xor(i, off, 64);
        if (ov > v || (ov == v && oi < i)) {
            v = ov;
            i = oi;
        }
    }
}

struct AxisCand {
    float len;
    f3 dir;
};

// std::sort of three (length, axis) pairs, descending (OBBtree.cpp:45-51):
// libstdc++ runs its insertion sort for three elements; reproduced branch
// for branch.
__device__ inline void sort3_desc(AxisCand c[3]) {
#pragma unroll
    for (int i = 1; i < 3; ++i) {
        AxisCand val = c[i];
        if (val.len > c[0].len) {
            for (int k = i; k > 0; --k) c[k] = c[k - 1];
            c[0] = val;
        } else {
            int j = i;
            while (val.len > c[j - 1].len) {
                c[j] = c[j - 1];
                --j;
            }
            c[j] = val;
        }
    }
}

// Triangle::GetMinMaxProjectionToAxis (Triangle.cpp:113-138) then the
// left/right rule of OBBtree.cpp:66-75, on a triangle held in registers.
__device__ inline bool goes_left(const float (&T)[9], f3 axis, float cproj) {
    float lo = INFINITY, hi = -INFINITY;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float q = dot(axis, mk(T[3 * k], T[3 * k + 1], T[3 * k + 2]));
        if (q < lo) lo = q;
        if (q > hi) hi = q;
    }
    float mid = (lo + hi) / 2.f;
    return mid <= cproj;
}

// The node scans below take GROUP wave-wide chunks per step and issue the
// loads of all of them before using any: a node's scans are latency bound
// (one wave per node), and a chunk at a time waits on memory once per chunk.
// Indices past the end are clamped to the last element rather than guarded:
// a guarded load is moved behind the guard of its use, back to one wait per
// chunk.
constexpr int GROUP = 8;

__device__ __forceinline__ void load_tris(const float* __restrict__ P, uint32_t c0, uint32_t n, int lane,
                                          float (&T)[GROUP][9]) {
#pragma unroll
    for (int j = 0; j < GROUP; ++j) {
        const size_t i = min(c0 + 64u * j + uint32_t(lane), n - 1);
#pragma unroll
        for (int t = 0; t < 9; ++t) T[j][t] = P[i * 9 + t];
    }
}

struct LevelShared {
    float raw[CHUNK * 3];
    double val[CHUNK * 6];
    double mean[3];
    double cov[6];
    double ax[9];
    float box[12];
};

// One node of a level: its box (OBB::CreateOBBfromPoints, OBB.cpp:33-89),
// then either its triangles parked as a leaf or its split into two
// children appended to the next level (OBBtreeSplitBuildNode::
// SplitOBBandCreateChildren, OBBtree.cpp:43-108).
__device__ __forceinline__ void level_node(const LevelArgs& a, uint32_t ni, uint32_t next_first,
                                           unsigned int* next_count, LevelShared& S) {
    const int lane = threadIdx.x;
    BNode* nd = &a.nodes[ni];
    const uint32_t beg = nd->beg, end = nd->end;
    const uint32_t n = end - beg, npts = 3 * n;
    const float* P = a.pos_in + size_t(beg) * 9;

    // more than one distinct point?  (the first wave-wide chunk that shows one ends the scan)
    bool multi = false;
    if (npts > 1) {
        const float x0 = P[0], y0 = P[1], z0 = P[2];
        for (uint32_t k0 = 1; k0 < npts && !multi; k0 += 64) {
            const uint32_t k = k0 + lane;
            const bool d = k < npts && !(P[3 * k] == x0 && P[3 * k + 1] == y0 && P[3 * k + 2] == z0);
            multi = __ballot(d) != 0;
        }
    }
    Box box;
    if (npts == 0 || !multi) {
        const float e = FLT_EPSILON;  // OBB::EmptyOBB (OBB.cpp:168-178)
        box.c = npts ? mk(P[0], P[1], P[2]) : mk(0.f, 0.f, 0.f);
        box.u = mk(e, 0.f, 0.f);
        box.v = mk(0.f, e, 0.f);
        box.w = mk(0.f, 0.f, e);
    } else {
        const double N = double(npts);
        double acc = stream_chain<3>(
            P, npts,
            [](const float* p, double* o) {
                o[0] = double(p[0]);
                o[1] = double(p[1]);
                o[2] = double(p[2]);
            },
            S.raw, S.val);
        if (lane < 3) S.mean[lane] = acc / N;
        wave_sync();
        // covariance order xx, yy, zz, xy, xz, yz (OBB.cpp:56-73)
        const double mx = S.mean[0], my = S.mean[1], mz = S.mean[2];
        acc = stream_chain<6>(
            P, npts,
            [mx, my, mz](const float* p, double* o) {
                const double dx = double(p[0]) - mx, dy = double(p[1]) - my, dz = double(p[2]) - mz;
                o[0] = dx * dx;
                o[1] = dy * dy;
                o[2] = dz * dz;
                o[3] = dx * dy;
                o[4] = dx * dz;
                o[5] = dy * dz;
            },
            S.raw, S.val);
        if (lane < 6) S.cov[lane] = acc / N;
        wave_sync();
        if (lane == 0) {
            const double xx = S.cov[0], yy = S.cov[1], zz = S.cov[2], xy = S.cov[3], xz = S.cov[4], yz = S.cov[5];
            double A[3][3] = {{xx, xy, xz}, {xy, yy, yz}, {xz, yz, zz}};
            double V[3][3], d[3];
            eig3(A, V, d);
            // OBB.cpp:80-87: the box axes are the ROWS of eig3's V
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) S.ax[r * 3 + c] = V[r][c];
        }
        wave_sync();
        double ax[9];
#pragma unroll
        for (int t = 0; t < 9; ++t) ax[t] = S.ax[t];
        // OBB::CreateAABBfromPoints (OBB.cpp:105-166); each lane visits its
        // points in increasing order, as the tie rules below assume
        double lo[3], hi[3];
        int loi[3], hii[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            lo[q] = INFINITY;
            hi[q] = -INFINITY;
            loi[q] = -1;
            hii[q] = -1;
        }
        for (uint32_t k0 = 0; k0 < npts; k0 += 64u * GROUP) {
            float px[GROUP], py[GROUP], pz[GROUP];
#pragma unroll
            for (int j = 0; j < GROUP; ++j) {
                const size_t k = min(k0 + 64u * j + uint32_t(lane), npts - 1);
                px[j] = P[3 * k];
                py[j] = P[3 * k + 1];
                pz[j] = P[3 * k + 2];
            }
#pragma unroll
            for (int j = 0; j < GROUP; ++j) {
                const uint32_t k = k0 + 64u * j + uint32_t(lane);
                if (k < npts) {
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        double pr = (ax[3 * q] * double(px[j]) + ax[3 * q + 1] * double(py[j])) +
                                    ax[3 * q + 2] * double(pz[j]);
                        if (!(lo[q] < pr)) { lo[q] = pr; loi[q] = int(k); }
                        if (hi[q] < pr) { hi[q] = pr; hii[q] = int(k); }
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            wave_min_last(lo[q], loi[q]);
            wave_max_first(hi[q], hii[q]);
        }
        if (lane == 0) {
            double cx = 0.0, cy = 0.0, cz = 0.0;
            float side[3][3];
            for (int q = 0; q < 3; ++q) {
                double delta = (hi[q] - lo[q]) + 2.0 * double(FLT_EPSILON);
                double mid = (hi[q] + lo[q]) / 2.0;
                cx = cx + mid * ax[3 * q];
                cy = cy + mid * ax[3 * q + 1];
                cz = cz + mid * ax[3 * q + 2];
                double half = delta / 2.0;
                side[q][0] = float(half * ax[3 * q]);
                side[q][1] = float(half * ax[3 * q + 1]);
                side[q][2] = float(half * ax[3 * q + 2]);
            }
            S.box[0] = float(cx);
            S.box[1] = float(cy);
            S.box[2] = float(cz);
            for (int q = 0; q < 3; ++q)
                for (int c = 0; c < 3; ++c) S.box[3 + 3 * q + c] = side[q][c];
        }
        wave_sync();
        box = box_from12(S.box);
    }
    if (lane == 0) {
        box_to12(box, nd->box);
        nd->surf = surface(box);
    }

    // ---- leaf: park the triangles in their final slots ----
    if (n <= a.leaf_max) {
        for (uint32_t t = lane; t < 9 * n; t += 64) a.pos_fin[size_t(beg) * 9 + t] = P[t];
        for (uint32_t t = lane; t < n; t += 64) a.id_fin[beg + t] = a.id_in[beg + t];
        if (lane == 0) {
            nd->flags = 1u;
            nd->icount = 0;
            nd->tcount = n;
        }
        return;
    }

    // ---- split (OBBtree.cpp:43-108): the longest axis whose centre
    // projection leaves both sides non-empty, else index halves; the three
    // candidates are counted in one scan ----
    AxisCand cand[3] = {{vlength(box.u), normalize(box.u)}, {vlength(box.v), normalize(box.v)},
                        {vlength(box.w), normalize(box.w)}};
    sort3_desc(cand);
    float cp[3];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) cp[ch] = dot(box.c, cand[ch].dir);
    uint32_t cnt[3] = {0, 0, 0};
    for (uint32_t c0 = 0; c0 < n; c0 += 64u * GROUP) {
        float T[GROUP][9];
        load_tris(P, c0, n, lane, T);
#pragma unroll
        for (int j = 0; j < GROUP; ++j) {
            const bool valid = c0 + 64u * j + uint32_t(lane) < n;
#pragma unroll
            for (int ch = 0; ch < 3; ++ch)
                cnt[ch] += uint32_t(__popcll(__ballot(valid && goes_left(T[j], cand[ch].dir, cp[ch]))));
        }
    }
    int pick = -1;
#pragma unroll
    for (int ch = 2; ch >= 0; --ch)
        if (cnt[ch] != 0 && cnt[ch] != n) pick = ch;
    const bool ok = pick >= 0;
    f3 axis = cand[0].dir;
    float cproj = 0.f;
    uint32_t nl = n / 2;  // index halves (OBBtree.cpp:81-93)
#pragma unroll
    for (int ch = 0; ch < 3; ++ch)
        if (ch == pick) {
            axis = cand[ch].dir;
            cproj = cp[ch];
            nl = cnt[ch];
        }
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint32_t runL = 0, runR = 0;
    for (uint32_t c0 = 0; c0 < n; c0 += 64u * GROUP) {
        float T[GROUP][9];
        load_tris(P, c0, n, lane, T);
        uint32_t idv[GROUP];
#pragma unroll
        for (int j = 0; j < GROUP; ++j) {
            const uint32_t i = c0 + 64u * j + uint32_t(lane);
            idv[j] = a.id_in[beg + min(i, n - 1)];
        }
#pragma unroll
        for (int j = 0; j < GROUP; ++j) {
            const uint32_t i = c0 + 64u * j + uint32_t(lane);
            const bool valid = i < n;
            const bool left = valid && (ok ? goes_left(T[j], axis, cproj) : (i < n / 2));
            const unsigned long long mL = __ballot(left);
            const unsigned long long mR = __ballot(valid && !left);
            if (valid) {
                const uint32_t dst = left ? beg + runL + uint32_t(__popcll(mL & below))
                                          : beg + nl + runR + uint32_t(__popcll(mR & below));
                float* o = a.pos_out + size_t(dst) * 9;
#pragma unroll
                for (int t = 0; t < 9; ++t) o[t] = T[j][t];
                a.id_out[dst] = idv[j];
            }
            runL += uint32_t(__popcll(mL));
            runR += uint32_t(__popcll(mR));
        }
    }
    if (lane == 0) {
        const uint32_t slot = next_first + atomicAdd(next_count, 2u);
        BNode* k0 = &a.nodes[slot];
        BNode* k1 = &a.nodes[slot + 1];
        k0->mesh = nd->mesh; k0->beg = beg;      k0->end = beg + nl; k0->level = nd->level + 1; k0->flags = 0;
        k1->mesh = nd->mesh; k1->beg = beg + nl; k1->end = end;      k1->level = nd->level + 1; k1->flags = 0;
        nd->kid[0] = slot;
        nd->kid[1] = slot + 1;
        nd->flags = 0;
        atomicMax(&a.mesh_depth[nd->mesh], nd->level + 1);
    }
}

// One level of the build: lv[0] = {first node, node count} of this level
// (written by the previous level's launch or by the host); the launch
// writes lv[1] = {first node, count} of the next, its count grown by two
// per split.  A fixed grid strides over the level's nodes, so levels can be
// launched back to back before their sizes are known on the host.
__global__ __launch_bounds__(64) void k_level(LevelArgs a, uint2* lv) {
    __shared__ LevelShared S;
    const uint32_t first = lv[0].x, count = lv[0].y;
    const uint32_t next_first = first + count;
    if (blockIdx.x == 0 && threadIdx.x == 0) lv[1].x = next_first;
    for (uint32_t b = blockIdx.x; b < count; b += gridDim.x) {
        level_node(a, first + b, next_first, &lv[1].y, S);
        wave_sync();  // the next node reuses the LDS
    }
}

// Left child = larger surface (OBBtree.cpp:101-107).
__global__ void k_order_kids(BNode* nodes, uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    BNode* nd = &nodes[i];
    if (nd->flags & 1u) return;
    const uint32_t l = nd->kid[0], r = nd->kid[1];
    if (nodes[r].surf > nodes[l].surf) {
        nd->kid[0] = r;
        nd->kid[1] = l;
    }
}

// Subtree sizes, one level at a time from the bottom.
__global__ void k_counts(BNode* nodes, uint32_t first, uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    BNode* nd = &nodes[first + i];
    if (nd->flags & 1u) return;
    const BNode& l = nodes[nd->kid[0]];
    const BNode& r = nodes[nd->kid[1]];
    nd->icount = 1 + l.icount + r.icount;
    nd->tcount = l.tcount + r.tcount;
}

// Preorder index and DFS triangle offset, one level at a time from the top
// (OBBtree.cpp:179-233, :385-394).
__global__ void k_preorder(BNode* nodes, uint32_t first, uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const BNode& nd = nodes[first + i];
    if (nd.flags & 1u) return;
    BNode& l = nodes[nd.kid[0]];
    BNode& r = nodes[nd.kid[1]];
    l.pre = nd.pre + 1;
    l.toff = nd.toff;
    r.pre = nd.pre + 1 + l.icount;
    r.toff = nd.toff + l.tcount;
}

// The three passes above in one workgroup, for builds of up to
// FLATTEN_SMALL nodes: one launch instead of one per level and pass.
// lv[l] = {first node, count} of level l.
constexpr uint32_t FLATTEN_SMALL = 1u << 16;
__global__ __launch_bounds__(1024) void k_flatten_small(BNode* nodes, uint32_t total, const uint2* lv, uint32_t nlev) {
    for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
        BNode* nd = &nodes[i];
        if (nd->flags & 1u) continue;
        const uint32_t l = nd->kid[0], r = nd->kid[1];
        if (nodes[r].surf > nodes[l].surf) {
            nd->kid[0] = r;
            nd->kid[1] = l;
        }
    }
    __syncthreads();
    for (uint32_t k = nlev; k-- > 0;) {
        for (uint32_t i = threadIdx.x; i < lv[k].y; i += blockDim.x) {
            BNode* nd = &nodes[lv[k].x + i];
            if (nd->flags & 1u) continue;
            const BNode& l = nodes[nd->kid[0]];
            const BNode& r = nodes[nd->kid[1]];
            nd->icount = 1 + l.icount + r.icount;
            nd->tcount = l.tcount + r.tcount;
        }
        __syncthreads();
    }
    for (uint32_t k = 0; k < nlev; ++k) {
        for (uint32_t i = threadIdx.x; i < lv[k].y; i += blockDim.x) {
            const BNode& nd = nodes[lv[k].x + i];
            if (nd.flags & 1u) continue;
            BNode& l = nodes[nd.kid[0]];
            BNode& r = nodes[nd.kid[1]];
            l.pre = nd.pre + 1;
            l.toff = nd.toff;
            r.pre = nd.pre + 1 + l.icount;
            r.toff = nd.toff + l.tcount;
        }
        __syncthreads();
    }
}

__global__ void k_emit_nodes(const BNode* nodes, uint32_t count, const uint32_t* node_base, NodeRec* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const BNode& nd = nodes[i];
    if (nd.flags & 1u) return;
    NodeRec rec;
    for (int s = 0; s < 2; ++s) {
        const BNode& k = nodes[nd.kid[s]];
        ChildRec& c = rec.c[s];
        for (int t = 0; t < 12; ++t) c.box[t] = k.box[t];
        c.surf = k.surf;
        const bool leaf = (k.flags & 1u) != 0u;
        c.link = leaf ? k.toff : k.pre;
        c.tris = k.tcount | (leaf ? LEAF_BIT : 0u);
        c.toff = k.toff;
    }
    out[node_base[nd.mesh] + nd.pre] = rec;
}

__global__ void k_emit_tris(const BNode* nodes, uint32_t count, const uint32_t* tri_base, const float* pos_fin,
                            const uint32_t* id_fin, float* tris, uint32_t* perm, uint32_t* leafof) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const BNode& nd = nodes[i];
    if (!(nd.flags & 1u)) return;
    const size_t dst0 = size_t(tri_base[nd.mesh]) + nd.toff;
    for (uint32_t t = 0; t < nd.tcount; ++t) {
        const float* s = pos_fin + size_t(nd.beg + t) * 9;
        float* o = tris + (dst0 + t) * 9;
        for (int q = 0; q < 9; ++q) o[q] = s[q];
        perm[dst0 + t] = id_fin[nd.beg + t];
        leafof[dst0 + t] = nd.toff;  // the leaf's first slot names the leaf
    }
}

// Containment slack of every box over its own triangles, and the mesh radius.
__global__ __launch_bounds__(64) void k_slack(const BNode* nodes, uint32_t count, const float* pos_fin,
                                              unsigned int* mesh_slack_bits, unsigned int* mesh_radius_bits) {
    const uint32_t ni = blockIdx.x;
    if (ni >= count) return;
    const BNode& nd = nodes[ni];
    const int lane = threadIdx.x;
    const double u[3] = {nd.box[3], nd.box[4], nd.box[5]};
    const double v[3] = {nd.box[6], nd.box[7], nd.box[8]};
    const double w[3] = {nd.box[9], nd.box[10], nd.box[11]};
    // a point's coordinates in the side vectors: its offset dotted with the
    // dual basis (v x w, w x u, u x v) / det(u, v, w)
    auto cross = [](const double* a, const double* b, double* o) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    double du[3], dv[3], dw[3];
    cross(v, w, du);
    cross(w, u, dv);
    cross(u, v, dw);
    const double D = u[0] * du[0] + u[1] * du[1] + u[2] * du[2];
    for (int k = 0; k < 3; ++k) {
        du[k] = du[k] / D;
        dv[k] = dv[k] / D;
        dw[k] = dw[k] / D;
    }
    const double lu = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double lv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double lw = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    double worst = 0.0, rad = 0.0;
    const uint32_t npts = 3 * (nd.end - nd.beg);
    const float* P = pos_fin + size_t(nd.beg) * 9;
    const double c[3] = {nd.box[0], nd.box[1], nd.box[2]};
    const bool root = nd.level == 0;
    for (uint32_t k0 = 0; k0 < npts; k0 += 64u * GROUP) {
        float px[GROUP], py[GROUP], pz[GROUP];
#pragma unroll
        for (int j = 0; j < GROUP; ++j) {
            const size_t k = min(k0 + 64u * j + uint32_t(lane), npts - 1);
            px[j] = P[3 * k];
            py[j] = P[3 * k + 1];
            pz[j] = P[3 * k + 2];
        }
#pragma unroll
        for (int j = 0; j < GROUP; ++j) {
            if (k0 + 64u * j + uint32_t(lane) >= npts) continue;
            const double p[3] = {px[j], py[j], pz[j]};
            const double r[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
            double ex;
            if (fabs(D) > 0.0) {
                const double al = r[0] * du[0] + r[1] * du[1] + r[2] * du[2];
                const double be = r[0] * dv[0] + r[1] * dv[1] + r[2] * dv[2];
                const double ga = r[0] * dw[0] + r[1] * dw[1] + r[2] * dw[2];
                ex = fmax(0.0, fabs(al) - 1.0) * lu + fmax(0.0, fabs(be) - 1.0) * lv + fmax(0.0, fabs(ga) - 1.0) * lw;
            } else {
                ex = INFINITY;
            }
            worst = fmax(worst, ex);
            if (root) rad = fmax(rad, p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);  // squared; roots only
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        worst = fmax(worst, __shfl_xor(worst, off, 64));
        rad = fmax(rad, __shfl_xor(rad, off, 64));
    }
    rad = sqrt(rad);  // sqrt is monotone: the root of the largest square is the largest root
    if (lane == 0) {
        // round up to float so the bound stays an upper bound
        float wf = float(worst);
        if (double(wf) < worst) wf = nextafterf(wf, INFINITY);
        float rf = float(rad);
        if (double(rf) < rad) rf = nextafterf(rf, INFINITY);
        // one of SLACK_REPS copies per mesh, so a mesh's many nodes do not queue on one address
        atomicMax(&mesh_slack_bits[size_t(nd.mesh) * SLACK_REPS + (ni % SLACK_REPS)], __float_as_uint(wf));
        if (nd.level == 0) atomicMax(&mesh_radius_bits[nd.mesh], __float_as_uint(rf));
    }
}

#define CK(x)                                                        \
    do {                                                             \
        hipError_t e_ = (x);                                         \
        if (e_ != hipSuccess) {                                      \
            res.err = std::string(#x ": ") + hipGetErrorString(e_);  \
            goto fail;                                               \
        }                                                            \
    } while (0)

static inline uint32_t blocks_for(uint64_t n, uint32_t b) { return uint32_t((n + b - 1) / b); }

constexpr uint32_t LEVEL_BATCH = 8;     // level launches per wait of the build
constexpr uint32_t LEVEL_GRID = 65536;  // workgroups of a level launch at most (larger levels are strided)

// Build every soup of the batch and append the trees to the context arenas.
BuildResult build_soups(cd_ctx* ctx, const float* tri9, const uint32_t* counts, uint32_t nmesh, bool tri9_on_device) {
    BuildResult res;
    hipStream_t st = ctx->stream;
    uint64_t T = 0, node_cap = 0;
    std::vector<uint64_t> tri_begin(nmesh);
    for (uint32_t m = 0; m < nmesh; ++m) {
        tri_begin[m] = T;
        T += counts[m];
        node_cap += counts[m] ? 2ull * counts[m] - 1 : 1;
    }
    if (T >= 0xFFFFFFF0ull || node_cap >= 0xFFFFFFF0ull) {
        res.err = "batch too large (triangle or node count exceeds 32-bit indexing)";
        return res;
    }
    // work buffers, kept in the context between builds of up to BUILD_KEEP bytes
    DevBuf* const W = ctx->build_scratch;
    DevBuf* const pos = W;      // pos[0], pos[1]
    DevBuf* const ids = W + 2;  // ids[0], ids[1]
    DevBuf &pos_fin = W[4], &id_fin = W[5], &nodes = W[6], &mesh_depth = W[7], &level_tab = W[8], &base_nodes = W[9],
           &base_tris = W[10], &slack_bits = W[11], &radius_bits = W[12], &level_list = W[13];
    std::vector<BNode> roots(nmesh);
    std::vector<std::pair<uint32_t, uint32_t>> levels;
    std::vector<uint32_t> host_ids;
    std::vector<uint32_t> node_base(nmesh), tri_base(nmesh), depth(nmesh), slack(size_t(nmesh) * SLACK_REPS),
        radius(nmesh);
    uint64_t total_internal = 0;
    uint32_t first = 0, count = nmesh, cur = 0;
    const size_t posb = std::max<uint64_t>(T, 1) * 9 * sizeof(float), idb = std::max<uint64_t>(T, 1) * sizeof(uint32_t);

    CK(pos[0].ensure(posb));
    CK(pos[1].ensure(posb));
    CK(pos_fin.ensure(posb));
    CK(ids[0].ensure(idb));
    CK(ids[1].ensure(idb));
    CK(id_fin.ensure(idb));
    CK(nodes.ensure(node_cap * sizeof(BNode)));
    CK(mesh_depth.ensure(nmesh * sizeof(uint32_t)));
    CK(slack_bits.ensure(size_t(nmesh) * SLACK_REPS * sizeof(uint32_t)));
    CK(radius_bits.ensure(nmesh * sizeof(uint32_t)));
    CK(level_tab.ensure((LEVEL_BATCH + 1) * sizeof(uint2)));
    CK(base_nodes.ensure(nmesh * sizeof(uint32_t)));
    CK(base_tris.ensure(nmesh * sizeof(uint32_t)));

    if (T)
        CK(hipMemcpyAsync(pos[0].p, tri9, T * 9 * sizeof(float),
                          tri9_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
    host_ids.resize(T);
    for (uint32_t m = 0; m < nmesh; ++m)
        for (uint32_t k = 0; k < counts[m]; ++k) host_ids[tri_begin[m] + k] = k;
    if (T) CK(hipMemcpyAsync(ids[0].p, host_ids.data(), T * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    for (uint32_t m = 0; m < nmesh; ++m) {
        BNode r{};
        r.mesh = m;
        r.beg = uint32_t(tri_begin[m]);
        r.end = uint32_t(tri_begin[m] + counts[m]);
        r.level = 0;
        r.pre = 0;
        r.toff = 0;
        roots[m] = r;
    }
    CK(hipMemcpyAsync(nodes.p, roots.data(), nmesh * sizeof(BNode), hipMemcpyHostToDevice, st));
    CK(hipMemsetAsync(mesh_depth.p, 0, nmesh * sizeof(uint32_t), st));
    CK(hipMemsetAsync(slack_bits.p, 0, size_t(nmesh) * SLACK_REPS * sizeof(uint32_t), st));
    CK(hipMemsetAsync(radius_bits.p, 0, nmesh * sizeof(uint32_t), st));

    // Levels go LEVEL_BATCH launches at a time with one wait per batch: each
    // launch takes its node range from the level table the previous one
    // filled, over a grid sized by an upper bound (a level at most doubles).
    while (count > 0) {
        std::vector<uint2> tab(LEVEL_BATCH + 1, make_uint2(0u, 0u));
        tab[0] = make_uint2(first, count);
        CK(hipMemcpyAsync(level_tab.p, tab.data(), tab.size() * sizeof(uint2), hipMemcpyHostToDevice, st));
        uint64_t bound = count;
        for (uint32_t j = 0; j < LEVEL_BATCH; ++j) {
            LevelArgs a;
            a.nodes = nodes.as<BNode>();
            a.pos_in = pos[cur].as<float>();
            a.id_in = ids[cur].as<uint32_t>();
            a.pos_out = pos[cur ^ 1].as<float>();
            a.id_out = ids[cur ^ 1].as<uint32_t>();
            a.pos_fin = pos_fin.as<float>();
            a.id_fin = id_fin.as<uint32_t>();
            a.mesh_depth = mesh_depth.as<unsigned int>();
            a.leaf_max = ctx->leaf_max;
            const uint32_t grid = uint32_t(std::min<uint64_t>(bound, LEVEL_GRID));
            hipLaunchKernelGGL(k_level, dim3(grid), dim3(64), 0, st, a, level_tab.as<uint2>() + j);
            CK(hipGetLastError());
            cur ^= 1;
            bound = std::min<uint64_t>(2 * bound, node_cap);
        }
        CK(hipMemcpyAsync(tab.data(), level_tab.p, tab.size() * sizeof(uint2), hipMemcpyDeviceToHost, st));
        CK(hipStreamSynchronize(st));
        for (uint32_t j = 0; j < LEVEL_BATCH && tab[j].y > 0; ++j) levels.push_back({tab[j].x, tab[j].y});
        first = tab[LEVEL_BATCH].x;
        count = tab[LEVEL_BATCH].y;
    }
    {
        const uint32_t total = first;
        std::vector<uint2> lv;  // read by an asynchronous copy until the wait below
        if (total <= FLATTEN_SMALL) {
            lv.resize(levels.size());
            for (size_t l = 0; l < levels.size(); ++l) lv[l] = make_uint2(levels[l].first, levels[l].second);
            CK(level_list.ensure(std::max<size_t>(lv.size(), 1) * sizeof(uint2)));
            CK(hipMemcpyAsync(level_list.p, lv.data(), lv.size() * sizeof(uint2), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_flatten_small, dim3(1), dim3(1024), 0, st, nodes.as<BNode>(), total,
                               level_list.as<uint2>(), uint32_t(lv.size()));
            CK(hipGetLastError());
        } else {
            hipLaunchKernelGGL(k_order_kids, dim3(blocks_for(total, 256)), dim3(256), 0, st, nodes.as<BNode>(), total);
            CK(hipGetLastError());
            for (size_t l = levels.size(); l-- > 0;) {
                hipLaunchKernelGGL(k_counts, dim3(blocks_for(levels[l].second, 256)), dim3(256), 0, st,
                                   nodes.as<BNode>(), levels[l].first, levels[l].second);
                CK(hipGetLastError());
            }
            for (size_t l = 0; l < levels.size(); ++l) {
                hipLaunchKernelGGL(k_preorder, dim3(blocks_for(levels[l].second, 256)), dim3(256), 0, st,
                                   nodes.as<BNode>(), levels[l].first, levels[l].second);
                CK(hipGetLastError());
            }
        }
        CK(hipMemcpyAsync(roots.data(), nodes.p, nmesh * sizeof(BNode), hipMemcpyDeviceToHost, st));
        CK(hipMemcpyAsync(depth.data(), mesh_depth.p, nmesh * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        CK(hipStreamSynchronize(st));
        for (uint32_t m = 0; m < nmesh; ++m) {
            node_base[m] = uint32_t(ctx->nodes_used + total_internal);
            total_internal += roots[m].icount;
            tri_base[m] = uint32_t(ctx->tris_used + tri_begin[m]);
        }
        // node indices travel as (node << 1) | side below a flag bit in the batched traversal
        if (ctx->nodes_used + total_internal >= (1ull << 30) || ctx->tris_used + T >= 0xFFFFFFF0ull) {
            res.err = "arena exceeds 32-bit indexing (at most 2^30 internal nodes)";
            goto fail;
        }
        // grow the arenas (keeping what is already registered)
        auto grow = [&](DevBuf& b, size_t need) -> hipError_t {
            if (need <= b.bytes) return hipSuccess;
            size_t want = std::max(need, b.bytes * 2);
            void* np = nullptr;
            hipError_t e = hipMalloc(&np, want);
            if (e != hipSuccess) return e;
            if (b.p) {
                e = hipMemcpyAsync(np, b.p, b.bytes, hipMemcpyDeviceToDevice, st);
                if (e != hipSuccess) return e;
                e = hipStreamSynchronize(st);
                if (e != hipSuccess) return e;
                (void)hipFree(b.p);
            }
            b.p = np;
            b.bytes = want;
            return hipSuccess;
        };
        CK(grow(ctx->d_nodes, (ctx->nodes_used + total_internal + 1) * sizeof(NodeRec)));
        CK(grow(ctx->d_tris, (ctx->tris_used + T + 1) * 9 * sizeof(float)));
        CK(grow(ctx->d_perm, (ctx->tris_used + T + 1) * sizeof(uint32_t)));
        CK(grow(ctx->d_leafof, (ctx->tris_used + T + 1) * sizeof(uint32_t)));
        CK(hipMemcpyAsync(base_nodes.p, node_base.data(), nmesh * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        CK(hipMemcpyAsync(base_tris.p, tri_base.data(), nmesh * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_emit_nodes, dim3(blocks_for(total, 256)), dim3(256), 0, st, nodes.as<BNode>(), total,
                           base_nodes.as<uint32_t>(), ctx->d_nodes.as<NodeRec>());
        CK(hipGetLastError());
        hipLaunchKernelGGL(k_emit_tris, dim3(blocks_for(total, 256)), dim3(256), 0, st, nodes.as<BNode>(), total,
                           base_tris.as<uint32_t>(), pos_fin.as<float>(), id_fin.as<uint32_t>(), ctx->d_tris.as<float>(),
                           ctx->d_perm.as<uint32_t>(), ctx->d_leafof.as<uint32_t>());
        CK(hipGetLastError());
        hipLaunchKernelGGL(k_slack, dim3(total), dim3(64), 0, st, nodes.as<BNode>(), total, pos_fin.as<float>(),
                           slack_bits.as<unsigned int>(), radius_bits.as<unsigned int>());
        CK(hipGetLastError());
        CK(hipMemcpyAsync(slack.data(), slack_bits.p, size_t(nmesh) * SLACK_REPS * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, st));
        CK(hipMemcpyAsync(radius.data(), radius_bits.p, nmesh * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        CK(hipStreamSynchronize(st));
        res.meshes.resize(nmesh);
        for (uint32_t m = 0; m < nmesh; ++m) {
            MeshRec& r = res.meshes[m];
            r.node_base = node_base[m];
            r.tri_base = tri_base[m];
            r.ntri = counts[m];
            r.nnodes = roots[m].icount;
            r.root = ChildRec{};
            std::memcpy(r.root.box, roots[m].box, sizeof r.root.box);
            r.root.surf = roots[m].surf;
            r.root.link = 0;  // node 0, or triangle 0 of a one-leaf tree
            r.root.tris = counts[m] | (roots[m].icount == 0 ? LEAF_BIT : 0u);
            r.root.toff = 0;
            r.depth = depth[m];
            uint32_t sb = 0;  // the largest of the mesh's copies (positive floats order as their bits)
            for (uint32_t k = 0; k < SLACK_REPS; ++k) sb = std::max(sb, slack[size_t(m) * SLACK_REPS + k]);
            std::memcpy(&r.slack, &sb, sizeof(float));
            std::memcpy(&r.radius, &radius[m], sizeof(float));
        }
        ctx->nodes_used += total_internal;
        ctx->tris_used += T;
    }
fail: {
        size_t kept = 0;
        for (uint32_t k = 0; k < BUILD_SCRATCH; ++k) kept += W[k].bytes;
        if (!res.err.empty() || kept > BUILD_KEEP)
            for (uint32_t k = 0; k < BUILD_SCRATCH; ++k) W[k].release();
    }
    return res;
}

}  // namespace cdk
