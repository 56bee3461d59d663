// Response: shooting the uncollide rays and splitting the delta
// (ShootUncollideRays::ExecuteShootUncollideRays,
//  inMyRoom_vulkan/src/CollisionDetection/ShootUncollideRays.cpp:14-179;
//  Ray::IntersectOBBtree, Ray.cpp:13-236; CollisionDetection.cpp:70-99, 143-150).
//
//  k_shoot    two lanes per uncollide ray.  Pass 1 shoots the ray at the
//             other entity's tree (closest hit, even lane); it succeeds when
//             that hit is a back face reached before the ray re-enters its own
//             tree (odd lane; the Hermann pass, :129-158).  On success the ray is reflected (origin
//             moved by the response, direction minus the hit's normal,
//             :104-110) and pass 2 shoots it the other way.  Each pass's force
//             response and push-out vector are kept per ray.
//  k_respond  one wavefront per entity pair: the force responses summed,
//             FindResponse (:160-169) over the push-out vectors, the
//             world-space delta, and its split by how far each entity's
//             average contact point moved since the previous frame.
//
// The closest-hit traversal keeps an explicit stack in LDS.  At an internal
// node both children's entry/exit distances are computed, the farther child
// is pushed below the nearer, and a child is entered only if its exit
// distance is not negative and, when it is popped, its entry distance is
// below the best hit so far: the nodes visited, and the order in which equal
// distances are met, are those of the reference's recursion.
#include "rayattr.hpp"

namespace cdk {

namespace {

constexpr float FLT_EPS = 1.1920928955078125e-07f;  // std::numeric_limits<float>::epsilon()
constexpr int RAY_STACK = RAY_STACK_ENTRIES;  // entries per lane: a tree's depth + 2

__device__ __forceinline__ float max_std(float a, float b) { return (a < b) ? b : a; }  // std::max
__device__ __forceinline__ float min_std(float a, float b) { return (b < a) ? b : a; }  // std::min

// glm::intersectRayTriangle from the origin (glm 0.9.9's Möller-Trumbore
// form); back = the negative-determinant branch, the backface output of the
// reference's glm fork.
__device__ __forceinline__ bool ray_triangle(f3 dir, f3 v0, f3 v1, f3 v2, float& bx, float& by, float& dist,
                                             bool& back) {
    const f3 e1 = sub(v1, v0), e2 = sub(v2, v0);
    const f3 p = cross(dir, e2);
    const float det = dot(e1, p);
    const f3 d0 = sub(mk(0.f, 0.f, 0.f), v0);
    f3 perp;
    if (det > FLT_EPS) {
        bx = dot(d0, p);
        if (bx < 0.f || bx > det) return false;
        perp = cross(d0, e1);
        by = dot(dir, perp);
        if (by < 0.f || bx + by > det) return false;
        back = false;
    } else if (det < -FLT_EPS) {
        bx = dot(d0, p);
        if (bx > 0.f || bx < det) return false;
        perp = cross(d0, e1);
        by = dot(dir, perp);
        if (by > 0.f || bx + by < det) return false;
        back = true;
    } else {
        return false;
    }
    const float inv = 1.f / det;
    dist = dot(e2, perp) * inv;
    bx = bx * inv;
    by = by * inv;
    return true;
}

// The ray in a tree's own frame.  C maps the mesh's coordinates to the
// ray's frame (origin at the ray's origin), so the mesh-frame ray
// x(t) = o + t d with o = -A^-1 c, d = A^-1 dir (A, c: C's linear part and
// translation) is mapped onto t dir: a slab crossing has the same parameter t
// in both frames.  The boxes are then tested where they are stored, with no
// transform and no normalised normal per test.
struct LocalRay {
    f3 o, d;
    float on, dn;   // |o|_1, |d|_1
    float kappa;    // relative rounding margin of the box tests, grown with A's condition number
    bool ok;        // A invertible: else every box is entered
};
__device__ __forceinline__ float norm1f(f3 v) { return (fabsf(v.x) + fabsf(v.y)) + fabsf(v.z); }
__device__ __forceinline__ LocalRay local_ray(const M4& C, f3 dir) {
    const f3 a0 = mk(C.m[0], C.m[1], C.m[2]), a1 = mk(C.m[4], C.m[5], C.m[6]), a2 = mk(C.m[8], C.m[9], C.m[10]);
    const f3 r0 = cross(a1, a2), r1 = cross(a2, a0), r2 = cross(a0, a1);  // rows of det(A) A^-1
    const float det = dot(a0, r0);
    const float inv = 1.f / det;
    const f3 c = mk(C.m[12], C.m[13], C.m[14]);
    LocalRay L;
    L.o = mk(-dot(r0, c) * inv, -dot(r1, c) * inv, -dot(r2, c) * inv);
    L.d = mk(dot(r0, dir) * inv, dot(r1, dir) * inv, dot(r2, dir) * inv);
    L.on = norm1f(L.o);
    L.dn = norm1f(L.d);
    // |A|_F |A^-1|_F is 3 for a rotation times a uniform scale
    const float fa = sqrtf((dot(a0, a0) + dot(a1, a1)) + dot(a2, a2));
    const float fi = sqrtf((dot(r0, r0) + dot(r1, r1)) + dot(r2, r2)) * fabsf(inv);
    const float cond = fa * fi * (1.f / 3.f);
    L.kappa = 0x1p-16f * (cond > 1.f ? cond : 1.f);
    L.ok = det != 0.f && __builtin_isfinite(inv) && __builtin_isfinite(L.kappa) && L.kappa < 0x1p-4f &&
           __builtin_isfinite(L.on) && __builtin_isfinite(L.dn);
    return L;
}
// Ray::IntersectParalgram from the origin (Ray.cpp:36-133): the three slabs
// of the box, normals v x w, w x u, u x v.
__device__ __forceinline__ bool ray_box(f3 dir, const Box& b, float& tmin, float& tmax) {
    tmin = -__builtin_huge_valf();
    tmax = __builtin_huge_valf();
    const f3 ro = sub(mk(0.f, 0.f, 0.f), b.c);
    const f3 side[3] = {b.u, b.v, b.w};
    const f3 pa[3] = {b.v, b.w, b.u}, pb[3] = {b.w, b.u, b.v};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const f3 n = normalize(cross(pa[k], pb[k]));
        const float d = -fabsf(dot(n, side[k]));
        const float hc = dot(n, ro);
        const float vn1 = hc + d;
        const float vn2 = -hc + d;
        const float vd = dot(n, dir);
        if (fabsf(vd) >= FLT_EPS) {
            const float inv = 1.f / vd;
            float t1 = -vn1 * inv;
            float t2 = vn2 * inv;
            if (t1 > t2) {
                const float t = t1;
                t1 = t2;
                t2 = t;
            }
            tmin = max_std(t1, tmin);
            tmax = min_std(t2, tmax);
            if (tmin > tmax || tmax < 0.f) return false;
        } else if (vn1 > 0.f || vn2 > 0.f) {
            return false;
        }
    }
    return true;
}

// The mesh-frame test of a stored box (centre c, half-sides u, v, w): the
// three slabs of Ray::IntersectParalgram (Ray.cpp:36-133), normals v x w,
// w x u, u x v, each slab's half-width the box's projection radius on its
// normal.  Every slab and crossing parameter is computed once and bounded by
// its rounding (kappa, and the relative error of the crossing's
// denominator), giving a widened and a shrunk interval.  BOX_OUT: even the
// widened interval misses, so the exact ray misses and so does the
// reference's float test, whose rounding is far inside these bounds.
// BOX_IN: even the shrunk interval meets the ray ahead of its origin.
// BOX_NEAR: the decision is within rounding (a box the ray grazes, one ending
// at the ray's origin, a flat box, a slab the ray runs along); the caller
// then runs the reference's own test.  tmin/tmax: the unwidened parameters,
// for the order of the children and the pruning against the best hit.
// The box-only terms (RaySlab), computed once per stored box.
__device__ __forceinline__ RaySlab make_slab(const float* bx) {
    const f3 c = mk(bx[0], bx[1], bx[2]);
    const f3 u = mk(bx[3], bx[4], bx[5]), v = mk(bx[6], bx[7], bx[8]), w = mk(bx[9], bx[10], bx[11]);
    const f3 nrm[3] = {cross(v, w), cross(w, u), cross(u, v)};
    RaySlab s;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const f3 n = nrm[k];
        s.n[3 * k] = n.x;
        s.n[3 * k + 1] = n.y;
        s.n[3 * k + 2] = n.z;
        s.h[k] = (fabsf(dot(n, u)) + fabsf(dot(n, v))) + fabsf(dot(n, w));
    }
    s.cr = norm1f(c);
    s.sr = (norm1f(u) + norm1f(v)) + norm1f(w);
    s.pad[0] = s.pad[1] = 0.f;
    return s;
}
enum : int { BOX_OUT = 0, BOX_IN = 1, BOX_NEAR = 2 };
__device__ __forceinline__ int box_local(const LocalRay& R, const float* bx, const RaySlab& sl, float& tmin,
                                         float& tmax) {
    tmin = -__builtin_huge_valf();
    tmax = __builtin_huge_valf();
    if (!R.ok) return BOX_NEAR;
    const f3 c = mk(bx[0], bx[1], bx[2]);
    const f3 q = sub(R.o, c);
    const float reach = ((norm1f(q) + R.on) + sl.cr) + sl.sr;
    const f3 nrm[3] = {mk(sl.n[0], sl.n[1], sl.n[2]), mk(sl.n[3], sl.n[4], sl.n[5]), mk(sl.n[6], sl.n[7], sl.n[8])};
    float lo_w = -__builtin_huge_valf(), hi_w = __builtin_huge_valf();
    float lo_s = -__builtin_huge_valf(), hi_s = __builtin_huge_valf();
    bool decided = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const f3 n = nrm[k];
        const float nn = norm1f(n);
        const float h = sl.h[k];
        const float e = R.kappa * nn * reach;  // bound on the rounding of h and num
        const float num = dot(n, q);
        const float den = dot(n, R.d);
        const float dl = R.kappa * nn * R.dn;  // bound on den's rounding
        if (fabsf(den) > 4.f * dl && fabsf(den) > 0x1p-100f) {
            const float r = __builtin_amdgcn_rcpf(den);  // within an ulp, far inside rho
            const float rho = 2.f * dl * fabsf(r) + R.kappa;  // relative error of a crossing parameter
            const float sg = r < 0.f ? -1.f : 1.f;  // the crossing through -h is the entry when den > 0
            const float t1 = (-h - num) * r, t2 = (h - num) * r;
            tmin = max_std(sg > 0.f ? t1 : t2, tmin);
            tmax = min_std(sg > 0.f ? t2 : t1, tmax);
            const float w1 = (-(h + e) - num) * r, w2 = ((h + e) - num) * r;
            const float a = sg > 0.f ? w1 : w2, b = sg > 0.f ? w2 : w1;
            lo_w = max_std(a - rho * fabsf(a), lo_w);
            hi_w = min_std(b + rho * fabsf(b), hi_w);
            if (h > e) {
                const float s1 = (-(h - e) - num) * r, s2 = ((h - e) - num) * r;
                const float as = sg > 0.f ? s1 : s2, bs = sg > 0.f ? s2 : s1;
                lo_s = max_std(as + rho * fabsf(as), lo_s);
                hi_s = min_std(bs - rho * fabsf(bs), hi_s);
            } else {
                decided = false;  // a slab thinner than its rounding
            }
        } else if (!(fabsf(num) + e < h)) {
            decided = false;  // runs along the slab, not clearly inside it
        }
    }
    if (!(lo_w <= hi_w && hi_w >= 0.f)) return BOX_OUT;
    if (decided && lo_s <= hi_s && hi_s >= 0.f) return BOX_IN;
    return BOX_NEAR;
}

struct Hit {
    bool hit, back;
    float dist, bx, by;
    uint32_t tri;  // mesh-local tree-order slot
};

// A leaf's triangles are fetched LEAF_FETCH at a time (their loads overlap),
// then tested in order.
constexpr uint32_t LEAF_FETCH = 4;
__device__ __forceinline__ void leaf_hits(const float* __restrict__ tris, const MeshRec& m, const M4& C, f3 dir, uint32_t off,
                          uint32_t cnt, Hit& best) {
    for (uint32_t i0 = off; i0 < off + cnt; i0 += LEAF_FETCH) {
        const uint32_t k = min(LEAF_FETCH, off + cnt - i0);
        float t9[LEAF_FETCH][9];
#pragma unroll
        for (uint32_t u = 0; u < LEAF_FETCH; ++u)
            if (u < k) {
                const float* src = tris + (size_t(m.tri_base) + i0 + u) * 9;
#pragma unroll
                for (int e = 0; e < 9; ++e) t9[u][e] = src[e];
            }
#pragma unroll
        for (uint32_t u = 0; u < LEAF_FETCH; ++u) {
            if (u >= k) break;
            const f3 P0 = mul4(C, mk(t9[u][0], t9[u][1], t9[u][2]), 1.f);
            const f3 P1 = mul4(C, mk(t9[u][3], t9[u][4], t9[u][5]), 1.f);
            const f3 P2 = mul4(C, mk(t9[u][6], t9[u][7], t9[u][8]), 1.f);
            float bx = 0.f, by = 0.f, dist = __builtin_huge_valf();
            bool back = false;
            if (ray_triangle(dir, P0, P1, P2, bx, by, dist, back) && dist > 0.f && dist < best.dist) {
                best.hit = true;
                best.back = back;
                best.dist = dist;
                best.tri = i0 + u;
                best.bx = bx;
                best.by = by;
            }
        }
    }
}

// Per-lane traversal stack in LDS, entry-major (entry e of lane l at
// e * 64 + l, so a wavefront's accesses fall in distinct banks).  An entry
// is a child still to visit: cnt 0 for internal node `link`, else a leaf of
// cnt triangles from `link`; t0 its entry distance.  A child whose exit
// distance is negative is never pushed (that check does not depend on the
// best hit, so making it at push time is the reference's check).
// link and cnt share one word (link < 2^RAY_LINK_BITS, cnt <= 64; the host
// checks the meshes): 8 bytes an entry, so six wavefronts fit a CU's LDS
// where 12-byte entries allowed four.
struct RayStack {
    uint32_t* lc;
    float* t0;
    int lane;
    int cap;  // entries per lane
    __device__ __forceinline__ void put(int i, uint32_t l, uint32_t c, float t) {
        lc[i * 64 + lane] = l | (c << RAY_LINK_BITS);
        t0[i * 64 + lane] = t;
    }
};

// The lanes' published best hits (LDS, one word per lane): relaxed atomic
// accesses, so that a lane reads its partner's latest value each time and the
// accesses stay LDS operations.
__device__ __forceinline__ void publish(float* sbest, int lane, float v) {
    __hip_atomic_store(&sbest[lane], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
__device__ __forceinline__ float partner(float* sbest, int lane) {
    return __hip_atomic_load(&sbest[lane ^ 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// Closest hit of the ray (o, dir) on mesh m moved by M (Ray.cpp:135-236).
// A stack that would overflow (the host sizes it from the tree depths, so
// this is unreachable unless that sizing is wrong) sets bit 0 of *err and
// ends the traversal; the host then fails the call with CD_E_STACK.
// sbest: per lane of the wavefront, the best hit distance so far of its walk,
// and when the walk is over its hit distance if that is a back face, else -1.
// bounded (the own-tree pass of a Hermann pair): the walk only decides
// whether p2.dist <= p3.dist + eps holds, with p2 the partner lane's closest
// hit, so an entry whose entry distance lies beyond the partner's best so far
// (never below its final one) plus eps, widened by BOUND_SLACK for rounding,
// is skipped, and the walk stops once the partner ends without a back-face
// hit (the pass fails whatever this walk finds).
constexpr float BOUND_SLACK = 1.f + 0x1p-12f;
// nslab: the slab records of this mesh's nodes' children (2 per node),
// rslab: its root's.
__device__ __forceinline__ Hit ray_tree(const MeshRec& m, const NodeRec* __restrict__ nodes,
                                        const RaySlab* __restrict__ nslab, const RaySlab& rslab,
                                        const float* __restrict__ tris, M4 C, f3 o, f3 dir, RayStack& S,
                                        unsigned long long* err, uint32_t& nbox, uint32_t& ntri, float* sbest,
                                        bool bounded, float eps) {
    C.m[12] = C.m[12] - o.x;  // centred on the ray origin (Ray.cpp:150-154)
    C.m[13] = C.m[13] - o.y;
    C.m[14] = C.m[14] - o.z;
    C.m[15] = C.m[15] - 0.f;
    Hit best = {false, false, __builtin_huge_valf(), 0.f, 0.f, 0xFFFFFFFFu};
    publish(sbest, S.lane, __builtin_huge_valf());
    const LocalRay R = local_ray(C, dir);
    // the mesh-frame test, or the reference's own where that one is within rounding
    auto enter = [&](const float* bx, const RaySlab& sl, float& t0, float& t1) -> bool {
        const int k = box_local(R, bx, sl, t0, t1);
        if (k != BOX_NEAR) return k == BOX_IN;
        return ray_box(dir, xform_box(C, box_from12(bx)), t0, t1) && t1 >= 0.f;
    };
    float r0, r1;
    ++nbox;
    auto finish = [&]() {  // the walk's end, as the partner lane reads it
        publish(sbest, S.lane, (best.hit && best.back) ? best.dist : -1.f);
        return best;
    };
    if (!enter(m.root.box, rslab, r0, r1)) return finish();
    if (m.nnodes == 0) {  // the root is a leaf
        ntri += m.ntri;
        leaf_hits(tris, m, C, dir, 0, m.ntri, best);
        return finish();
    }
    const NodeRec* nb = nodes + m.node_base;
    int sp = 0;
    // One step enters at most one internal node and tests at most one leaf,
    // then pops the next entry: the lanes of a wavefront, each at its own
    // place in its own walk, all make progress every step, where a lane
    // working through a run of leaves would otherwise hold the others.  Each
    // lane's sequence of pops, tests and best-hit updates is the walk's own.
    bool at_node = true;
    uint32_t node = 0;  // the root, entered without a distance check
    uint32_t leaf_link = 0, leaf_cnt = 0;
    for (;;) {
        if (at_node) {
            const NodeRec& n = nb[node];
            float l0, l1, q0, q1;
            const RaySlab* ns = nslab + 2 * size_t(node);
            const bool lh = enter(n.c[0].box, ns[0], l0, l1);
            const bool rh = enter(n.c[1].box, ns[1], q0, q1);
            nbox += 2;
            if (sp + 2 > S.cap) {  // deeper than the host allows (it sizes the stack from the tree depths)
                atomicOr(err, 1ull);
                best.hit = false;
                break;
            }
            const uint32_t cl = child_cnt(n.c[0]), cr = child_cnt(n.c[1]);
            // the farther below the nearer (Ray.cpp:175-199); order set by the entry distances alone
            const bool left_first = !(lh && rh) || l0 < q0;
            if (lh && rh) {
                if (left_first) {
                    S.put(sp++, n.c[1].link, cr, q0);
                    S.put(sp++, n.c[0].link, cl, l0);
                } else {
                    S.put(sp++, n.c[0].link, cl, l0);
                    S.put(sp++, n.c[1].link, cr, q0);
                }
            } else if (lh) {
                S.put(sp++, n.c[0].link, cl, l0);
            } else if (rh) {
                S.put(sp++, n.c[1].link, cr, q0);
            }
            at_node = false;
        }
        if (leaf_cnt) {
            ntri += leaf_cnt;
            leaf_hits(tris, m, C, dir, leaf_link, leaf_cnt, best);
            leaf_cnt = 0;
            publish(sbest, S.lane, best.dist);
        }
        // the next entry the best hit so far does not prune
        bool more = false;
        bool stop = false;
        const float pb = bounded ? partner(sbest, S.lane) : __builtin_huge_valf();
        if (pb < 0.f) stop = true;  // the partner ended without a back-face hit
        const float bound = (pb + eps) * BOUND_SLACK;
        while (sp > 0 && !stop) {
            --sp;
            const uint32_t ew = S.lc[sp * 64 + S.lane];
            const uint32_t el = ew & ((1u << RAY_LINK_BITS) - 1u), ec = ew >> RAY_LINK_BITS;
            const float et = S.t0[sp * 64 + S.lane];
            if (!(et < best.dist)) continue;
            if (et > bound) continue;  // no hit in it can make the pass fail
            if (ec == 0u) {
                node = el;
                at_node = true;
            } else {
                leaf_link = el;
                leaf_cnt = ec;
            }
            more = true;
            break;
        }
        if (!more) break;
    }
    return finish();
}

// Ray::MoveOriginEpsilonTowardsDirection (Ray.cpp:13-21).
__device__ __forceinline__ f3 move_eps(f3 o, f3 d, float factor) {
    float big = fabsf(o.x);
    if (big < fabsf(o.y)) big = fabsf(o.y);
    if (big < fabsf(o.z)) big = fabsf(o.z);
    const float e = big * FLT_EPS;
    return add(o, scale(d, factor * e));
}

struct Pass {
    bool ok;
    f3 response, normal;
};

struct MeshView {
    const MeshRec* meshes;
    const NodeRec* nodes;
    const RaySlab* slabs;  // 2 per node, node arena order
    const RaySlab* roots;  // per mesh
    const float* tris;
    const uint2* attrs;
    const uint32_t* vid;
    const float* nrm;
    const uint32_t* perm;
};

// HermannPass (ShootUncollideRays.cpp:129-158), on two lanes: the even lane
// shoots the ray at the other entity's tree (p2), the odd lane shoots the
// ray moved 4 epsilon along itself at the entity's own tree (p3), which does
// not depend on p2; both lanes then hold both hits and decide the same way.
// The reference skips p3 when p2 fails; computing it anyway changes nothing.
__device__ __forceinline__ Hit shfl_hit(const Hit& h, int src) {
    Hit r;
    r.hit = __shfl(int(h.hit), src, 64) != 0;
    r.back = __shfl(int(h.back), src, 64) != 0;
    r.dist = __shfl(h.dist, src, 64);
    r.bx = __shfl(h.bx, src, 64);
    r.by = __shfl(h.by, src, 64);
    r.tri = uint32_t(__shfl(int(h.tri), src, 64));
    return r;
}

__device__ __forceinline__ Pass hermann(const MeshView& V, uint32_t meshA, const M4& Amat, uint32_t meshB, const M4& Bmat,
                        const M3& Bn, f3 o, f3 d, RayStack& S, unsigned long long* err, uint32_t& nbox,
                        uint32_t& ntri, float* sbest) {
    Pass r = {false, mk(0.f, 0.f, 0.f), mk(0.f, 0.f, 0.f)};
    const bool own = (threadIdx.x & 1u) != 0;
    const f3 o2 = move_eps(o, d, 4.f);
    M4 C;  // one call with lane-selected arguments, so the two traversals run as one instruction stream
#pragma unroll
    for (int e = 0; e < 16; ++e) C.m[e] = own ? Amat.m[e] : Bmat.m[e];
    const float eps = vlength(sub(o2, o));
    const uint32_t mesh = own ? meshA : meshB;
    const MeshRec& mw = V.meshes[mesh];
    const Hit h = ray_tree(mw, V.nodes, V.slabs + 2 * size_t(mw.node_base), V.roots[mesh], V.tris, C, own ? o2 : o, d, S,
                           err, nbox, ntri, sbest, own, eps);
    const int even = int(threadIdx.x & ~1u);
    const Hit p2 = shfl_hit(h, even), p3 = shfl_hit(h, even + 1);
    if (!(p2.hit && p2.back)) return r;
    if (p2.dist <= p3.dist + eps) {
        const MeshRec mb = V.meshes[meshB];
        r.ok = true;
        r.response = scale(d, p2.dist);
        const uint2 attr = V.attrs[meshB];
        const float* t9 = V.tris + (size_t(mb.tri_base) + p2.tri) * 9;
        f3 nk[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) nk[k] = corner_normal(attr, V.nrm, vertex_id(mb, attr, V.vid, V.perm, p2.tri, k), t9);
        const f3 ni = add(add(scale(nk[0], (1.f - p2.bx) - p2.by), scale(nk[1], p2.bx)), scale(nk[2], p2.by));
        r.normal = normalize(mul3(Bn, ni));
    }
    return r;
}

// CalcForceResponse (:118-127).
__device__ __forceinline__ f3 force_response(const Pass& p) {
    const f3 dir = normalize(p.response);
    const float len = vlength(p.response);
    const float c = dot(p.normal, dir);
    const float c2 = c * c;
    return scale(dir, c2 * len);
}

__device__ __forceinline__ f3 neg(f3 v) { return mk(-v.x, -v.y, -v.z); }  // glm's unary minus
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ void put3(float* o, f3 v) {
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
}
__device__ __forceinline__ f3 get3(const float* o) { return mk(o[0], o[1], o[2]); }

}  // namespace

// Per-ray result: for each of the two passes, the contribution to the summed
// force response and the push-out vector, as the reference accumulates them
// (a first-to-second pass subtracts its force and pushes -response, a
// second-to-first pass adds and pushes +response).
// pbegin: per side (at + stride), the first ray of every pair
// (non-decreasing; a pair without rays has the next pair's begin)
// dc: the step's count block, when the ray and pair counts are read on the
// device (nrays, nq then the capacities of the launch)
// DEEP: the stacks live in HBM, a block's 2 stack_entries 64 words at
// gstack + 2 b stack_entries 64 (link words, then entry distances), for
// trees deeper than the LDS stack holds; the walk is the same.
template <bool DEEP>
__global__ __launch_bounds__(64) void k_shoot(const cd_ray* __restrict__ rays, uint32_t nrays, uint32_t base1,
                                              const uint32_t* __restrict__ pbegin, uint32_t nq, uint32_t stride,
                                              const uint32_t* __restrict__ dc,
                                              const QueryRec* __restrict__ qs, MeshView V,
                                              RayShot* __restrict__ out, int stack_entries,
                                              RunCounters* __restrict__ ctr, uint32_t* __restrict__ gstack) {
    unsigned long long* err = &ctr->err;
    // the stacks are sized by the launch: the deepest tree of the call + 2
    // entries per lane, so shallow trees leave LDS for more wavefronts
    extern __shared__ uint32_t s_stack[];
    __shared__ float s_best[64];  // the lanes' best hits, for the own-tree walks' bound
    uint32_t* base = DEEP ? gstack + size_t(blockIdx.x) * 2 * size_t(stack_entries) * 64 : s_stack;
    RayStack S = {base, reinterpret_cast<float*>(base + stack_entries * 64), int(threadIdx.x), stack_entries};
    const uint32_t r = (blockIdx.x * blockDim.x + threadIdx.x) >> 1;  // two lanes per ray
    uint32_t nbox = 0, ntri = 0;  // boxes tested, ray-triangle tests: the work counted in ctr->sat, ctr->tritests
    if (dc) {
        base1 = dc[DC_RAYS0];
        nrays = min(nrays, dc[DC_RAYS0] + dc[DC_RAYS1]);
        nq = min(nq, dc[DC_NQ]);
    }
    if (r < nrays) {  // both lanes of a ray take part or neither
        const int side = r >= base1 ? 1 : 0;
        const uint32_t local = r - (side ? base1 : 0u);
        const uint32_t* begin = pbegin + size_t(stride) * side;
        // the pair: the last one whose rays begin at or before this ray (of equal
        // begins only the last can hold rays)
        uint32_t lo = 0, hi = nq;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (begin[mid] <= local) lo = mid;
            else hi = mid;
        }
        const QueryRec q = qs[lo];
        M4 M, I;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            M.m[e] = q.rel[e];
            I.m[e] = (e % 5 == 0) ? 1.f : 0.f;
        }
        const M3 Mn = adjoint_transpose(q.rel);
        M3 I3;
#pragma unroll
        for (int e = 0; e < 9; ++e) I3.m[e] = (e % 4 == 0) ? 1.f : 0.f;
        RayShot s = {};
        f3 o = get3(rays[r].origin), d = get3(rays[r].direction);
        // side 0: first-to-second, then the reflected ray second-to-first; side 1 the other way round
#pragma unroll 1
        for (int pass = 0; pass < 2; ++pass) {
            const bool f2s = (pass == 0) == (side == 0);
            const Pass p = f2s ? hermann(V, q.mesh_a, I, q.mesh_b, M, Mn, o, d, S, err, nbox, ntri, s_best)
                               : hermann(V, q.mesh_b, M, q.mesh_a, I, I3, o, d, S, err, nbox, ntri, s_best);
            if (!p.ok) break;
            const f3 F = force_response(p);
            put3(s.force[pass], f2s ? neg(F) : F);  // sum - F == sum + (-F) exactly
            put3(s.push[pass], f2s ? neg(p.response) : p.response);
            s.ok |= 1u << pass;
            o = add(o, p.response);  // ReflectHermannResult (:104-110)
            d = neg(p.normal);
        }
        s.pair = lo;
        if ((threadIdx.x & 1u) == 0) out[r] = s;
    }
    // per wavefront, one atomic each
    const unsigned long long bsum = wave_sum_u64(nbox), tsum = wave_sum_u64(ntri);
    if ((threadIdx.x & 63u) == 0) {
        atomicAdd(&ctr->sat, bsum);
        atomicAdd(&ctr->tritests, tsum);
    }
}

__device__ __forceinline__ float smoother_step(float a, float b, float x) {  // :171-175
    float t = (x - a) / (b - a);
    t = (t < 0.f) ? 0.f : ((1.f < t) ? 1.f : t);
    return t * t * t * (t * (t * 6.f - 15.f) + 10.f);
}
__device__ __forceinline__ bool same_matrix(const float* a, const float* b) {
    for (int i = 0; i < 16; ++i)
        if (!(a[i] == b[i])) return false;
    return true;
}

__device__ __forceinline__ f3 shfl_down3(f3 v, int off) {
    return mk(__shfl_down(v.x, off, 64), __shfl_down(v.y, off, 64), __shfl_down(v.z, off, 64));
}
// One wavefront per pair.  The force responses are summed lane-strided
// (each shot's first pass, then its second) and then over a fixed shuffle
// tree: a sum that is the same on every run.  The reference sums them in its
// ray order, which follows an unordered_map walk no GPU reproduces.  The
// maximum of FindResponse does not depend on the order.
// qw: per side (at + 2 stride) the pairs' ray counts, then (at + stride)
// their first rays; avg: per side (at + 3 stride) the average origins; dc as
// for k_shoot; nshots: the shots there are (a pair whose rays would lie
// beyond them, which only a step over its ray capacity can give, is left
// unshot for the step to run again).
__global__ __launch_bounds__(64) void k_respond(const RayShot* __restrict__ shots, uint32_t base1,
                                                const uint32_t* __restrict__ qw, uint32_t nq, uint32_t stride,
                                                const uint32_t* __restrict__ dc, uint64_t nshots,
                                                const float* __restrict__ avg, const QueryRec* __restrict__ qs,
                                                const uint2* __restrict__ pair_ids, const cd_entity* __restrict__ ents,
                                                float cos_start, float cos_finish, float bias,
                                                cd_pair_response* __restrict__ out) {
    const uint32_t q = blockIdx.x;
    const int lane = threadIdx.x;
    if (dc) {
        base1 = dc[DC_RAYS0];
        nq = min(nq, dc[DC_NQ]);
    }
    if (q >= nq) return;
    const uint2 ids = pair_ids[q];
    cd_pair_response res = {};
    res.first = ids.x;
    res.second = ids.y;
    const float* c1 = ents[ids.x].current;
    const float* c2 = ents[ids.y].current;
    const float* p1 = ents[ids.x].previous;
    const float* p2 = ents[ids.y].previous;
    const uint32_t n0 = qw[q], b0 = qw[stride + q];
    const uint32_t n1 = qw[2 * size_t(stride) + q], b1 = base1 + qw[3 * size_t(stride) + q];
    const bool beyond = (n0 && uint64_t(b0) + n0 > nshots) || (n1 && uint64_t(b1) + n1 > nshots);
    if ((n0 == 0 && n1 == 0) || beyond || (same_matrix(c1, p1) && same_matrix(c2, p2))) {  // :63, :80-81
        if (lane == 0) out[q] = res;
        return;
    }
    res.shot = 1;
    const uint32_t n = n0 + n1;
    f3 sum = mk(0.f, 0.f, 0.f);
    for (uint32_t i = uint32_t(lane); i < n; i += 64) {
        const RayShot& s = shots[i < n0 ? b0 + i : b1 + (i - n0)];
        for (int pass = 0; pass < 2; ++pass)
            if (s.ok & (1u << pass)) sum = add(sum, get3(s.force[pass]));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sum = add(sum, shfl_down3(sum, off));
    sum = mk(__shfl(sum.x, 0, 64), __shfl(sum.y, 0, 64), __shfl(sum.z, 0, 64));
    f3 delta = mk(0.f, 0.f, 0.f);
    if (!(sum.x == 0.f && sum.y == 0.f && sum.z == 0.f)) {  // :87
        const f3 nf = normalize(sum);
        float max_len = 0.f;  // FindResponse (:160-169)
        for (uint32_t i = uint32_t(lane); i < n; i += 64) {
            const RayShot& s = shots[i < n0 ? b0 + i : b1 + (i - n0)];
            for (int pass = 0; pass < 2; ++pass) {
                if (!(s.ok & (1u << pass))) continue;
                const f3 r = get3(s.push[pass]);
                const float c = dot(normalize(r), nf);
                const float need = vlength(r) / c;
                max_len = max_std(max_len, smoother_step(cos_finish, cos_start, c) * need);
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) max_len = max_std(max_len, __shfl_down(max_len, off, 64));
        M4 F;
        for (int e = 0; e < 16; ++e) F.m[e] = c1[e];
        delta = mul4(F, scale(scale(nf, max_len), bias), 0.f);
    }
    if (lane != 0) return;
    // the split by each entity's movement (CollisionDetection.cpp:86-97, 143-150)
    M4 C1, P1, C2, P2;
    for (int e = 0; e < 16; ++e) {
        C1.m[e] = c1[e];
        P1.m[e] = p1[e];
        C2.m[e] = c2[e];
        P2.m[e] = p2[e];
    }
    const f3 a1 = get3(avg + 3 * size_t(q)), a2 = get3(avg + 3 * (size_t(stride) + q));
    const float m1 = vlength(sub(mul4(C1, a1, 1.f), mul4(P1, a1, 1.f)));
    const float m2 = vlength(sub(mul4(C2, a2, 1.f), mul4(P2, a2, 1.f)));
    const float total = m1 + m2;
    put3(res.delta, delta);
    put3(res.delta_first, scale(neg(delta), m1 / total));
    put3(res.delta_second, scale(delta, m2 / total));
    out[q] = res;
}

// Slab records: one thread per child of count nodes, and one for mesh m's root.
__global__ void k_ray_slabs(const NodeRec* __restrict__ nodes, uint32_t count, RaySlab* __restrict__ out,
                            const MeshRec* __restrict__ meshes, uint32_t m, RaySlab* __restrict__ root_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2 * count) out[i] = make_slab(nodes[i >> 1].c[i & 1u].box);
    if (i == 0) *root_out = make_slab(meshes[m].root.box);
}
hipError_t launch_ray_slabs(const NodeRec* nodes, uint32_t count, RaySlab* out, const MeshRec* meshes, uint32_t m,
                            RaySlab* root_out, hipStream_t st) {
    hipLaunchKernelGGL(k_ray_slabs, dim3((2 * count + 255) / 256 + 1), dim3(256), 0, st, nodes, count, out, meshes, m,
                       root_out);
    return hipGetLastError();
}
hipError_t launch_shoot(const cd_ray* rays, uint32_t nrays, uint32_t base1, const uint32_t* pbegin, uint32_t nq,
                        const QueryRec* qs, const MeshRec* meshes, const NodeRec* nodes, const RaySlab* slabs,
                        const RaySlab* roots, const float* tris,
                        const uint2* attrs, const uint32_t* vid, const float* nrm, const uint32_t* perm, RayShot* shots,
                        uint32_t stack_entries, RunCounters* ctr, hipStream_t st, uint32_t* gstack, uint32_t stride,
                        const uint32_t* dc) {
    if (!nrays) return hipSuccess;
    if (stack_entries < 2) return hipErrorInvalidValue;
    if (!stride) stride = nq;
    const MeshView V = {meshes, nodes, slabs, roots, tris, attrs, vid, nrm, perm};
    const uint64_t lanes = 2ull * nrays;
    const dim3 grid(uint32_t((lanes + 63) / 64));
    if (stack_entries <= uint32_t(RAY_STACK)) {
        const size_t lds = size_t(stack_entries) * 64 * (sizeof(uint32_t) + sizeof(float));
        hipLaunchKernelGGL(k_shoot<false>, grid, dim3(64), lds, st, rays, nrays, base1, pbegin, nq, stride, dc, qs, V,
                           shots, int(stack_entries), ctr, (uint32_t*)nullptr);
    } else {
        if (!gstack) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_shoot<true>, grid, dim3(64), 0, st, rays, nrays, base1, pbegin, nq, stride, dc, qs, V,
                           shots, int(stack_entries), ctr, gstack);
    }
    return hipGetLastError();
}
hipError_t launch_respond(const RayShot* shots, uint32_t base1, const uint32_t* qw, uint32_t nq, const float* avg,
                          const QueryRec* qs, const uint2* pair_ids, const cd_entity* ents, float cos_start,
                          float cos_finish, float bias, cd_pair_response* out, hipStream_t st, uint32_t stride,
                          const uint32_t* dc, uint64_t nshots) {
    if (!nq) return hipSuccess;
    if (!stride) stride = nq;
    hipLaunchKernelGGL(k_respond, dim3(nq), dim3(64), 0, st, shots, base1, qw, nq, stride, dc, nshots, avg, qs, pair_ids,
                       ents, cos_start, cos_finish, bias, out);
    return hipGetLastError();
}

}  // namespace cdk
