// Exact-sign orientation predicate for float32 inputs and the exact-semantics
// triangle-triangle test built on it.
//
// side(a,b,c,d) = sign( ((a-c) x (b-c)) . (d-c) ).
// Stage 1: float32 evaluation with a forward error bound (semi-static filter
//          in the style of Shewchuk 1997, "Adaptive Precision Floating-Point
//          Arithmetic and Fast Robust Geometric Predicates");
// Stage 2: the same expression in float64 with its bound;
// Stage 3: exact — the determinant expanded over the original coordinates is
//          a sum of 24 products of three floats; each product is two doubles
//          exactly (the float*float product is exact in double, the third
//          factor goes through an FMA-based two-product), and the sign of the
//          48 doubles' exact sum is found by certified error-free vector sums
//          in registers, backed by a non-overlapping expansion with zero
//          elimination (Shewchuk's GROW-EXPANSION).
// Only stage 1 runs for almost every predicate; FP64 is confined to the
// predicates whose float32 value is inside its error bound.  Streaming
// kernels run the stages as separate passes (orient_at's levels): a test
// that stage 1 cannot decide is deferred to a float64 pass and from there
// to an exact pass, so no wavefront mixes the three kinds of arithmetic.
#pragma once
#pragma clang fp contract(off)

#include "../../include/mi355x_collide.h"
#include "geom.hpp"

namespace cdk {

struct PredStats {
    uint32_t fp64 = 0, exact = 0;
};

__device__ __forceinline__ void two_sum(double a, double b, double& x, double& y) {
    x = a + b;
    double bv = x - a;
    double av = x - bv;
    double br = b - bv;
    double ar = a - av;
    y = ar + br;
}

// The determinant over the original coordinates as 48 doubles whose exact
// sum it is: [a-c, b-c, d-c] = [a,b,d] - [a,b,c] - [a,c,d] + [b,c,d], each
// triple product six terms +-u_i*v_j*w_k, each term the double product
// u_i*v_j (exact: two floats) times w_k split into its rounded value and
// the FMA remainder.  Every term is a multiple of 2^-447 (three factors
// that are multiples of 2^-149) and below 2^384, so nothing underflows or
// overflows here or in any exact sum of them.
__device__ __forceinline__ void det_terms(f3 a, f3 b, f3 c, f3 d, double (&t)[48]) {
    int n = 0;
    auto triple = [&](f3 u, f3 v, f3 w, double sg) {
        // u . (v x w) = ux(vy wz - vz wy) + uy(vz wx - vx wz) + uz(vx wy - vy wx)
        const double ux = u.x, uy = u.y, uz = u.z, vx = v.x, vy = v.y, vz = v.z, wx = w.x, wy = w.y, wz = w.z;
        const double q[6][3] = {{ux * vy, wz, 1}, {ux * vz, wy, -1}, {uy * vz, wx, 1},
                                {uy * vx, wz, -1}, {uz * vx, wy, 1}, {uz * vy, wx, -1}};
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const double p = q[k][0] * q[k][1];
            const double err = fma(q[k][0], q[k][1], -p);
            const double s = sg * q[k][2];
            t[n++] = s * p;
            t[n++] = s * err;
        }
    };
    triple(a, b, d, 1.0);
    triple(a, b, c, -1.0);
    triple(a, c, d, -1.0);
    triple(b, c, d, 1.0);
}

// Sign of the exact sum by growing a non-overlapping expansion with zero
// elimination (Shewchuk's GROW-EXPANSION); its array lives in scratch
// memory, so it only backs up orient_exact.
__device__ __noinline__ int orient_exact_slow(f3 a, f3 b, f3 c, f3 d) {
    double t[48];
    det_terms(a, b, c, d, t);
    double e[48];
    int n = 0;
    for (int k = 0; k < 48; ++k) {
        double q = t[k];
        int m = 0;
        for (int i = 0; i < n; ++i) {
            double s, h;
            two_sum(q, e[i], s, h);
            q = s;
            if (h != 0.0) e[m++] = h;
        }
        if (q != 0.0) e[m++] = q;
        n = m;
    }
    for (int i = n - 1; i >= 0; --i)
        if (e[i] != 0.0) return e[i] > 0.0 ? 1 : -1;
    return 0;
}

// Exact sign, in registers: iterated error-free vector sums (VecSum,
// Ogita, Rump & Oishi 2005, "Accurate sum and dot product", SIAM J. Sci.
// Comput. 26(6), Algorithm 4.3).  A pass replaces p[i-1], p[i] by the sum
// and the rounding error of the pair, walking up the array, so the terms
// keep their exact sum and p[47] becomes its floating-point value.  Once
// |p[47]| exceeds the magnitude of the other 47 terms together, their sum
// cannot change its sign: the float sum r of their magnitudes is within
// gamma_46 < 5.2e-15 of the true one, so |p[47]| > r (1 + 1e-13) certifies
// it; r == 0 leaves p[47] as the exact sum.  Degenerate inputs (shared
// vertices, points on edges, 1-ulp offsets) certify within three passes;
// a case still open after eight takes the expansion.
__device__ __forceinline__ int orient_exact(f3 a, f3 b, f3 c, f3 d) {
    double p[48];
    det_terms(a, b, c, d, p);
#pragma unroll 1
    for (int pass = 0; pass < 8; ++pass) {
#pragma unroll
        for (int i = 1; i < 48; ++i) {
            double s, e;
            two_sum(p[i], p[i - 1], s, e);
            p[i] = s;
            p[i - 1] = e;
        }
        double r = 0.0;
#pragma unroll
        for (int i = 0; i < 47; ++i) r += fabs(p[i]);
        if (r == 0.0 || fabs(p[47]) > r * (1.0 + 1e-13)) return p[47] > 0.0 ? 1 : (p[47] < 0.0 ? -1 : 0);
    }
    return orient_exact_slow(a, b, c, d);
}

// Float64 stage: +1 / -1, or UNDECIDED when |det| is inside its bound.
constexpr int UNDECIDED = 2;
__device__ __forceinline__ int orient_f64(f3 a, f3 b, f3 c, f3 d) {
    const double ax = double(a.x) - c.x, ay = double(a.y) - c.y, az = double(a.z) - c.z;
    const double bx = double(b.x) - c.x, by = double(b.y) - c.y, bz = double(b.z) - c.z;
    const double dx = double(d.x) - c.x, dy = double(d.y) - c.y, dz = double(d.z) - c.z;
    const double p0 = ay * bz, p1 = by * az, p2 = az * bx, p3 = bz * ax, p4 = ax * by, p5 = bx * ay;
    const double det = (p0 - p1) * dx + (p2 - p3) * dy + (p4 - p5) * dz;
    const double perm = (fabs(p0) + fabs(p1)) * fabs(dx) + (fabs(p2) + fabs(p3)) * fabs(dy) +
                        (fabs(p4) + fabs(p5)) * fabs(dz);
    const double bound = 1.7763568394002505e-15 * perm;  // 16 * 2^-53
    if (det > bound) return 1;
    if (det < -bound) return -1;
    return UNDECIDED;
}

// Float32 stage: +1 / -1, or UNDECIDED.
__device__ __forceinline__ int orient_f32(f3 a, f3 b, f3 c, f3 d) {
    const float ax = a.x - c.x, ay = a.y - c.y, az = a.z - c.z;
    const float bx = b.x - c.x, by = b.y - c.y, bz = b.z - c.z;
    const float dx = d.x - c.x, dy = d.y - c.y, dz = d.z - c.z;
    const float p0 = ay * bz, p1 = by * az, p2 = az * bx, p3 = bz * ax, p4 = ax * by, p5 = bx * ay;
    const float det = (p0 - p1) * dx + (p2 - p3) * dy + (p4 - p5) * dz;
    const float perm = (fabsf(p0) + fabsf(p1)) * fabsf(dx) + (fabsf(p2) + fabsf(p3)) * fabsf(dy) +
                       (fabsf(p4) + fabsf(p5)) * fabsf(dz);
    const float bound = 9.5367431640625e-07f * perm + 1e-30f;  // 16 * 2^-24
    if (det > bound) return 1;
    if (det < -bound) return -1;
    return UNDECIDED;
}

// Stages 2 and 3 out of line, for kernels where the fallback is rare and
// the float32 path should stay small.  Returns (sign + 1) | 4 when the exact
// stage was needed.
__device__ __noinline__ int orient_slow(f3 a, f3 b, f3 c, f3 d) {
    const int r = orient_f64(a, b, c, d);
    if (r != UNDECIDED) return r + 1;
    return (orient_exact(a, b, c, d) + 1) | 4;
}

// Orientation at a precision level:
//   0  float32 only (UNDECIDED when inside the float32 bound),
//   1  float32, then float64 inline (UNDECIDED when inside both bounds),
//   2  float32, float64, exact: always decided.
// st.fp64 counts predicates that left the float32 stage, st.exact those that
// reached the exact stage.
template <int L>
__device__ __forceinline__ int orient_at(f3 a, f3 b, f3 c, f3 d, PredStats& st) {
    const int r = orient_f32(a, b, c, d);
    if (L == 0 || r != UNDECIDED) return r;
    ++st.fp64;
    if (L == 1) return orient_f64(a, b, c, d);
    const int s = orient_slow(a, b, c, d);
    if (s & 4) ++st.exact;
    return (s & 3) - 1;
}

__device__ __forceinline__ int orient_filtered(f3 a, f3 b, f3 c, f3 d, PredStats& st) {
    return orient_at<2>(a, b, c, d, st);
}

// Guigue & Devillers 2003 decision structure ("Fast and robust
// triangle-triangle overlap test using orientation predicates", JGT 8(1)),
// every sign exact.  Returns 1 = intersect (closed triangles, not coplanar),
// 0 = disjoint, 2 = coplanar.
//
// The published test is a tree of cases that permute the two triangles into
// a canonical form and then evaluate two final orientations.  Written as
// that tree, every case inlines its own copy of the predicate and the
// function no longer fits the instruction cache; here the same cases are
// reduced to two small permutations (a cyclic rotation of T1 plus an
// optional q/r flip, a q/r swap of T2 plus a cyclic rotation) and the final
// two predicates are evaluated once.
//
// dsides packs the exact sides (+1 each, 2 bits): U's vertices vs V's plane
// in bits 0-5, V's vertices vs U's plane in bits 6-11.
//
// At precision level L < 2 (see orient_at) a predicate can stay undecided;
// the test then stops and returns DEFERRED, for a rerun at a higher level.
__device__ __forceinline__ int side_of(uint32_t packed, int k) { return int((packed >> (2 * k)) & 3u) - 1; }
__device__ __forceinline__ f3 pick3(f3 a, f3 b, f3 c, int i) { return i == 0 ? a : (i == 1 ? b : c); }

constexpr int DEFERRED = -1;
#define CD_DECIDE(v) \
    if (L < 2 && (v) == UNDECIDED) return DEFERRED
//
// The test runs in three parts that streaming kernels can also run as
// separate batches (most pairs end in the first or second part):
//   tt_part1  T1's vertices against T2's plane: 0 (all on one side) or MORE;
//   tt_part2  T2's vertices against T1's plane: 0, 2 (coplanar) or MORE;
//   tt_part3  the canonical form and its two final orientations: 0 or 1.
// Each may also return DEFERRED at a level below 2.
// At level 0 the three orientations of a part are evaluated before one
// combined check: with a check after each, the compiler moves the loads of
// the later vertices behind it, and a test waits for memory three times.
constexpr int TT_MORE = 3;
template <int L>
__device__ __forceinline__ bool any_undecided(int a, int b, int c) {
    return L < 2 && ((a == UNDECIDED) | (b == UNDECIDED) | (c == UNDECIDED));
}
template <int L>
__device__ __forceinline__ int tt_part1(f3 p1, f3 q1, f3 r1, f3 p2, f3 q2, f3 r2, uint32_t& dsides, PredStats& st) {
    int dp1, dq1, dr1;
    if (L == 0) {
        dp1 = orient_at<L>(p2, q2, r2, p1, st);
        dq1 = orient_at<L>(p2, q2, r2, q1, st);
        dr1 = orient_at<L>(p2, q2, r2, r1, st);
        if (any_undecided<L>(dp1, dq1, dr1)) return DEFERRED;
    } else {
        dp1 = orient_at<L>(p2, q2, r2, p1, st);
        CD_DECIDE(dp1);
        dq1 = orient_at<L>(p2, q2, r2, q1, st);
        CD_DECIDE(dq1);
        dr1 = orient_at<L>(p2, q2, r2, r1, st);
        CD_DECIDE(dr1);
    }
    dsides = (uint32_t(dp1 + 1) << 6) | (uint32_t(dq1 + 1) << 8) | (uint32_t(dr1 + 1) << 10);
    if (dp1 * dq1 > 0 && dp1 * dr1 > 0) return 0;
    return TT_MORE;
}
template <int L>
__device__ __forceinline__ int tt_part2(f3 p1, f3 q1, f3 r1, f3 p2, f3 q2, f3 r2, uint32_t& dsides, PredStats& st) {
    const int dp1 = side_of(dsides, 3), dq1 = side_of(dsides, 4), dr1 = side_of(dsides, 5);
    int dp2, dq2, dr2;
    if (L == 0) {
        dp2 = orient_at<L>(q1, r1, p1, p2, st);
        dq2 = orient_at<L>(q1, r1, p1, q2, st);
        dr2 = orient_at<L>(q1, r1, p1, r2, st);
        if (any_undecided<L>(dp2, dq2, dr2)) return DEFERRED;
    } else {
        dp2 = orient_at<L>(q1, r1, p1, p2, st);
        CD_DECIDE(dp2);
        dq2 = orient_at<L>(q1, r1, p1, q2, st);
        CD_DECIDE(dq2);
        dr2 = orient_at<L>(q1, r1, p1, r2, st);
        CD_DECIDE(dr2);
    }
    dsides |= uint32_t(dp2 + 1) | (uint32_t(dq2 + 1) << 2) | (uint32_t(dr2 + 1) << 4);
    if (dp2 * dq2 > 0 && dp2 * dr2 > 0) return 0;
    if ((dp1 == 0 && dq1 == 0 && dr1 == 0) || (dp2 == 0 && dq2 == 0 && dr2 == 0)) return 2;
    return TT_MORE;
}
template <int L>
__device__ __forceinline__ int tt_part3(f3 p1, f3 q1, f3 r1, f3 p2, f3 q2, f3 r2, uint32_t dsides, PredStats& st) {
    const int dp1 = side_of(dsides, 3), dq1 = side_of(dsides, 4), dr1 = side_of(dsides, 5);
    const int dp2 = side_of(dsides, 0), dq2 = side_of(dsides, 1), dr2 = side_of(dsides, 2);
    // stage 1: rotate T1 so that its first vertex is alone on its side of
    // T2's plane (rot1 = index of that vertex); swap2 exchanges q2 and r2
    int rot1, swap2;
    if (dp1 > 0) {
        if (dq1 > 0) { rot1 = 2; swap2 = 1; }
        else if (dr1 > 0) { rot1 = 1; swap2 = 1; }
        else { rot1 = 0; swap2 = 0; }
    } else if (dp1 < 0) {
        if (dq1 < 0) { rot1 = 2; swap2 = 0; }
        else if (dr1 < 0) { rot1 = 1; swap2 = 0; }
        else { rot1 = 0; swap2 = 1; }
    } else {
        if (dq1 < 0) {
            if (dr1 >= 0) { rot1 = 1; swap2 = 1; } else { rot1 = 0; swap2 = 0; }
        } else if (dq1 > 0) {
            if (dr1 > 0) { rot1 = 0; swap2 = 1; } else { rot1 = 1; swap2 = 0; }
        } else {
            rot1 = 2;
            swap2 = dr1 > 0 ? 0 : 1;  // dr1 == 0 is the coplanar case, excluded above
        }
    }
    // T2 after the optional swap, with its signs
    const f3 s_q2 = swap2 ? r2 : q2, s_r2 = swap2 ? q2 : r2;
    const int e_p = dp2, e_q = swap2 ? dr2 : dq2, e_r = swap2 ? dq2 : dr2;
    // stage 2: rotate T2 the same way (rot2), flip1 exchanges T1's q and r
    int rot2, flip1;
    if (e_p > 0) {
        if (e_q > 0) { flip1 = 1; rot2 = 2; }
        else if (e_r > 0) { flip1 = 1; rot2 = 1; }
        else { flip1 = 0; rot2 = 0; }
    } else if (e_p < 0) {
        if (e_q < 0) { flip1 = 0; rot2 = 2; }
        else if (e_r < 0) { flip1 = 0; rot2 = 1; }
        else { flip1 = 1; rot2 = 0; }
    } else {
        if (e_q < 0) {
            if (e_r >= 0) { flip1 = 1; rot2 = 1; } else { flip1 = 0; rot2 = 0; }
        } else if (e_q > 0) {
            if (e_r > 0) { flip1 = 1; rot2 = 0; } else { flip1 = 0; rot2 = 1; }
        } else {
            if (e_r == 0) return 2;
            flip1 = e_r > 0 ? 0 : 1;
            rot2 = 2;
        }
    }
    const f3 P1 = pick3(p1, q1, r1, rot1);
    const f3 b1 = pick3(p1, q1, r1, rot1 == 2 ? 0 : rot1 + 1);
    const f3 c1 = pick3(p1, q1, r1, rot1 == 0 ? 2 : rot1 - 1);
    const f3 Q1 = flip1 ? c1 : b1, R1 = flip1 ? b1 : c1;
    const f3 P2 = pick3(p2, s_q2, s_r2, rot2);
    const f3 Q2 = pick3(p2, s_q2, s_r2, rot2 == 2 ? 0 : rot2 + 1);
    const f3 R2 = pick3(p2, s_q2, s_r2, rot2 == 0 ? 2 : rot2 - 1);
    // the two final orientations (the check of the canonical form)
    const int f1 = orient_at<L>(P2, P1, Q1, Q2, st);
    CD_DECIDE(f1);
    if (f1 > 0) return 0;
    const int f2 = orient_at<L>(P2, R1, P1, R2, st);
    CD_DECIDE(f2);
    if (f2 > 0) return 0;
    return 1;
}
#undef CD_DECIDE

template <int L = 2>
__device__ inline int tri_tri_exact(f3 p1, f3 q1, f3 r1, f3 p2, f3 q2, f3 r2, uint32_t& dsides, PredStats& st) {
    int r = tt_part1<L>(p1, q1, r1, p2, q2, r2, dsides, st);
    if (r != TT_MORE) return r;
    r = tt_part2<L>(p1, q1, r1, p2, q2, r2, dsides, st);
    if (r != TT_MORE) return r;
    return tt_part3<L>(p1, q1, r1, p2, q2, r2, dsides, st);
}

// Contact segment of an intersecting, non-coplanar pair, with Möller's
// interval construction (Triangle.cpp:764-1001) driven by the exact sides:
// a float distance whose sign disagrees with the exact side is replaced by a
// tiny value of the exact sign, an exact zero becomes 0.  When the float
// intervals fail to overlap (a touching pair), the segment collapses to the
// midpoint of the closest ends.
__device__ inline float signed_like(float v, int s) {
    if (s == 0) return 0.0f;
    if (s > 0) return v > 0.0f ? v : 1e-30f;
    return v < 0.0f ? v : -1e-30f;
}
__device__ inline void exact_segment(f3 V0, f3 V1, f3 V2, f3 U0, f3 U1, f3 U2, uint32_t ds, f3& s, f3& t) {
    const f3 N1 = cross(sub(V1, V0), sub(V2, V0));
    const float d1 = -dot(N1, V0);
    const float du0 = signed_like(dot(N1, U0) + d1, side_of(ds, 0));
    const float du1 = signed_like(dot(N1, U1) + d1, side_of(ds, 1));
    const float du2 = signed_like(dot(N1, U2) + d1, side_of(ds, 2));
    const f3 N2 = cross(sub(U1, U0), sub(U2, U0));
    const float d2 = -dot(N2, U0);
    const float dv0 = signed_like(dot(N2, V0) + d2, side_of(ds, 3));
    const float dv1 = signed_like(dot(N2, V1) + d2, side_of(ds, 4));
    const float dv2 = signed_like(dot(N2, V2) + d2, side_of(ds, 5));
    const f3 D = cross(N1, N2);
    float mx = fabsf(D.x);
    int index = 0;
    if (fabsf(D.y) > mx) { mx = fabsf(D.y); index = 1; }
    if (fabsf(D.z) > mx) { index = 2; }
    float i10, i11, i20, i21;
    f3 A1, A2, B1, B2;
    if (intervals(V0, V1, V2, comp(V0, index), comp(V1, index), comp(V2, index), dv0, dv1, dv2, dv0 * dv1, dv0 * dv2,
                  i10, i11, A1, A2) ||
        intervals(U0, U1, U2, comp(U0, index), comp(U1, index), comp(U2, index), du0, du1, du2, du0 * du1, du0 * du2,
                  i20, i21, B1, B2)) {
        s = t = V0;
        return;
    }
    int sm1 = 0, sm2 = 0;
    if (i10 > i11) { float q = i10; i10 = i11; i11 = q; sm1 = 1; }
    if (i20 > i21) { float q = i20; i20 = i21; i21 = q; sm2 = 1; }
    if (i11 < i20 || i21 < i10) {
        const f3 a = (i11 < i20) ? (sm1 == 0 ? A2 : A1) : (sm2 == 0 ? B2 : B1);
        const f3 b = (i11 < i20) ? (sm2 == 0 ? B1 : B2) : (sm1 == 0 ? A1 : A2);
        s = t = mk(0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z));
        return;
    }
    if (i20 < i10) {
        s = sm1 == 0 ? A1 : A2;
        t = (i21 < i11) ? (sm2 == 0 ? B2 : B1) : (sm1 == 0 ? A2 : A1);
    } else {
        s = sm2 == 0 ? B1 : B2;
        t = (i21 > i11) ? (sm1 == 0 ? A2 : A1) : (sm2 == 0 ? B2 : B1);
    }
}

// Contact record of a colliding pair, written in place.  Inline: as a call
// it would keep the caller's live registers across the call boundary.
__device__ __forceinline__ void write_contact(cd_contact* dst, f3 A0, f3 A1, f3 A2, f3 B0, f3 B1, f3 B2, uint32_t ds,
                                           uint32_t q, uint32_t ta, uint32_t tb) {
    f3 s, t;
    exact_segment(A0, A1, A2, B0, B1, B2, ds, s, t);
    const PlaneF pa = plane_of(A0, A1, A2), pb = plane_of(B0, B1, B2);
    cd_contact c;
    c.query = q;
    c.tri_a = ta;
    c.tri_b = tb;
    c.flags = outside_flags(pa, pb, A0, A1, A2, B0, B1, B2);
    c.s[0] = s.x; c.s[1] = s.y; c.s[2] = s.z;
    c.t[0] = t.x; c.t[1] = t.y; c.t[2] = t.z;
    c.na[0] = pa.n.x; c.na[1] = pa.n.y; c.na[2] = pa.n.z;
    c.nb[0] = pb.n.x; c.nb[1] = pb.n.y; c.nb[2] = pb.n.z;
    *dst = c;
}

}  // namespace cdk
