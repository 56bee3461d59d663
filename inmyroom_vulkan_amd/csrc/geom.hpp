// Device geometry for the collision path on gfx950.
//
// Every routine here reproduces one function of inMyRoom_vulkan's collision
// code with the same operation order, so that under -ffp-contract=off (and
// the pragma below) each float/double operation is rounded exactly once in
// the order written and the GPU reproduces the reference bits.
//
// Third-party math the reference takes from glm (fork thesmallcreeper/glm,
// public 0.9.9 semantics): dot = (x*x' + y*y') + z*z', cross, length =
// sqrt(dot), normalize = v * (1/sqrt(dot)), mat4*vec4 = (m0x + m1y) + (m2z +
// m3w), mat4*mat4 = ((A0b0 + A1b1) + A2b2) + A3b3, cofactor inverse.
#pragma once
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cdk {

struct f3 { float x, y, z; };
struct Box { f3 c, u, v, w; };  // centre + three half-extent side vectors (48 B)

#define CD_HD __host__ __device__ __forceinline__

CD_HD f3 mk(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
CD_HD f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
CD_HD f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
CD_HD float dot(f3 a, f3 b) {
    float px = a.x * b.x;
    float py = a.y * b.y;
    float pz = a.z * b.z;
    return (px + py) + pz;
}
CD_HD f3 cross(f3 a, f3 b) {
    return mk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
CD_HD float vlength(f3 a) { return sqrtf(dot(a, a)); }
CD_HD f3 normalize(f3 a) {
    float s = 1.0f / sqrtf(dot(a, a));
    return mk(a.x * s, a.y * s, a.z * s);
}

// Column-major 4x4 (glm layout): m[c*4 + r].
struct M4 { float m[16]; };

// xyz of M * (p, w) in glm's pairwise order (Paralgram.cpp:8-12, Triangle.cpp:73-75).
CD_HD f3 mul4(const M4& M, f3 p, float w) {
    float o[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        float a = M.m[0 + r] * p.x;
        float b = M.m[4 + r] * p.y;
        float c = M.m[8 + r] * p.z;
        float d = M.m[12 + r] * w;
        o[r] = (a + b) + (c + d);
    }
    return mk(o[0], o[1], o[2]);
}
CD_HD f3 xform_point(const M4& M, f3 p) { return mul4(M, p, 1.f); }

CD_HD M4 mat_mul(const M4& A, const M4& B) {
    M4 R;
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float t0 = A.m[0 * 4 + r] * B.m[c * 4 + 0];
            float t1 = A.m[1 * 4 + r] * B.m[c * 4 + 1];
            float t2 = A.m[2 * 4 + r] * B.m[c * 4 + 2];
            float t3 = A.m[3 * 4 + r] * B.m[c * 4 + 3];
            R.m[c * 4 + r] = ((t0 + t1) + t2) + t3;
        }
    return R;
}

// Node transforms (NodeDataCompEntity::GetGlobalMatrix, src/ECS/
// GeneralCompEntities/NodeDataCompEntity.cpp:63-72) with glm's builders
// applied to the identity exactly as written there, zero signs included:
// translate: column 3 = ((I0 v.x + I1 v.y) + I2 v.z) + I3; scale: column i =
// Ii v[i]; toMat4 of a quaternion (x, y, z, w) = mat3_cast's nine terms.
CD_HD M4 identity4() {
    M4 I;
#pragma unroll
    for (int i = 0; i < 16; ++i) I.m[i] = (i % 5 == 0) ? 1.f : 0.f;
    return I;
}
CD_HD M4 translate_identity(float x, float y, float z) {
    M4 R = identity4();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const float i0 = r == 0 ? 1.f : 0.f, i1 = r == 1 ? 1.f : 0.f, i2 = r == 2 ? 1.f : 0.f, i3 = r == 3 ? 1.f : 0.f;
        R.m[12 + r] = ((i0 * x + i1 * y) + i2 * z) + i3;
    }
    return R;
}
CD_HD M4 scale_identity(float x, float y, float z) {
    M4 R = identity4();
    const float v[3] = {x, y, z};
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r) R.m[c * 4 + r] = (r == c ? 1.f : 0.f) * v[c];
    return R;
}
CD_HD M4 quat_to_mat4(float qx, float qy, float qz, float qw) {
    const float qxx = qx * qx, qyy = qy * qy, qzz = qz * qz;
    const float qxz = qx * qz, qxy = qx * qy, qyz = qy * qz;
    const float qwx = qw * qx, qwy = qw * qy, qwz = qw * qz;
    M4 R = identity4();
    R.m[0] = 1.f - 2.f * (qyy + qzz);
    R.m[1] = 2.f * (qxy + qwz);
    R.m[2] = 2.f * (qxz - qwy);
    R.m[4] = 2.f * (qxy - qwz);
    R.m[5] = 1.f - 2.f * (qxx + qzz);
    R.m[6] = 2.f * (qyz + qwx);
    R.m[8] = 2.f * (qxz + qwy);
    R.m[9] = 2.f * (qyz - qwx);
    R.m[10] = 1.f - 2.f * (qxx + qyy);
    return R;
}
// gT * parent * T * R * S, multiplied left to right as the reference writes it.
// n: scale[3], rotation (x, y, z, w), translation[3], global translation[3].
CD_HD M4 node_global(const float* n, const M4& parent) {
    const M4 S = scale_identity(n[0], n[1], n[2]);
    const M4 R = quat_to_mat4(n[3], n[4], n[5], n[6]);
    const M4 T = translate_identity(n[7], n[8], n[9]);
    const M4 gT = translate_identity(n[10], n[11], n[12]);
    return mat_mul(mat_mul(mat_mul(mat_mul(gT, parent), T), R), S);
}

// glm 0.9.9 compute_inverse<4,4> (used by OBBtreesCollision.cpp:15).
CD_HD M4 mat_inverse(const M4& M) {
#define MM(c, r) M.m[(c) * 4 + (r)]
    float c00 = MM(2, 2) * MM(3, 3) - MM(3, 2) * MM(2, 3);
    float c02 = MM(1, 2) * MM(3, 3) - MM(3, 2) * MM(1, 3);
    float c03 = MM(1, 2) * MM(2, 3) - MM(2, 2) * MM(1, 3);
    float c04 = MM(2, 1) * MM(3, 3) - MM(3, 1) * MM(2, 3);
    float c06 = MM(1, 1) * MM(3, 3) - MM(3, 1) * MM(1, 3);
    float c07 = MM(1, 1) * MM(2, 3) - MM(2, 1) * MM(1, 3);
    float c08 = MM(2, 1) * MM(3, 2) - MM(3, 1) * MM(2, 2);
    float c10 = MM(1, 1) * MM(3, 2) - MM(3, 1) * MM(1, 2);
    float c11 = MM(1, 1) * MM(2, 2) - MM(2, 1) * MM(1, 2);
    float c12 = MM(2, 0) * MM(3, 3) - MM(3, 0) * MM(2, 3);
    float c14 = MM(1, 0) * MM(3, 3) - MM(3, 0) * MM(1, 3);
    float c15 = MM(1, 0) * MM(2, 3) - MM(2, 0) * MM(1, 3);
    float c16 = MM(2, 0) * MM(3, 2) - MM(3, 0) * MM(2, 2);
    float c18 = MM(1, 0) * MM(3, 2) - MM(3, 0) * MM(1, 2);
    float c19 = MM(1, 0) * MM(2, 2) - MM(2, 0) * MM(1, 2);
    float c20 = MM(2, 0) * MM(3, 1) - MM(3, 0) * MM(2, 1);
    float c22 = MM(1, 0) * MM(3, 1) - MM(3, 0) * MM(1, 1);
    float c23 = MM(1, 0) * MM(2, 1) - MM(2, 0) * MM(1, 1);
    const float f0[4] = {c00, c00, c02, c03}, f1[4] = {c04, c04, c06, c07};
    const float f2[4] = {c08, c08, c10, c11}, f3v[4] = {c12, c12, c14, c15};
    const float f4[4] = {c16, c16, c18, c19}, f5[4] = {c20, c20, c22, c23};
    const float g0[4] = {MM(1, 0), MM(0, 0), MM(0, 0), MM(0, 0)};
    const float g1[4] = {MM(1, 1), MM(0, 1), MM(0, 1), MM(0, 1)};
    const float g2[4] = {MM(1, 2), MM(0, 2), MM(0, 2), MM(0, 2)};
    const float g3[4] = {MM(1, 3), MM(0, 3), MM(0, 3), MM(0, 3)};
    const float sA[4] = {+1.f, -1.f, +1.f, -1.f}, sB[4] = {-1.f, +1.f, -1.f, +1.f};
    float inv[4][4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float i0 = (g1[k] * f0[k] - g2[k] * f1[k]) + g3[k] * f2[k];
        float i1 = (g0[k] * f0[k] - g2[k] * f3v[k]) + g3[k] * f4[k];
        float i2 = (g0[k] * f1[k] - g1[k] * f3v[k]) + g3[k] * f5[k];
        float i3 = (g0[k] * f2[k] - g1[k] * f4[k]) + g2[k] * f5[k];
        inv[0][k] = i0 * sA[k];
        inv[1][k] = i1 * sB[k];
        inv[2][k] = i2 * sA[k];
        inv[3][k] = i3 * sB[k];
    }
    float d0 = MM(0, 0) * inv[0][0], d1 = MM(0, 1) * inv[1][0];
    float d2 = MM(0, 2) * inv[2][0], d3 = MM(0, 3) * inv[3][0];
    float det = (d0 + d1) + (d2 + d3);
    float one_over = 1.0f / det;
    M4 R;
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r) R.m[c * 4 + r] = inv[c][r] * one_over;
    return R;
#undef MM
}

// ---------------------------------------------------------------------------
// Paralgram / OBB (inMyRoom_vulkan/src/Geometry/Paralgram.cpp)
// ---------------------------------------------------------------------------
CD_HD Box xform_box(const M4& M, const Box& b) {  // :4-15
    Box r;
    r.c = mul4(M, b.c, 1.f);
    r.u = mul4(M, b.u, 0.f);
    r.v = mul4(M, b.v, 0.f);
    r.w = mul4(M, b.w, 0.f);
    return r;
}
CD_HD void project(const Box& b, f3 a, float& lo, float& hi) {  // :175-195
    float c = dot(b.c, a);
    float su = fabsf(dot(a, b.u)), sv = fabsf(dot(a, b.v)), sw = fabsf(dot(a, b.w));
    float s = (su + sv) + sw;
    lo = c - s;
    hi = c + s;
}
CD_HD bool axis_overlaps(const Box& l, const Box& r, f3 a) {  // :197-201
    float l0, l1, r0, r1;
    project(l, a, l0, l1);
    project(r, a, r0, r1);
    return (l1 >= r0) && (r1 >= l0);
}
// :17-173 — 15 unnormalised axes in a fixed order with early out.
CD_HD bool sat(const Box& l, const Box& r) {
    if (!axis_overlaps(l, r, cross(l.v, l.w))) return false;
    if (!axis_overlaps(l, r, cross(l.u, l.w))) return false;
    if (!axis_overlaps(l, r, cross(l.u, l.v))) return false;
    if (!axis_overlaps(l, r, cross(r.v, r.w))) return false;
    if (!axis_overlaps(l, r, cross(r.u, r.w))) return false;
    if (!axis_overlaps(l, r, cross(r.u, r.v))) return false;
    if (!axis_overlaps(l, r, cross(l.u, r.u))) return false;
    if (!axis_overlaps(l, r, cross(l.u, r.v))) return false;
    if (!axis_overlaps(l, r, cross(l.u, r.w))) return false;
    if (!axis_overlaps(l, r, cross(l.v, r.u))) return false;
    if (!axis_overlaps(l, r, cross(l.v, r.v))) return false;
    if (!axis_overlaps(l, r, cross(l.v, r.w))) return false;
    if (!axis_overlaps(l, r, cross(l.w, r.u))) return false;
    if (!axis_overlaps(l, r, cross(l.w, r.v))) return false;
    if (!axis_overlaps(l, r, cross(l.w, r.w))) return false;
    return true;
}
CD_HD float surface(const Box& b) {  // :203-210
    float uv = vlength(cross(b.u, b.v));
    float uw = vlength(cross(b.u, b.w));
    float vw = vlength(cross(b.v, b.w));
    return 2.f * ((uv + uw) + vw);
}

// ---------------------------------------------------------------------------
// Möller 1997 triangle-triangle test with intersection line
// (inMyRoom_vulkan/src/Geometry/Triangle.cpp:323-1003)
// ---------------------------------------------------------------------------
CD_HD float comp(f3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// :402-420
CD_HD bool edge_edge(float Ax, float Ay, f3 V0, f3 U0, f3 U1, int i0, int i1) {
    float Bx = comp(U0, i0) - comp(U1, i0);
    float By = comp(U0, i1) - comp(U1, i1);
    float Cx = comp(V0, i0) - comp(U0, i0);
    float Cy = comp(V0, i1) - comp(U0, i1);
    float f = Ay * Bx - Ax * By;
    float d = By * Cx - Bx * Cy;
    if ((f > 0 && d >= 0 && d <= f) || (f < 0 && d <= 0 && d >= f)) {
        float e = Ax * Cy - Ay * Cx;
        if (f > 0) {
            if (e >= 0 && e <= f) return true;
        } else {
            if (e <= 0 && e >= f) return true;
        }
    }
    return false;
}
// :422-433
CD_HD bool edge_vs_tri(f3 V0, f3 V1, f3 U0, f3 U1, f3 U2, int i0, int i1) {
    float Ax = comp(V1, i0) - comp(V0, i0);
    float Ay = comp(V1, i1) - comp(V0, i1);
    return edge_edge(Ax, Ay, V0, U0, U1, i0, i1) || edge_edge(Ax, Ay, V0, U1, U2, i0, i1) ||
           edge_edge(Ax, Ay, V0, U2, U0, i0, i1);
}
// :435-458
CD_HD bool point_in_tri(f3 V0, f3 U0, f3 U1, f3 U2, int i0, int i1) {
    float a = comp(U1, i1) - comp(U0, i1);
    float b = -(comp(U1, i0) - comp(U0, i0));
    float c = -a * comp(U0, i0) - b * comp(U0, i1);
    float d0 = a * comp(V0, i0) + b * comp(V0, i1) + c;
    a = comp(U2, i1) - comp(U1, i1);
    b = -(comp(U2, i0) - comp(U1, i0));
    c = -a * comp(U1, i0) - b * comp(U1, i1);
    float d1 = a * comp(V0, i0) + b * comp(V0, i1) + c;
    a = comp(U0, i1) - comp(U2, i1);
    b = -(comp(U0, i0) - comp(U2, i0));
    c = -a * comp(U2, i0) - b * comp(U2, i1);
    float d2 = a * comp(V0, i0) + b * comp(V0, i1) + c;
    return (d0 * d1 > 0.0f) && (d0 * d2 > 0.0f);
}
// :460-507
CD_HD int coplanar_test(f3 N, f3 V0, f3 V1, f3 V2, f3 U0, f3 U1, f3 U2) {
    float A0 = fabsf(N.x), A1 = fabsf(N.y), A2 = fabsf(N.z);
    int i0, i1;
    if (A0 > A1) {
        if (A0 > A2) { i0 = 1; i1 = 2; } else { i0 = 0; i1 = 1; }
    } else {
        if (A2 > A1) { i0 = 0; i1 = 1; } else { i0 = 0; i1 = 2; }
    }
    if (edge_vs_tri(V0, V1, U0, U1, U2, i0, i1)) return 1;
    if (edge_vs_tri(V1, V2, U0, U1, U2, i0, i1)) return 1;
    if (edge_vs_tri(V2, V0, U0, U1, U2, i0, i1)) return 1;
    if (point_in_tri(V0, U0, U1, U2, i0, i1)) return 1;
    if (point_in_tri(U0, V0, V1, V2, i0, i1)) return 1;
    return 0;
}
// :764-778
CD_HD void interval_ends(f3 X0, f3 X1, f3 X2, float W0, float W1, float W2, float D0, float D1, float D2,
                         float& is0, float& is1, f3& P0, f3& P1) {
    float tmp = D0 / (D0 - D1);
    is0 = W0 + (W1 - W0) * tmp;
    f3 df = sub(X1, X0);
    df = mk(tmp * df.x, tmp * df.y, tmp * df.z);
    P0 = add(df, X0);
    tmp = D0 / (D0 - D2);
    is1 = W0 + (W2 - W0) * tmp;
    df = sub(X2, X0);
    df = mk(tmp * df.x, tmp * df.y, tmp * df.z);
    P1 = add(X0, df);
}
// :795-830 — true when all three signed distances are zero.
CD_HD bool intervals(f3 X0, f3 X1, f3 X2, float W0, float W1, float W2, float D0, float D1, float D2, float D0D1,
                     float D0D2, float& is0, float& is1, f3& P0, f3& P1) {
    if (D0D1 > 0.0f) interval_ends(X2, X0, X1, W2, W0, W1, D2, D0, D1, is0, is1, P0, P1);
    else if (D0D2 > 0.0f) interval_ends(X1, X0, X2, W1, W0, W2, D1, D0, D2, is0, is1, P0, P1);
    else if (D1 * D2 > 0.0f || D0 != 0.0f) interval_ends(X0, X1, X2, W0, W1, W2, D0, D1, D2, is0, is1, P0, P1);
    else if (D1 != 0.0f) interval_ends(X1, X0, X2, W1, W0, W2, D1, D0, D2, is0, is1, P0, P1);
    else if (D2 != 0.0f) interval_ends(X2, X0, X1, W2, W0, W1, D2, D0, D1, is0, is1, P0, P1);
    else return true;
    return false;
}

// |x| < 1e-6 compared in double (Triangle.cpp:332, :899-901, :922-924).
CD_HD float snap(float x) { return ((double)fabsf(x) < 0.000001) ? 0.0f : x; }

// Result bits of the triangle test.
enum : uint32_t { TT_HIT = 1u, TT_COPLANAR = 2u, TT_UB = 4u };

// The plane stage of the test (:870-904): each triangle's snapped signed
// distances to the other's plane.  False at the test's first two exits (one
// triangle wholly on one side of the other's plane); those pairs do not
// intersect, and the rest go on to the interval stage.
struct TTPlanes {
    f3 N1, N2;
    float du0, du1, du2, du0du1, du0du2;
    float dv0, dv1, dv2, dv0dv1, dv0dv2;
};
CD_HD bool tri_tri_planes(f3 V0, f3 V1, f3 V2, f3 U0, f3 U1, f3 U2, TTPlanes& p) {
    p.N1 = cross(sub(V1, V0), sub(V2, V0));
    float d1 = -dot(p.N1, V0);
    p.du0 = snap(dot(p.N1, U0) + d1);
    p.du1 = snap(dot(p.N1, U1) + d1);
    p.du2 = snap(dot(p.N1, U2) + d1);
    p.du0du1 = p.du0 * p.du1;
    p.du0du2 = p.du0 * p.du2;
    if (p.du0du1 > 0.0f && p.du0du2 > 0.0f) return false;
    p.N2 = cross(sub(U1, U0), sub(U2, U0));
    float d2 = -dot(p.N2, U0);
    p.dv0 = snap(dot(p.N2, V0) + d2);
    p.dv1 = snap(dot(p.N2, V1) + d2);
    p.dv2 = snap(dot(p.N2, V2) + d2);
    p.dv0dv1 = p.dv0 * p.dv1;
    p.dv0dv2 = p.dv0 * p.dv2;
    return !(p.dv0dv1 > 0.0f && p.dv0dv2 > 0.0f);
}

// :866-1003.  The case the reference leaves undefined (U's distances all
// snap to zero while V's do not, :960-964) is defined as coplanar and
// decided by the 2D coplanar test; TT_UB flags it.
CD_HD uint32_t tri_tri_isect(f3 V0, f3 V1, f3 V2, f3 U0, f3 U1, f3 U2, f3& s, f3& t) {
    TTPlanes p;
    if (!tri_tri_planes(V0, V1, V2, U0, U1, U2, p)) return 0u;
    const f3 N1 = p.N1, N2 = p.N2;
    const float du0 = p.du0, du1 = p.du1, du2 = p.du2, du0du1 = p.du0du1, du0du2 = p.du0du2;
    const float dv0 = p.dv0, dv1 = p.dv1, dv2 = p.dv2, dv0dv1 = p.dv0dv1, dv0dv2 = p.dv0dv2;
    f3 D = cross(N1, N2);
    float mx = fabsf(D.x);
    int index = 0;
    float b = fabsf(D.y), c = fabsf(D.z);
    if (b > mx) { mx = b; index = 1; }
    if (c > mx) { mx = c; index = 2; }
    float vp0 = comp(V0, index), vp1 = comp(V1, index), vp2 = comp(V2, index);
    float up0 = comp(U0, index), up1 = comp(U1, index), up2 = comp(U2, index);
    float i10, i11, i20, i21;
    f3 A1, A2, B1, B2;
    if (intervals(V0, V1, V2, vp0, vp1, vp2, dv0, dv1, dv2, dv0dv1, dv0dv2, i10, i11, A1, A2))
        return TT_COPLANAR | (coplanar_test(N1, V0, V1, V2, U0, U1, U2) ? TT_HIT : 0u);
    if (intervals(U0, U1, U2, up0, up1, up2, du0, du1, du2, du0du1, du0du2, i20, i21, B1, B2))
        return TT_COPLANAR | TT_UB | (coplanar_test(N1, V0, V1, V2, U0, U1, U2) ? TT_HIT : 0u);
    int sm1 = 0, sm2 = 0;
    if (i10 > i11) { float q = i10; i10 = i11; i11 = q; sm1 = 1; }
    if (i20 > i21) { float q = i20; i20 = i21; i21 = q; sm2 = 1; }
    if (i11 < i20 || i21 < i10) return 0u;
    if (i20 < i10) {
        s = sm1 == 0 ? A1 : A2;
        if (i21 < i11) t = sm2 == 0 ? B2 : B1;
        else t = sm1 == 0 ? A2 : A1;
    } else {
        s = sm2 == 0 ? B1 : B2;
        if (i21 > i11) t = sm1 == 0 ? A2 : A1;
        else t = sm2 == 0 ? B2 : B1;
    }
    return TT_HIT;
}

// Plane of a triangle (Plane.cpp:5-20, Triangle.cpp:140-146) and the
// outside test (Plane.cpp:22-28).
struct PlaneF { f3 n; float d; };
CD_HD PlaneF plane_of(f3 P0, f3 P1, f3 P2) {
    f3 n0 = normalize(cross(sub(P1, P0), sub(P2, P0)));
    float d0 = -dot(P0, n0);
    float len = vlength(n0);
    PlaneF r;
    r.n = mk(n0.x / len, n0.y / len, n0.z / len);
    r.d = d0 / len;
    return r;
}
CD_HD bool outside(const PlaneF& pl, f3 p) { return (dot(p, pl.n) + pl.d) > 0; }

// Contact flags (CreateUncollideRays.cpp:102-112): bit k = A's vertex k
// outside B's plane, bit 3+k = B's vertex k outside A's plane.
CD_HD uint32_t outside_flags(const PlaneF& pa, const PlaneF& pb, f3 A0, f3 A1, f3 A2, f3 B0, f3 B1, f3 B2) {
    uint32_t f = 0;
    if (outside(pb, A0)) f |= 1u;
    if (outside(pb, A1)) f |= 2u;
    if (outside(pb, A2)) f |= 4u;
    if (outside(pa, B0)) f |= 8u;
    if (outside(pa, B1)) f |= 16u;
    if (outside(pa, B2)) f |= 32u;
    return f;
}

}  // namespace cdk
