// TEST INFRASTRUCTURE ONLY — the oracle (see oracle.hpp for the rules).
//
// Every function below names the reference file:line whose behaviour it
// restates.  Paths are relative to the reference repository root.
#include "oracle.hpp"

#include <algorithm>
#include <functional>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <utility>

namespace orc {

// ---------------------------------------------------------------------------
// glm 0.9.9 formulas (third-party; see header comment)
// ---------------------------------------------------------------------------
static inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline float dot(V3 a, V3 b) {
    float px = a.x * b.x, py = a.y * b.y, pz = a.z * b.z;
    return (px + py) + pz;
}
static inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
static inline float length(V3 a) { return std::sqrt(dot(a, a)); }
static inline V3 normalize(V3 a) {
    float s = 1.0f / std::sqrt(dot(a, a));
    return {a.x * s, a.y * s, a.z * s};
}
// mat4 * vec4, glm's pairwise association; returns xyz.
static inline V3 mul4(const M4& M, V3 p, float w) {
    float o[3];
    for (int r = 0; r < 3; ++r) {
        float a = M.m[0 + r] * p.x;
        float b = M.m[4 + r] * p.y;
        float c = M.m[8 + r] * p.z;
        float d = M.m[12 + r] * w;
        o[r] = (a + b) + (c + d);
    }
    return {o[0], o[1], o[2]};
}
static M4 mat_mul(const M4& A, const M4& B) {
    M4 R;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            float t0 = A.m[0 * 4 + r] * B.m[c * 4 + 0];
            float t1 = A.m[1 * 4 + r] * B.m[c * 4 + 1];
            float t2 = A.m[2 * 4 + r] * B.m[c * 4 + 2];
            float t3 = A.m[3 * 4 + r] * B.m[c * 4 + 3];
            R.m[c * 4 + r] = ((t0 + t1) + t2) + t3;
        }
    return R;
}
// glm compute_inverse<4,4>: 18 2x2 cofactors, signed 3x3 minors, 1/det.
static M4 mat_inverse(const M4& M) {
    auto m = [&](int c, int r) { return M.m[c * 4 + r]; };
    float c00 = m(2, 2) * m(3, 3) - m(3, 2) * m(2, 3);
    float c02 = m(1, 2) * m(3, 3) - m(3, 2) * m(1, 3);
    float c03 = m(1, 2) * m(2, 3) - m(2, 2) * m(1, 3);
    float c04 = m(2, 1) * m(3, 3) - m(3, 1) * m(2, 3);
    float c06 = m(1, 1) * m(3, 3) - m(3, 1) * m(1, 3);
    float c07 = m(1, 1) * m(2, 3) - m(2, 1) * m(1, 3);
    float c08 = m(2, 1) * m(3, 2) - m(3, 1) * m(2, 2);
    float c10 = m(1, 1) * m(3, 2) - m(3, 1) * m(1, 2);
    float c11 = m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2);
    float c12 = m(2, 0) * m(3, 3) - m(3, 0) * m(2, 3);
    float c14 = m(1, 0) * m(3, 3) - m(3, 0) * m(1, 3);
    float c15 = m(1, 0) * m(2, 3) - m(2, 0) * m(1, 3);
    float c16 = m(2, 0) * m(3, 2) - m(3, 0) * m(2, 2);
    float c18 = m(1, 0) * m(3, 2) - m(3, 0) * m(1, 2);
    float c19 = m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2);
    float c20 = m(2, 0) * m(3, 1) - m(3, 0) * m(2, 1);
    float c22 = m(1, 0) * m(3, 1) - m(3, 0) * m(1, 1);
    float c23 = m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1);
    const float f0[4] = {c00, c00, c02, c03}, f1[4] = {c04, c04, c06, c07};
    const float f2[4] = {c08, c08, c10, c11}, f3[4] = {c12, c12, c14, c15};
    const float f4[4] = {c16, c16, c18, c19}, f5[4] = {c20, c20, c22, c23};
    const float g0[4] = {m(1, 0), m(0, 0), m(0, 0), m(0, 0)};
    const float g1[4] = {m(1, 1), m(0, 1), m(0, 1), m(0, 1)};
    const float g2[4] = {m(1, 2), m(0, 2), m(0, 2), m(0, 2)};
    const float g3[4] = {m(1, 3), m(0, 3), m(0, 3), m(0, 3)};
    const float sA[4] = {+1.f, -1.f, +1.f, -1.f}, sB[4] = {-1.f, +1.f, -1.f, +1.f};
    float inv[4][4];
    for (int k = 0; k < 4; ++k) {
        float i0 = (g1[k] * f0[k] - g2[k] * f1[k]) + g3[k] * f2[k];
        float i1 = (g0[k] * f0[k] - g2[k] * f3[k]) + g3[k] * f4[k];
        float i2 = (g0[k] * f1[k] - g1[k] * f3[k]) + g3[k] * f5[k];
        float i3 = (g0[k] * f2[k] - g1[k] * f4[k]) + g2[k] * f5[k];
        inv[0][k] = i0 * sA[k];
        inv[1][k] = i1 * sB[k];
        inv[2][k] = i2 * sA[k];
        inv[3][k] = i3 * sB[k];
    }
    float d0 = m(0, 0) * inv[0][0], d1 = m(0, 1) * inv[1][0];
    float d2 = m(0, 2) * inv[2][0], d3 = m(0, 3) * inv[3][0];
    float det = (d0 + d1) + (d2 + d3);
    float one_over = 1.0f / det;
    M4 R;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) R.m[c * 4 + r] = inv[c][r] * one_over;
    return R;
}

// ---------------------------------------------------------------------------
// eig3: Householder tridiagonalisation + implicit QL (eig3/eig3.cpp:21-265).
// The operation sequence is kept identical to the reference so that every
// rounding matches; the algorithm is JAMA/EISPACK tred2 + tql2.
// ---------------------------------------------------------------------------
static void tridiagonalize(double V[3][3], double d[3], double e[3]) {
    const int n = 3;
    for (int j = 0; j < n; ++j) d[j] = V[n - 1][j];
    for (int i = n - 1; i > 0; --i) {
        double scale = 0.0, h = 0.0;
        for (int k = 0; k < i; ++k) scale = scale + std::fabs(d[k]);
        if (scale == 0.0) {
            e[i] = d[i - 1];
            for (int j = 0; j < i; ++j) {
                d[j] = V[i - 1][j];
                V[i][j] = 0.0;
                V[j][i] = 0.0;
            }
        } else {
            for (int k = 0; k < i; ++k) {
                d[k] /= scale;
                h += d[k] * d[k];
            }
            double f = d[i - 1];
            double g = std::sqrt(h);
            if (f > 0) g = -g;
            e[i] = scale * g;
            h = h - f * g;
            d[i - 1] = f - g;
            for (int j = 0; j < i; ++j) e[j] = 0.0;
            for (int j = 0; j < i; ++j) {
                f = d[j];
                V[j][i] = f;
                g = e[j] + V[j][j] * f;
                for (int k = j + 1; k <= i - 1; ++k) {
                    g += V[k][j] * d[k];
                    e[k] += V[k][j] * f;
                }
                e[j] = g;
            }
            f = 0.0;
            for (int j = 0; j < i; ++j) {
                e[j] /= h;
                f += e[j] * d[j];
            }
            double hh = f / (h + h);
            for (int j = 0; j < i; ++j) e[j] -= hh * d[j];
            for (int j = 0; j < i; ++j) {
                f = d[j];
                g = e[j];
                for (int k = j; k <= i - 1; ++k) V[k][j] -= (f * e[k] + g * d[k]);
                d[j] = V[i - 1][j];
                V[i][j] = 0.0;
            }
        }
        d[i] = h;
    }
    for (int i = 0; i < n - 1; ++i) {
        V[n - 1][i] = V[i][i];
        V[i][i] = 1.0;
        double h = d[i + 1];
        if (h != 0.0) {
            for (int k = 0; k <= i; ++k) d[k] = V[k][i + 1] / h;
            for (int j = 0; j <= i; ++j) {
                double g = 0.0;
                for (int k = 0; k <= i; ++k) g += V[k][i + 1] * V[k][j];
                for (int k = 0; k <= i; ++k) V[k][j] -= g * d[k];
            }
        }
        for (int k = 0; k <= i; ++k) V[k][i + 1] = 0.0;
    }
    for (int j = 0; j < n; ++j) {
        d[j] = V[n - 1][j];
        V[n - 1][j] = 0.0;
    }
    V[n - 1][n - 1] = 1.0;
    e[0] = 0.0;
}

static inline double hyp(double x, double y) { return std::sqrt(x * x + y * y); }

static void ql_implicit(double V[3][3], double d[3], double e[3]) {
    const int n = 3;
    for (int i = 1; i < n; ++i) e[i - 1] = e[i];
    e[n - 1] = 0.0;
    double f = 0.0, tst1 = 0.0;
    const double eps = std::ldexp(1.0, -52);
    for (int l = 0; l < n; ++l) {
        double cand = std::fabs(d[l]) + std::fabs(e[l]);
        tst1 = (tst1 > cand) ? tst1 : cand;
        int m = l;
        while (m < n) {
            if (std::fabs(e[m]) <= eps * tst1) break;
            ++m;
        }
        if (m > l) {
            do {
                double g = d[l];
                double p = (d[l + 1] - g) / (2.0 * e[l]);
                double r = hyp(p, 1.0);
                if (p < 0) r = -r;
                d[l] = e[l] / (p + r);
                d[l + 1] = e[l] * (p + r);
                double dl1 = d[l + 1];
                double h = g - d[l];
                for (int i = l + 2; i < n; ++i) d[i] -= h;
                f = f + h;
                p = d[m];
                double c = 1.0, c2 = c, c3 = c;
                double el1 = e[l + 1];
                double s = 0.0, s2 = 0.0;
                for (int i = m - 1; i >= l; --i) {
                    c3 = c2;
                    c2 = c;
                    s2 = s;
                    g = c * e[i];
                    h = c * p;
                    r = hyp(p, e[i]);
                    e[i + 1] = s * r;
                    s = e[i] / r;
                    c = p / r;
                    p = c * d[i] - s * g;
                    d[i + 1] = h + s * (c * g + s * d[i]);
                    for (int k = 0; k < n; ++k) {
                        h = V[k][i + 1];
                        V[k][i + 1] = s * V[k][i] + c * h;
                        V[k][i] = c * V[k][i] - s * h;
                    }
                }
                p = -s * s2 * c3 * el1 * e[l] / dl1;
                e[l] = s * p;
                d[l] = c * p;
            } while (std::fabs(e[l]) > eps * tst1);
        }
        d[l] = d[l] + f;
        e[l] = 0.0;
    }
    for (int i = 0; i < n - 1; ++i) {
        int k = i;
        double p = d[i];
        for (int j = i + 1; j < n; ++j)
            if (d[j] < p) {
                k = j;
                p = d[j];
            }
        if (k != i) {
            d[k] = d[i];
            d[i] = p;
            for (int j = 0; j < n; ++j) {
                p = V[j][i];
                V[j][i] = V[j][k];
                V[j][k] = p;
            }
        }
    }
}

void eig3(const double A[3][3], double V[3][3], double d[3]) {
    double e[3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) V[i][j] = A[i][j];
    tridiagonalize(V, d, e);
    ql_implicit(V, d, e);
}

// ---------------------------------------------------------------------------
// OBB construction (inMyRoom_vulkan/src/Geometry/OBB.cpp:33-178)
// ---------------------------------------------------------------------------
static Box empty_box() {  // OBB.cpp:168-178
    const float e = std::numeric_limits<float>::epsilon();
    return {{0.f, 0.f, 0.f}, {e, 0.f, 0.f}, {0.f, e, 0.f}, {0.f, 0.f, e}};
}

static inline bool same_point(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// Distinct-point count capped at 4 (OBB.cpp:35-41): a std::unordered_set<vec3>
// keyed by float ==, so +0/-0 coincide and a point holding NaN is never equal
// to anything.
static int distinct_upto4(const V3* p, size_t n) {
    V3 seen[4];
    int k = 0;
    for (size_t i = 0; i < n; ++i) {
        bool dup = false;
        for (int j = 0; j < k; ++j)
            if (same_point(seen[j], p[i])) { dup = true; break; }
        if (!dup) {
            seen[k++] = p[i];
            if (k > 3) break;
        }
    }
    return k;
}

// OBB.cpp:105-166: extents along three double axes.  std::min(x, m) keeps m
// when m < x; std::max(x, m) keeps m unless m < x.
static Box box_along_axes(const V3* p, size_t n, const D3 ax[3]) {
    Box out;
    D3 ctr = {0.0, 0.0, 0.0};
    V3* sides[3] = {&out.u, &out.v, &out.w};
    for (int a = 0; a < 3; ++a) {
        double lo = std::numeric_limits<double>::infinity();
        double hi = -std::numeric_limits<double>::infinity();
        for (size_t i = 0; i < n; ++i) {
            double q = (ax[a].x * double(p[i].x) + ax[a].y * double(p[i].y)) + ax[a].z * double(p[i].z);
            lo = (lo < q) ? lo : q;
            hi = (hi < q) ? q : hi;
        }
        double delta = (hi - lo) + 2.0 * double(std::numeric_limits<float>::epsilon());
        double mid = (hi + lo) / 2.0;
        ctr.x = ctr.x + mid * ax[a].x;
        ctr.y = ctr.y + mid * ax[a].y;
        ctr.z = ctr.z + mid * ax[a].z;
        double half = delta / 2.0;
        *sides[a] = {float(half * ax[a].x), float(half * ax[a].y), float(half * ax[a].z)};
    }
    out.c = {float(ctr.x), float(ctr.y), float(ctr.z)};
    return out;
}

// OBB.cpp:33-89: PCA box of a point list.
Box box_from_points(const V3* p, size_t n) {
    int uniq = distinct_upto4(p, n);
    if (uniq == 0) return empty_box();
    if (uniq == 1) {
        Box b = empty_box();
        b.c = p[0];
        return b;
    }
    D3 sum = {0.0, 0.0, 0.0};
    for (size_t i = 0; i < n; ++i) {
        sum.x = sum.x + double(p[i].x);
        sum.y = sum.y + double(p[i].y);
        sum.z = sum.z + double(p[i].z);
    }
    const double N = double(n);
    D3 mean = {sum.x / N, sum.y / N, sum.z / N};
    double sxx = 0., syy = 0., szz = 0., sxy = 0., sxz = 0., syz = 0.;
    for (size_t i = 0; i < n; ++i) {
        double dx = double(p[i].x) - mean.x;
        double dy = double(p[i].y) - mean.y;
        double dz = double(p[i].z) - mean.z;
        sxx = sxx + dx * dx;
        syy = syy + dy * dy;
        szz = szz + dz * dz;
        sxy = sxy + dx * dy;
        sxz = sxz + dx * dz;
        syz = syz + dy * dz;
    }
    double A[3][3] = {{sxx / N, sxy / N, sxz / N}, {sxy / N, syy / N, syz / N}, {sxz / N, syz / N, szz / N}};
    double V[3][3], d[3];
    eig3(A, V, d);
    // OBB.cpp:80-87: glm's column-major storage is handed to eig3's row-major
    // arrays, so the box axes are the ROWS of V.
    D3 ax[3] = {{V[0][0], V[0][1], V[0][2]}, {V[1][0], V[1][1], V[1][2]}, {V[2][0], V[2][1], V[2][2]}};
    return box_along_axes(p, n, ax);
}

// ---------------------------------------------------------------------------
// Paralgram operations (inMyRoom_vulkan/src/Geometry/Paralgram.cpp)
// ---------------------------------------------------------------------------
static inline Box xform_box(const M4& M, const Box& b) {  // :4-15
    return {mul4(M, b.c, 1.f), mul4(M, b.u, 0.f), mul4(M, b.v, 0.f), mul4(M, b.w, 0.f)};
}
static inline void project(const Box& b, V3 a, float& lo, float& hi) {  // :175-195
    float c = dot(b.c, a);
    float su = std::fabs(dot(a, b.u)), sv = std::fabs(dot(a, b.v)), sw = std::fabs(dot(a, b.w));
    float s = (su + sv) + sw;
    lo = c - s;
    hi = c + s;
}
static inline bool axis_overlaps(const Box& l, const Box& r, V3 a) {  // :197-201
    float l0, l1, r0, r1;
    project(l, a, l0, l1);
    project(r, a, r0, r1);
    return (l1 >= r0) && (r1 >= l0);
}
// :17-173: 15 unnormalised axes, fixed order, early out.
bool sat(const Box& l, const Box& r) {
    if (!axis_overlaps(l, r, cross(l.v, l.w))) return false;
    if (!axis_overlaps(l, r, cross(l.u, l.w))) return false;
    if (!axis_overlaps(l, r, cross(l.u, l.v))) return false;
    if (!axis_overlaps(l, r, cross(r.v, r.w))) return false;
    if (!axis_overlaps(l, r, cross(r.u, r.w))) return false;
    if (!axis_overlaps(l, r, cross(r.u, r.v))) return false;
    const V3 ls[3] = {l.u, l.v, l.w}, rs[3] = {r.u, r.v, r.w};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (!axis_overlaps(l, r, cross(ls[i], rs[j]))) return false;
    return true;
}
float surface(const Box& b) {  // :203-210
    float uv = length(cross(b.u, b.v));
    float uw = length(cross(b.u, b.w));
    float vw = length(cross(b.v, b.w));
    return 2.f * ((uv + uw) + vw);
}

// ---------------------------------------------------------------------------
// OBB tree build + flatten (inMyRoom_vulkan/src/Geometry/OBBtree.cpp:8-394)
// ---------------------------------------------------------------------------
static inline V3 tri_pt(const float* tri9, uint32_t t, int k) {
    const float* q = tri9 + size_t(t) * 9 + k * 3;
    return {q[0], q[1], q[2]};
}

struct BuildNode {
    Box box;
    bool leaf = true;
    std::vector<uint32_t> tris;
    std::unique_ptr<BuildNode> kid[2];
};

// std::sort of three (half-length, axis) pairs by length, descending: for
// three elements libstdc++ runs its insertion sort (threshold 16), which this
// reproduces branch for branch so ties and NaNs land identically.
struct AxisCand { float len; V3 dir; };
static void sort3_desc(AxisCand c[3]) {
    auto before = [](const AxisCand& a, const AxisCand& b) { return a.len > b.len; };
    for (int i = 1; i < 3; ++i) {
        AxisCand val = c[i];
        if (before(val, c[0])) {
            for (int k = i; k > 0; --k) c[k] = c[k - 1];
            c[0] = val;
        } else {
            int j = i;
            while (before(val, c[j - 1])) {
                c[j] = c[j - 1];
                --j;
            }
            c[j] = val;
        }
    }
}

static std::unique_ptr<BuildNode> build_rec(const float* tri9, std::vector<uint32_t>&& tris, unsigned leaf_max) {
    auto nd = std::make_unique<BuildNode>();
    nd->tris = std::move(tris);
    {  // OBB.cpp:91-103: three points per triangle, in triangle order
        std::vector<V3> pts;
        pts.reserve(nd->tris.size() * 3);
        for (uint32_t t : nd->tris)
            for (int k = 0; k < 3; ++k) pts.push_back(tri_pt(tri9, t, k));
        nd->box = box_from_points(pts.data(), pts.size());
    }
    if (nd->tris.size() <= leaf_max) return nd;  // OBBtree.cpp:16-17

    // OBBtree.cpp:43-108
    const Box& b = nd->box;
    AxisCand cand[3] = {{length(b.u), normalize(b.u)}, {length(b.v), normalize(b.v)}, {length(b.w), normalize(b.w)}};
    sort3_desc(cand);
    std::vector<uint32_t> L, R;
    size_t chosen = 0;
    do {
        L.clear();
        R.clear();
        V3 axis = cand[chosen].dir;
        float cproj = dot(b.c, axis);
        for (uint32_t t : nd->tris) {
            float lo = std::numeric_limits<float>::infinity();
            float hi = -std::numeric_limits<float>::infinity();
            for (int k = 0; k < 3; ++k) {  // Triangle.cpp:113-138
                float q = dot(axis, tri_pt(tri9, t, k));
                if (q < lo) lo = q;
                if (q > hi) hi = q;
            }
            float mid = (lo + hi) / 2.f;
            (mid <= cproj ? L : R).push_back(t);
        }
        ++chosen;
    } while ((L.empty() || R.empty()) && chosen < 3);
    if (L.empty() || R.empty()) {
        L.clear();
        R.clear();
        const size_t n = nd->tris.size();
        for (size_t i = 0; i < n; ++i) (i < n / 2 ? L : R).push_back(nd->tris[i]);
    }
    nd->tris.clear();
    nd->leaf = false;
    nd->kid[0] = build_rec(tri9, std::move(L), leaf_max);
    nd->kid[1] = build_rec(tri9, std::move(R), leaf_max);
    if (surface(nd->kid[1]->box) > surface(nd->kid[0]->box)) std::swap(nd->kid[0], nd->kid[1]);
    return nd;
}

static int subtree_depth(const BuildNode* n) {
    if (n->leaf) return 0;
    return 1 + std::max(subtree_depth(n->kid[0].get()), subtree_depth(n->kid[1].get()));
}

// OBBtree.cpp:179-233 and :385-394: DFS preorder of internal nodes, leaves
// append their triangles in visiting order.
static void flatten_rec(const BuildNode* n, uint32_t parent, int side, Tree& T,
                        std::vector<std::pair<const Box*, std::pair<uint32_t, uint32_t>>>& ranges) {
    uint32_t first = uint32_t(T.perm.size());
    if (!n->leaf) {
        uint32_t idx = uint32_t(T.nodes.size());
        T.nodes.emplace_back();
        T.nodes[parent].link[side] = idx;
        T.nodes[parent].cnt[side] = 0;
        T.nodes[parent].obb[side] = n->box;
        flatten_rec(n->kid[0].get(), idx, 0, T, ranges);
        flatten_rec(n->kid[1].get(), idx, 1, T, ranges);
    } else {
        T.nodes[parent].link[side] = first;
        T.nodes[parent].cnt[side] = uint16_t(n->tris.size());
        T.nodes[parent].obb[side] = n->box;
        for (uint32_t t : n->tris) T.perm.push_back(t);
    }
    ranges.push_back({&n->box, {first, uint32_t(T.perm.size())}});
}

// Largest distance by which a vertex sticks out of a box that should hold it
// (the box is a parallelepiped c + a*u + b*v + g*w, |a|,|b|,|g| <= 1).
static double excess_of(const Box& b, const float* pos, uint32_t t0, uint32_t t1) {
    const double u[3] = {b.u.x, b.u.y, b.u.z}, v[3] = {b.v.x, b.v.y, b.v.z}, w[3] = {b.w.x, b.w.y, b.w.z};
    auto det3 = [](const double* a, const double* bb, const double* c) {
        return a[0] * (bb[1] * c[2] - bb[2] * c[1]) - a[1] * (bb[0] * c[2] - bb[2] * c[0]) + a[2] * (bb[0] * c[1] - bb[1] * c[0]);
    };
    double D = det3(u, v, w);
    if (!(std::fabs(D) > 0.0)) return std::numeric_limits<double>::infinity();
    const double lu = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double lv = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double lw = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    double worst = 0.0;
    for (uint32_t t = t0; t < t1; ++t)
        for (int k = 0; k < 3; ++k) {
            const float* q = pos + size_t(t) * 9 + k * 3;
            double r[3] = {double(q[0]) - b.c.x, double(q[1]) - b.c.y, double(q[2]) - b.c.z};
            double a = det3(r, v, w) / D, be = det3(u, r, w) / D, g = det3(u, v, r) / D;
            double ex = std::max(0.0, std::fabs(a) - 1.0) * lu + std::max(0.0, std::fabs(be) - 1.0) * lv +
                        std::max(0.0, std::fabs(g) - 1.0) * lw;
            worst = std::max(worst, ex);
        }
    return worst;
}

Tree* tree_build(const float* tri9, size_t ntri, unsigned leaf_max) {
    auto* T = new Tree();
    std::vector<uint32_t> all(ntri);
    for (size_t i = 0; i < ntri; ++i) all[i] = uint32_t(i);
    std::unique_ptr<BuildNode> root = build_rec(tri9, std::move(all), leaf_max);
    T->root = root->box;
    T->depth = subtree_depth(root.get());
    std::vector<std::pair<const Box*, std::pair<uint32_t, uint32_t>>> ranges;
    if (root->leaf) {
        T->perm = root->tris;
    } else {
        T->nodes.emplace_back();
        flatten_rec(root->kid[0].get(), 0, 0, *T, ranges);
        flatten_rec(root->kid[1].get(), 0, 1, *T, ranges);
    }
    T->pos.resize(T->perm.size() * 9);
    double rad = 0.0;
    for (size_t s = 0; s < T->perm.size(); ++s) {
        std::memcpy(&T->pos[s * 9], tri9 + size_t(T->perm[s]) * 9, 9 * sizeof(float));
        for (int k = 0; k < 3; ++k) {
            const float* q = &T->pos[s * 9 + k * 3];
            rad = std::max(rad, std::sqrt(double(q[0]) * q[0] + double(q[1]) * q[1] + double(q[2]) * q[2]));
        }
    }
    T->radius = rad;
    double slack = excess_of(T->root, T->pos.data(), 0, uint32_t(T->perm.size()));
    for (auto& r : ranges) slack = std::max(slack, excess_of(*r.first, T->pos.data(), r.second.first, r.second.second));
    T->slack = slack;
    return T;
}

// ---------------------------------------------------------------------------
// Traversal (inMyRoom_vulkan/src/Geometry/OBBtree.cpp:235-319, :396-477)
// ---------------------------------------------------------------------------
struct Trav {
    const Box* box;
    bool leaf;
    uint32_t node, off, cnt;
};
static Trav root_trav(const Tree& T) {
    if (!T.nodes.empty()) return {&T.root, false, 0, 0, 0};
    return {&T.root, true, 0, 0, uint32_t(T.perm.size())};
}
static Trav child_trav(const Tree& T, const Trav& p, int side) {
    const Node& n = T.nodes[p.node];
    if (n.cnt[side] == 0) return {&n.obb[side], false, n.link[side], 0, 0};
    return {&n.obb[side], true, 0, n.link[side], n.cnt[side]};
}

static void traverse_rec(const Tree& A, const Tree& B, const Trav& f, const Trav& s, const M4& M,
                         std::vector<LeafPair>& out, Counters& k) {
    ++k.sat;
    Box fb = *f.box;
    Box sb = xform_box(M, *s.box);
    if (!sat(fb, sb)) return;
    if (!f.leaf && !s.leaf) {
        if (surface(fb) >= surface(sb)) {
            traverse_rec(A, B, child_trav(A, f, 0), s, M, out, k);
            traverse_rec(A, B, child_trav(A, f, 1), s, M, out, k);
        } else {
            traverse_rec(A, B, f, child_trav(B, s, 0), M, out, k);
            traverse_rec(A, B, f, child_trav(B, s, 1), M, out, k);
        }
    } else if (f.leaf && !s.leaf) {
        traverse_rec(A, B, f, child_trav(B, s, 0), M, out, k);
        traverse_rec(A, B, f, child_trav(B, s, 1), M, out, k);
    } else if (!f.leaf && s.leaf) {
        traverse_rec(A, B, child_trav(A, f, 0), s, M, out, k);
        traverse_rec(A, B, child_trav(A, f, 1), s, M, out, k);
    } else {
        out.push_back({f.off, f.cnt, s.off, s.cnt});
    }
}

void traverse(const Tree& A, const Tree& B, const M4& M, std::vector<LeafPair>& out, Counters& k) {
    traverse_rec(A, B, root_trav(A), root_trav(B), M, out, k);
}

// ---------------------------------------------------------------------------
// Möller 1997 triangle-triangle test with intersection line
// (inMyRoom_vulkan/src/Geometry/Triangle.cpp:323-1003).
// ---------------------------------------------------------------------------
struct Tri { float p[3][3]; };

static inline void vsub(float d[3], const float a[3], const float b[3]) { d[0] = a[0] - b[0]; d[1] = a[1] - b[1]; d[2] = a[2] - b[2]; }
static inline void vcross(float d[3], const float a[3], const float b[3]) {
    d[0] = a[1] * b[2] - a[2] * b[1];
    d[1] = a[2] * b[0] - a[0] * b[2];
    d[2] = a[0] * b[1] - a[1] * b[0];
}
static inline float vdot(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// :402-420 (one edge of V against one edge of U, projected on axes i0,i1)
static inline bool edge_edge(float Ax, float Ay, const float* V0, const float* U0, const float* U1, int i0, int i1) {
    float Bx = U0[i0] - U1[i0];
    float By = U0[i1] - U1[i1];
    float Cx = V0[i0] - U0[i0];
    float Cy = V0[i1] - U0[i1];
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
static inline bool edge_vs_tri(const float* V0, const float* V1, const float* U0, const float* U1, const float* U2, int i0, int i1) {
    float Ax = V1[i0] - V0[i0];
    float Ay = V1[i1] - V0[i1];
    return edge_edge(Ax, Ay, V0, U0, U1, i0, i1) || edge_edge(Ax, Ay, V0, U1, U2, i0, i1) ||
           edge_edge(Ax, Ay, V0, U2, U0, i0, i1);
}
// :435-458
static inline bool point_in_tri(const float* V0, const float* U0, const float* U1, const float* U2, int i0, int i1) {
    float a, b, c;
    a = U1[i1] - U0[i1];
    b = -(U1[i0] - U0[i0]);
    c = -a * U0[i0] - b * U0[i1];
    float d0 = a * V0[i0] + b * V0[i1] + c;
    a = U2[i1] - U1[i1];
    b = -(U2[i0] - U1[i0]);
    c = -a * U1[i0] - b * U1[i1];
    float d1 = a * V0[i0] + b * V0[i1] + c;
    a = U0[i1] - U2[i1];
    b = -(U0[i0] - U2[i0]);
    c = -a * U2[i0] - b * U2[i1];
    float d2 = a * V0[i0] + b * V0[i1] + c;
    if (d0 * d1 > 0.0f) {
        if (d0 * d2 > 0.0f) return true;
    }
    return false;
}
// :460-507
static int coplanar_test(const float N[3], const float* V0, const float* V1, const float* V2, const float* U0,
                         const float* U1, const float* U2) {
    float A0 = std::fabs(N[0]), A1 = std::fabs(N[1]), A2 = std::fabs(N[2]);
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
static inline void interval_ends(const float* X0, const float* X1, const float* X2, float W0, float W1, float W2,
                                 float D0, float D1, float D2, float& is0, float& is1, float P0[3], float P1[3]) {
    float tmp = D0 / (D0 - D1);
    is0 = W0 + (W1 - W0) * tmp;
    float df[3];
    vsub(df, X1, X0);
    df[0] = tmp * df[0]; df[1] = tmp * df[1]; df[2] = tmp * df[2];
    P0[0] = df[0] + X0[0]; P0[1] = df[1] + X0[1]; P0[2] = df[2] + X0[2];
    tmp = D0 / (D0 - D2);
    is1 = W0 + (W2 - W0) * tmp;
    vsub(df, X2, X0);
    df[0] = tmp * df[0]; df[1] = tmp * df[1]; df[2] = tmp * df[2];
    P1[0] = X0[0] + df[0]; P1[1] = X0[1] + df[1]; P1[2] = X0[2] + df[2];
}
// :795-830; returns true when all three distances are zero (coplanar).
static inline bool intervals(const float* X0, const float* X1, const float* X2, float W0, float W1, float W2,
                             float D0, float D1, float D2, float D0D1, float D0D2, float& is0, float& is1,
                             float P0[3], float P1[3]) {
    if (D0D1 > 0.0f) interval_ends(X2, X0, X1, W2, W0, W1, D2, D0, D1, is0, is1, P0, P1);
    else if (D0D2 > 0.0f) interval_ends(X1, X0, X2, W1, W0, W2, D1, D0, D2, is0, is1, P0, P1);
    else if (D1 * D2 > 0.0f || D0 != 0.0f) interval_ends(X0, X1, X2, W0, W1, W2, D0, D1, D2, is0, is1, P0, P1);
    else if (D1 != 0.0f) interval_ends(X1, X0, X2, W1, W0, W2, D1, D0, D2, is0, is1, P0, P1);
    else if (D2 != 0.0f) interval_ends(X2, X0, X1, W2, W0, W1, D2, D0, D1, is0, is1, P0, P1);
    else return true;
    return false;
}

// :866-1003.  Returns 1 on intersection.  *coplanar as in the reference.
// *ub = 1 marks the case the reference leaves undefined (triangle U's
// distances all snap to zero while V's do not, :960-964); it is defined here
// as coplanar, decided by the 2D coplanar test.
int tri_tri_isect(const float* V0, const float* V1, const float* V2, const float* U0, const float* U1,
                  const float* U2, int* coplanar, int* ub, float s[3], float t[3]) {
    *coplanar = 0;
    *ub = 0;
    float E1[3], E2[3], N1[3], N2[3], D[3];
    vsub(E1, V1, V0);
    vsub(E2, V2, V0);
    vcross(N1, E1, E2);
    float d1 = -vdot(N1, V0);
    float du0 = vdot(N1, U0) + d1;
    float du1 = vdot(N1, U1) + d1;
    float du2 = vdot(N1, U2) + d1;
    const double EPS = 0.000001;  // compared in double, :332, :899-901
    if (std::fabs(du0) < EPS) du0 = 0.0;
    if (std::fabs(du1) < EPS) du1 = 0.0;
    if (std::fabs(du2) < EPS) du2 = 0.0;
    float du0du1 = du0 * du1, du0du2 = du0 * du2;
    if (du0du1 > 0.0f && du0du2 > 0.0f) return 0;
    vsub(E1, U1, U0);
    vsub(E2, U2, U0);
    vcross(N2, E1, E2);
    float d2 = -vdot(N2, U0);
    float dv0 = vdot(N2, V0) + d2;
    float dv1 = vdot(N2, V1) + d2;
    float dv2 = vdot(N2, V2) + d2;
    if (std::fabs(dv0) < EPS) dv0 = 0.0;
    if (std::fabs(dv1) < EPS) dv1 = 0.0;
    if (std::fabs(dv2) < EPS) dv2 = 0.0;
    float dv0dv1 = dv0 * dv1, dv0dv2 = dv0 * dv2;
    if (dv0dv1 > 0.0f && dv0dv2 > 0.0f) return 0;
    vcross(D, N1, N2);
    float mx = std::fabs(D[0]);
    int index = 0;
    float b = std::fabs(D[1]), c = std::fabs(D[2]);
    if (b > mx) { mx = b; index = 1; }
    if (c > mx) { mx = c; index = 2; }
    float vp0 = V0[index], vp1 = V1[index], vp2 = V2[index];
    float up0 = U0[index], up1 = U1[index], up2 = U2[index];
    float is1[2], is2[2], A1[3], A2[3], B1[3], B2[3];
    if (intervals(V0, V1, V2, vp0, vp1, vp2, dv0, dv1, dv2, dv0dv1, dv0dv2, is1[0], is1[1], A1, A2)) {
        *coplanar = 1;
        return coplanar_test(N1, V0, V1, V2, U0, U1, U2);
    }
    if (intervals(U0, U1, U2, up0, up1, up2, du0, du1, du2, du0du1, du0du2, is2[0], is2[1], B1, B2)) {
        *coplanar = 1;
        *ub = 1;
        return coplanar_test(N1, V0, V1, V2, U0, U1, U2);
    }
    int sm1 = 0, sm2 = 0;
    if (is1[0] > is1[1]) { std::swap(is1[0], is1[1]); sm1 = 1; }
    if (is2[0] > is2[1]) { std::swap(is2[0], is2[1]); sm2 = 1; }
    if (is1[1] < is2[0] || is2[1] < is1[0]) return 0;
    auto set = [](float* d, const float* s) { d[0] = s[0]; d[1] = s[1]; d[2] = s[2]; };
    if (is2[0] < is1[0]) {
        set(s, sm1 == 0 ? A1 : A2);
        if (is2[1] < is1[1]) set(t, sm2 == 0 ? B2 : B1);
        else set(t, sm1 == 0 ? A2 : A1);
    } else {
        set(s, sm2 == 0 ? B1 : B2);
        if (is2[1] > is1[1]) set(t, sm1 == 0 ? A2 : A1);
        else set(t, sm2 == 0 ? B2 : B1);
    }
    return 1;
}

// ---------------------------------------------------------------------------
// Narrow phase: candidate leaf pairs -> contacts
// (inMyRoom_vulkan/src/CollisionDetection/CreateUncollideRays.cpp:60-128,
//  inMyRoom_vulkan/src/Geometry/Plane.cpp:5-28, Triangle.cpp:69-78, :140-146)
// ---------------------------------------------------------------------------
struct PlaneF { V3 n; float d; };
static PlaneF plane_of(const V3 P[3]) {
    V3 n0 = normalize(cross(sub(P[1], P[0]), sub(P[2], P[0])));
    float d0 = -dot(P[0], n0);
    float len = length(n0);
    return {{n0.x / len, n0.y / len, n0.z / len}, d0 / len};
}
static inline bool outside(const PlaneF& pl, V3 p) { return (dot(p, pl.n) + pl.d) > 0; }

void narrow(const Tree& A, const Tree& B, const M4& M, const std::vector<LeafPair>& lps, uint32_t query,
            std::vector<Contact>& out, Counters& k) {
    for (const LeafPair& lp : lps) {
        k.tritests += uint64_t(lp.cntA) * lp.cntB;
        for (uint32_t i = 0; i < lp.cntA; ++i)
            for (uint32_t j = 0; j < lp.cntB; ++j) {
                const uint32_t ta = lp.offA + i, tb = lp.offB + j;
                V3 PA[3], PB[3];
                for (int q = 0; q < 3; ++q) {
                    const float* a = &A.pos[size_t(ta) * 9 + q * 3];
                    const float* b = &B.pos[size_t(tb) * 9 + q * 3];
                    PA[q] = {a[0], a[1], a[2]};
                    PB[q] = mul4(M, V3{b[0], b[1], b[2]}, 1.f);
                }
                int cop = 0, ub = 0;
                float s[3], t[3];
                int hit = tri_tri_isect(&PA[0].x, &PA[1].x, &PA[2].x, &PB[0].x, &PB[1].x, &PB[2].x, &cop, &ub, s, t);
                k.ub += ub;
                if (!(hit && !cop)) continue;
                PlaneF pa = plane_of(PA), pb = plane_of(PB);
                Contact c;
                c.query = query;
                c.triA = ta;
                c.triB = tb;
                c.flags = 0;
                for (int q = 0; q < 3; ++q) {
                    if (outside(pb, PA[q])) c.flags |= 1u << q;
                    if (outside(pa, PB[q])) c.flags |= 1u << (3 + q);
                }
                std::memcpy(c.s, s, sizeof s);
                std::memcpy(c.t, t, sizeof t);
                c.na[0] = pa.n.x; c.na[1] = pa.n.y; c.na[2] = pa.n.z;
                c.nb[0] = pb.n.x; c.nb[1] = pb.n.y; c.nb[2] = pb.n.z;
                out.push_back(c);
                ++k.contacts;
            }
    }
}

// ---------------------------------------------------------------------------
// Ray generation: CreateUncollideRays::ExecuteCreateUncollideRays
// (inMyRoom_vulkan/src/CollisionDetection/CreateUncollideRays.cpp:60-200)
// with Plane (src/Geometry/Plane.cpp:5-28) and the triangle helpers of
// src/Geometry/Triangle.cpp:140-212.  The candidate maps and the emplaced
// vertex set are std::unordered_map / std::unordered_set keyed by uint32_t,
// filled in the reference's order, so their iteration order is the one a
// libstdc++ build of the reference sees.
// glm formulas used here beyond the header's list (parity unpinned: no glm
// source is present): mat3(mat4) = upper-left 3x3; adjointTranspose = the
// cofactor matrix, each entry a 2x2 minor a*b - c*d; mat3 * vec3 =
// (m0*x + m1*y) + m2*z per row.
// ---------------------------------------------------------------------------
struct Cand {  // TriangleCandidateRays (CreateUncollideRays.cpp:13-58)
    uint8_t notout = 7;  // std::bitset<3> point_is_not_outside = -1
    V3 wp = {0.f, 0.f, 0.f};
    float w = 0.f;
    bool null() const { return w == 0.f; }
    void merge(const Cand& r) {
        notout &= r.notout;
        wp = add(wp, r.wp);
        w += r.w;
    }
};
struct M3 { float m[9]; };  // column-major: m[c*3 + r]
static M3 adjoint_transpose(const M4& M) {  // glm::adjointTranspose(mat3(M)) (Triangle.cpp:175-178)
    auto a = [&](int c, int r) { return M.m[c * 4 + r]; };
    M3 R;
    R.m[0 * 3 + 0] = +(a(1, 1) * a(2, 2) - a(2, 1) * a(1, 2));
    R.m[0 * 3 + 1] = -(a(1, 0) * a(2, 2) - a(2, 0) * a(1, 2));
    R.m[0 * 3 + 2] = +(a(1, 0) * a(2, 1) - a(2, 0) * a(1, 1));
    R.m[1 * 3 + 0] = -(a(0, 1) * a(2, 2) - a(2, 1) * a(0, 2));
    R.m[1 * 3 + 1] = +(a(0, 0) * a(2, 2) - a(2, 0) * a(0, 2));
    R.m[1 * 3 + 2] = -(a(0, 0) * a(2, 1) - a(2, 0) * a(0, 1));
    R.m[2 * 3 + 0] = +(a(0, 1) * a(1, 2) - a(1, 1) * a(0, 2));
    R.m[2 * 3 + 1] = -(a(0, 0) * a(1, 2) - a(1, 0) * a(0, 2));
    R.m[2 * 3 + 2] = +(a(0, 0) * a(1, 1) - a(1, 0) * a(0, 1));
    return R;
}
static V3 mul3(const M3& M, V3 v) {
    float o[3];
    for (int r = 0; r < 3; ++r) o[r] = (M.m[0 + r] * v.x + M.m[3 + r] * v.y) + M.m[6 + r] * v.z;
    return {o[0], o[1], o[2]};
}
static V3 scale(V3 v, float s) { return {v.x * s, v.y * s, v.z * s}; }
static V3 divs(V3 v, float s) { return {v.x / s, v.y / s, v.z / s}; }

// Face normal of a triangle (Triangle.cpp:140-146).
static V3 face_normal(const V3 P[3]) { return normalize(cross(sub(P[1], P[0]), sub(P[2], P[0]))); }

// Barycentric (u, v) of a point (Triangle.cpp:148-163).
static void barycentric(const V3 P[3], V3 point, float& bx, float& by) {
    V3 v0 = sub(P[1], P[0]), v1 = sub(P[2], P[0]), v2 = sub(point, P[0]);
    float d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(v2, v0), d21 = dot(v2, v1);
    float denom = d00 * d11 - d01 * d01;
    bx = (d11 * d20 - d01 * d21) / denom;
    by = (d00 * d21 - d01 * d20) / denom;
}

static V3 corner_normal(const Tree& T, uint32_t t, int k) {
    const float* n = &T.nrm[size_t(t) * 9 + size_t(k) * 3];
    return {n[0], n[1], n[2]};
}

PairRays uncollide_rays(const Tree& A, const Tree& B, const M4& M, const std::vector<LeafPair>& lps) {
    const M3 NM = adjoint_transpose(M);
    std::unordered_map<uint32_t, Cand> first_map, second_map;
    for (const LeafPair& lp : lps) {
        std::vector<Cand> fc(lp.cntA), sc(lp.cntB);
        for (uint32_t i = 0; i < lp.cntA; ++i)
            for (uint32_t j = 0; j < lp.cntB; ++j) {
                const uint32_t ta = lp.offA + i, tb = lp.offB + j;
                V3 PA[3], PB[3];
                for (int q = 0; q < 3; ++q) {
                    const float* a = &A.pos[size_t(ta) * 9 + q * 3];
                    const float* b = &B.pos[size_t(tb) * 9 + q * 3];
                    PA[q] = {a[0], a[1], a[2]};
                    PB[q] = mul4(M, V3{b[0], b[1], b[2]}, 1.f);
                }
                int cop = 0, ub = 0;
                float s[3], t[3];
                int hit = tri_tri_isect(&PA[0].x, &PA[1].x, &PA[2].x, &PB[0].x, &PB[1].x, &PB[2].x, &cop, &ub, s, t);
                if (!(hit && !cop)) continue;  // :88
                PlaneF pa = plane_of(PA), pb = plane_of(PB);
                V3 S = {s[0], s[1], s[2]}, Tt = {t[0], t[1], t[2]};
                float weight = length(sub(S, Tt));                  // :93
                V3 avg = divs(scale(add(S, Tt), weight), 2.f);       // :94  weight * (s + t) / 2.f
                fc[i].wp = add(fc[i].wp, avg);
                fc[i].w += weight;
                sc[j].wp = add(sc[j].wp, avg);
                sc[j].w += weight;
                for (int q = 0; q < 3; ++q) {  // :102-112
                    if (outside(pb, PA[q])) fc[i].notout &= uint8_t(~(1u << q));
                    if (outside(pa, PB[q])) sc[j].notout &= uint8_t(~(1u << q));
                }
            }
        for (uint32_t i = 0; i < lp.cntA; ++i)  // :117-121
            if (!fc[i].null()) {
                auto r = first_map.emplace(lp.offA + i, fc[i]);
                if (!r.second) r.first->second.merge(fc[i]);
            }
        for (uint32_t j = 0; j < lp.cntB; ++j)  // :123-127
            if (!sc[j].null()) {
                auto r = second_map.emplace(lp.offB + j, sc[j]);
                if (!r.second) r.first->second.merge(sc[j]);
            }
    }
    auto find_rays = [&](const std::unordered_map<uint32_t, Cand>& cands, const Tree& T, bool xf) {  // :131-175
        std::vector<RayRec> rays;
        std::unordered_set<uint32_t> emplaced;
        for (const auto& kv : cands) {
            const uint32_t t = kv.first;
            const Cand& c = kv.second;
            V3 P[3];
            for (int q = 0; q < 3; ++q) {
                const float* p = &T.pos[size_t(t) * 9 + q * 3];
                P[q] = xf ? mul4(M, V3{p[0], p[1], p[2]}, 1.f) : V3{p[0], p[1], p[2]};
            }
            if (c.notout != 0) {
                for (int q = 0; q < 3; ++q) {
                    if (!(c.notout & (1u << q))) continue;
                    if (!emplaced.insert(T.vid[size_t(t) * 3 + q]).second) continue;
                    V3 n = corner_normal(T, t, q);
                    V3 nn = xf ? normalize(mul3(NM, n)) : normalize(n);
                    rays.push_back(RayRec{{P[q].x, P[q].y, P[q].z}, {-nn.x, -nn.y, -nn.z}});
                }
            } else {
                V3 pos = divs(c.wp, c.w);
                float bx, by;
                barycentric(P, pos, bx, by);
                V3 n0 = corner_normal(T, t, 0), n1 = corner_normal(T, t, 1), n2 = corner_normal(T, t, 2);
                V3 in = add(add(scale(n0, (1.f - bx) - by), scale(n1, bx)), scale(n2, by));
                V3 nn = xf ? normalize(mul3(NM, in)) : normalize(in);
                rays.push_back(RayRec{{pos.x, pos.y, pos.z}, {-nn.x, -nn.y, -nn.z}});
            }
        }
        return rays;
    };
    PairRays out;
    auto average = [](const std::vector<RayRec>& rays, float* avg) {  // :183-188, std::accumulate then / float(size)
        V3 acc = {0.f, 0.f, 0.f};
        for (const RayRec& r : rays) acc = add(acc, V3{r.o[0], r.o[1], r.o[2]});
        V3 a = divs(acc, float(rays.size()));
        avg[0] = a.x; avg[1] = a.y; avg[2] = a.z;
    };
    out.from_first = find_rays(first_map, A, false);
    average(out.from_first, out.avg_first);
    out.from_second = find_rays(second_map, B, true);
    average(out.from_second, out.avg_second);
    V3 back = mul4(mat_inverse(M), V3{out.avg_second[0], out.avg_second[1], out.avg_second[2]}, 1.f);  // :196-197
    out.avg_second[0] = back.x; out.avg_second[1] = back.y; out.avg_second[2] = back.z;
    return out;
}

// Tree attributes for rays (Triangle.cpp:214-280): vertex normals through the
// index buffer, or the face normal when the mesh has none; vertex ids are the
// index-buffer values (a soup's are 3 t + k).
// ---------------------------------------------------------------------------
// Response: shooting the uncollide rays (ShootUncollideRays.cpp:14-179,
// Ray.cpp:13-236) and the movement split (CollisionDetection.cpp:80-99,
// :143-150).
// ---------------------------------------------------------------------------

// glm::intersectRayTriangle with the ray origin at 0 (every call in the
// reference is made with a centred ray, Ray.cpp:137-155).  glm 0.9.9's
// Möller-Trumbore form; the reference's glm fork adds a backface output,
// restated here as the negative-determinant branch (the ray leaves through
// the triangle's back: its direction points along the winding normal).
bool ray_triangle(V3 dir, V3 v0, V3 v1, V3 v2, float& bx, float& by, float& dist, bool& back) {
    const V3 e1 = sub(v1, v0), e2 = sub(v2, v0);
    const V3 p = cross(dir, e2);
    const float det = dot(e1, p);
    const float eps = std::numeric_limits<float>::epsilon();
    const V3 d0 = sub(V3{0.f, 0.f, 0.f}, v0);  // orig - vert0
    V3 perp{0.f, 0.f, 0.f};
    if (det > eps) {
        bx = dot(d0, p);
        if (bx < 0.f || bx > det) return false;
        perp = cross(d0, e1);
        by = dot(dir, perp);
        if (by < 0.f || bx + by > det) return false;
        back = false;
    } else if (det < -eps) {
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

// Ray::IntersectParalgram (Ray.cpp:36-133), origin 0: entry and exit
// distances of the ray through the box's three slabs (Haines).
static inline float fmax_std(float a, float b) { return (a < b) ? b : a; }  // std::max(a, b)
static inline float fmin_std(float a, float b) { return (b < a) ? b : a; }  // std::min(a, b)
bool ray_box(V3 dir, const Box& b, float& tmin, float& tmax) {
    tmin = -std::numeric_limits<float>::infinity();
    tmax = +std::numeric_limits<float>::infinity();
    const V3 ro = sub(V3{0.f, 0.f, 0.f}, b.c);
    const V3 sides[3] = {b.u, b.v, b.w};
    const V3 pa[3] = {b.v, b.w, b.u}, pb[3] = {b.w, b.u, b.v};  // slab normals v x w, w x u, u x v
    for (int k = 0; k < 3; ++k) {
        const V3 n = normalize(cross(pa[k], pb[k]));
        const float d = -std::fabs(dot(n, sides[k]));
        const float hc = dot(n, ro);
        const float vn1 = hc + d;
        const float vn2 = -hc + d;
        const float vd = dot(n, dir);
        if (std::fabs(vd) >= std::numeric_limits<float>::epsilon()) {
            const float inv = 1.f / vd;
            float t1 = -vn1 * inv;
            float t2 = vn2 * inv;
            if (t1 > t2) std::swap(t1, t2);
            tmin = fmax_std(t1, tmin);
            tmax = fmin_std(t2, tmax);
            if (tmin > tmax || tmax < 0.f) return false;
        } else if (vn1 > 0 || vn2 > 0) {
            return false;
        }
    }
    return true;
}

// Closest hit of a ray on a tree (Ray::IntersectOBBtree, Ray.cpp:135-236):
// the tree moved by M with the ray origin subtracted from M's translation,
// children visited nearer entry first, each only while its entry distance
// is below the best hit so far.
static void ray_tree_rec(const Tree& T, const Trav& t, const M4& M, V3 dir, RayHit& best) {
    if (!t.leaf) {
        const Trav L = child_trav(T, t, 0), R = child_trav(T, t, 1);
        float l0, l1, r0, r1;
        const bool lh = ray_box(dir, xform_box(M, *L.box), l0, l1);
        const bool rh = ray_box(dir, xform_box(M, *R.box), r0, r1);
        auto visit = [&](const Trav& c, float c0, float c1) {
            if (c0 < best.dist && c1 >= 0.f) ray_tree_rec(T, c, M, dir, best);
        };
        if (lh && rh) {
            if (l0 < r0) {
                visit(L, l0, l1);
                visit(R, r0, r1);
            } else {
                visit(R, r0, r1);
                visit(L, l0, l1);
            }
        } else if (lh) {
            visit(L, l0, l1);
        } else if (rh) {
            visit(R, r0, r1);
        }
        return;
    }
    for (uint32_t i = t.off; i != t.off + t.cnt; ++i) {
        V3 P[3];
        for (int k = 0; k < 3; ++k) P[k] = mul4(M, tri_pt(T.pos.data(), i, k), 1.f);
        float bx = 0.f, by = 0.f, dist = std::numeric_limits<float>::infinity();
        bool back = false;
        if (ray_triangle(dir, P[0], P[1], P[2], bx, by, dist, back) && dist > 0.f && dist < best.dist) {
            best.hit = true;
            best.back = back;
            best.dist = dist;
            best.tri = i;
            best.bx = bx;
            best.by = by;
        }
    }
}
RayHit ray_tree(const Tree& T, const M4& M, V3 o, V3 d) {
    M4 C = M;  // centred: the ray starts at 0 (Ray.cpp:150-154)
    C.m[12] = C.m[12] - o.x;
    C.m[13] = C.m[13] - o.y;
    C.m[14] = C.m[14] - o.z;
    C.m[15] = C.m[15] - 0.f;
    RayHit best;
    const Trav root = root_trav(T);
    float t0, t1;
    if (ray_box(d, xform_box(C, *root.box), t0, t1) && t1 >= 0.f) ray_tree_rec(T, root, C, d, best);
    return best;
}

// Ray::MoveOriginEpsilonTowardsDirection (Ray.cpp:13-21).
static inline V3 neg(V3 v) { return {-v.x, -v.y, -v.z}; }  // glm's unary minus
static V3 move_eps(V3 o, V3 d, float factor) {
    const float big = std::max({std::fabs(o.x), std::fabs(o.y), std::fabs(o.z)});
    const float e = big * std::numeric_limits<float>::epsilon();
    return add(o, scale(d, factor * e));
}

struct Hermann {
    bool ok = false;
    V3 response{0.f, 0.f, 0.f}, normal{0.f, 0.f, 0.f};
};
// ShootUncollideRays::HermannPass (:129-158): the ray must leave B through a
// back face before it re-enters A (within the origin's epsilon step).
static Hermann hermann(const Tree& A, const M4& Amat, const Tree& B, const M4& Bmat, const M3& Bn, V3 o, V3 d) {
    Hermann h;
    const RayHit p2 = ray_tree(B, Bmat, o, d);
    if (!(p2.hit && p2.back)) return h;
    const V3 o2 = move_eps(o, d, 4.f);
    const float eps = length(sub(o2, o));
    const RayHit p3 = ray_tree(A, Amat, o2, d);
    if (p2.dist <= p3.dist + eps) {
        h.ok = true;
        h.response = scale(d, p2.dist);
        const V3 n0 = corner_normal(B, p2.tri, 0), n1 = corner_normal(B, p2.tri, 1), n2 = corner_normal(B, p2.tri, 2);
        const float w0 = (1.f - p2.bx) - p2.by;
        const V3 ni = add(add(scale(n0, w0), scale(n1, p2.bx)), scale(n2, p2.by));
        h.normal = normalize(mul3(Bn, ni));
    }
    return h;
}
// :118-127
static V3 force_response(const Hermann& h) {
    const V3 dir = normalize(h.response);
    const float len = length(h.response);
    const float c = dot(h.normal, dir);
    const float c2 = c * c;
    return scale(dir, c2 * len);
}
// :171-175
static float smoother_step(float a, float b, float x) {
    float t = (x - a) / (b - a);
    t = (t < 0.f) ? 0.f : ((1.f < t) ? 1.f : t);  // std::clamp(t, 0, 1)
    return t * t * t * (t * (t * 6.f - 15.f) + 10.f);
}

// ExecuteShootUncollideRays (:14-102) for one pair, in the first entity's
// model space until the final transform by the first's current matrix.
V3 shoot_pair(const Tree& first, const Tree& second, const M4& first_cur, const M4& M, const PairRays& pr,
              float cos_start, float cos_finish, float bias, ShootStats* stats) {
    M4 I{};
    I.m[0] = I.m[5] = I.m[10] = I.m[15] = 1.f;
    M3 I3{};
    I3.m[0] = I3.m[4] = I3.m[8] = 1.f;
    const M3 Mn = adjoint_transpose(M);
    V3 avg{0.f, 0.f, 0.f};
    std::vector<V3> responses;
    // returns whether the pass succeeded and, if so, the reflected ray
    auto f2s = [&](V3 o, V3 d, V3& ro, V3& rd) {
        const Hermann h = hermann(first, I, second, M, Mn, o, d);
        if (stats) ++stats->passes;
        if (!h.ok) return false;
        avg = sub(avg, force_response(h));
        responses.push_back(neg(h.response));
        ro = add(o, h.response);
        rd = neg(h.normal);
        return true;
    };
    auto s2f = [&](V3 o, V3 d, V3& ro, V3& rd) {
        const Hermann h = hermann(second, M, first, I, I3, o, d);
        if (stats) ++stats->passes;
        if (!h.ok) return false;
        avg = add(avg, force_response(h));
        responses.push_back(h.response);
        ro = add(o, h.response);
        rd = neg(h.normal);
        return true;
    };
    for (const RayRec& r : pr.from_first) {
        V3 ro, rd, xo, xd;
        if (f2s({r.o[0], r.o[1], r.o[2]}, {r.d[0], r.d[1], r.d[2]}, ro, rd)) s2f(ro, rd, xo, xd);
    }
    for (const RayRec& r : pr.from_second) {
        V3 ro, rd, xo, xd;
        if (s2f({r.o[0], r.o[1], r.o[2]}, {r.d[0], r.d[1], r.d[2]}, ro, rd)) f2s(ro, rd, xo, xd);
    }
    if (stats) stats->successes += responses.size();
    if (avg.x == 0.f && avg.y == 0.f && avg.z == 0.f) return {0.f, 0.f, 0.f};
    const V3 n = normalize(avg);
    float max_len = 0.f;  // FindResponse (:160-169)
    for (const V3& r : responses) {
        const V3 rn = normalize(r);
        const float c = dot(rn, n);
        const float need = length(r) / c;
        const float f = smoother_step(cos_finish, cos_start, c);
        max_len = fmax_std(max_len, f * need);
    }
    const V3 local = scale(scale(n, max_len), bias);  // bias * (max_len * n): glm scalar * vec3, per component
    return mul4(first_cur, local, 0.f);
}

// PointMovementBetweenFrames (CollisionDetection.cpp:143-150).
static float point_movement(V3 p, const M4& cur, const M4& prev) {
    return length(sub(mul4(cur, p, 1.f), mul4(prev, p, 1.f)));
}
static bool mat_equal(const M4& a, const M4& b) {
    for (int i = 0; i < 16; ++i)
        if (!(a.m[i] == b.m[i])) return false;
    return true;
}
PairResponse respond(const Tree& first, const Tree& second, const M4& first_cur, const M4& first_prev,
                     const M4& second_cur, const M4& second_prev, const PairRays& pr, float cos_start,
                     float cos_finish, float bias, ShootStats* stats) {
    PairResponse out;
    out.first = pr.first;
    out.second = pr.second;
    if (mat_equal(first_cur, first_prev) && mat_equal(second_cur, second_prev)) return out;  // :80-81
    out.shot = true;
    const M4 M = mat_mul(mat_inverse(first_cur), second_cur);
    const V3 delta = shoot_pair(first, second, first_cur, M, pr, cos_start, cos_finish, bias, stats);
    const float m1 = point_movement({pr.avg_first[0], pr.avg_first[1], pr.avg_first[2]}, first_cur, first_prev);
    const float m2 = point_movement({pr.avg_second[0], pr.avg_second[1], pr.avg_second[2]}, second_cur, second_prev);
    const float total = m1 + m2;
    const V3 nd = neg(delta);
    const V3 d1 = scale(nd, m1 / total), d2 = scale(delta, m2 / total);
    for (int k = 0; k < 3; ++k) {
        out.delta[k] = (&delta.x)[k];
        out.delta_first[k] = (&d1.x)[k];
        out.delta_second[k] = (&d2.x)[k];
    }
    return out;
}

void set_attrs(Tree& T, const float* nrm9, const uint32_t* vid3) {
    const size_t n = T.perm.size();
    T.nrm.assign(n * 9, 0.f);
    T.vid.assign(n * 3, 0u);
    for (size_t s = 0; s < n; ++s) {
        const uint32_t t = T.perm[s];
        V3 P[3];
        for (int q = 0; q < 3; ++q) P[q] = {T.pos[s * 9 + q * 3], T.pos[s * 9 + q * 3 + 1], T.pos[s * 9 + q * 3 + 2]};
        const V3 fn = face_normal(P);
        for (int q = 0; q < 3; ++q) {
            V3 nq = nrm9 ? V3{nrm9[size_t(t) * 9 + q * 3], nrm9[size_t(t) * 9 + q * 3 + 1], nrm9[size_t(t) * 9 + q * 3 + 2]} : fn;
            T.nrm[s * 9 + q * 3] = nq.x;
            T.nrm[s * 9 + q * 3 + 1] = nq.y;
            T.nrm[s * 9 + q * 3 + 2] = nq.z;
            T.vid[s * 3 + q] = vid3 ? vid3[size_t(t) * 3 + q] : uint32_t(3 * t + q);
        }
    }
}

// ---------------------------------------------------------------------------
// Broad phase: sweep and prune on a fixed rotated frame
// (inMyRoom_vulkan/src/CollisionDetection/CollisionDetection.cpp:9-13,
//  inMyRoom_vulkan/src/CollisionDetection/SweepAndPrune.cpp:15-88)
// Ties on an interval's min are ordered by entry index (the reference's
// std::sort leaves them unspecified).
// ---------------------------------------------------------------------------
void sap_axes(V3 ax[3]) {
    V3 U = normalize(V3{0.8f, -0.2f, 0.f});
    V3 W = normalize(cross(U, V3{0.f, -1.f, 0.f}));
    V3 V = normalize(cross(W, U));
    ax[0] = U;
    ax[1] = V;
    ax[2] = W;
}

std::vector<std::pair<uint32_t, uint32_t>> sap(const M4* cur, const Box* const* roots, const uint8_t* cb, size_t n) {
    V3 ax[3];
    sap_axes(ax);
    struct Ent { float lo, hi; uint32_t idx; bool cb; };
    std::vector<Ent> lists[3];
    for (size_t i = 0; i < n; ++i) {
        Box P = xform_box(cur[i], *roots[i]);
        for (int a = 0; a < 3; ++a) {
            Ent e;
            project(P, ax[a], e.lo, e.hi);
            e.idx = uint32_t(i);
            e.cb = cb[i] != 0;
            lists[a].push_back(e);
        }
    }
    for (int a = 0; a < 3; ++a)
        std::sort(lists[a].begin(), lists[a].end(), [](const Ent& x, const Ent& y) {
            if (x.lo < y.lo) return true;
            if (y.lo < x.lo) return false;
            return x.idx < y.idx;
        });
    std::map<std::pair<uint32_t, uint32_t>, std::pair<int, std::pair<uint32_t, uint32_t>>> count;  // unordered key
    for (int a = 0; a < 3; ++a) {
        std::vector<Ent> active;
        for (const Ent& e : lists[a]) {
            std::vector<Ent> keep;
            for (const Ent& act : active) {
                if (act.hi < e.lo) continue;  // retired (:58-59)
                keep.push_back(act);
                if (!(act.cb || e.cb)) continue;
                auto key = std::make_pair(std::min(act.idx, e.idx), std::max(act.idx, e.idx));
                if (a == 0) count[key] = {1, {act.idx, e.idx}};
                else {
                    auto it = count.find(key);
                    if (it != count.end()) it->second.first++;
                }
            }
            keep.push_back(e);
            active.swap(keep);
        }
    }
    std::vector<std::pair<uint32_t, uint32_t>> out;
    for (auto& kv : count)
        if (kv.second.first == 3) out.push_back(kv.second.second);
    std::sort(out.begin(), out.end());
    return out;
}

// ---------------------------------------------------------------------------
// Exact orientation (bignum) and the exact-semantics triangle test.
// side(a,b,c,d) = sign(((a-c) x (b-c)) . (d-c)) evaluated exactly: every
// float is m*2^e with |m| < 2^24, so after scaling by the smallest exponent
// all 12 coordinates are integers and the determinant is an integer.
// ---------------------------------------------------------------------------
struct Big {
    static const int L = 40;  // 1280 bits: enough for 3 x 302-bit factors
    int sgn = 0;              // -1, 0, +1
    int n = 0;                // limbs in use
    uint32_t d[L];
};
static int mag_cmp(const Big& a, const Big& b) {
    if (a.n != b.n) return a.n < b.n ? -1 : 1;
    for (int i = a.n - 1; i >= 0; --i)
        if (a.d[i] != b.d[i]) return a.d[i] < b.d[i] ? -1 : 1;
    return 0;
}
static void trim(Big& a) {
    while (a.n > 0 && a.d[a.n - 1] == 0) --a.n;
    if (a.n == 0) a.sgn = 0;
}
static Big mag_add(const Big& a, const Big& b) {
    Big r;
    uint64_t carry = 0;
    int n = std::max(a.n, b.n);
    for (int i = 0; i < n; ++i) {
        uint64_t s = carry + (i < a.n ? a.d[i] : 0) + (i < b.n ? b.d[i] : 0);
        r.d[i] = uint32_t(s);
        carry = s >> 32;
    }
    r.n = n;
    if (carry) r.d[r.n++] = uint32_t(carry);
    r.sgn = 1;
    trim(r);
    return r;
}
static Big mag_sub(const Big& a, const Big& b) {  // |a| >= |b|
    Big r;
    int64_t borrow = 0;
    for (int i = 0; i < a.n; ++i) {
        int64_t s = int64_t(a.d[i]) - (i < b.n ? b.d[i] : 0) - borrow;
        borrow = s < 0;
        if (s < 0) s += (int64_t(1) << 32);
        r.d[i] = uint32_t(s);
    }
    r.n = a.n;
    r.sgn = 1;
    trim(r);
    return r;
}
static Big big_add(const Big& a, const Big& b) {
    if (a.sgn == 0) return b;
    if (b.sgn == 0) return a;
    if (a.sgn == b.sgn) {
        Big r = mag_add(a, b);
        r.sgn = r.n ? a.sgn : 0;
        return r;
    }
    int c = mag_cmp(a, b);
    if (c == 0) return Big();
    Big r = c > 0 ? mag_sub(a, b) : mag_sub(b, a);
    r.sgn = r.n ? (c > 0 ? a.sgn : b.sgn) : 0;
    return r;
}
static Big big_neg(Big a) { a.sgn = -a.sgn; return a; }
static Big big_sub(const Big& a, const Big& b) { return big_add(a, big_neg(b)); }
static Big big_mul(const Big& a, const Big& b) {
    Big r;
    if (a.sgn == 0 || b.sgn == 0) return r;
    r.n = a.n + b.n;
    for (int i = 0; i < r.n; ++i) r.d[i] = 0;
    for (int i = 0; i < a.n; ++i) {
        uint64_t carry = 0;
        for (int j = 0; j < b.n; ++j) {
            uint64_t cur = uint64_t(a.d[i]) * b.d[j] + r.d[i + j] + carry;
            r.d[i + j] = uint32_t(cur);
            carry = cur >> 32;
        }
        int k = i + b.n;
        while (carry) {
            uint64_t cur = uint64_t(r.d[k]) + carry;
            r.d[k] = uint32_t(cur);
            carry = cur >> 32;
            ++k;
        }
    }
    r.sgn = a.sgn * b.sgn;
    trim(r);
    return r;
}
// integer m * 2^shift (m signed, |m| < 2^25, shift >= 0)
static Big big_from(int64_t m, int shift) {
    Big r;
    if (m == 0) return r;
    r.sgn = m < 0 ? -1 : 1;
    uint64_t mag = uint64_t(m < 0 ? -m : m);
    int limb = shift / 32, bit = shift % 32;
    for (int i = 0; i < limb; ++i) r.d[i] = 0;
    unsigned __int128 v = (unsigned __int128)mag << bit;
    r.n = limb;
    while (v) {
        r.d[r.n++] = uint32_t(v);
        v >>= 32;
    }
    return r;
}

static int orient_bignum(const float* a, const float* b, const float* c, const float* d) {
    const float* pts[4] = {a, b, c, d};
    int64_t man[4][3];
    int ex[4][3];
    int emin = std::numeric_limits<int>::max();
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 3; ++k) {
            float x = pts[i][k];
            int e;
            float f = std::frexp(x, &e);  // x = f * 2^e, 0.5 <= |f| < 1
            man[i][k] = int64_t(std::ldexp(f, 24));
            ex[i][k] = e - 24;
            if (man[i][k] != 0) emin = std::min(emin, ex[i][k]);
        }
    if (emin == std::numeric_limits<int>::max()) return 0;
    Big q[4][3];
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 3; ++k) q[i][k] = man[i][k] ? big_from(man[i][k], ex[i][k] - emin) : Big();
    Big A[3], B[3], D[3];
    for (int k = 0; k < 3; ++k) {
        A[k] = big_sub(q[0][k], q[2][k]);
        B[k] = big_sub(q[1][k], q[2][k]);
        D[k] = big_sub(q[3][k], q[2][k]);
    }
    Big cx = big_sub(big_mul(A[1], B[2]), big_mul(B[1], A[2]));
    Big cy = big_sub(big_mul(A[2], B[0]), big_mul(B[2], A[0]));
    Big cz = big_sub(big_mul(A[0], B[1]), big_mul(B[0], A[1]));
    Big det = big_add(big_add(big_mul(cx, D[0]), big_mul(cy, D[1])), big_mul(cz, D[2]));
    return det.sgn;
}

static uint64_t g_bignum_calls = 0;

// Double evaluation with a forward error bound; bignum only when the bound
// cannot certify the sign.
int orient_exact(const float* a, const float* b, const float* c, const float* d) {
    double ax = double(a[0]) - c[0], ay = double(a[1]) - c[1], az = double(a[2]) - c[2];
    double bx = double(b[0]) - c[0], by = double(b[1]) - c[1], bz = double(b[2]) - c[2];
    double dx = double(d[0]) - c[0], dy = double(d[1]) - c[1], dz = double(d[2]) - c[2];
    double det = (ay * bz - by * az) * dx + (az * bx - bz * ax) * dy + (ax * by - bx * ay) * dz;
    double perm = (std::fabs(ay * bz) + std::fabs(by * az)) * std::fabs(dx) +
                  (std::fabs(az * bx) + std::fabs(bz * ax)) * std::fabs(dy) +
                  (std::fabs(ax * by) + std::fabs(bx * ay)) * std::fabs(dz);
    double bound = 16.0 * std::ldexp(1.0, -53) * perm;
    if (std::isfinite(det) && std::isfinite(perm) && perm > 1e-280) {
        if (det > bound) return 1;
        if (det < -bound) return -1;
    }
    ++g_bignum_calls;
    return orient_bignum(a, b, c, d);
}

// Guigue & Devillers 2003 ("Fast and robust triangle-triangle overlap test
// using orientation predicates", JGT 8(1)) decision structure, with every
// sign taken from orient_exact.  Closed triangles; a pair with all three
// vertices of either triangle on the other's plane is "coplanar".
// Returns 1 = intersect (non-coplanar), 0 = disjoint, 2 = coplanar.
static int gd_check(const float* p1, const float* q1, const float* r1, const float* p2, const float* q2, const float* r2) {
    if (orient_exact(p2, p1, q1, q2) > 0) return 0;
    if (orient_exact(p2, r1, p1, r2) > 0) return 0;
    return 1;
}
static int gd_canon2(const float* p1, const float* q1, const float* r1, const float* p2, const float* q2,
                     const float* r2, int dp2, int dq2, int dr2) {
    if (dp2 > 0) {
        if (dq2 > 0) return gd_check(p1, r1, q1, r2, p2, q2);
        if (dr2 > 0) return gd_check(p1, r1, q1, q2, r2, p2);
        return gd_check(p1, q1, r1, p2, q2, r2);
    }
    if (dp2 < 0) {
        if (dq2 < 0) return gd_check(p1, q1, r1, r2, p2, q2);
        if (dr2 < 0) return gd_check(p1, q1, r1, q2, r2, p2);
        return gd_check(p1, r1, q1, p2, q2, r2);
    }
    if (dq2 < 0) {
        if (dr2 >= 0) return gd_check(p1, r1, q1, q2, r2, p2);
        return gd_check(p1, q1, r1, p2, q2, r2);
    }
    if (dq2 > 0) {
        if (dr2 > 0) return gd_check(p1, r1, q1, p2, q2, r2);
        return gd_check(p1, q1, r1, q2, r2, p2);
    }
    if (dr2 > 0) return gd_check(p1, q1, r1, r2, p2, q2);
    if (dr2 < 0) return gd_check(p1, r1, q1, r2, p2, q2);
    return 2;
}
int tri_tri_exact(const float* p1, const float* q1, const float* r1, const float* p2, const float* q2, const float* r2) {
    int dp1 = orient_exact(p2, q2, r2, p1), dq1 = orient_exact(p2, q2, r2, q1), dr1 = orient_exact(p2, q2, r2, r1);
    if (dp1 * dq1 > 0 && dp1 * dr1 > 0) return 0;
    int dp2 = orient_exact(q1, r1, p1, p2), dq2 = orient_exact(q1, r1, p1, q2), dr2 = orient_exact(q1, r1, p1, r2);
    if (dp2 * dq2 > 0 && dp2 * dr2 > 0) return 0;
    if ((dp1 == 0 && dq1 == 0 && dr1 == 0) || (dp2 == 0 && dq2 == 0 && dr2 == 0)) return 2;
    if (dp1 > 0) {
        if (dq1 > 0) return gd_canon2(r1, p1, q1, p2, r2, q2, dp2, dr2, dq2);
        if (dr1 > 0) return gd_canon2(q1, r1, p1, p2, r2, q2, dp2, dr2, dq2);
        return gd_canon2(p1, q1, r1, p2, q2, r2, dp2, dq2, dr2);
    }
    if (dp1 < 0) {
        if (dq1 < 0) return gd_canon2(r1, p1, q1, p2, q2, r2, dp2, dq2, dr2);
        if (dr1 < 0) return gd_canon2(q1, r1, p1, p2, q2, r2, dp2, dq2, dr2);
        return gd_canon2(p1, q1, r1, p2, r2, q2, dp2, dr2, dq2);
    }
    if (dq1 < 0) {
        if (dr1 >= 0) return gd_canon2(q1, r1, p1, p2, r2, q2, dp2, dr2, dq2);
        return gd_canon2(p1, q1, r1, p2, q2, r2, dp2, dq2, dr2);
    }
    if (dq1 > 0) {
        if (dr1 > 0) return gd_canon2(p1, q1, r1, p2, r2, q2, dp2, dr2, dq2);
        return gd_canon2(q1, r1, p1, p2, q2, r2, dp2, dq2, dr2);
    }
    if (dr1 > 0) return gd_canon2(r1, p1, q1, p2, q2, r2, dp2, dq2, dr2);
    if (dr1 < 0) return gd_canon2(r1, p1, q1, p2, r2, q2, dp2, dr2, dq2);
    return 2;
}

// Exact-semantics query: conservative traversal (double SAT with every
// interval widened by the containment slack of both trees plus the rounding
// of B's float transform), then the exact test on every candidate pair.
struct DBox { double c[3], s[3][3]; };
static DBox dbox_of(const Box& b) {
    return {{b.c.x, b.c.y, b.c.z}, {{b.u.x, b.u.y, b.u.z}, {b.v.x, b.v.y, b.v.z}, {b.w.x, b.w.y, b.w.z}}};
}
static DBox dxform(const M4& M, const DBox& b) {
    DBox r;
    for (int i = 0; i < 3; ++i) {
        r.c[i] = double(M.m[0 + i]) * b.c[0] + double(M.m[4 + i]) * b.c[1] + double(M.m[8 + i]) * b.c[2] + double(M.m[12 + i]);
        for (int k = 0; k < 3; ++k)
            r.s[k][i] = double(M.m[0 + i]) * b.s[k][0] + double(M.m[4 + i]) * b.s[k][1] + double(M.m[8 + i]) * b.s[k][2];
    }
    return r;
}
static bool dsat(const DBox& l, const DBox& r, double delta) {
    auto crossd = [](const double* a, const double* b, double* o) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    auto test = [&](const double* ax) {
        auto dd = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
        double n = std::sqrt(dd(ax, ax));
        if (!(n > 0)) return true;
        double cl = dd(ax, l.c), cr = dd(ax, r.c);
        double rl = std::fabs(dd(ax, l.s[0])) + std::fabs(dd(ax, l.s[1])) + std::fabs(dd(ax, l.s[2]));
        double rr = std::fabs(dd(ax, r.s[0])) + std::fabs(dd(ax, r.s[1])) + std::fabs(dd(ax, r.s[2]));
        double tol = delta * n + 1e-12 * (std::fabs(cl) + rl + std::fabs(cr) + rr);
        return !((cl + rl + tol < cr - rr) || (cr + rr + tol < cl - rl));
    };
    double ax[3];
    const int fl[3][2] = {{1, 2}, {0, 2}, {0, 1}};
    for (int i = 0; i < 3; ++i) {
        crossd(l.s[fl[i][0]], l.s[fl[i][1]], ax);
        if (!test(ax)) return false;
    }
    for (int i = 0; i < 3; ++i) {
        crossd(r.s[fl[i][0]], r.s[fl[i][1]], ax);
        if (!test(ax)) return false;
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            crossd(l.s[i], r.s[j], ax);
            if (!test(ax)) return false;
        }
    return true;
}
static double dsurface(const DBox& b) {
    auto cl = [](const double* a, const double* c) {
        double o0 = a[1] * c[2] - a[2] * c[1], o1 = a[2] * c[0] - a[0] * c[2], o2 = a[0] * c[1] - a[1] * c[0];
        return std::sqrt(o0 * o0 + o1 * o1 + o2 * o2);
    };
    return 2.0 * (cl(b.s[0], b.s[1]) + cl(b.s[0], b.s[2]) + cl(b.s[1], b.s[2]));
}
static void exact_rec(const Tree& A, const Tree& B, const Trav& f, const Trav& s, const M4& M, double delta,
                      std::vector<LeafPair>& out, Counters& k) {
    ++k.sat;
    DBox fb = dbox_of(*f.box), sb = dxform(M, dbox_of(*s.box));
    if (!dsat(fb, sb, delta)) return;
    if (!f.leaf && !s.leaf) {
        if (dsurface(fb) >= dsurface(sb)) {
            exact_rec(A, B, child_trav(A, f, 0), s, M, delta, out, k);
            exact_rec(A, B, child_trav(A, f, 1), s, M, delta, out, k);
        } else {
            exact_rec(A, B, f, child_trav(B, s, 0), M, delta, out, k);
            exact_rec(A, B, f, child_trav(B, s, 1), M, delta, out, k);
        }
    } else if (f.leaf && !s.leaf) {
        exact_rec(A, B, f, child_trav(B, s, 0), M, delta, out, k);
        exact_rec(A, B, f, child_trav(B, s, 1), M, delta, out, k);
    } else if (!f.leaf && s.leaf) {
        exact_rec(A, B, child_trav(A, f, 0), s, M, delta, out, k);
        exact_rec(A, B, child_trav(A, f, 1), s, M, delta, out, k);
    } else {
        out.push_back({f.off, f.cnt, s.off, s.cnt});
    }
}

// The intersection segment of two intersecting, non-coplanar triangles in
// float64 from their float corners (the checker of the exact mode's contact
// segments, which have no reference counterpart): each triangle's boundary
// meets the other's plane in one or two points of the common line (a corner
// on the plane, or an edge whose ends lie strictly on opposite sides); the
// segment is the overlap of the two intervals those points span along the
// line.  s gets the point at the low end, t the point at the high end.
// Which side of the other plane a corner lies on is its exact orientation
// sign; a float64 distance of the wrong sign (a corner within rounding of the
// plane) is replaced by a tiny one of the exact sign, so the crossing point
// falls on that corner.  NaN when no corner or edge meets the other plane
// (the pair does not intersect).
static void segment_f64(const float PA[3][3], const float PB[3][3], float s[3], float t[3]) {
    auto sub = [](const double* a, const double* b, double* o) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; };
    auto crs = [](const double* a, const double* b, double* o) {
        o[0] = a[1] * b[2] - a[2] * b[1]; o[1] = a[2] * b[0] - a[0] * b[2]; o[2] = a[0] * b[1] - a[1] * b[0];
    };
    auto dt = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    double a[3][3], b[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) { a[i][j] = PA[i][j]; b[i][j] = PB[i][j]; }
    double e1[3], e2[3], n1[3], n2[3], D[3];
    sub(a[1], a[0], e1); sub(a[2], a[0], e2); crs(e1, e2, n1);
    sub(b[1], b[0], e1); sub(b[2], b[0], e2); crs(e1, e2, n2);
    crs(n1, n2, D);
    // [lo point, hi point] of one triangle X against the plane (n, p0)
    // sides[k]: exact side of X's corner k of the other triangle's plane
    auto span = [&](const double X[3][3], const double* n, const double* p0, const int* sides, double lo[3],
                    double hi[3], double& tlo, double& thi) {
        double d[3], w[3];
        for (int k = 0; k < 3; ++k) {
            sub(X[k], p0, w);
            d[k] = dt(n, w);
            if (sides[k] == 0) d[k] = 0.0;
            else if ((d[k] > 0.0) != (sides[k] > 0)) d[k] = sides[k] * 1e-300;
        }
        tlo = std::numeric_limits<double>::infinity();
        thi = -tlo;
        auto take = [&](const double* q) {
            const double tq = dt(D, q);
            if (tq < tlo) { tlo = tq; std::memcpy(lo, q, 3 * sizeof(double)); }
            if (tq > thi) { thi = tq; std::memcpy(hi, q, 3 * sizeof(double)); }
        };
        for (int k = 0; k < 3; ++k) {
            if (d[k] == 0.0) take(X[k]);
            const int j = (k + 1) % 3;
            if ((d[k] < 0.0 && d[j] > 0.0) || (d[k] > 0.0 && d[j] < 0.0)) {
                const double u = d[k] / (d[k] - d[j]);
                double q[3];
                for (int c = 0; c < 3; ++c) q[c] = X[k][c] + (X[j][c] - X[k][c]) * u;
                take(q);
            }
        }
    };
    int sa[3], sb[3];
    for (int k = 0; k < 3; ++k) {
        sa[k] = orient_exact(PB[1], PB[2], PB[0], PA[k]);  // sign of n2 . (a_k - b_0)
        sb[k] = orient_exact(PA[1], PA[2], PA[0], PB[k]);  // sign of n1 . (b_k - a_0)
    }
    double alo[3], ahi[3], blo[3], bhi[3], tal, tah, tbl, tbh;
    span(a, n2, b[0], sa, alo, ahi, tal, tah);
    span(b, n1, a[0], sb, blo, bhi, tbl, tbh);
    if (!(tal <= tah) || !(tbl <= tbh)) {
        for (int c = 0; c < 3; ++c) s[c] = t[c] = std::numeric_limits<float>::quiet_NaN();
        return;
    }
    const double* lo = tal >= tbl ? alo : blo;
    const double* hi = tah <= tbh ? ahi : bhi;
    for (int c = 0; c < 3; ++c) { s[c] = float(lo[c]); t[c] = float(hi[c]); }
}

void query_exact(const Tree& A, const Tree& B, const M4& M, uint32_t query, std::vector<Contact>& out, Counters& k) {
    // rounding of B's float transform, measured on B's own vertices
    double dev = 0.0;
    for (size_t v = 0; v < B.pos.size() / 3; ++v) {
        V3 p = {B.pos[v * 3], B.pos[v * 3 + 1], B.pos[v * 3 + 2]};
        V3 q = mul4(M, p, 1.f);
        double e[3];
        for (int i = 0; i < 3; ++i) {
            double ex = double(M.m[0 + i]) * p.x + double(M.m[4 + i]) * p.y + double(M.m[8 + i]) * p.z + double(M.m[12 + i]);
            e[i] = double((&q.x)[i]) - ex;
        }
        dev = std::max(dev, std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]));
    }
    double fro = 0.0;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) fro += double(M.m[c * 4 + r]) * M.m[c * 4 + r];
    double delta = 2.0 * (A.slack + std::sqrt(fro) * B.slack + dev) + 1e-9 * (1.0 + A.radius + B.radius);
    std::vector<LeafPair> lps;
    exact_rec(A, B, root_trav(A), root_trav(B), M, delta, lps, k);
    k.leafpairs += lps.size();
    for (const LeafPair& lp : lps) {
        k.tritests += uint64_t(lp.cntA) * lp.cntB;
        for (uint32_t i = 0; i < lp.cntA; ++i)
            for (uint32_t j = 0; j < lp.cntB; ++j) {
                const uint32_t ta = lp.offA + i, tb = lp.offB + j;
                float PA[3][3], PB[3][3];
                for (int q = 0; q < 3; ++q) {
                    const float* a = &A.pos[size_t(ta) * 9 + q * 3];
                    const float* b = &B.pos[size_t(tb) * 9 + q * 3];
                    PA[q][0] = a[0]; PA[q][1] = a[1]; PA[q][2] = a[2];
                    V3 t = mul4(M, V3{b[0], b[1], b[2]}, 1.f);
                    PB[q][0] = t.x; PB[q][1] = t.y; PB[q][2] = t.z;
                }
                if (tri_tri_exact(PA[0], PA[1], PA[2], PB[0], PB[1], PB[2]) == 1) {
                    Contact c{};
                    c.query = query;
                    c.triA = ta;
                    c.triB = tb;
                    // the contact filter's planes and vertex-outside flags, as the
                    // reference forms them for a colliding pair (CreateUncollideRays.cpp:88-112)
                    V3 VA[3], VB[3];
                    for (int q = 0; q < 3; ++q) {
                        VA[q] = {PA[q][0], PA[q][1], PA[q][2]};
                        VB[q] = {PB[q][0], PB[q][1], PB[q][2]};
                    }
                    const PlaneF pa = plane_of(VA), pb = plane_of(VB);
                    for (int q = 0; q < 3; ++q) {
                        if (outside(pb, VA[q])) c.flags |= 1u << q;
                        if (outside(pa, VB[q])) c.flags |= 1u << (3 + q);
                    }
                    c.na[0] = pa.n.x; c.na[1] = pa.n.y; c.na[2] = pa.n.z;
                    c.nb[0] = pb.n.x; c.nb[1] = pb.n.y; c.nb[2] = pb.n.z;
                    segment_f64(PA, PB, c.s, c.t);
                    out.push_back(c);
                    ++k.contacts;
                }
            }
    }
}

}  // namespace orc

// ===========================================================================
// C ABI for ctypes (tests, smoke(), bench.py cpu_baseline only)
// ===========================================================================
using namespace orc;

struct OrcResult {
    std::vector<LeafPair> leafpairs;
    std::vector<uint32_t> lp_query;
    std::vector<Contact> contacts;
    std::vector<uint64_t> per_query;  // nq x 4: sat, leafpairs, tritests, contacts
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    Counters total;
};

// ---------------------------------------------------------------------------
// Node transforms (NodeDataCompEntity::GetGlobalMatrix, inMyRoom_vulkan/src/
// ECS/GeneralCompEntities/NodeDataCompEntity.cpp:63-72), glm's builders on the
// identity written out as glm writes them: translate sets column 3 to
// ((I0 x + I1 y) + I2 z) + I3, scale sets column i to Ii v[i], toMat4 is
// mat3_cast's nine terms (glm gtc/quaternion).
// ---------------------------------------------------------------------------
static M4 ident4() {
    M4 I;
    for (int i = 0; i < 16; ++i) I.m[i] = (i % 5 == 0) ? 1.f : 0.f;
    return I;
}
static M4 glm_translate_identity(float x, float y, float z) {
    const M4 I = ident4();
    M4 R = I;
    for (int r = 0; r < 4; ++r) R.m[12 + r] = ((I.m[0 + r] * x + I.m[4 + r] * y) + I.m[8 + r] * z) + I.m[12 + r];
    return R;
}
static M4 glm_scale_identity(float x, float y, float z) {
    const M4 I = ident4();
    M4 R = I;
    const float v[3] = {x, y, z};
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 4; ++r) R.m[c * 4 + r] = I.m[c * 4 + r] * v[c];
    return R;
}
static M4 glm_to_mat4(float x, float y, float z, float w) {
    const float qxx = x * x, qyy = y * y, qzz = z * z, qxz = x * z, qxy = x * y, qyz = y * z;
    const float qwx = w * x, qwy = w * y, qwz = w * z;
    M4 R = ident4();
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
// trs13: scale[3], rotation (x, y, z, w), translation[3], global translation[3]
static M4 get_global_matrix(const float* n, const M4& parent) {
    const M4 S = glm_scale_identity(n[0], n[1], n[2]);
    const M4 R = glm_to_mat4(n[3], n[4], n[5], n[6]);
    const M4 T = glm_translate_identity(n[7], n[8], n[9]);
    const M4 gT = glm_translate_identity(n[10], n[11], n[12]);
    return mat_mul(mat_mul(mat_mul(mat_mul(gT, parent), T), R), S);
}

extern "C" {

// Global matrices of n nodes (parents computed first, recursively);
// parent ~0u = root.  Returns 0, or -1 on a cycle.
int orc_node_globals(const float* trs13, const uint32_t* parent, int64_t n, float* out16) {
    std::vector<int> state(static_cast<size_t>(n), 0);  // 0 todo, 1 on the stack, 2 done
    std::vector<M4> G(static_cast<size_t>(n));
    std::function<bool(int64_t)> visit = [&](int64_t i) -> bool {
        if (state[size_t(i)] == 2) return true;
        if (state[size_t(i)] == 1) return false;
        state[size_t(i)] = 1;
        M4 P = ident4();
        if (parent[i] != 0xFFFFFFFFu) {
            if (!visit(int64_t(parent[i]))) return false;
            P = G[parent[i]];
        }
        G[size_t(i)] = get_global_matrix(trs13 + 13 * i, P);
        state[size_t(i)] = 2;
        return true;
    };
    for (int64_t i = 0; i < n; ++i)
        if (!visit(i)) return -1;
    for (int64_t i = 0; i < n; ++i) std::memcpy(out16 + 16 * i, G[size_t(i)].m, 16 * sizeof(float));
    return 0;
}

void orc_eig3(const double* A9, double* V9, double* d3) {
    double A[3][3], V[3][3];
    for (int i = 0; i < 9; ++i) A[i / 3][i % 3] = A9[i];
    eig3(A, V, d3);
    for (int i = 0; i < 9; ++i) V9[i] = V[i / 3][i % 3];
}

void orc_box_from_points(const float* pts, int64_t n, float* box12) {
    Box b = box_from_points(reinterpret_cast<const V3*>(pts), size_t(n));
    std::memcpy(box12, &b, sizeof b);
}

// float64 intersection segments of n intersecting, non-coplanar triangle
// pairs (9 floats each), 6 floats per pair: the low end, then the high end.
void orc_segment_f64(const float* a9, const float* b9, int64_t n, float* out6) {
    for (int64_t i = 0; i < n; ++i) {
        float PA[3][3], PB[3][3];
        std::memcpy(PA, a9 + 9 * i, sizeof PA);
        std::memcpy(PB, b9 + 9 * i, sizeof PB);
        segment_f64(PA, PB, out6 + 6 * i, out6 + 6 * i + 3);
    }
}

// The split's axis order (OBBtree.cpp:45-51): three half-lengths sorted
// descending; order3 receives the original positions (0 = u, 1 = v, 2 = w)
// in sorted order.
void orc_sort3_desc(const float* len3, int32_t* order3) {
    AxisCand c[3];
    for (int k = 0; k < 3; ++k) c[k] = {len3[k], {float(k), 0.f, 0.f}};
    sort3_desc(c);
    for (int k = 0; k < 3; ++k) order3[k] = int32_t(c[k].dir.x);
}

int orc_sat(const float* l12, const float* r12) {
    Box l, r;
    std::memcpy(&l, l12, sizeof l);
    std::memcpy(&r, r12, sizeof r);
    return sat(l, r) ? 1 : 0;
}
float orc_surface(const float* b12) {
    Box b;
    std::memcpy(&b, b12, sizeof b);
    return surface(b);
}
void orc_xform_box(const float* m16, const float* b12, float* out12) {
    M4 M;
    std::memcpy(M.m, m16, sizeof M.m);
    Box b;
    std::memcpy(&b, b12, sizeof b);
    Box r = xform_box(M, b);
    std::memcpy(out12, &r, sizeof r);
}
void orc_mat_inverse(const float* m16, float* out16) {
    M4 M;
    std::memcpy(M.m, m16, sizeof M.m);
    M4 R = mat_inverse(M);
    std::memcpy(out16, R.m, sizeof R.m);
}
void orc_mat_mul(const float* a16, const float* b16, float* out16) {
    M4 A, B;
    std::memcpy(A.m, a16, sizeof A.m);
    std::memcpy(B.m, b16, sizeof B.m);
    M4 R = mat_mul(A, B);
    std::memcpy(out16, R.m, sizeof R.m);
}
void orc_sap_axes(float* out9) {
    V3 ax[3];
    sap_axes(ax);
    std::memcpy(out9, ax, sizeof ax);
}

// flags: bit0 = intersect, bit1 = coplanar, bit2 = undefined-in-reference case
void orc_tritri(const float* a9, const float* b9, int64_t n, int32_t* flags, float* seg6) {
    for (int64_t i = 0; i < n; ++i) {
        const float* a = a9 + i * 9;
        const float* b = b9 + i * 9;
        int cop, ub;
        float s[3] = {0, 0, 0}, t[3] = {0, 0, 0};
        int hit = tri_tri_isect(a, a + 3, a + 6, b, b + 3, b + 6, &cop, &ub, s, t);
        flags[i] = (hit ? 1 : 0) | (cop ? 2 : 0) | (ub ? 4 : 0);
        if (seg6) {
            float* o = seg6 + i * 6;
            if (hit && !cop) {
                o[0] = s[0]; o[1] = s[1]; o[2] = s[2];
                o[3] = t[0]; o[4] = t[1]; o[5] = t[2];
            } else {
                for (int k = 0; k < 6; ++k) o[k] = 0.f;
            }
        }
    }
}

// exact semantics: 1 = intersect, 0 = disjoint, 2 = coplanar
void orc_tritri_exact(const float* a9, const float* b9, int64_t n, int32_t* out) {
    for (int64_t i = 0; i < n; ++i) {
        const float* a = a9 + i * 9;
        const float* b = b9 + i * 9;
        out[i] = tri_tri_exact(a, a + 3, a + 6, b, b + 3, b + 6);
    }
}
int orc_orient_exact(const float* a, const float* b, const float* c, const float* d) { return orient_exact(a, b, c, d); }
int orc_orient_bignum(const float* a, const float* b, const float* c, const float* d) { return orient_bignum(a, b, c, d); }
uint64_t orc_bignum_calls(void) { return g_bignum_calls; }

void* orc_tree_build(const float* tri9, int64_t ntri, int leaf_max) {
    return tree_build(tri9, size_t(ntri), unsigned(leaf_max));
}
void orc_tree_free(void* t) { delete static_cast<Tree*>(t); }
int64_t orc_tree_nodes(void* t) { return int64_t(static_cast<Tree*>(t)->nodes.size()); }
int64_t orc_tree_tris(void* t) { return int64_t(static_cast<Tree*>(t)->perm.size()); }
int32_t orc_tree_depth(void* t) { return static_cast<Tree*>(t)->depth; }
double orc_tree_slack(void* t) { return static_cast<Tree*>(t)->slack; }
// node_obbs: N x 24 floats (left box, right box); node_links: N x 4 (linkL, linkR, cntL, cntR)
void orc_tree_export(void* tp, float* root12, float* node_obbs, uint32_t* node_links, uint32_t* perm) {
    const Tree& T = *static_cast<Tree*>(tp);
    std::memcpy(root12, &T.root, sizeof(Box));
    for (size_t i = 0; i < T.nodes.size(); ++i) {
        std::memcpy(node_obbs + i * 24, &T.nodes[i].obb[0], sizeof(Box));
        std::memcpy(node_obbs + i * 24 + 12, &T.nodes[i].obb[1], sizeof(Box));
        node_links[i * 4 + 0] = T.nodes[i].link[0];
        node_links[i * 4 + 1] = T.nodes[i].link[1];
        node_links[i * 4 + 2] = T.nodes[i].cnt[0];
        node_links[i * 4 + 3] = T.nodes[i].cnt[1];
    }
    std::memcpy(perm, T.perm.data(), T.perm.size() * sizeof(uint32_t));
}

// ---- ray generation ----
void orc_tree_set_attrs(void* tp, const float* nrm9, const uint32_t* vid3) {
    set_attrs(*static_cast<Tree*>(tp), nrm9, vid3);
}
struct OrcRays { std::vector<PairRays> pairs; };
static Tree& with_attrs(void* tp) {  // a tree without attributes gets face normals and soup vertex ids
    Tree& T = *static_cast<Tree*>(tp);
    if (T.nrm.size() != T.perm.size() * 9) set_attrs(T, nullptr, nullptr);
    return T;
}
// Per query (mesh pair + relative matrix): CDentriesUncollideRays of every
// query, rays or not.
void* orc_rays_queries(void** trees, const uint32_t* qa, const uint32_t* qb, const float* m16, int64_t nq) {
    auto* R = new OrcRays();
    for (int64_t q = 0; q < nq; ++q) {
        const Tree& A = with_attrs(trees[qa[q]]);
        const Tree& B = with_attrs(trees[qb[q]]);
        M4 M;
        std::memcpy(M.m, m16 + q * 16, sizeof M.m);
        Counters k;
        std::vector<LeafPair> lps;
        traverse(A, B, M, lps, k);
        PairRays pr = uncollide_rays(A, B, M, lps);
        pr.first = qa[q];
        pr.second = qb[q];
        R->pairs.push_back(std::move(pr));
    }
    return R;
}
// A frame (CollisionDetection.cpp:38-67): broad phase, mid phase, rays; pairs
// without a ray are dropped (:63).
void* orc_rays_frame(void** trees, const uint32_t* mesh_of, const float* cur16, const uint8_t* cb, int64_t n) {
    auto* R = new OrcRays();
    if (n < 2) return R;
    std::vector<M4> cur(static_cast<size_t>(n));
    std::vector<const Box*> roots(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
        std::memcpy(cur[size_t(i)].m, cur16 + i * 16, sizeof(float) * 16);
        roots[size_t(i)] = &static_cast<Tree*>(trees[mesh_of[i]])->root;
    }
    for (const auto& pr : sap(cur.data(), roots.data(), cb, size_t(n))) {
        const Tree& A = with_attrs(trees[mesh_of[pr.first]]);
        const Tree& B = with_attrs(trees[mesh_of[pr.second]]);
        M4 M = mat_mul(mat_inverse(cur[pr.first]), cur[pr.second]);
        Counters k;
        std::vector<LeafPair> lps;
        traverse(A, B, M, lps, k);
        if (lps.empty()) continue;  // :52
        PairRays r = uncollide_rays(A, B, M, lps);
        if (r.from_first.empty() && r.from_second.empty()) continue;  // :63
        r.first = pr.first;
        r.second = pr.second;
        R->pairs.push_back(std::move(r));
    }
    return R;
}
// Closed-form probes of the ray pieces.  out6: hit, back, dist, bx, by, (unused)
void orc_ray_triangle(const float* dir3, const float* tri9, float* out6) {
    float bx = 0.f, by = 0.f, dist = std::numeric_limits<float>::infinity();
    bool back = false;
    const bool hit = ray_triangle({dir3[0], dir3[1], dir3[2]}, {tri9[0], tri9[1], tri9[2]}, {tri9[3], tri9[4], tri9[5]},
                                  {tri9[6], tri9[7], tri9[8]}, bx, by, dist, back);
    out6[0] = hit ? 1.f : 0.f;
    out6[1] = back ? 1.f : 0.f;
    out6[2] = dist;
    out6[3] = bx;
    out6[4] = by;
    out6[5] = 0.f;
}
int orc_ray_box(const float* dir3, const float* b12, float* out2) {
    Box b;
    std::memcpy(&b, b12, sizeof b);
    return ray_box({dir3[0], dir3[1], dir3[2]}, b, out2[0], out2[1]) ? 1 : 0;
}
// out8: hit, back, dist, tri, bx, by
void orc_ray_tree(void* tp, const float* m16, const float* o3, const float* d3, float* out8) {
    M4 M;
    std::memcpy(M.m, m16, sizeof M.m);
    const RayHit h = ray_tree(*static_cast<Tree*>(tp), M, {o3[0], o3[1], o3[2]}, {d3[0], d3[1], d3[2]});
    out8[0] = h.hit ? 1.f : 0.f;
    out8[1] = h.back ? 1.f : 0.f;
    out8[2] = h.dist;
    out8[3] = float(h.tri == 0xFFFFFFFFu ? -1.0 : double(h.tri));
    out8[4] = h.bx;
    out8[5] = h.by;
    out8[6] = out8[7] = 0.f;
}
// A frame's response: broad and mid phase, rays, then for every pair with
// rays the shot delta and its split (CollisionDetection.cpp:38-99).
struct OrcShoot {
    std::vector<PairResponse> pairs;
    ShootStats stats;
};
void* orc_shoot_frame(void** trees, const uint32_t* mesh_of, const float* cur16, const float* prev16, const uint8_t* cb,
                      int64_t n, float cos_start, float cos_finish, float bias) {
    auto* S = new OrcShoot();
    auto* R = static_cast<OrcRays*>(orc_rays_frame(trees, mesh_of, cur16, cb, n));
    for (const PairRays& pr : R->pairs) {
        M4 c1, p1, c2, p2;
        std::memcpy(c1.m, cur16 + size_t(pr.first) * 16, sizeof c1.m);
        std::memcpy(p1.m, prev16 + size_t(pr.first) * 16, sizeof p1.m);
        std::memcpy(c2.m, cur16 + size_t(pr.second) * 16, sizeof c2.m);
        std::memcpy(p2.m, prev16 + size_t(pr.second) * 16, sizeof p2.m);
        S->pairs.push_back(respond(with_attrs(trees[mesh_of[pr.first]]), with_attrs(trees[mesh_of[pr.second]]), c1, p1,
                                   c2, p2, pr, cos_start, cos_finish, bias, &S->stats));
    }
    delete R;
    return S;
}
// out: per pair first, second, shot (u32 x 3) and delta, delta_first, delta_second (f32 x 9); stats2: passes, successes
int64_t orc_shoot_get(void* sp, uint32_t* ids3, float* d9, uint64_t* stats2) {
    const OrcShoot& S = *static_cast<OrcShoot*>(sp);
    if (ids3 && d9)
        for (size_t i = 0; i < S.pairs.size(); ++i) {
            const PairResponse& r = S.pairs[i];
            ids3[i * 3 + 0] = r.first;
            ids3[i * 3 + 1] = r.second;
            ids3[i * 3 + 2] = r.shot ? 1u : 0u;
            for (int k = 0; k < 3; ++k) {
                d9[i * 9 + k] = r.delta[k];
                d9[i * 9 + 3 + k] = r.delta_first[k];
                d9[i * 9 + 6 + k] = r.delta_second[k];
            }
        }
    if (stats2) {
        stats2[0] = S.stats.passes;
        stats2[1] = S.stats.successes;
    }
    return int64_t(S.pairs.size());
}
void orc_shoot_free(void* sp) { delete static_cast<OrcShoot*>(sp); }
// what: 0 pairs, 1 rays (both directions, all pairs)
int64_t orc_rays_count(void* rp, int what) {
    const OrcRays& R = *static_cast<OrcRays*>(rp);
    if (what == 0) return int64_t(R.pairs.size());
    int64_t n = 0;
    for (const auto& p : R.pairs) n += int64_t(p.from_first.size() + p.from_second.size());
    return n;
}
// pairs4: first, second, rays from first, rays from second; avg6: average
// point of each side (first's and second's model space); rays6: per pair the
// first's rays then the second's (origin, direction)
void orc_rays_get(void* rp, uint32_t* pairs4, float* avg6, float* rays6) {
    const OrcRays& R = *static_cast<OrcRays*>(rp);
    size_t k = 0;
    for (size_t i = 0; i < R.pairs.size(); ++i) {
        const PairRays& p = R.pairs[i];
        pairs4[i * 4 + 0] = p.first;
        pairs4[i * 4 + 1] = p.second;
        pairs4[i * 4 + 2] = uint32_t(p.from_first.size());
        pairs4[i * 4 + 3] = uint32_t(p.from_second.size());
        for (int c = 0; c < 3; ++c) {
            avg6[i * 6 + c] = p.avg_first[c];
            avg6[i * 6 + 3 + c] = p.avg_second[c];
        }
        for (const auto* v : {&p.from_first, &p.from_second})
            for (const RayRec& r : *v) {
                std::memcpy(rays6 + k * 6, r.o, sizeof r.o);
                std::memcpy(rays6 + k * 6 + 3, r.d, sizeof r.d);
                ++k;
            }
    }
}
void orc_rays_free(void* rp) { delete static_cast<OrcRays*>(rp); }

// Batch of mesh-pair queries.  mode 0 = reference semantics (Möller), mode 1
// = exact semantics.  keep_leafpairs stores the candidate list (mode 0).
void* orc_batch_run(void** trees, const uint32_t* qa, const uint32_t* qb, const float* m16, int64_t nq, int mode,
                    int nthreads, int keep_leafpairs) {
    auto* R = new OrcResult();
    R->per_query.assign(size_t(nq) * 4, 0);
    if (nthreads < 1) nthreads = 1;
    std::vector<OrcResult> part{static_cast<size_t>(nthreads)};
    auto work = [&](int tid) {
        OrcResult& P = part[size_t(tid)];
        for (int64_t q = tid; q < nq; q += nthreads) {
            const Tree& A = *static_cast<Tree*>(trees[qa[q]]);
            const Tree& B = *static_cast<Tree*>(trees[qb[q]]);
            M4 M;
            std::memcpy(M.m, m16 + q * 16, sizeof M.m);
            Counters k;
            if (mode == 0) {
                std::vector<LeafPair> lps;
                traverse(A, B, M, lps, k);
                k.leafpairs = lps.size();
                narrow(A, B, M, lps, uint32_t(q), P.contacts, k);
                if (keep_leafpairs)
                    for (auto& lp : lps) {
                        P.leafpairs.push_back(lp);
                        P.lp_query.push_back(uint32_t(q));
                    }
            } else {
                query_exact(A, B, M, uint32_t(q), P.contacts, k);
            }
            uint64_t* pq = &R->per_query[size_t(q) * 4];
            pq[0] = k.sat; pq[1] = k.leafpairs; pq[2] = k.tritests; pq[3] = k.contacts;
            P.total.sat += k.sat; P.total.leafpairs += k.leafpairs; P.total.tritests += k.tritests;
            P.total.contacts += k.contacts; P.total.ub += k.ub;
        }
    };
    if (nthreads == 1) work(0);
    else {
        std::vector<std::thread> th;
        for (int t = 0; t < nthreads; ++t) th.emplace_back(work, t);
        for (auto& t : th) t.join();
    }
    for (auto& P : part) {
        R->contacts.insert(R->contacts.end(), P.contacts.begin(), P.contacts.end());
        R->leafpairs.insert(R->leafpairs.end(), P.leafpairs.begin(), P.leafpairs.end());
        R->lp_query.insert(R->lp_query.end(), P.lp_query.begin(), P.lp_query.end());
        R->total.sat += P.total.sat; R->total.leafpairs += P.total.leafpairs; R->total.tritests += P.total.tritests;
        R->total.contacts += P.total.contacts; R->total.ub += P.total.ub;
    }
    return R;
}

// Frame: broad + mid + narrow (CollisionDetection.cpp:38-67).  Entities give a
// current global matrix (16 floats), a tree and a shouldCallback flag.
void* orc_frame_run(void** trees, const uint32_t* mesh_of, const float* cur16, const uint8_t* cb, int64_t n, int mode) {
    auto* R = new OrcResult();
    if (n < 2) return R;
    std::vector<M4> cur(static_cast<size_t>(n));
    std::vector<const Box*> roots(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
        std::memcpy(cur[size_t(i)].m, cur16 + i * 16, sizeof(float) * 16);
        roots[size_t(i)] = &static_cast<Tree*>(trees[mesh_of[i]])->root;
    }
    R->pairs = sap(cur.data(), roots.data(), cb, size_t(n));
    R->per_query.assign(R->pairs.size() * 4, 0);
    for (size_t p = 0; p < R->pairs.size(); ++p) {
        uint32_t a = R->pairs[p].first, b = R->pairs[p].second;
        const Tree& A = *static_cast<Tree*>(trees[mesh_of[a]]);
        const Tree& B = *static_cast<Tree*>(trees[mesh_of[b]]);
        M4 M = mat_mul(mat_inverse(cur[a]), cur[b]);  // OBBtreesCollision.cpp:15
        Counters k;
        if (mode == 0) {
            std::vector<LeafPair> lps;
            traverse(A, B, M, lps, k);
            k.leafpairs = lps.size();
            narrow(A, B, M, lps, uint32_t(p), R->contacts, k);
        } else {
            query_exact(A, B, M, uint32_t(p), R->contacts, k);
        }
        uint64_t* pq = &R->per_query[p * 4];
        pq[0] = k.sat; pq[1] = k.leafpairs; pq[2] = k.tritests; pq[3] = k.contacts;
        R->total.sat += k.sat; R->total.leafpairs += k.leafpairs; R->total.tritests += k.tritests;
        R->total.contacts += k.contacts; R->total.ub += k.ub;
    }
    return R;
}

void orc_result_free(void* r) { delete static_cast<OrcResult*>(r); }
// what: 0 contacts, 1 leafpairs, 2 pairs, 3 total sat, 4 total leafpairs, 5 total tritests, 6 total ub
int64_t orc_result_count(void* rp, int what) {
    const OrcResult& R = *static_cast<OrcResult*>(rp);
    switch (what) {
        case 0: return int64_t(R.contacts.size());
        case 1: return int64_t(R.leafpairs.size());
        case 2: return int64_t(R.pairs.size());
        case 3: return int64_t(R.total.sat);
        case 4: return int64_t(R.total.leafpairs);
        case 5: return int64_t(R.total.tritests);
        case 6: return int64_t(R.total.ub);
    }
    return -1;
}
// contacts as 16 x 32-bit words each (Contact layout)
void orc_result_contacts(void* rp, void* out) {
    const OrcResult& R = *static_cast<OrcResult*>(rp);
    std::memcpy(out, R.contacts.data(), R.contacts.size() * sizeof(Contact));
}
void orc_result_leafpairs(void* rp, uint32_t* out4, uint32_t* query) {
    const OrcResult& R = *static_cast<OrcResult*>(rp);
    std::memcpy(out4, R.leafpairs.data(), R.leafpairs.size() * sizeof(LeafPair));
    if (query) std::memcpy(query, R.lp_query.data(), R.lp_query.size() * sizeof(uint32_t));
}
void orc_result_pairs(void* rp, uint32_t* out2) {
    const OrcResult& R = *static_cast<OrcResult*>(rp);
    for (size_t i = 0; i < R.pairs.size(); ++i) {
        out2[2 * i] = R.pairs[i].first;
        out2[2 * i + 1] = R.pairs[i].second;
    }
}
void orc_result_per_query(void* rp, uint64_t* out) {
    const OrcResult& R = *static_cast<OrcResult*>(rp);
    std::memcpy(out, R.per_query.data(), R.per_query.size() * sizeof(uint64_t));
}

}  // extern "C"
