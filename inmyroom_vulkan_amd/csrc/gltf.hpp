// glTF 2.0 reader for collision meshes (host code): .gltf with external or
// data: URI buffers, and .glb.  It reads what the reference's mesh ingest
// reads and nothing else: per mesh, per primitive, the draw mode, POSITION,
// NORMAL and the index list, and whether the primitive is skinned or morphed.
//
// Reference behaviour restated here (inMyRoom_vulkan/src):
//   - draw mode: absent -> triangles, line loop -> line strip
//     (Graphics/Meshes/PrimitivesOfMeshes.cpp:47-54)
//   - indices: unsigned short or unsigned int scalars (:57-68); positions and
//     normals: float VEC3, stored as (x, -y, -z) (:70-86, :121-137)
//   - an accessor is read as tightly packed elements starting at
//     accessor.byteOffset + bufferView.byteOffset (:695-752); byteStride is
//     not consulted, so a strided (interleaved) view is rejected here rather
//     than read differently
//   - skinned (JOINTS_0, JOINTS_1, ...) or position-morphed primitives give
//     no collision triangles (:528-552, :96-110, :642)
//   - primitives are concatenated in the order MeshesOfNodes sorts them
//     (Graphics/Meshes/MeshesOfNodes.cpp:41-43), see tree_order()
// Departures, each where the reference's result is undefined: a primitive
// without indices gets 0 .. nvert-1 (the reference sizes that list by the
// 4-float position array, :660-663, and reads past the vertices); strips and
// fans of fewer than 2 indices and line strips of none give no triangles
// (the reference's size() - 2 wraps around, Geometry/Triangle.cpp:31-55).
#pragma once

#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "geom.hpp"
#include "json.hpp"

namespace cdk {
namespace gltf {

enum : uint32_t { POINTS = 0, LINES = 1, LINE_LOOP = 2, LINE_STRIP = 3, TRIANGLES = 4, TRIANGLE_STRIP = 5, TRIANGLE_FAN = 6 };

struct Primitive {
    int32_t gltf_mode = -1;      // as written, -1 when absent
    uint32_t mode = TRIANGLES;   // draw mode after the line-loop fallback
    bool excluded = false;       // skinned or morphed: no collision triangles
    bool skin_attr = false;      // JOINTS_0 or WEIGHTS_0
    bool targets = false;        // morph targets of any attribute
    uint32_t nvert = 0;
    std::vector<float> pos;      // nvert x 3, (x, -y, -z)
    std::vector<float> nrm;      // nvert x 3, (x, -y, -z), or empty
    std::vector<uint32_t> idx;   // index list
    uint64_t ntri = 0;           // triangles its mode makes from idx
};

struct Mesh {
    std::string name;
    std::vector<Primitive> prims;  // in the order the tree concatenates them
    uint64_t ntri = 0, nvert = 0;  // over the primitives that are not excluded
    uint32_t with_normals = 0, without_normals = 0;  // included primitives with / without NORMAL
    bool skinned = false;  // some primitive has JOINTS_0 or WEIGHTS_0 (PrimitivesOfMeshes.cpp:1189-1198)
    bool morphed = false;  // some primitive has morph targets (:1200-1210)
};

// A node as its NodeData component holds it after import
// (GameImporter::CreatePositionInitMap, src/GameImporter.cpp:503-552, then
// NodeDataCompEntity::CreateComponentEntityByMap, NodeDataCompEntity.cpp:19-61).
struct Node {
    float scale[3] = {1.f, 1.f, 1.f};
    float rotation[4] = {0.f, 0.f, 0.f, 1.f};  // unit quaternion (x, y, z, w)
    float translation[3] = {0.f, 0.f, 0.f};
    uint32_t parent = 0xFFFFFFFFu;
    int32_t mesh = -1;
    bool collides = false;  // gets a ModelCollision component (GameImporter.cpp:320-326)
};

struct Document {
    std::vector<Mesh> meshes;
    std::vector<Node> nodes;
};

// Triangles Triangle::CreateIndicesTriplets makes from n indices in a mode
// (src/Geometry/Triangle.cpp:10-62).
inline uint64_t triangles_of(uint32_t mode, uint64_t n) {
    switch (mode) {
        case POINTS: return n;
        case LINES: return n / 2;
        case LINE_STRIP: return n ? n - 1 : 0;
        case TRIANGLES: return n / 3;
        case TRIANGLE_STRIP:
        case TRIANGLE_FAN: return n >= 2 ? n - 2 : 0;
        default: return 0;
    }
}

// The concatenation order of MeshesOfNodes::AddMeshesOfModel: std::sort with
// the comparator "lhs is triangles-mode (or has no mode)" (MeshesOfNodes.cpp:
// 41-43).  That comparator depends on its left operand only, so it is not a
// strict weak ordering; libstdc++ sorts up to 16 elements by insertion, and
// insertion under it moves every triangles-mode primitive after the first
// to the front, one by one: the result is those primitives in reverse order,
// then the first primitive, then the others in their order.  The same rule
// is applied to longer lists, where the reference's behaviour is undefined.
inline std::vector<uint32_t> tree_order(const std::vector<int32_t>& gltf_modes) {
    const size_t n = gltf_modes.size();
    std::vector<uint32_t> front, back;
    for (size_t i = 1; i < n; ++i) {
        const bool tri = gltf_modes[i] == int32_t(TRIANGLES) || gltf_modes[i] == -1;
        (tri ? front : back).push_back(uint32_t(i));
    }
    std::vector<uint32_t> order(front.rbegin(), front.rend());
    if (n) order.push_back(0);
    order.insert(order.end(), back.begin(), back.end());
    return order;
}

namespace detail {

inline bool read_file(const std::string& path, std::vector<uint8_t>& out) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    out.clear();
    uint8_t buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) out.insert(out.end(), buf, buf + n);
    const bool ok = !std::ferror(f);
    std::fclose(f);
    return ok;
}

inline int b64(char c) {
    if (c >= 'A' && c <= 'Z') return c - 'A';
    if (c >= 'a' && c <= 'z') return c - 'a' + 26;
    if (c >= '0' && c <= '9') return c - '0' + 52;
    if (c == '+' || c == '-') return 62;
    if (c == '/' || c == '_') return 63;
    return -1;
}

inline bool base64(const std::string& s, size_t from, std::vector<uint8_t>& out) {
    uint32_t acc = 0;
    int bits = 0;
    for (size_t i = from; i < s.size(); ++i) {
        if (s[i] == '=') break;
        const int v = b64(s[i]);
        if (v < 0) return false;
        acc = (acc << 6) | uint32_t(v);
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back(uint8_t(acc >> bits));
        }
    }
    return true;
}

inline std::string percent_decode(const std::string& s) {
    std::string o;
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] == '%' && i + 2 < s.size()) {
            const int h = std::isxdigit(static_cast<unsigned char>(s[i + 1])), l = std::isxdigit(static_cast<unsigned char>(s[i + 2]));
            if (h && l) {
                o += char(std::strtol(s.substr(i + 1, 2).c_str(), nullptr, 16));
                i += 2;
                continue;
            }
        }
        o += s[i];
    }
    return o;
}

inline bool get_int(const json::Value& o, const char* key, int64_t& out, int64_t dflt, bool required, std::string& err,
                    const std::string& where) {
    const json::Value* v = o.find(key);
    if (!v) {
        if (required) {
            err = where + ": missing \"" + key + "\"";
            return false;
        }
        out = dflt;
        return true;
    }
    if (!v->is_int()) {
        err = where + ": \"" + key + "\" is not an integer";
        return false;
    }
    out = int64_t(v->num);
    return true;
}

struct Ctx {
    const json::Value* root;
    std::vector<std::vector<uint8_t>> buffers;
    std::string err;
};

inline const json::Value* nth(const json::Value& root, const char* list, int64_t i) {
    const json::Value* a = root.find(list);
    if (!a || a->kind != json::Value::Array || i < 0 || uint64_t(i) >= a->arr.size()) return nullptr;
    return &a->arr[size_t(i)];
}

// One accessor's elements as a packed run inside its buffer, checked against
// the buffer view and the buffer.
inline bool accessor(Ctx& c, int64_t ai, uint32_t want_comps, const std::vector<int64_t>& comp_types, const uint8_t*& p,
                     uint64_t& count, int64_t& comp_type, const std::string& what) {
    const std::string where = what + " (accessor " + std::to_string(ai) + ")";
    const json::Value* a = nth(*c.root, "accessors", ai);
    if (!a) { c.err = where + ": no such accessor"; return false; }
    if (a->find("sparse")) { c.err = where + ": sparse accessors are not read by the mesh ingest"; return false; }
    int64_t bvi, aoff, cnt;
    if (!get_int(*a, "bufferView", bvi, -1, false, c.err, where)) return false;
    if (bvi < 0) { c.err = where + ": no buffer view"; return false; }
    if (!get_int(*a, "byteOffset", aoff, 0, false, c.err, where) || !get_int(*a, "count", cnt, 0, true, c.err, where) ||
        !get_int(*a, "componentType", comp_type, 0, true, c.err, where))
        return false;
    bool ok_type = false;
    for (int64_t t : comp_types) ok_type |= comp_type == t;
    if (!ok_type) { c.err = where + ": component type " + std::to_string(comp_type) + " is not supported"; return false; }
    const json::Value* ty = a->find("type");
    const char* names[] = {"", "SCALAR", "VEC2", "VEC3", "VEC4"};
    if (!ty || ty->kind != json::Value::String || want_comps > 4 || ty->str != names[want_comps]) {
        c.err = where + ": expected type " + names[want_comps];
        return false;
    }
    const json::Value* bv = nth(*c.root, "bufferViews", bvi);
    if (!bv) { c.err = where + ": no such buffer view"; return false; }
    int64_t bi, boff, blen, stride;
    if (!get_int(*bv, "buffer", bi, 0, true, c.err, where) || !get_int(*bv, "byteOffset", boff, 0, false, c.err, where) ||
        !get_int(*bv, "byteLength", blen, 0, true, c.err, where) || !get_int(*bv, "byteStride", stride, 0, false, c.err, where))
        return false;
    const uint64_t csize = (comp_type == 5120 || comp_type == 5121) ? 1 : (comp_type == 5122 || comp_type == 5123) ? 2 : 4;
    const uint64_t esize = csize * want_comps;
    if (stride != 0 && uint64_t(stride) != esize) {
        c.err = where + ": interleaved buffer view (byteStride " + std::to_string(stride) +
                "); the mesh ingest reads packed elements only";
        return false;
    }
    if (bi < 0 || uint64_t(bi) >= c.buffers.size()) { c.err = where + ": no such buffer"; return false; }
    if (cnt < 0 || aoff < 0 || boff < 0 || blen < 0) { c.err = where + ": negative size or offset"; return false; }
    if (cnt > (int64_t(1) << 40)) { c.err = where + ": count too large"; return false; }
    const uint64_t need = uint64_t(aoff) + uint64_t(cnt) * esize;
    if (need > uint64_t(blen)) { c.err = where + ": runs past its buffer view"; return false; }
    if (uint64_t(boff) + uint64_t(blen) > c.buffers[size_t(bi)].size()) {
        c.err = where + ": buffer view runs past its buffer";
        return false;
    }
    p = c.buffers[size_t(bi)].data() + boff + aoff;
    count = uint64_t(cnt);
    return true;
}

inline bool vec3_attr(Ctx& c, int64_t ai, std::vector<float>& out, uint64_t& count, const std::string& what) {
    const uint8_t* p;
    int64_t ct;
    if (!accessor(c, ai, 3, {5126}, p, count, ct, what)) return false;
    out.resize(size_t(count) * 3);
    if (count) std::memcpy(out.data(), p, size_t(count) * 12);  // an empty vector's data() may be null
    for (uint64_t v = 0; v < count; ++v) {  // (x, y, z) -> (x, -y, -z)
        out[size_t(v) * 3 + 1] = -out[size_t(v) * 3 + 1];
        out[size_t(v) * 3 + 2] = -out[size_t(v) * 3 + 2];
    }
    return true;
}

inline bool primitive(Ctx& c, const json::Value& pv, Primitive& pr, const std::string& where) {
    int64_t mode, ind;
    if (!get_int(pv, "mode", mode, -1, false, c.err, where) || !get_int(pv, "indices", ind, -1, false, c.err, where))
        return false;
    if (mode < -1 || mode > 6) { c.err = where + ": invalid mode " + std::to_string(mode); return false; }
    pr.gltf_mode = int32_t(mode);
    pr.mode = mode == -1 ? TRIANGLES : mode == LINE_LOOP ? LINE_STRIP : uint32_t(mode);
    const json::Value* attrs = pv.find("attributes");
    if (!attrs || attrs->kind != json::Value::Object) { c.err = where + ": no attributes"; return false; }
    const bool skinned = attrs->find("JOINTS_0") != nullptr;
    bool morphed = false;
    pr.skin_attr = skinned || attrs->find("WEIGHTS_0") != nullptr;
    if (const json::Value* t = pv.find("targets")) {
        morphed = t->kind == json::Value::Array && !t->arr.empty() && t->arr[0].find("POSITION") != nullptr;
        pr.targets = t->kind == json::Value::Array && !t->arr.empty() && t->arr[0].kind == json::Value::Object &&
                     !t->arr[0].obj.empty();
    }
    if (skinned || morphed) {
        pr.excluded = true;
        return true;
    }
    int64_t pa, na;
    if (!get_int(*attrs, "POSITION", pa, -1, true, c.err, where) || !get_int(*attrs, "NORMAL", na, -1, false, c.err, where))
        return false;
    uint64_t nv;
    if (!vec3_attr(c, pa, pr.pos, nv, where + " POSITION")) return false;
    if (nv > 0xFFFFFFFFull) { c.err = where + ": too many vertices"; return false; }
    pr.nvert = uint32_t(nv);
    for (float x : pr.pos)
        if (!(x - x == 0.f)) { c.err = where + ": non-finite position"; return false; }
    if (na >= 0) {
        uint64_t nn;
        if (!vec3_attr(c, na, pr.nrm, nn, where + " NORMAL")) return false;
        if (nn != nv) { c.err = where + ": NORMAL count differs from POSITION count"; return false; }
        for (float x : pr.nrm)
            if (!(x - x == 0.f)) { c.err = where + ": non-finite normal"; return false; }
    }
    if (ind >= 0) {
        const uint8_t* p;
        uint64_t n;
        int64_t ct;
        if (!accessor(c, ind, 1, {5123, 5125}, p, n, ct, where + " indices")) return false;
        pr.idx.resize(size_t(n));
        if (ct == 5123) {
            for (uint64_t i = 0; i < n; ++i) {
                uint16_t v;
                std::memcpy(&v, p + 2 * i, 2);
                pr.idx[size_t(i)] = v;
            }
        } else {
            if (n) std::memcpy(pr.idx.data(), p, size_t(n) * 4);  // an empty vector's data() may be null
        }
        for (uint32_t v : pr.idx)
            if (v >= pr.nvert) { c.err = where + ": index " + std::to_string(v) + " out of range"; return false; }
    } else {
        pr.idx.resize(pr.nvert);
        for (uint32_t v = 0; v < pr.nvert; ++v) pr.idx[v] = v;
    }
    pr.ntri = triangles_of(pr.mode, pr.idx.size());
    return true;
}

// glm::normalize of a quaternion: 1 / sqrt((w w + x x) + (y y + z z)) times
// each component; (1, 0, 0, 0) when the length is not positive.
inline void quat_normalize(float q[4]) {  // q = (x, y, z, w)
    const float d = (q[3] * q[3] + q[0] * q[0]) + (q[1] * q[1] + q[2] * q[2]);
    const float len = std::sqrt(d);
    if (!(len > 0.f)) {
        q[0] = q[1] = q[2] = 0.f;
        q[3] = 1.f;
        return;
    }
    const float inv = 1.f / len;
    const float w = q[3] * inv, x = q[0] * inv, y = q[1] * inv, z = q[2] * inv;
    q[0] = x;
    q[1] = y;
    q[2] = z;
    q[3] = w;
}

// glm::decompose (gtx/matrix_decompose, after Thomas, "Decomposing a matrix
// into simple transformations", Graphics Gems II) of an affine float matrix,
// keeping the parts the node uses: scale, rotation, translation.  Returns
// false, leaving the outputs alone, where glm gives up (w or the 3x3
// determinant within FLT_EPSILON of zero).
inline bool decompose(M4 L, float scale[3], float q[4], float t[3]) {
    const float eps = 1.1920928955078125e-7f;
    if (std::fabs(L.m[15]) < eps) return false;
    for (int i = 0; i < 16; ++i) L.m[i] = L.m[i] / L.m[15];
    const f3 c0 = mk(L.m[0], L.m[1], L.m[2]), c1 = mk(L.m[4], L.m[5], L.m[6]), c2 = mk(L.m[8], L.m[9], L.m[10]);
    if (std::fabs(dot(c0, cross(c1, c2))) < eps) return false;
    if (std::fabs(L.m[3]) >= eps || std::fabs(L.m[7]) >= eps || std::fabs(L.m[11]) >= eps) {
        L.m[3] = L.m[7] = L.m[11] = 0.f;  // perspective part: solved for and dropped
        L.m[15] = 1.f;
    }
    t[0] = L.m[12];
    t[1] = L.m[13];
    t[2] = L.m[14];
    f3 row[3] = {c0, c1, c2};
    auto unit = [](f3 v) {
        const float len = vlength(v);
        return mk(v.x / len, v.y / len, v.z / len);
    };
    auto combine = [](f3 a, f3 b, float bs) { return mk(a.x + b.x * bs, a.y + b.y * bs, a.z + b.z * bs); };
    scale[0] = vlength(row[0]);
    row[0] = unit(row[0]);
    float skew_z = dot(row[0], row[1]);
    row[1] = combine(row[1], row[0], -skew_z);
    scale[1] = vlength(row[1]);
    row[1] = unit(row[1]);
    float skew_y = dot(row[0], row[2]);
    row[2] = combine(row[2], row[0], -skew_y);
    float skew_x = dot(row[1], row[2]);
    row[2] = combine(row[2], row[1], -skew_x);
    scale[2] = vlength(row[2]);
    row[2] = unit(row[2]);
    if (dot(row[0], cross(row[1], row[2])) < 0.f) {
        for (int i = 0; i < 3; ++i) {
            scale[i] = scale[i] * -1.f;
            row[i] = mk(row[i].x * -1.f, row[i].y * -1.f, row[i].z * -1.f);
        }
    }
    auto at = [&](int i, int j) { const f3& r = row[i]; return j == 0 ? r.x : (j == 1 ? r.y : r.z); };
    const float trace = (row[0].x + row[1].y) + row[2].z;
    if (trace > 0.f) {
        float root = std::sqrt(trace + 1.f);
        q[3] = 0.5f * root;
        root = 0.5f / root;
        q[0] = root * (row[1].z - row[2].y);
        q[1] = root * (row[2].x - row[0].z);
        q[2] = root * (row[0].y - row[1].x);
    } else {
        const int next[3] = {1, 2, 0};
        int i = 0;
        if (row[1].y > row[0].x) i = 1;
        if (row[2].z > at(i, i)) i = 2;
        const int j = next[i], k = next[j];
        float root = std::sqrt(((at(i, i) - at(j, j)) - at(k, k)) + 1.f);
        q[i] = 0.5f * root;
        root = 0.5f / root;
        q[j] = root * (at(i, j) + at(j, i));
        q[k] = root * (at(i, k) + at(k, i));
        q[3] = root * (at(j, k) - at(k, j));
    }
    return true;
}

inline bool number_array(const json::Value& o, const char* key, size_t n, double* out, std::string& err,
                         const std::string& where) {
    const json::Value* v = o.find(key);
    if (!v) return false;
    if (v->kind != json::Value::Array || v->arr.size() != n) {
        err = where + ": \"" + key + "\" is not an array of " + std::to_string(n) + " numbers";
        return true;
    }
    for (size_t i = 0; i < n; ++i) {
        if (v->arr[i].kind != json::Value::Number) {
            err = where + ": \"" + key + "\" holds a non-number";
            return true;
        }
        out[i] = v->arr[i].num;
    }
    return true;
}

// The node's NodeData after import: a matrix is turned into (x, -y, -z)
// space by diag(1, -1, -1, 1) on both sides and decomposed; otherwise TRS with
// the rotation (w, x, -y, -z) and the translation (x, -y, -z); the rotation
// is then normalized (NodeDataCompEntity.cpp:33-41).
inline bool node(const json::Value& nv, Node& nd, std::string& err, const std::string& where) {
    double m[16], v[4];
    err.clear();
    if (number_array(nv, "matrix", 16, m, err, where)) {
        if (!err.empty()) return false;
        M4 A, F = identity4();
        for (int i = 0; i < 16; ++i) A.m[i] = float(m[i]);
        F.m[5] = -1.f;
        F.m[10] = -1.f;
        const M4 TRS = mat_mul(mat_mul(F, A), F);
        decompose(TRS, nd.scale, nd.rotation, nd.translation);
    } else {
        if (number_array(nv, "scale", 3, v, err, where)) {
            if (!err.empty()) return false;
            for (int i = 0; i < 3; ++i) nd.scale[i] = float(v[i]);
        }
        if (number_array(nv, "rotation", 4, v, err, where)) {
            if (!err.empty()) return false;
            nd.rotation[0] = float(v[0]);
            nd.rotation[1] = float(-v[1]);
            nd.rotation[2] = float(-v[2]);
            nd.rotation[3] = float(v[3]);
        }
        if (number_array(nv, "translation", 3, v, err, where)) {
            if (!err.empty()) return false;
            nd.translation[0] = float(v[0]);
            nd.translation[1] = float(-v[1]);
            nd.translation[2] = float(-v[2]);
        }
    }
    quat_normalize(nd.rotation);
    return true;
}

inline bool load_buffers(Ctx& c, const std::string& dir, const std::vector<uint8_t>* glb_bin) {
    const json::Value* bufs = c.root->find("buffers");
    if (!bufs) return true;
    if (bufs->kind != json::Value::Array) { c.err = "\"buffers\" is not an array"; return false; }
    for (size_t i = 0; i < bufs->arr.size(); ++i) {
        const json::Value& b = bufs->arr[i];
        const std::string where = "buffer " + std::to_string(i);
        int64_t len;
        if (!get_int(b, "byteLength", len, 0, true, c.err, where)) return false;
        std::vector<uint8_t> data;
        const json::Value* uri = b.find("uri");
        if (!uri) {
            if (i != 0 || !glb_bin) { c.err = where + ": no uri"; return false; }
            data = *glb_bin;
            if (data.size() < uint64_t(len) || data.size() > uint64_t(len) + 3) {
                c.err = where + ": BIN chunk size differs from byteLength";
                return false;
            }
            data.resize(size_t(len));
        } else {
            if (uri->kind != json::Value::String) { c.err = where + ": uri is not a string"; return false; }
            const std::string& u = uri->str;
            if (u.compare(0, 5, "data:") == 0) {
                const size_t comma = u.find(',');
                if (comma == std::string::npos || u.rfind(";base64", comma) == std::string::npos) {
                    c.err = where + ": only base64 data URIs are supported";
                    return false;
                }
                if (!base64(u, comma + 1, data)) { c.err = where + ": bad base64"; return false; }
            } else {
                const std::string path = dir + percent_decode(u);
                if (!read_file(path, data)) { c.err = where + ": cannot read " + path; return false; }
            }
            if (data.size() != uint64_t(len)) { c.err = where + ": size differs from byteLength"; return false; }
        }
        c.buffers.push_back(std::move(data));
    }
    return true;
}

inline uint32_t le32(const uint8_t* p) { return uint32_t(p[0]) | uint32_t(p[1]) << 8 | uint32_t(p[2]) << 16 | uint32_t(p[3]) << 24; }

}  // namespace detail

// Reads a .gltf or .glb file (told apart by the GLB magic) into doc.
inline bool load(const std::string& path, Document& doc, std::string& err) {
    using namespace detail;
    std::vector<uint8_t> file;
    if (!read_file(path, file)) {
        err = "cannot read " + path;
        return false;
    }
    const size_t slash = path.find_last_of('/');
    const std::string dir = slash == std::string::npos ? std::string() : path.substr(0, slash + 1);
    const char* jtext = reinterpret_cast<const char*>(file.data());
    size_t jlen = file.size();
    std::vector<uint8_t> bin;
    bool glb = false;
    if (file.size() >= 12 && le32(file.data()) == 0x46546C67u) {  // "glTF"
        glb = true;
        if (le32(file.data() + 4) != 2) { err = "GLB: version is not 2"; return false; }
        const uint64_t total = le32(file.data() + 8);
        if (total > file.size()) { err = "GLB: length exceeds the file"; return false; }
        uint64_t off = 12;
        bool have_json = false;
        while (off + 8 <= total) {
            const uint64_t clen = le32(file.data() + off), ctype = le32(file.data() + off + 4);
            if (off + 8 + clen > total) { err = "GLB: chunk runs past the end"; return false; }
            const uint8_t* cdata = file.data() + off + 8;
            if (!have_json) {
                if (ctype != 0x4E4F534Au) { err = "GLB: first chunk is not JSON"; return false; }
                jtext = reinterpret_cast<const char*>(cdata);
                jlen = size_t(clen);
                have_json = true;
            } else if (ctype == 0x004E4942u && bin.empty()) {
                bin.assign(cdata, cdata + clen);
            }
            off += 8 + clen;
        }
        if (!have_json) { err = "GLB: no JSON chunk"; return false; }
    }
    json::Value root;
    if (!json::parse(jtext, jlen, root, err)) return false;
    if (root.kind != json::Value::Object) { err = "glTF: the document is not an object"; return false; }
    Ctx c;
    c.root = &root;
    if (!load_buffers(c, dir, glb ? &bin : nullptr)) { err = c.err; return false; }
    doc.meshes.clear();
    doc.nodes.clear();
    const json::Value* meshes = root.find("meshes");
    if (meshes && meshes->kind != json::Value::Array) { err = "\"meshes\" is not an array"; return false; }
    for (size_t m = 0; meshes && m < meshes->arr.size(); ++m) {
        const json::Value& mv = meshes->arr[m];
        Mesh mesh;
        if (const json::Value* nm = mv.find("name"))
            if (nm->kind == json::Value::String) mesh.name = nm->str;
        const json::Value* prims = mv.find("primitives");
        if (!prims || prims->kind != json::Value::Array) { err = "mesh " + std::to_string(m) + ": no primitives"; return false; }
        std::vector<Primitive> in(prims->arr.size());
        std::vector<int32_t> modes(in.size());
        for (size_t p = 0; p < in.size(); ++p) {
            if (!primitive(c, prims->arr[p], in[p], "mesh " + std::to_string(m) + " primitive " + std::to_string(p))) {
                err = c.err;
                return false;
            }
            modes[p] = in[p].gltf_mode;
        }
        for (uint32_t p : tree_order(modes)) {
            Primitive& pr = in[p];
            mesh.skinned = mesh.skinned || pr.skin_attr;
            mesh.morphed = mesh.morphed || pr.targets;
            if (!pr.excluded) {
                mesh.ntri += pr.ntri;
                mesh.nvert += pr.nvert;
                (pr.nrm.empty() ? mesh.without_normals : mesh.with_normals) += 1;
            }
            mesh.prims.push_back(std::move(pr));
        }
        doc.meshes.push_back(std::move(mesh));
    }
    if (const json::Value* nodes = root.find("nodes")) {
        if (nodes->kind != json::Value::Array) { err = "\"nodes\" is not an array"; return false; }
        doc.nodes.resize(nodes->arr.size());
        for (size_t i = 0; i < nodes->arr.size(); ++i) {
            const json::Value& nv = nodes->arr[i];
            const std::string where = "node " + std::to_string(i);
            Node& nd = doc.nodes[i];
            if (!node(nv, nd, err, where)) return false;
            int64_t mi, sk;
            if (!get_int(nv, "mesh", mi, -1, false, err, where) || !get_int(nv, "skin", sk, -1, false, err, where))
                return false;
            if (mi >= int64_t(doc.meshes.size()) || mi < -1) { err = where + ": no such mesh"; return false; }
            nd.mesh = int32_t(mi);
            nd.collides = mi >= 0 && !doc.meshes[size_t(mi)].skinned && !doc.meshes[size_t(mi)].morphed;
        }
        for (size_t i = 0; i < nodes->arr.size(); ++i) {
            const json::Value* ch = nodes->arr[i].find("children");
            if (!ch) continue;
            if (ch->kind != json::Value::Array) { err = "node " + std::to_string(i) + ": bad children"; return false; }
            for (const json::Value& cv : ch->arr) {
                if (!cv.is_int() || cv.num < 0 || cv.num >= double(doc.nodes.size())) {
                    err = "node " + std::to_string(i) + ": child out of range";
                    return false;
                }
                Node& kid = doc.nodes[size_t(cv.num)];
                if (kid.parent != 0xFFFFFFFFu || size_t(cv.num) == i) {
                    err = "node " + std::to_string(size_t(cv.num)) + ": more than one parent";
                    return false;
                }
                kid.parent = uint32_t(i);
            }
        }
    }
    return true;
}

}  // namespace gltf
}  // namespace cdk
