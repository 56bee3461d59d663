// Host-only driver of the glTF reader (inmyroom_vulkan_amd/csrc/gltf.hpp) for
// the sanitizer and fuzz tests: loads every path given, prints one line per
// document ("ok <meshes> <nodes> <triangles>" or "err <message>") and checks
// what the library relies on in a document the reader accepts.  Exit status
// 2 when such a check fails; a sanitizer report aborts the process.
#include <cstdio>
#include <string>

#include "gltf.hpp"

namespace {

bool check(const cdk::gltf::Document& d, std::string& why) {
    for (size_t m = 0; m < d.meshes.size(); ++m) {
        for (const cdk::gltf::Primitive& p : d.meshes[m].prims) {
            if (p.excluded) continue;
            if (p.pos.size() != size_t(p.nvert) * 3) { why = "positions do not match the vertex count"; return false; }
            if (!p.nrm.empty() && p.nrm.size() != p.pos.size()) { why = "normals do not match the positions"; return false; }
            for (uint32_t i : p.idx)
                if (i >= p.nvert) { why = "index past the vertices"; return false; }
            if (p.ntri != cdk::gltf::triangles_of(p.mode, p.idx.size())) { why = "triangle count"; return false; }
        }
    }
    for (const cdk::gltf::Node& n : d.nodes) {
        if (n.parent != 0xFFFFFFFFu && n.parent >= d.nodes.size()) { why = "parent out of range"; return false; }
        if (n.mesh >= int32_t(d.meshes.size())) { why = "node mesh out of range"; return false; }
    }
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
        cdk::gltf::Document doc;
        std::string err;
        if (!cdk::gltf::load(argv[i], doc, err)) {
            std::printf("err %s\n", err.c_str());
            continue;
        }
        std::string why;
        if (!check(doc, why)) {
            std::printf("bad %s: %s\n", argv[i], why.c_str());
            return 2;
        }
        uint64_t tris = 0;
        for (const auto& m : doc.meshes)
            for (const auto& p : m.prims)
                if (!p.excluded) tris += p.ntri;
        std::printf("ok %zu %zu %llu\n", doc.meshes.size(), doc.nodes.size(), static_cast<unsigned long long>(tris));
    }
    return 0;
}
