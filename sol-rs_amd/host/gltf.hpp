// gltf.hpp -- minimal glTF 2.0 reader (JSON .gltf with embedded / external
// buffers, binary .glb): what scene::load_scene needs from the `gltf` crate
// (src/scene/mod.rs:138-341): meshes / primitives / accessors, materials,
// the node tree with matrix or TRS transforms, perspective cameras.
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace sol {
namespace gltf {

struct Json {
    enum Type { Null, Bool, Num, Str, Arr, Obj } t = Null;
    bool b = false;
    double n = 0.0;
    std::string s;
    std::vector<Json> a;
    std::vector<std::pair<std::string, Json>> o;

    const Json *get(const std::string &key) const;
    double num(const std::string &key, double def) const;
    int64_t integer(const std::string &key, int64_t def) const;
    std::string str(const std::string &key, const std::string &def = "") const;
};

Json parse_json(const std::string &text);

struct Accessor {
    int64_t view = -1;
    uint64_t offset = 0;
    int component = 0;    // 5120..5126
    uint64_t count = 0;
    int ncomp = 1;        // SCALAR 1, VEC2 2, VEC3 3, VEC4 4, MAT4 16
    bool normalized = false;
};
struct View {
    int64_t buffer = 0;
    uint64_t offset = 0, length = 0, stride = 0;
};
struct Primitive {
    int64_t position = -1, normal = -1, texcoord0 = -1, color0 = -1, indices = -1, material = -1;
};
struct Node {
    int64_t mesh = -1, camera = -1;
    std::vector<int64_t> children;
    float matrix[16];     // column-major local transform
};
struct Camera {
    bool perspective = false;
    float yfov = 0.0f, znear = 0.0f;
    bool has_zfar = false;
    float zfar = 0.0f;
};
struct Material {
    float base_color[4] = {1, 1, 1, 1};
    float emissive[3] = {0, 0, 0};
    float metallic = 1.0f, roughness = 1.0f;
};

struct Document {
    std::vector<std::vector<uint8_t>> buffers;
    std::vector<View> views;
    std::vector<Accessor> accessors;
    std::vector<std::vector<Primitive>> meshes;
    std::vector<std::string> mesh_names;
    std::vector<Node> nodes;
    std::vector<Camera> cameras;
    std::vector<Material> materials;

    // accessor -> float[count * ncomp] (normalized integers scaled to [0,1])
    std::vector<float> read_floats(int64_t accessor) const;
    // accessor -> u32 indices
    std::vector<uint32_t> read_indices(int64_t accessor) const;
};

// throws std::runtime_error on malformed input or missing files
Document load(const std::string &path);

}  // namespace gltf
}  // namespace sol
