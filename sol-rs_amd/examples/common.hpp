// common.hpp -- shared pieces of the headless examples: argument parsing,
// SceneUniforms (examples/{4-ray-ao,5-pathtrace}.rs:19-31), image writers and
// a dump of the exact inputs handed to the ray path (for parity checks).
#pragma once

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../host/sol.hpp"

namespace example {

inline bool has_flag(int argc, char **argv, const char *f)
{
    for (int i = 1; i < argc; ++i)
        if (std::strcmp(argv[i], f) == 0) return true;
    return false;
}
inline std::string arg(int argc, char **argv, const char *f, const std::string &def = "")
{
    for (int i = 1; i + 1 < argc; ++i)
        if (std::strcmp(argv[i], f) == 0) return argv[i + 1];
    return def;
}
inline void parse_size(const std::string &s, uint32_t &w, uint32_t &h)
{
    if (std::sscanf(s.c_str(), "%ux%u", &w, &h) != 2) {
        std::fprintf(stderr, "bad --size %s\n", s.c_str());
        std::exit(2);
    }
}

// SceneUniforms::from(camera, frame) (examples/5-pathtrace.rs:19-31)
inline solhip_scene_uniforms scene_uniforms(const sol::scene::Camera &camera, uint32_t w, uint32_t h, uint32_t tick)
{
    solhip_scene_uniforms u{};
    const sol::Mat4 view = camera.view_matrix(), proj = camera.perspective_matrix();
    const sol::Mat4 model = sol::Mat4::identity(), vp = proj * view;
    std::memcpy(u.model, model.data(), 64);
    std::memcpy(u.view, view.data(), 64);
    std::memcpy(u.view_inverse, view.inverse().data(), 64);
    std::memcpy(u.projection, proj.data(), 64);
    std::memcpy(u.projection_inverse, proj.inverse().data(), 64);
    std::memcpy(u.model_view_projection, vp.data(), 64);
    u.frame[0] = w;
    u.frame[1] = h;
    u.frame[2] = tick;
    return u;
}

inline void write_file(const std::string &path, const void *data, size_t bytes)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(1);
    }
    if (bytes) std::fwrite(data, 1, bytes, f);
    std::fclose(f);
}

// binary PPM of an RGBA8 image
inline void write_ppm(const std::string &path, const std::vector<uint8_t> &rgba, uint32_t w, uint32_t h)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return;
    std::fprintf(f, "P6\n%u %u\n255\n", w, h);
    for (size_t i = 0; i < (size_t)w * h; ++i) std::fwrite(&rgba[4 * i], 1, 3, f);
    std::fclose(f);
}

// the exact inputs of the ray path, one file per array:
//   blas_<k>_vertices.bin (ModelVertex) / blas_<k>_indices.bin (u32), per BLAS
//   instances.bin (64-B TLAS records, blas = BLAS index), scene_instances.bin
//   (144-B SceneInstance), materials.bin (48-B MaterialInfo per instance),
//   uniforms_<f>.bin (400-B SceneUniforms per frame), manifest.txt
struct Dump {
    std::string dir;
    explicit Dump(std::string d) : dir(std::move(d)) { mkdir(dir.c_str(), 0755); }
    void scene(const sol::scene::Scene &scene, const sol::ray::SceneDescription &sd) const
    {
        size_t k = 0;
        std::vector<sol::scene::MaterialInfo> mats;
        for (const sol::scene::Mesh &m : scene.meshes)
            for (const sol::scene::PrimitiveSection &ps : m.primitive_sections) {
                if (!ps.indices) continue;
                const sol::scene::ModelVertex *v = m.host_vertices.data() + ps.vertices.offset;
                const uint32_t *ix = m.host_indices.data() + ps.indices->offset;
                write_file(dir + "/blas_" + std::to_string(k) + "_vertices.bin", v,
                           ps.vertices.element_count * sizeof(sol::scene::ModelVertex));
                write_file(dir + "/blas_" + std::to_string(k) + "_indices.bin", ix,
                           ps.indices->element_count * sizeof(uint32_t));
                sol::scene::MaterialInfo mi{};
                if (ps.material_index && *ps.material_index < scene.materials.size())
                    mi = scene.materials[*ps.material_index];
                else if (!scene.materials.empty())
                    mi = scene.materials[0];
                mats.push_back(mi);
                ++k;
            }
        std::vector<solhip_instance> inst = sol::ray::TLAS::create_instances(sd.blas());
        for (size_t i = 0; i < inst.size(); ++i) inst[i].blas = i;
        write_file(dir + "/instances.bin", inst.data(), inst.size() * sizeof(solhip_instance));
        write_file(dir + "/scene_instances.bin", sd.instances().data(), sd.instances().size() * 144);
        write_file(dir + "/materials.bin", mats.data(), mats.size() * sizeof(sol::scene::MaterialInfo));
        FILE *f = std::fopen((dir + "/manifest.txt").c_str(), "w");
        if (f) {
            std::fprintf(f, "blas %zu\n", k);
            std::fclose(f);
        }
    }
    void uniforms(uint32_t frame, const solhip_scene_uniforms &u) const
    {
        write_file(dir + "/uniforms_" + std::to_string(frame) + ".bin", &u, sizeof(u));
    }
};

}  // namespace example
