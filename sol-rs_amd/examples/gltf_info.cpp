// gltf_info.cpp -- host-only check of scene decoding: prints, as JSON, what
// scene::decode_scene made of a glTF file (meshes, primitive sections,
// triangle counts, global transforms, materials, camera) and optionally the
// RGBA8 decode of a PNG texture.  No device is touched.  --arrays writes the
// decoded arrays per mesh (mesh_<m>_vertices.bin: 64-B ModelVertex,
// mesh_<m>_indices.bin: u32) and materials.bin (48-B MaterialInfo).
//
//   gltf_info scene.gltf [--arrays dir] [--png texture.png --png-out texels.rgba8]
#include <cstdio>

#include "../host/png.hpp"
#include "common.hpp"

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: gltf_info scene.gltf [--arrays dir] [--png in.png --png-out out.rgba8]\n");
        return 2;
    }
    try {
        const std::string png = example::arg(argc, argv, "--png");
        if (!png.empty()) {
            sol::ImageRgba8 img = sol::load_png_rgba8_flipv(png);
            example::write_file(example::arg(argc, argv, "--png-out", "texels.rgba8"), img.rgba.data(), img.rgba.size());
            std::printf("{\"png\": [%u, %u]}\n", img.width, img.height);
            return 0;
        }
        const sol::scene::HostScene hs = sol::scene::decode_scene(argv[1]);
        const std::string arrays = example::arg(argc, argv, "--arrays");
        if (!arrays.empty()) {
            mkdir(arrays.c_str(), 0755);
            for (size_t m = 0; m < hs.meshes.size(); ++m) {
                const auto &mesh = hs.meshes[m];
                example::write_file(arrays + "/mesh_" + std::to_string(m) + "_vertices.bin", mesh.vertices.data(),
                                    mesh.vertices.size() * sizeof(sol::scene::ModelVertex));
                example::write_file(arrays + "/mesh_" + std::to_string(m) + "_indices.bin", mesh.indices.data(),
                                    mesh.indices.size() * sizeof(uint32_t));
            }
            example::write_file(arrays + "/materials.bin", hs.materials.data(),
                                hs.materials.size() * sizeof(sol::scene::MaterialInfo));
        }
        std::printf("{\"meshes\": [");
        size_t tris = 0;
        for (size_t m = 0; m < hs.meshes.size(); ++m) {
            const auto &mesh = hs.meshes[m];
            std::printf("%s{\"name\": \"%s\", \"vertices\": %zu, \"indices\": %zu, \"transform\": [", m ? ", " : "",
                        mesh.name.c_str(), mesh.vertices.size(), mesh.indices.size());
            for (int i = 0; i < 16; ++i) std::printf("%s%.9g", i ? ", " : "", mesh.transform.data()[i]);
            std::printf("], \"sections\": [");
            for (size_t s = 0; s < mesh.primitive_sections.size(); ++s) {
                const auto &ps = mesh.primitive_sections[s];
                std::printf("%s[%zu, %zu, %zu, %zu, %lld]", s ? ", " : "", ps.vertices.offset,
                            ps.vertices.element_count, ps.indices ? ps.indices->offset : 0,
                            ps.indices ? ps.indices->element_count : 0,
                            ps.material_index ? (long long)*ps.material_index : -1LL);
                if (ps.indices) tris += ps.indices->element_count / 3;
            }
            std::printf("], \"first_vertex\": [%.9g, %.9g, %.9g]}", mesh.vertices.empty() ? 0.0 : mesh.vertices[0].pos.x,
                        mesh.vertices.empty() ? 0.0 : mesh.vertices[0].pos.y,
                        mesh.vertices.empty() ? 0.0 : mesh.vertices[0].pos.z);
        }
        std::printf("], \"triangles\": %zu, \"materials\": [", tris);
        for (size_t i = 0; i < hs.materials.size(); ++i) {
            const auto &mi = hs.materials[i];
            std::printf("%s{\"base\": [%.9g, %.9g, %.9g, %.9g], \"emissive\": [%.9g, %.9g, %.9g], \"metallic\": %.9g, "
                        "\"roughness\": %.9g}",
                        i ? ", " : "", mi.base_color.x, mi.base_color.y, mi.base_color.z, mi.base_color.w,
                        mi.emissive_factor.x, mi.emissive_factor.y, mi.emissive_factor.z, mi.metallic_factor,
                        mi.roughness_factor);
        }
        std::printf("], \"camera\": ");
        if (hs.camera) {
            const sol::Mat4 v = hs.camera->view_matrix();
            std::printf("{\"view\": [");
            for (int i = 0; i < 16; ++i) std::printf("%s%.9g", i ? ", " : "", v.data()[i]);
            std::printf("]}");
        } else {
            std::printf("null");
        }
        std::printf("}\n");
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
