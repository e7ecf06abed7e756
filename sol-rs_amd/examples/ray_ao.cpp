// ray_ao.cpp -- examples/4-ray-ao.rs, headless, over the sol C++ mirror:
// load a glTF scene (optionally with the example's transform override), build
// the scene description, the ao.* pipeline and its SBT, then trace frames with
// progressive accumulation into the RGBA32F target.  The blue-noise texture is
// the decoded assets/textures/HDR_RGBA_0.png (--blue) or seeded white noise.
//
//   ray_ao --model scene.gltf [--blue HDR_RGBA_0.png] [--toycar-transform]
//          [--size 900x600] [--frames 16] [--accum accum.f32] [--dump dir]
#include <chrono>
#include <cmath>

#include "../host/png.hpp"
#include "common.hpp"

int main(int argc, char **argv)
{
    using namespace sol;
    const std::string model = example::arg(argc, argv, "--model");
    if (model.empty()) {
        std::fprintf(stderr, "no gltf file given (--model)\n");
        return 2;
    }
    uint32_t W = 900, H = 600;   // examples/4-ray-ao.rs:327
    example::parse_size(example::arg(argc, argv, "--size", "900x600"), W, H);
    const uint32_t frames = (uint32_t)std::atoi(example::arg(argc, argv, "--frames", "16").c_str());
    const std::string dump_dir = example::arg(argc, argv, "--dump");
    try {
        auto ctx = std::make_shared<Context>(std::atoi(example::arg(argc, argv, "--device", "0").c_str()));
        scene::Scene sc = scene::load_scene(ctx, model);
        if (example::has_flag(argc, argv, "--toycar-transform")) {
            // examples/4-ray-ao.rs:82-86
            for (scene::Mesh &m : sc.meshes)
                m.transform = Mat4::from_scale(Vec3::splat(0.01f)) * Mat4::from_rotation_x(1.57079632679f);
        }
        ray::SceneDescription sd = ray::SceneDescription::from_scene(ctx, sc);
        scene::Camera camera(Vec2{(float)W, (float)H});
        camera.look_at(Vec3{4.0f, 1.0f, 4.0f}, Vec3{0.0f, 0.5f, 0.0f}, Vec3{0.0f, -1.0f, 0.0f});   // :89-90
        // blue noise (examples/4-ray-ao.rs setup; src/texture.rs:487-496)
        ImageRgba8 blue;
        const std::string blue_path = example::arg(argc, argv, "--blue");
        if (!blue_path.empty()) {
            blue = load_png_rgba8_flipv(blue_path);
        } else {
            blue.width = blue.height = 256;
            blue.rgba.resize(256 * 256 * 4);
            uint32_t s = 7u;
            for (uint8_t &b : blue.rgba) {
                s = s * 747796405u + 2891336453u;
                b = (uint8_t)(s >> 24);
            }
        }
        if (!dump_dir.empty()) {
            example::Dump(dump_dir).scene(sc, sd);
            example::write_file(dump_dir + "/blue_noise.bin", blue.rgba.data(), blue.rgba.size());
        }
        Buffer blue_buf = Buffer::from_data(ctx, blue.rgba);
        ray::Pipeline pipeline(ctx, ray::PipelineInfo()
                                        .shader("glsl/ao.rgen", ray::ShaderStage::Raygen)
                                        .shader("glsl/ao.rmiss", ray::ShaderStage::Miss)
                                        .shader("glsl/ao.rchit", ray::ShaderStage::ClosestHit));
        ray::ShaderBindingTable sbt(ctx, pipeline, ray::ShaderBindingTableInfo().raygen(0).miss(1).hitgroup(2));
        Buffer target(ctx, (uint64_t)W * H * 16);
        std::vector<float> zeros((size_t)W * H * 4, 0.0f);
        target.upload(zeros.data(), zeros.size() * 4, 0);
        // the push constant is set every frame and the target keeps its
        // contents (the reference sets it on reset frames only and
        // re-transitions the image from UNDEFINED, examples/4-ray-ao.rs:258-276)
        const uint32_t accumulation_start_frame = 0;
        const auto t0 = std::chrono::steady_clock::now();
        for (uint32_t tick = 0; tick < frames; ++tick) {
            ray::CommandBuffer cmd;
            const solhip_scene_uniforms u = example::scene_uniforms(camera, W, H, tick);
            if (!dump_dir.empty()) example::Dump(dump_dir).uniforms(tick, u);
            cmd.set_uniforms(u);
            cmd.push_constants((int32_t)accumulation_start_frame);
            sd.tlas_regenerate(cmd);
            cmd.bind_pipeline(pipeline);
            cmd.bind_scene(sd);
            ray::Images im;
            im.accum = &target;
            im.blue_noise = &blue_buf;
            im.blue_w = blue.width;
            im.blue_h = blue.height;
            cmd.bind_images(im);
            sbt.cmd_trace_rays(cmd, ray::Extent3D{W, H, 1});
        }
        ctx->wait_idle();
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        std::vector<float> a((size_t)W * H * 4);
        target.download(a.data(), a.size() * 4);
        const std::string accum_out = example::arg(argc, argv, "--accum", "ao.f32");
        example::write_file(accum_out, a.data(), a.size() * 4);
        std::vector<uint8_t> rgba((size_t)W * H * 4);
        for (size_t i = 0; i < rgba.size(); ++i) {
            float v = a[i];
            v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
            rgba[i] = (uint8_t)std::floor(v * 255.0f + 0.5f);
        }
        example::write_ppm(example::arg(argc, argv, "--out", "ao.ppm"), rgba, W, H);
        solhip_counters c{};
        check(solhip_stats(ctx->handle(), &c, 0), "solhip_stats");
        std::printf("{\"frames\": %u, \"ms\": %.3f, \"rays\": %llu, \"overflow\": %llu}\n", frames, ms,
                    (unsigned long long)c.rays, (unsigned long long)c.overflow);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
