// ray_debug.cpp -- examples/3-ray-debug.rs, headless, over the sol C++ mirror:
// load a glTF scene, build the scene description, the debug.* pipeline
// (debug.rgen / debug.rmiss / debug.rchit: barycentrics of the closest hit,
// the ray direction on a miss) and its SBT, and trace the frames into the
// target.  The camera is the example's Camera::new(size) looking at the
// origin from (5, 5, 5) with -Y up (examples/3-ray-debug.rs:77-78); the size
// defaults to its 900x600 window (:257-260).
//
//   ray_debug --model scene.gltf [--size 900x600] [--frames 1] [--target debug.f32]
//             [--out debug.ppm] [--dump dir]
#include <chrono>
#include <cmath>

#include "common.hpp"

int main(int argc, char **argv)
{
    using namespace sol;
    const std::string model = example::arg(argc, argv, "--model");
    if (model.empty()) {
        std::fprintf(stderr, "no gltf file given (--model)\n");
        return 2;
    }
    uint32_t W = 900, H = 600;
    example::parse_size(example::arg(argc, argv, "--size", "900x600"), W, H);
    const uint32_t frames = (uint32_t)std::atoi(example::arg(argc, argv, "--frames", "1").c_str());
    const std::string dump_dir = example::arg(argc, argv, "--dump");
    try {
        auto ctx = std::make_shared<Context>(std::atoi(example::arg(argc, argv, "--device", "0").c_str()));
        scene::Scene sc = scene::load_scene(ctx, model);
        scene::Camera camera(Vec2{(float)W, (float)H});
        camera.look_at(Vec3::splat(5.0f), Vec3{0.0f, 0.0f, 0.0f}, Vec3{0.0f, -1.0f, 0.0f});
        ray::SceneDescription sd = ray::SceneDescription::from_scene(ctx, sc);
        if (!dump_dir.empty()) example::Dump(dump_dir).scene(sc, sd);
        ray::Pipeline pipeline(ctx, ray::PipelineInfo()
                                        .shader("glsl/debug.rgen", ray::ShaderStage::Raygen)
                                        .shader("glsl/debug.rmiss", ray::ShaderStage::Miss)
                                        .shader("glsl/debug.rchit", ray::ShaderStage::ClosestHit));
        ray::ShaderBindingTable sbt(ctx, pipeline, ray::ShaderBindingTableInfo().raygen(0).miss(1).hitgroup(2));
        Buffer target(ctx, (uint64_t)W * H * 16);
        const auto t0 = std::chrono::steady_clock::now();
        for (uint32_t tick = 0; tick < frames; ++tick) {
            ray::CommandBuffer cmd;
            const solhip_scene_uniforms u = example::scene_uniforms(camera, W, H, tick);
            if (!dump_dir.empty()) example::Dump(dump_dir).uniforms(tick, u);
            cmd.set_uniforms(u);
            cmd.bind_pipeline(pipeline);
            cmd.bind_scene(sd);
            ray::Images im;
            im.accum = &target;
            cmd.bind_images(im);
            sbt.cmd_trace_rays(cmd, ray::Extent3D{W, H, 1});
        }
        ctx->wait_idle();
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        std::vector<float> a((size_t)W * H * 4);
        target.download(a.data(), a.size() * 4);
        example::write_file(example::arg(argc, argv, "--target", "debug.f32"), a.data(), a.size() * 4);
        // the reference's target is R8G8B8A8_UNORM (examples/3-ray-debug.rs:53)
        std::vector<uint8_t> rgba((size_t)W * H * 4);
        for (size_t i = 0; i < rgba.size(); ++i) {
            float v = a[i];
            v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
            rgba[i] = (uint8_t)std::floor(v * 255.0f + 0.5f);
        }
        example::write_ppm(example::arg(argc, argv, "--out", "debug.ppm"), rgba, W, H);
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
