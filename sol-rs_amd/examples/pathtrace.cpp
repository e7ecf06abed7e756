// pathtrace.cpp -- examples/5-pathtrace.rs, headless, over the sol C++
// mirror: load a glTF scene, build the scene description, the pathtrace
// pipeline (specialization constant 0 = --sky) and its SBT, then trace
// frames with progressive accumulation and write the RGBA8 image.
//
//   pathtrace --model scene.gltf [--sky] [--size 1280x720] [--frames 16]
//             [--out image.ppm] [--accum accum.f32] [--dump dir] [--device 0]
//             [--wavefront [--nee] [--rr N] [--spp N] [--depth N] [--streams]]
//
// --wavefront runs the wavefront path tracer (same image as the default
// program without --nee / --rr / --streams); --nee adds next-event
// estimation, --rr N Russian roulette from bounce N, --streams a random
// stream per sample (many samples of every pixel traced at once).
#include <chrono>

#include "common.hpp"

int main(int argc, char **argv)
{
    using namespace sol;
    const std::string model = example::arg(argc, argv, "--model");
    if (model.empty()) {
        std::fprintf(stderr, "no gltf file given (--model)\n");
        return 2;
    }
    uint32_t W = 1280, H = 720;   // examples/5-pathtrace.rs:374
    example::parse_size(example::arg(argc, argv, "--size", "1280x720"), W, H);
    const uint32_t frames = (uint32_t)std::atoi(example::arg(argc, argv, "--frames", "16").c_str());
    const bool enable_sky = example::has_flag(argc, argv, "--sky");
    const std::string dump_dir = example::arg(argc, argv, "--dump");
    try {
        auto ctx = std::make_shared<Context>(std::atoi(example::arg(argc, argv, "--device", "0").c_str()));
        // setup (examples/5-pathtrace.rs:118-248)
        scene::Scene sc = scene::load_scene(ctx, model);
        ray::SceneDescription sd = ray::SceneDescription::from_scene(ctx, sc);
        scene::Camera camera = sc.camera ? *sc.camera : scene::Camera(Vec2{(float)W, (float)H});
        camera.set_window_size(Vec2{(float)W, (float)H});
        const bool wavefront = example::has_flag(argc, argv, "--wavefront");
        solhip_spec spec;
        solhip_spec_default(wavefront ? SOLHIP_PROG_PATHTRACE : SOLHIP_PROG_PATHTRACE_GLSL, &spec);
        spec.enable_sky = enable_sky ? 1u : 0u;
        spec.nee = example::has_flag(argc, argv, "--nee") ? 1u : 0u;
        spec.rr_start = (uint32_t)std::atoi(example::arg(argc, argv, "--rr", "0").c_str());
        spec.spp_per_frame = (uint32_t)std::atoi(example::arg(argc, argv, "--spp", "8").c_str());
        spec.max_depth = (uint32_t)std::atoi(example::arg(argc, argv, "--depth", "32").c_str());
        spec.sample_streams = example::has_flag(argc, argv, "--streams") ? 1u : 0u;
        ray::Pipeline pipeline(ctx,
                               ray::PipelineInfo()
                                   .shader(wavefront ? "glsl/pathtrace_wavefront.rgen" : "glsl/pathtrace.rgen",
                                           ray::ShaderStage::Raygen)
                                   .shader("glsl/pathtrace.rmiss", ray::ShaderStage::Miss)
                                   .shader("glsl/pathtrace.rchit", ray::ShaderStage::ClosestHit)
                                   .specialization(enable_sky ? 1u : 0u, 0)
                                   .name("AO_mat"),
                               &spec);
        ray::ShaderBindingTable sbt(ctx, pipeline, ray::ShaderBindingTableInfo().raygen(0).miss(1).hitgroup(2));
        Buffer accum(ctx, (uint64_t)W * H * 16), render(ctx, (uint64_t)W * H * 4);
        std::vector<float> zeros((size_t)W * H * 4, 0.0f);
        accum.upload(zeros.data(), zeros.size() * 4, 0);
        if (!dump_dir.empty()) example::Dump(dump_dir).scene(sc, sd);
        const uint32_t accumulation_start_frame = 0;
        const auto t0 = std::chrono::steady_clock::now();
        for (uint32_t tick = 0; tick < frames; ++tick) {
            // render (examples/5-pathtrace.rs:294-369)
            ray::CommandBuffer cmd;
            const solhip_scene_uniforms u = example::scene_uniforms(camera, W, H, tick);
            if (!dump_dir.empty()) example::Dump(dump_dir).uniforms(tick, u);
            cmd.set_uniforms(u);
            cmd.push_constants((int32_t)accumulation_start_frame);
            sd.tlas_regenerate(cmd);
            cmd.bind_pipeline(pipeline);
            cmd.bind_scene(sd);
            ray::Images im;
            im.accum = &accum;
            im.render = &render;
            cmd.bind_images(im);
            sbt.cmd_trace_rays(cmd, ray::Extent3D{W, H, 1});
        }
        ctx->wait_idle();
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        std::vector<uint8_t> rgba((size_t)W * H * 4);
        render.download(rgba.data(), rgba.size());
        example::write_ppm(example::arg(argc, argv, "--out", "pathtrace.ppm"), rgba, W, H);
        const std::string accum_out = example::arg(argc, argv, "--accum");
        if (!accum_out.empty()) {
            std::vector<float> a((size_t)W * H * 4);
            accum.download(a.data(), a.size() * 4);
            example::write_file(accum_out, a.data(), a.size() * 4);
        }
        solhip_counters c{};
        check(solhip_stats(ctx->handle(), &c, 0), "solhip_stats");
        std::printf("{\"frames\": %u, \"ms\": %.3f, \"rays\": %llu, \"overflow\": %llu}\n", frames, ms,
                    (unsigned long long)c.rays, (unsigned long long)c.overflow);
    } catch (const sol::Error &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
