"""Benchmark: ray-traced AO on the 2.1M-triangle heightfield at 1080p, 1 GPU.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--grid 1024] [--res 1920x1080] [--no-cpu]

A step is one frame of the plain AO program (SOLHIP_PROG_AO): one jittered
primary ray per pixel, 8 cosine AO rays per primary hit (closest hit, range
[1e-4, 0.1] x extent), running-mean accumulation.  Inputs (BVH, tables,
accumulation buffer) are resident in HBM before the timed region.  Prints
one JSON line.

roofline: the dominant kernel (the AO-ray kernel) against the 8 TB/s HBM
peak, with algorithmic bytes per ray
    B_ray = 32 (ray) + 20 (hit) + 80 N_node + 36 N_tri + 64 N_inst
where the fetch counts come from the counting variant of the same kernels on
the same frame, and the launch duration from hipEvents recorded by the
library on the launch stream during the timed region.

cpu_baseline: the oracle (oracle/, a C restatement of the reference GLSL with
its own binary BVH, OpenMP over rows) on whole frames of the same workload,
as many as fit ~10 s; the GPU accumulation of the same frames is compared
with it bit for bit.

    python bench.py --workload pathtrace-c0|pathtrace-c2|server-c3 [--steps K] [--warmup W]

other configurations of the path, one JSON line each (not the headline
metric): the wavefront path tracer on the 36-triangle Cornell box (512^2,
16 spp, depth 4, next-event estimation) and on 10,000 instances of 5 BLAS
(1080p, 64 spp, depth 8, Russian roulette from bounce 3, gradient sky, every
BLAS and the TLAS rebuilt each frame); the render server under a Poisson
request load (p50 / p99 latency beside throughput).

    python bench.py --copy-calibration BYTES

runs one plain device-to-device copy of BYTES bytes (torch), the reference
kernel whose FETCH_SIZE / WRITE_SIZE readings scale the counters of the AO
kernel to bytes.
"""
import argparse
import json
import os
import re
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "sol-rs_amd", "python")]

HBM_PEAK_GBS = 8000.0
B_RAY_FIXED = 32 + 20
B_NODE, B_TRI, B_INST = 80, 36, 64
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_ao_rays_heightfield1024_1080p.txt")
C2_PMC_FILE = os.path.join(ROOT, "profiles", "pmc_pt_extend_instanced10k_1080p.txt")


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--grid", type=int, default=1024)
    ap.add_argument("--res", default="1920x1080")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--copy-calibration", type=int, default=0)
    ap.add_argument("--workload", default="ao",
                    choices=["ao", "pathtrace-c0", "pathtrace-c2", "server-c3", "cache-c4"])
    ap.add_argument("--jobs", type=int, default=40, help="server-c3: requests of the configuration's burst")
    ap.add_argument("--cap-jobs", type=int, default=160, help="server-c3: requests of the capacity burst")
    ap.add_argument("--load-jobs", type=int, default=400, help="server-c3: requests per Poisson load")
    ap.add_argument("--rate", type=float, default=200.0, help="server-c3: requests per second")
    ap.add_argument("--streams", type=int, default=None, choices=(0, 1),
                    help="1 = a random stream per sample, 0 = one per pixel (pathtrace.rgen); default 1 for "
                         "pathtrace-c0/c2, 0 for server-c3 (its jobs already fill the GPU; extra slots per pixel "
                         "only delay admissions)")
    ap.add_argument("--per-pass", type=int, default=0, help="pathtrace-c0/c2: samples per pass (0 = library default)")
    ap.add_argument("--slices", type=int, default=0, help="pathtrace-c0/c2: image slices (0 = library default)")
    ap.add_argument("--no-tune", action="store_true", help="pathtrace-c2: library default launch shape")
    ap.add_argument("--tune-cache", default=os.path.join(tempfile.gettempdir(), "solhip_autotune.txt"),
                    help="pathtrace-c2: the autotuner's cache file")
    ap.add_argument("--any-hit", action="store_true",
                    help="AO rays stop at the first hit (same image; hit records are not the closest)")
    return ap.parse_args()


def copy_calibration(nbytes):
    import torch
    n = nbytes // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4}))


def bytes_per_ray(c):
    if c.rays == 0:
        return 0.0
    return B_RAY_FIXED + (B_NODE * c.node_fetches + B_TRI * c.tri_fetches + B_INST * c.inst_fetches) / c.rays


def read_profile(path):
    """the 'key: number' lines of a committed rocprofv3 --pmc summary
    (traffic_bytes_per_launch, l2_hit_rate, valu_wave_instructions_per_launch,
    valu_lanes_active, valu_peak_wave_instr_per_s), {} without the file"""
    try:
        with open(path) as f:
            text = f.read()
    except OSError:
        return {}
    return {k: float(v) for k, v in re.findall(r"^([a-z0-9_]+):\s*([0-9.eE+-]+)\s*$", text, flags=re.M)}


def read_traffic():
    """HBM bytes per launch of the AO-ray kernel from the committed rocprofv3
    --pmc summary (line 'traffic_bytes_per_launch: <int>'), else None."""
    return read_profile(TRAFFIC_FILE).get("traffic_bytes_per_launch")


def limiter(path, avg_launch_s):
    """What binds the traversal kernel, from the counters of the committed
    --pmc summary at path and this run's launch time: the VALU issue rate
    against the fp32-FMA issue rate the probe sustains (the binding limit:
    the kernels issue VALU instructions at most of that rate with about half
    the lanes active), next to the measured HBM rate and the L2 hit rate."""
    p = read_profile(path)
    if "valu_wave_instructions_per_launch" not in p:
        return None
    out = {"kind": "VALU issue", "source": os.path.relpath(path, ROOT)}
    rate = p["valu_wave_instructions_per_launch"] / avg_launch_s
    if "valu_peak_wave_instr_per_s" in p:
        out["valu_issue_frac"] = round(rate / p["valu_peak_wave_instr_per_s"], 3)
    out["valu_lanes_active"] = p.get("valu_lanes_active")
    if "traffic_bytes_per_launch" in p:
        gbs = p["traffic_bytes_per_launch"] / avg_launch_s / 1e9
        out["hbm_measured_gbs"] = round(gbs, 1)
        out["hbm_measured_frac"] = round(gbs / HBM_PEAK_GBS, 4)
    out["l2_hit_rate"] = p.get("l2_hit_rate")
    return out


def main():
    a = parse_args()
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.gpus != 1:
        print("bench.py measures one GPU only (the library has no multi-GPU path); got --gpus %d" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    if a.workload.startswith("pathtrace"):
        bench_pathtrace(a)
        return
    if a.workload == "server-c3":
        bench_server(a)
        return
    if a.workload == "cache-c4":
        bench_cache(a)
        return
    W, H = (int(v) for v in a.res.lower().split("x"))
    import solhip
    from solhip import scenes

    sd = scenes.heightfield(a.grid)
    ctx = solhip.Context(0)
    t0 = time.perf_counter()
    # one-time build: binned SAH (the PREFER_FAST_TRACE build of BLAS::new)
    gs = solhip.GpuScene(ctx, sd, build_flags=solhip.BUILD_PREFER_FAST_TRACE)
    ctx.sync()
    upload_build_s = time.perf_counter() - t0
    # BVH build alone (rebuild in place on the resident vertex buffers)
    ctx.sync()
    t0 = time.perf_counter()
    gs.rebuild_blas(0)
    ctx.sync()
    blas_build_ms = (time.perf_counter() - t0) * 1e3
    # a TLAS is built over its BLASes: rebuilt after them
    gs.update_instances(sd.instance_descs)
    ctx.sync()
    info = ctx.blas_info(gs.blas[0])

    spec = solhip.spec_default(solhip.PROG_AO)
    spec.n_ao = 8
    spec.ao_tmin = 1e-4 * sd.extent
    spec.ao_tmax = 0.1 * sd.extent
    spec.ao_any_hit = 1 if a.any_hit else 0
    prog = ctx.program(solhip.PROG_AO, spec)
    acc = ctx.upload(np.zeros((H, W, 4), np.float32))
    img = solhip.Images(acc.h, 0, 0, 0, 0, 0)

    # counting frame (fetch counts per ray, same frame as the timed ones)
    ctx.profile(solhip.api.PROFILE_COUNT)
    for k in ("ao_primary", "ao_rays"):
        ctx.kernel_stats(k, reset=True)
    ctx.trace(prog, gs.tlas, gs.tables, sd.uniforms(W, H, 0), 0, img, W, H)
    cp = ctx.kernel_stats("ao_primary", reset=True).counters
    ca = ctx.kernel_stats("ao_rays", reset=True).counters
    ctx.profile(0)

    # warmup, then the timed region
    for i in range(a.warmup):
        ctx.trace(prog, gs.tlas, gs.tables, sd.uniforms(W, H, i), 0, img, W, H)
    ctx.sync()
    ctx.stats(reset=True)
    for k in ("ao_primary", "ao_rays"):
        ctx.kernel_stats(k, reset=True)
    ctx.profile(solhip.api.PROFILE_TIME)
    unis = [sd.uniforms(W, H, a.warmup + i) for i in range(a.steps)]
    ctx.sync()
    t0 = time.perf_counter()
    for i in range(a.steps):
        ctx.trace(prog, gs.tlas, gs.tables, unis[i], a.warmup, img, W, H)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    ctx.profile(0)
    kp = ctx.kernel_stats("ao_primary")
    ka = ctx.kernel_stats("ao_rays")
    st = ctx.stats(reset=True)
    assert st.overflow == 0, "traversal stack overflow"
    rays = st.rays
    mrays = rays / elapsed / 1e6

    # roofline of the dominant kernel
    dom_name, dom, dom_c = ("ao_rays", ka, ca) if ka.total_ms >= kp.total_ms else ("ao_primary", kp, cp)
    avg_launch_s = dom.total_ms / max(1, dom.launches) / 1e3
    rays_per_launch = dom.counters.rays / max(1, dom.launches)
    b_ray = bytes_per_ray(dom_c)
    achieved = b_ray * rays_per_launch / avg_launch_s / 1e9
    traffic = read_traffic()

    result = {
        "metric": "Mrays/s on 1 GPU",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed / a.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "plain ray-traced AO, %dx%d-quad fBm heightfield (%d triangles, one BLAS), %dx%d, "
                        "1 primary + %d cosine AO rays per primary hit, %s, AO range [%.0e, %.2f]"
                        % (a.grid, a.grid, sd.tri_count, W, H, spec.n_ao,
                           "first hit" if a.any_hit else "closest hit", spec.ao_tmin, spec.ao_tmax),
            "triangles": sd.tri_count,
            "resolution": [W, H],
            "n_ao": spec.n_ao,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_ray": round(b_ray, 1),
            "fetches_per_ray": {
                "node": round(dom_c.node_fetches / max(1, dom_c.rays), 3),
                "tri": round(dom_c.tri_fetches / max(1, dom_c.rays), 3),
                "inst": round(dom_c.inst_fetches / max(1, dom_c.rays), 3),
            },
            "avg_launch_ms": round(avg_launch_s * 1e3, 4),
            "rays_per_launch": int(rays_per_launch),
            # the algorithmic HBM roofline above is not what binds this
            # kernel: its node and triangle data come mostly from L2
            "limiter": limiter(TRAFFIC_FILE, avg_launch_s) if dom_name == "ao_rays" else None,
        },
        "extra": {
            "rays_per_frame": int(rays / a.steps),
            "primary_hit_fraction": round((ca.rays / spec.n_ao) / (W * H), 4),
            "kernel_ms_per_frame": {"ao_primary": round(kp.total_ms / max(1, kp.launches), 4),
                                    "ao_rays": round(ka.total_ms / max(1, ka.launches), 4)},
            "blas_build_ms": round(blas_build_ms, 2),
            "scene_upload_and_build_s": round(upload_build_s, 3),
            "bvh": {"nodes": info.nodes, "depth": info.depth, "bytes": info.bytes},
            "primary_bytes_per_ray": round(bytes_per_ray(cp), 1),
        },
    }

    if not a.no_cpu:
        result["cpu_baseline"] = cpu_baseline(ctx, gs, sd, spec, W, H, a.cpu_seconds)
    ctx.close()
    print(json.dumps(result))


def bench_pathtrace(a):
    """the wavefront path tracer on C0 (Cornell, 512^2, 16 spp, depth 4, NEE)
    or C2 (10k instances, 1080p, 64 spp, depth 8, RR from 3, gradient sky,
    all BLAS and the TLAS rebuilt every frame); one JSON line"""
    import solhip
    from solhip import scenes
    c0 = a.workload == "pathtrace-c0"
    sd = scenes.cornell_box() if c0 else scenes.instanced(10000)
    W, H = (512, 512) if c0 else (1920, 1080)
    ctx = solhip.Context(0)
    # C0 is static: binned SAH; C2 rebuilds every BLAS each frame: Morton order
    gs = solhip.GpuScene(ctx, sd, build_flags=solhip.BUILD_PREFER_FAST_TRACE if c0 else 0)
    spec = solhip.spec_default(solhip.PROG_PATHTRACE)
    if c0:
        spec.enable_sky, spec.spp_per_frame, spec.max_depth, spec.nee, spec.rr_start = 0, 16, 4, 1, 0
    else:
        spec.enable_sky, spec.spp_per_frame, spec.max_depth, spec.nee, spec.rr_start = 2, 64, 8, 1, 3
    spec.wf_graph = 1
    # one random stream per sample (sample_streams = 1) lets a pass trace many
    # samples of every pixel at once; 0 keeps pathtrace.rgen's one stream per
    # pixel, whose samples must follow one another
    spec.sample_streams = 1 if a.streams is None else a.streams
    spec.wf_samples_per_pass = a.per_pass
    spec.wf_slices = a.slices
    acc = ctx.upload(np.zeros((H, W, 4), np.float32))
    img = solhip.Images(acc.h, 0, 0, 0, 0, 0)
    tuned = None
    if not c0 and not a.no_tune:
        # the launch shape (block, persistent grid, graph, image slices) the
        # library's autotuner picks for this workload; cached per workload
        print("pathtrace: autotuning the launch shape", file=sys.stderr, flush=True)
        best, best_ms = ctx.autotune_pathtrace(gs.tlas, gs.tables, sd.uniforms(W, H, 0), img, W, H, spec, 1,
                                               a.tune_cache)
        tuned = {"wf_block": best.wf_block, "wf_blocks_per_cu": best.wf_blocks_per_cu, "wf_graph": best.wf_graph,
                 "wf_slices": best.wf_slices, "frame_ms": round(best_ms, 3)}
        spec = best
    prog = ctx.program(solhip.PROG_PATHTRACE, spec)

    def rebuild():
        # C2: every BLAS and the TLAS again (a per-frame animated scene)
        for k in range(len(sd.blas)):
            gs.rebuild_blas(k)
        gs.update_instances(sd.instance_descs)

    for i in range(a.warmup):
        if not c0:
            rebuild()
        ctx.trace(prog, gs.tlas, gs.tables, sd.uniforms(W, H, i), 0, img, W, H)
    ctx.sync()
    ctx.stats(reset=True)
    unis = [sd.uniforms(W, H, a.warmup + i) for i in range(a.steps)]
    t0 = time.perf_counter()
    for i in range(a.steps):
        if not c0:
            rebuild()
        ctx.trace(prog, gs.tlas, gs.tables, unis[i], a.warmup, img, W, H)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    st = ctx.stats(reset=True)
    assert st.overflow == 0
    roofline = None if c0 else extend_roofline(ctx, gs, sd, spec, img, W, H, rebuild, a.warmup + a.steps)
    result = {
        "metric": "Mrays/s on 1 GPU", "value": round(st.rays / elapsed / 1e6, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(elapsed / a.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": ("wavefront path tracer, Cornell box (36 triangles), 512x512, 16 spp, depth 4, "
                                "next-event estimation" if c0 else
                                "wavefront path tracer, 10000 instances of 5 BLAS (%d triangles), 1920x1080, 64 spp, "
                                "depth 8, Russian roulette from bounce 3, next-event estimation (no emitters), "
                                "gradient sky, all BLAS + TLAS rebuilt per frame" % sd.tri_count),
                   "resolution": [W, H], "spp": spec.spp_per_frame, "max_depth": spec.max_depth,
                   "sample_streams": spec.sample_streams, "parallelism": "single GPU"},
        "extra": {"rays_per_frame": int(st.rays / a.steps), "hipgraph": bool(spec.wf_graph),
                  "samples_per_pass": a.per_pass or "default", "slices": spec.wf_slices or "default",
                  "autotuned": tuned},
    }
    if roofline:
        result["roofline"] = roofline
    if not c0 and not a.no_cpu:
        result["cpu_baseline"] = cpu_baseline_pathtrace(sd, spec, W, H, a.cpu_seconds)
    ctx.close()
    print(json.dumps(result))


def extend_roofline(ctx, gs, sd, spec, img, W, H, rebuild, frame0):
    """roofline of the wavefront tracer's dominant kernel (the extension
    kernel k_pt_extend): fetch counts from its counting variant on one frame,
    launch durations from hipEvents on the launch stream over two more
    frames.  Events are per launch, so these frames run eagerly on one image
    slice (the captured graph and the slices of the timed region would
    hide the launches), on the library's default persistent grid (the
    launch shape of the committed counter summary); the kernel is the same."""
    import solhip
    # the library's default persistent grid, which fills the GPU on its own:
    # the autotuned shape is chosen for image slices running side by side and
    # leaves the GPU half empty in a one-slice launch
    spec = type(spec).from_buffer_copy(spec)
    spec.wf_block, spec.wf_blocks_per_cu = 0, 0
    rebuild()
    ctx.profile(solhip.api.PROFILE_COUNT)
    for k in ("pt_extend", "pt_shade", "pt_shadow"):
        ctx.kernel_stats(k, reset=True)
    ctx.trace(ctx_prog(ctx, spec), gs.tlas, gs.tables, sd.uniforms(W, H, frame0), frame0, img, W, H)
    cnt = ctx.kernel_stats("pt_extend", reset=True).counters
    ctx.profile(solhip.api.PROFILE_TIME)
    for k in ("pt_generate", "pt_extend", "pt_shade", "pt_shadow", "pt_finalize"):
        ctx.kernel_stats(k, reset=True)
    for f in range(2):
        rebuild()
        ctx.trace(ctx_prog(ctx, spec), gs.tlas, gs.tables, sd.uniforms(W, H, frame0 + 1 + f), frame0, img, W, H)
    ctx.sync()
    ctx.profile(0)
    ks = {k: ctx.kernel_stats(k) for k in ("pt_generate", "pt_extend", "pt_shade", "pt_shadow", "pt_finalize")}
    ext = ks["pt_extend"]
    b_ray = bytes_per_ray(cnt)
    achieved = b_ray * ext.counters.rays / (ext.total_ms / 1e3) / 1e9
    avg_launch_s = ext.total_ms / max(1, ext.launches) / 1e3
    return {"bound": "hbm", "kernel": "pt_extend", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": read_profile(C2_PMC_FILE).get("traffic_bytes_per_launch"),
            "limiter": limiter(C2_PMC_FILE, avg_launch_s),
            "bytes_per_ray": round(b_ray, 1),
            "fetches_per_ray": {"node": round(cnt.node_fetches / max(1, cnt.rays), 3),
                                "tri": round(cnt.tri_fetches / max(1, cnt.rays), 3),
                                "inst": round(cnt.inst_fetches / max(1, cnt.rays), 3)},
            "avg_launch_ms": round(ext.total_ms / max(1, ext.launches), 4),
            "rays_per_launch": int(ext.counters.rays / max(1, ext.launches)),
            "kernel_ms_per_frame": {k: round(v.total_ms / 2, 3) for k, v in ks.items()},
            "timing": "hipEvents per launch over 2 eager one-slice frames after the timed region, default persistent grid"}


_progs = {}


def ctx_prog(ctx, spec):
    """one program per spec (cached: the wavefront buffers are kept)"""
    import solhip
    key = bytes(spec)
    if key not in _progs:
        _progs[key] = ctx.program(solhip.PROG_PATHTRACE, spec)
    return _progs[key]


def cpu_baseline_pathtrace(sd, spec, W, H, target_s):
    """the oracle's restatement of the same estimator (sr_render_pathtrace,
    OpenMP over rows) on a band of rows of frame 0, sized to about target_s"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import solref
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    sc = solref.Scene(sd)
    cfg = solref.PTConfig(spec.enable_sky, spec.spp_per_frame, spec.max_depth, spec.rr_start, spec.nee,
                          spec.sample_streams)
    acc = np.zeros((H, W, 4), np.float32)
    u = sd.uniforms(W, H, 0)
    y0 = H // 2
    t0 = time.perf_counter()
    n = sc.render_pathtrace(u, cfg, 0, acc, None, W, H, y0, y0 + 4, nthreads=threads)
    dt = time.perf_counter() - t0
    rows = int(max(4, min(H - y0, 4 * target_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    n = sc.render_pathtrace(u, cfg, 0, acc, None, W, H, y0, y0 + rows, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": round(n / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": "rows %d-%d of frame 0 (%d x %d pixels x %d spp, %d rays, %.1f s wall, %d OpenMP threads)"
                      % (y0, y0 + rows - 1, W, rows, spec.spp_per_frame, n, dt, threads)}


def bench_server(a):
    """The render server over the C0, C1 and C2 scenes under the C3 request
    mix.  One GPU serves far fewer than the configuration's 200 requests/s, so
    there is no steady state at that rate (latency grows without bound with
    the number of requests).  Measured instead: the capacity (a saturating
    burst of --cap-jobs requests: jobs/s and Mrays/s), then Poisson loads at
    0.5x and 0.9x that capacity (--load-jobs requests each), the steady
    states whose p50 / p99 latency the line reports (value = p50 at 0.9x);
    the 200 req/s burst of --jobs requests is reported beside them with the
    factor by which it exceeds the capacity.  One JSON line."""
    import solhip
    from solhip import scenes, serve
    ctx = solhip.Context(0)
    sds = [scenes.cornell_box(), scenes.heightfield(1024), scenes.instanced(10000)]
    # the served scenes are static: binned SAH builds
    gss = [solhip.GpuScene(ctx, sd, build_flags=solhip.BUILD_PREFER_FAST_TRACE) for sd in sds]
    ctx.sync()
    warm = serve.draw_requests(4, 1000.0, seed=1, resolutions=[(256, 256)], spp_max=4)
    streams = 0 if a.streams is None else a.streams
    serve.run_load(ctx, gss, sds, warm, streams=streams)

    def brief(st, rate):
        keys = ("jobs", "makespan_s", "jobs_per_s", "mrays_per_s", "latency_ms_p50", "latency_ms_p99", "queue_ms_p50",
                "queue_ms_p99")
        out = {k: round(st[k], 3) if isinstance(st[k], float) else st[k] for k in keys}
        out["offered_req_per_s"] = round(rate, 3)
        return out

    print("server: capacity burst of %d requests" % a.cap_jobs, file=sys.stderr, flush=True)
    cap = serve.run_load(ctx, gss, sds, serve.draw_requests(a.cap_jobs, 1e9, seed=11), streams=streams)
    capacity = cap["jobs_per_s"]
    loads = []
    for frac, seed in ((0.5, 12), (0.9, 13)):
        rate = frac * capacity
        print("server: Poisson load at %.1fx capacity (%.2f req/s, %d requests)" % (frac, rate, a.load_jobs),
              file=sys.stderr, flush=True)
        st = serve.run_load(ctx, gss, sds, serve.draw_requests(a.load_jobs, rate, seed=seed), streams=streams)
        row = brief(st, rate)
        row["load_vs_capacity"] = frac
        loads.append(row)
    print("server: the configuration's %.0f req/s, %d requests" % (a.rate, a.jobs), file=sys.stderr, flush=True)
    c3 = serve.run_load(ctx, gss, sds, serve.draw_requests(a.jobs, a.rate, seed=3), streams=streams)
    steady = loads[-1]
    result = {
        "metric": "render-request latency on 1 GPU", "value": round(steady["latency_ms_p50"], 2),
        "unit": "ms (p50, Poisson load at 0.9x the measured capacity)",
        "n_gpus": 1, "steps": steady["jobs"], "warmup": 4, "ms_per_step": round(steady["makespan_s"] * 1e3 /
                                                                                 steady["jobs"], 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "render server, requests over Cornell / 2.1M-triangle heightfield / 10k instances, "
                               "resolutions 256^2..1080p, spp log-uniform 1-256, depth 4-8, NEE + roulette; "
                               "Poisson arrivals at 0.5x and 0.9x the measured capacity (the configuration's "
                               "%.0f req/s exceeds one GPU by %.1fx)" % (a.rate, a.rate / capacity),
                   "sample_streams": streams, "parallelism": "single GPU"},
        "capacity_jobs_per_s": round(capacity, 3),
        "capacity_mrays_per_s": round(cap["mrays_per_s"], 1),
        "latency_vs_load": loads,
        "offered_200_burst": dict(brief(c3, a.rate), exceeds_capacity_by=round(a.rate / capacity, 1)),
        "extra": {"capacity_run": brief(cap, float("inf"))},
    }
    ctx.close()
    print(json.dumps(result))


def bench_cache(a):
    """learned radiance cache against the plain tracer on the C2 scene:
    relative RMSE against a 4096-spp plain reference as a function of wall
    time, and the time the cache needs to reach the plain tracer's 1024-spp
    error (if it does); one JSON line"""
    import solhip
    from solhip import scenes
    W, H = (int(v) for v in a.res.lower().split("x"))
    sd = scenes.instanced(10000)
    ctx = solhip.Context(0)
    gs = solhip.GpuScene(ctx, sd)

    def spec(nrc, spp, streams=0, lr=0.0, steps=0):
        s = solhip.spec_default(solhip.PROG_PATHTRACE)
        s.enable_sky, s.spp_per_frame, s.max_depth, s.rr_start, s.nee, s.nrc = 2, spp, 8, 3, 1, nrc
        s.sample_streams = streams
        s.nrc_learning_rate = lr      # 0: the library default, Adam lr 1e-2 (the configuration's)
        s.nrc_train_steps = steps     # 0: 4 Adam steps per frame
        return s

    acc = ctx.upload(np.zeros((H, W, 4), np.float32))
    img = solhip.Images(acc.h, 0, 0, 0, 0, 0)
    # reference: 4096 spp of the plain tracer (frames of 64 spp with per-sample
    # streams, running mean; the same expectation as the estimators compared)
    pr = ctx.program(solhip.PROG_PATHTRACE, spec(0, 64, streams=1))
    for f in range(64):
        ctx.trace(pr, gs.tlas, gs.tables, sd.uniforms(W, H, 100000 + f), 100000, img, W, H)
    ref = acc.download(np.float32).reshape(H, W, 4)[..., :3].astype(np.float64)
    ctx.program_destroy(pr)
    norm = float(np.sqrt((ref ** 2).mean()))

    def run(nrc, frames, checkpoints, warm=0, lr=0.0, steps=0):
        """frames of 1 spp into a running mean; with warm > 0 the first warm
        frames train the cache without being kept (their time counts)"""
        p = ctx.program(solhip.PROG_PATHTRACE, spec(nrc, 1, lr=lr, steps=steps))
        out = []
        t = 0.0
        for f in range(warm + frames):
            ctx.sync()
            t0 = time.perf_counter()
            start = f if f < warm else warm   # alpha = 1 while warming up
            ctx.trace(p, gs.tlas, gs.tables, sd.uniforms(W, H, f), start, img, W, H)
            ctx.sync()
            t += time.perf_counter() - t0
            n = f + 1 - warm
            if n in checkpoints:
                im = acc.download(np.float32).reshape(H, W, 4)[..., :3].astype(np.float64)
                out.append({"spp": n, "seconds": round(t, 4),
                            "rel_rmse": round(float(np.sqrt(((im - ref) ** 2).mean())) / norm, 5)})
        ctx.program_destroy(p)
        return out

    def floor(curve):
        """least-squares fit of rel_rmse^2 = bias^2 + var / spp over the
        checkpoints from 16 spp on: the error a run keeps however long it
        runs (bias) and the per-sample variance"""
        pts = [(c["spp"], c["rel_rmse"] ** 2) for c in curve if c["spp"] >= 16]
        A = np.array([[1.0, 1.0 / n] for n, _ in pts])
        y = np.array([e for _, e in pts])
        (b2, v), *_ = np.linalg.lstsq(A, y, rcond=None)
        return {"bias": round(float(np.sqrt(max(b2, 0.0))), 5), "sigma_1spp": round(float(np.sqrt(max(v, 0.0))), 4)}

    cps = [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024]
    plain = run(0, 1024, cps)
    cache = run(1, 1024, cps, warm=64)
    target = plain[-1]["rel_rmse"]
    reach = next((c for c in cache if c["rel_rmse"] <= target), None)
    # experiments on the cache's error floor: the step size, and training
    # four times as long per frame
    experiments = []
    for name, lr, steps in (("adam lr 1e-3", 1e-3, 0), ("16 Adam steps per frame", 0.0, 16)):
        print("cache-c4: experiment %s" % name, file=sys.stderr, flush=True)
        c = run(1, 1024, cps, warm=64, lr=lr, steps=steps)
        experiments.append({"name": name, "curve": c, "fit": floor(c),
                            "rel_rmse_1024": c[-1]["rel_rmse"], "seconds_1024": c[-1]["seconds"]})
    result = {
        "metric": "learned-cache time to the plain tracer's 1024-spp error, fraction of the plain time",
        "value": round(reach["seconds"] / plain[-1]["seconds"], 4) if reach else None,
        "unit": "fraction", "n_gpus": 1, "steps": 1024, "warmup": 0,
        "ms_per_step": round(cache[-1]["seconds"] / 1024 * 1e3, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (network) / f32",
        "data": "synthetic",
        "config": {"workload": "10000 instances of 5 BLAS, %dx%d, 1 spp per frame, depth 8, RR from 3, NEE, "
                               "gradient sky; cache: 64-64x4-3 ReLU MLP, BF16 MFMA, 4 Adam steps (lr 1e-2) per frame on "
                               "<= 65536 rows from 1-in-R training paths, 64 warm-up frames (timed, not "
                               "accumulated); reference 4096 spp plain" % (W, H),
                   "parallelism": "single GPU"},
        "extra": {"plain": plain, "cache": cache, "plain_1024spp_rel_rmse": target,
                  "fit": {"plain": floor(plain), "cache": floor(cache)}, "experiments": experiments},
    }
    ctx.close()
    print(json.dumps(result))


def cpu_baseline(ctx, gs, sd, spec, W, H, target_s):
    """the oracle on whole frames 0..F-1 of the same workload (F sized to about
    target_s of CPU work, at most 30 frames); the GPU renders the same frames
    into a fresh accumulation buffer and the two are compared bit for bit"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import solref
    import solhip

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    sc = solref.Scene(sd)
    params = solref.AOParams(spec.n_ao, spec.ao_tmin, spec.ao_tmax, 0)
    acc = np.zeros((H, W, 4), np.float32)
    t0 = time.perf_counter()
    nrays, _, _ = sc.render_ao(sd.uniforms(W, H, 0), params, 0, acc, W, H, nthreads=threads)
    t_first = time.perf_counter() - t0
    frames = int(max(1, min(30, round(target_s / max(t_first, 1e-3)))))
    t0 = time.perf_counter()
    for f in range(1, frames):
        n, _, _ = sc.render_ao(sd.uniforms(W, H, f), params, 0, acc, W, H, nthreads=threads)
        nrays += n
    dt = t_first + (time.perf_counter() - t0)
    prog = ctx.program(solhip.PROG_AO, spec)
    gacc = ctx.upload(np.zeros((H, W, 4), np.float32))
    img = solhip.Images(gacc.h, 0, 0, 0, 0, 0)
    for f in range(frames):
        ctx.trace(prog, gs.tlas, gs.tables, sd.uniforms(W, H, f), 0, img, W, H)
    g = gacc.download(np.float32).reshape(H, W, 4)
    parity = bool(np.array_equal(g.view(np.uint32), acc.view(np.uint32)))
    return {
        "value": round(nrays / dt / 1e6, 3),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": "frames 0-%d of the same workload (%d rays, %.1f s wall, %d OpenMP threads); the GPU "
                  "accumulation of the same frames is bit-identical: %s" % (frames - 1, nrays, dt, threads, parity),
        "gpu_parity_on_sample": parity,
    }


if __name__ == "__main__":
    main()
