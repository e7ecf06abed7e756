"""The wavefront path tracer (SOLHIP_PROG_PATHTRACE) on the device.

* Without extensions it must reproduce SOLHIP_PROG_PATHTRACE_GLSL (the
  pathtrace.* megakernel, itself bit-exact against the oracle) bit for bit:
  accumulation, RGBA8 and ray count.
* With next-event estimation and Russian roulette it must reproduce the
  oracle's per-pixel restatement of the same estimator bit for bit.
* Next-event estimation and roulette keep the expectation: on the device, at
  sample counts the CPU cannot afford, block means agree with the plain
  estimator within 5 standard errors."""
import numpy as np
import pytest

import solhip
from solhip import scenes

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _spec(kind, sky=0, spp=4, depth=6, rr=0, nee=0, block=0, per_cu=0):
    s = solhip.spec_default(kind)
    s.enable_sky = sky
    s.spp_per_frame = spp
    s.max_depth = depth
    if kind == solhip.PROG_PATHTRACE:
        s.rr_start = rr
        s.nee = nee
        s.wf_block = block
        s.wf_blocks_per_cu = per_cu
    return s


def _render(ctx, gs, sd, kind, spec, W, H, frames, accum_start=0):
    """frames rendered into one accumulation target; returns (accum, rgba8, rays)"""
    prog = ctx.program(kind, spec)
    acc = ctx.upload(np.zeros((H, W, 4), np.float32))
    out = ctx.upload(np.zeros((H, W, 4), np.uint8))
    img = solhip.Images(acc.h, out.h, 0, 0, 0, 0)
    ctx.stats(reset=True)
    for f in range(frames):
        ctx.trace(prog, gs.tlas, gs.tables, sd.uniforms(W, H, f), accum_start, img, W, H)
    rays = ctx.stats().rays
    a, o = acc.download(np.float32).reshape(H, W, 4), out.download(np.uint8).reshape(H, W, 4)
    ctx.program_destroy(prog)
    return a, o, rays


def _mk(kind, **kw):
    return kind, _spec(kind, **kw)


SCENES = {"cornell": scenes.cornell_box, "instances": scenes.random_instances}


@pytest.mark.parametrize("name,sky,W,H", [("cornell", 0, 40, 32), ("cornell", 1, 37, 21), ("instances", 1, 45, 27)])
def test_plain_wavefront_equals_glsl_megakernel(ctx, name, sky, W, H):
    sd = SCENES[name]()
    gs = solhip.GpuScene(ctx, sd)
    print("wavefront vs megakernel:", name, sky, W, H, flush=True)
    a1, o1, n1 = _render(ctx, gs, sd, *_mk(solhip.PROG_PATHTRACE_GLSL, sky=sky, spp=3, depth=6), W, H, 2)
    a2, o2, n2 = _render(ctx, gs, sd, *_mk(solhip.PROG_PATHTRACE, sky=sky, spp=3, depth=6), W, H, 2)
    assert (_bits(a1) != _bits(a2)).sum() == 0
    assert np.array_equal(o1, o2)
    assert n1 == n2 and n1 > 0


@pytest.mark.parametrize("block,per_cu", [(64, 0), (128, 4), (256, 2)])
def test_launch_shapes_do_not_change_the_image(ctx, block, per_cu):
    sd = scenes.cornell_box()
    gs = solhip.GpuScene(ctx, sd)
    W, H = 33, 29
    print("launch shape:", block, per_cu, flush=True)
    a0, _, n0 = _render(ctx, gs, sd, *_mk(solhip.PROG_PATHTRACE, spp=2, depth=4, nee=1, rr=2), W, H, 1)
    a1, _, n1 = _render(ctx, gs, sd, *_mk(solhip.PROG_PATHTRACE, spp=2, depth=4, nee=1, rr=2, block=block,
                                         per_cu=per_cu), W, H, 1)
    assert (_bits(a0) != _bits(a1)).sum() == 0 and n0 == n1


@pytest.mark.parametrize("slices,graph", [(2, 0), (3, 0), (16, 0), (3, 1), (16, 1)])
def test_image_slices_do_not_change_the_image(ctx, slices, graph):
    """the frame split into tile slices on their own streams (and captured
    as one graph with a branch per slice) equals the one-stream frame, on
    an image whose 24 tiles do not split evenly"""
    sd = scenes.random_instances()
    gs = solhip.GpuScene(ctx, sd)
    W, H = 45, 27
    print("slices:", slices, graph, flush=True)
    s0 = _spec(solhip.PROG_PATHTRACE, sky=1, spp=3, depth=6, nee=1, rr=2)
    s0.wf_slices = 1
    s1 = _spec(solhip.PROG_PATHTRACE, sky=1, spp=3, depth=6, nee=1, rr=2)
    s1.wf_slices, s1.wf_graph = slices, graph
    a0, o0, n0 = _render(ctx, gs, sd, solhip.PROG_PATHTRACE, s0, W, H, 3)
    a1, o1, n1 = _render(ctx, gs, sd, solhip.PROG_PATHTRACE, s1, W, H, 3)
    assert (_bits(a0) != _bits(a1)).sum() == 0 and np.array_equal(o0, o1) and n0 == n1 > 0


@pytest.mark.parametrize("name,sky,nee,rr,depth", [("cornell", 0, 1, 0, 4), ("cornell", 0, 1, 2, 8),
                                                  ("cornell", 1, 0, 1, 5), ("instances", 2, 1, 3, 8)])
def test_extended_wavefront_matches_oracle(ctx, oracle, name, sky, nee, rr, depth):
    sd = SCENES[name]()
    gs = solhip.GpuScene(ctx, sd)
    W, H, spp, F = 36, 28, 3, 2
    print("wavefront vs oracle:", name, sky, nee, rr, depth, flush=True)
    a, o, n = _render(ctx, gs, sd, *_mk(solhip.PROG_PATHTRACE, sky=sky, spp=spp, depth=depth, rr=rr, nee=nee), W, H, F)
    sc = oracle.Scene(sd)
    ref = np.zeros((H, W, 4), np.float32)
    ref8 = np.zeros((H, W, 4), np.uint8)
    nref = 0
    for f in range(F):
        nref += sc.render_pathtrace(sd.uniforms(W, H, f), oracle.PTConfig(sky, spp, depth, rr, nee), 0, ref, ref8, W, H)
    assert (_bits(a) != _bits(ref)).sum() == 0
    assert np.abs(o.astype(int) - ref8.astype(int)).max() <= 1
    assert n == nref


@pytest.mark.parametrize("name,per_pass,slices,graph", [("cornell", 1, 1, 0), ("cornell", 2, 1, 0), ("cornell", 5, 3, 1),
                                                        ("instances", 0, 1, 0), ("instances", 3, 2, 1)])
def test_sample_streams_match_oracle(ctx, oracle, name, per_pass, slices, graph):
    """sample_streams = 1: every pass traces per_pass samples of every pixel
    at once (5 samples: a short last pass for 2 and 3); the image equals
    the oracle's per-pixel loop with the same seeding, bit for bit"""
    sd = SCENES[name]()
    gs = solhip.GpuScene(ctx, sd)
    W, H, spp, F = 36, 28, 5, 2
    sky, nee, rr, depth = (0, 1, 2, 6) if name == "cornell" else (2, 1, 3, 8)
    print("sample streams vs oracle:", name, per_pass, slices, graph, flush=True)
    s = _spec(solhip.PROG_PATHTRACE, sky=sky, spp=spp, depth=depth, rr=rr, nee=nee)
    s.sample_streams, s.wf_samples_per_pass, s.wf_slices, s.wf_graph = 1, per_pass, slices, graph
    a, o, n = _render(ctx, gs, sd, solhip.PROG_PATHTRACE, s, W, H, F)
    sc = oracle.Scene(sd)
    ref = np.zeros((H, W, 4), np.float32)
    ref8 = np.zeros((H, W, 4), np.uint8)
    nref = 0
    for f in range(F):
        nref += sc.render_pathtrace(sd.uniforms(W, H, f), oracle.PTConfig(sky, spp, depth, rr, nee, 1), 0, ref, ref8,
                                    W, H)
    assert (_bits(a) != _bits(ref)).sum() == 0
    assert np.abs(o.astype(int) - ref8.astype(int)).max() <= 1
    assert n == nref


def test_depth_zero_and_single_sample(ctx, oracle):
    sd = scenes.cornell_box()
    gs = solhip.GpuScene(ctx, sd)
    W, H = 24, 24
    a, _, n = _render(ctx, gs, sd, *_mk(solhip.PROG_PATHTRACE, spp=1, depth=0, nee=1), W, H, 1)
    ref = np.zeros((H, W, 4), np.float32)
    nref = oracle.Scene(sd).render_pathtrace(sd.uniforms(W, H, 0), oracle.PTConfig(0, 1, 0, 0, 1), 0, ref, None, W, H)
    assert (_bits(a) != _bits(ref)).sum() == 0 and n == nref == W * H


def test_emitter_free_scene_with_nee(ctx, oracle):
    """no emitter: the table is empty and next-event estimation is a no-op"""
    sd = scenes.random_instances()
    gs = solhip.GpuScene(ctx, sd)
    W, H = 20, 20
    a0, _, _ = _render(ctx, gs, sd, *_mk(solhip.PROG_PATHTRACE, sky=1, spp=2, depth=4), W, H, 1)
    a1, _, _ = _render(ctx, gs, sd, *_mk(solhip.PROG_PATHTRACE, sky=1, spp=2, depth=4, nee=1), W, H, 1)
    assert (_bits(a0) != _bits(a1)).sum() == 0


def test_bad_specs_are_rejected(ctx):
    for kw in ({"depth": 5000}, {"sky": 3}, {"block": 96}, {"block": 512}, {"spp": 0}):
        s = _spec(solhip.PROG_PATHTRACE, **kw)
        with pytest.raises(solhip.SolhipError):
            ctx.program(solhip.PROG_PATHTRACE, s)
    for field, value in (("wf_slices", 17), ("sample_streams", 2), ("wf_samples_per_pass", 4)):
        s = _spec(solhip.PROG_PATHTRACE)
        setattr(s, field, value)
        with pytest.raises(solhip.SolhipError):
            ctx.program(solhip.PROG_PATHTRACE, s)


def _block_means(ctx, gs, sd, spec, W, H, frames, blk=8):
    """per-frame 8x8 block means (alpha = 1 each frame: accum_start = frame)"""
    prog = ctx.program(solhip.PROG_PATHTRACE, spec)
    acc = ctx.upload(np.zeros((H, W, 4), np.float32))
    img = solhip.Images(acc.h, 0, 0, 0, 0, 0)
    out = []
    for f in range(frames):
        ctx.trace(prog, gs.tlas, gs.tables, sd.uniforms(W, H, f), f, img, W, H)
        a = acc.download(np.float32).reshape(H // blk, blk, W // blk, blk, 4)[..., :3].astype(np.float64)
        out.append(a.mean(axis=(1, 3)))
    ctx.program_destroy(prog)
    return np.array(out)


@pytest.mark.parametrize("nee,rr", [(1, 0), (1, 3)])
def test_extensions_converge_to_the_plain_estimator(ctx, nee, rr):
    """Cornell box, depth 4, 64 x 64 pixels, 16 frames x 256 samples: every
    8 x 8 block mean within 5 standard errors of the plain estimator's (and
    the whole-image mean within 4)"""
    sd = scenes.cornell_box()
    gs = solhip.GpuScene(ctx, sd)
    W = H = 64
    F = 16
    print("convergence:", nee, rr, flush=True)
    p = _block_means(ctx, gs, sd, _spec(solhip.PROG_PATHTRACE, spp=256, depth=4), W, H, F)
    e = _block_means(ctx, gs, sd, _spec(solhip.PROG_PATHTRACE, spp=256, depth=4, nee=nee, rr=rr), W, H, F)
    se = np.sqrt(p.var(0, ddof=1) / F + e.var(0, ddof=1) / F) + 1e-7
    z = np.abs(p.mean(0) - e.mean(0)) / se
    assert z.max() < 5.0, z.max()
    pm, em = p.mean(axis=(1, 2)), e.mean(axis=(1, 2))
    zs = np.abs(pm.mean(0) - em.mean(0)) / np.sqrt(pm.var(0, ddof=1) / F + em.var(0, ddof=1) / F)
    assert zs.max() < 4.0, zs


def test_graph_replays_equal_eager_frames(ctx):
    """wf_graph = 1: every replay (new uniforms each frame, other eager work
    and syncs in between, a TLAS rebuild midway) equals the eager launch"""
    sd = scenes.random_instances()
    sd_lit = scenes.cornell_box()
    W, H = 40, 24
    for scene in (sd, sd_lit):
        gs = solhip.GpuScene(ctx, scene)
        se = _spec(solhip.PROG_PATHTRACE, sky=1, spp=2, depth=5, nee=1, rr=2)
        sg = _spec(solhip.PROG_PATHTRACE, sky=1, spp=2, depth=5, nee=1, rr=2)
        sg.wf_graph = 1
        pe, pg = ctx.program(solhip.PROG_PATHTRACE, se), ctx.program(solhip.PROG_PATHTRACE, sg)
        acc_e = ctx.upload(np.zeros((H, W, 4), np.float32))
        acc_g = ctx.upload(np.zeros((H, W, 4), np.float32))
        out_e = ctx.upload(np.zeros((H, W, 4), np.uint8))
        out_g = ctx.upload(np.zeros((H, W, 4), np.uint8))
        ie = solhip.Images(acc_e.h, out_e.h, 0, 0, 0, 0)
        ig = solhip.Images(acc_g.h, out_g.h, 0, 0, 0, 0)
        other = ctx.upload(np.zeros((H, W, 4), np.float32))
        io = solhip.Images(other.h, 0, 0, 0, 0, 0)
        for f in range(6):
            print("graph replay frame", f, flush=True)
            if f == 3:
                descs = scene.instance_descs.copy()
                descs["transform"][:, 3] += 0.05
                gs.update_instances(descs)
            u = scene.uniforms(W, H, f)
            ctx.trace(pe, gs.tlas, gs.tables, u, 0, ie, W, H)
            ctx.trace(pg, gs.tlas, gs.tables, u, 0, ig, W, H)
            ctx.trace(pe, gs.tlas, gs.tables, scene.uniforms(W, H, 100 + f), 0, io, W, H)   # other eager work
            ctx.sync()
            assert (_bits(acc_e.download(np.float32)) != _bits(acc_g.download(np.float32))).sum() == 0, f
            assert np.array_equal(out_e.download(np.uint8), out_g.download(np.uint8)), f
        ctx.program_destroy(pe)
        ctx.program_destroy(pg)


def test_autotuner_picks_and_caches_a_shape(ctx, tmp_path):
    sd = scenes.cornell_box()
    gs = solhip.GpuScene(ctx, sd)
    W, H = 96, 64
    acc = ctx.upload(np.zeros((H, W, 4), np.float32))
    img = solhip.Images(acc.h, 0, 0, 0, 0, 0)
    spec = _spec(solhip.PROG_PATHTRACE, spp=2, depth=4, nee=1, rr=2)
    cache = str(tmp_path / "tune.txt")
    print("autotune", flush=True)
    best, ms = ctx.autotune_pathtrace(gs.tlas, gs.tables, sd.uniforms(W, H, 0), img, W, H, spec, 2, cache)
    assert best.wf_block in (64, 128, 256) and best.wf_blocks_per_cu > 0 and best.wf_graph in (0, 1) and ms > 0
    assert best.nee == 1 and best.rr_start == 2 and best.max_depth == 4
    lines = open(cache).read().splitlines()
    assert len(lines) == 1 and "spp=2 depth=4 nee=1 rr=2" in lines[0]
    again, ms2 = ctx.autotune_pathtrace(gs.tlas, gs.tables, sd.uniforms(W, H, 0), img, W, H, spec, 2, cache)
    assert (again.wf_block, again.wf_blocks_per_cu, again.wf_graph) == (best.wf_block, best.wf_blocks_per_cu,
                                                                           best.wf_graph)
    assert len(open(cache).read().splitlines()) == 1
    # the tuned shape renders the same image as the default one
    a0, _, _ = _render(ctx, gs, sd, solhip.PROG_PATHTRACE, spec, W, H, 1)
    a1, _, _ = _render(ctx, gs, sd, solhip.PROG_PATHTRACE, best, W, H, 1)
    assert (_bits(a0) != _bits(a1)).sum() == 0


def test_autotuner_picks_slices_with_sample_streams(ctx, tmp_path):
    """with per-sample streams the tuner also times image slices; its pick
    renders the same image as the default shape"""
    sd = scenes.random_instances()
    gs = solhip.GpuScene(ctx, sd)
    W, H = 80, 48
    acc = ctx.upload(np.zeros((H, W, 4), np.float32))
    img = solhip.Images(acc.h, 0, 0, 0, 0, 0)
    spec = _spec(solhip.PROG_PATHTRACE, sky=2, spp=4, depth=5, nee=1, rr=2)
    spec.sample_streams = 1
    cache = str(tmp_path / "tune.txt")
    print("autotune with sample streams", flush=True)
    best, ms = ctx.autotune_pathtrace(gs.tlas, gs.tables, sd.uniforms(W, H, 0), img, W, H, spec, 1, cache)
    assert best.wf_slices in (0, 1, 2, 4, 8) and best.sample_streams == 1 and ms > 0
    lines = open(cache).read().splitlines()
    assert len(lines) == 1 and "streams=1" in lines[0] and "slices=" in lines[0]
    again, _ = ctx.autotune_pathtrace(gs.tlas, gs.tables, sd.uniforms(W, H, 0), img, W, H, spec, 1, cache)
    assert again.wf_slices == best.wf_slices and len(open(cache).read().splitlines()) == 1
    a0, _, _ = _render(ctx, gs, sd, solhip.PROG_PATHTRACE, spec, W, H, 1)
    a1, _, _ = _render(ctx, gs, sd, solhip.PROG_PATHTRACE, best, W, H, 1)
    assert (_bits(a0) != _bits(a1)).sum() == 0
