"""GPU parity of the ray-tracing programs against the oracle's restatement of
the GLSL (same scenes, cameras, frame seeds and accumulation).

Bar: accumulation buffers (RGBA32F) and hit records bit-identical; the RGBA8
gamma output within 1 LSB (pow differs between the device and host math
libraries; nothing downstream reads it)."""
import numpy as np
import pytest

import solhip
from solhip import scenes
from solhip.records import HIT

pytestmark = pytest.mark.gpu


def _images(ctx, W, H, n_hits=0, blue=None, rgba8=False):
    acc = ctx.upload(np.zeros((H, W, 4), np.float32))
    out = ctx.upload(np.zeros((H, W, 4), np.uint8)) if rgba8 else None
    hits = ctx.buffer(n_hits * HIT.itemsize) if n_hits else None
    bn = ctx.upload(blue) if blue is not None else None
    img = solhip.Images(acc.h, out.h if out else 0, bn.h if bn else 0,
                        blue.shape[1] if blue is not None else 0, blue.shape[0] if blue is not None else 0,
                        hits.h if hits else 0)
    return img, acc, out, hits


def _diff(a, b):
    a = np.ascontiguousarray(a).view(np.uint32).reshape(-1)
    b = np.ascontiguousarray(b).view(np.uint32).reshape(-1)
    return int((a != b).sum())


AO_SCENES = {
    "cornell": (scenes.cornell_box, 48, 40),
    "heightfield128": (lambda: scenes.heightfield(128), 64, 36),
    "instances": (scenes.random_instances, 48, 32),
}


@pytest.mark.parametrize("build", [0, solhip.BUILD_PREFER_FAST_TRACE], ids=["morton", "sah"])
@pytest.mark.parametrize("name", sorted(AO_SCENES))
def test_ao_matches_oracle_three_frames(ctx, oracle, name, build):
    f, W, H = AO_SCENES[name]
    sd = f()
    gs = solhip.GpuScene(ctx, sd, build_flags=build)
    spec = solhip.spec_default(solhip.PROG_AO)
    prog = ctx.program(solhip.PROG_AO, spec)
    img, acc, _, hits = _images(ctx, W, H, n_hits=W * H * (1 + spec.n_ao))
    sc = oracle.Scene(sd)
    ref = np.zeros((H, W, 4), np.float32)
    params = oracle.AOParams(spec.n_ao, spec.ao_tmin, spec.ao_tmax, 0)
    for frame in range(3):
        u = sd.uniforms(W, H, frame)
        ctx.stats(reset=True)
        ctx.trace(prog, gs.tlas, gs.tables, u, 0, img, W, H)
        st = ctx.stats()
        nrays, ph, ah = sc.render_ao(u, params, 0, ref, W, H, want_hits=True)
        g_hits = hits.download(HIT)
        assert _diff(g_hits, np.concatenate([ph, ah])) == 0, "hit records differ (frame %d)" % frame
        assert _diff(acc.download(np.float32), ref) == 0, "AO accumulation differs (frame %d)" % frame
        assert st.rays == nrays and st.overflow == 0


def test_ao_any_hit_same_image(ctx, oracle):
    sd = scenes.heightfield(128)
    W, H = 64, 36
    gs = solhip.GpuScene(ctx, sd)
    spec = solhip.spec_default(solhip.PROG_AO)
    spec.ao_any_hit = 1
    prog = ctx.program(solhip.PROG_AO, spec)
    img, acc, _, _ = _images(ctx, W, H)
    u = sd.uniforms(W, H, 7)
    ctx.trace(prog, gs.tlas, gs.tables, u, 7, img, W, H)
    ref = np.zeros((H, W, 4), np.float32)
    oracle.Scene(sd).render_ao(u, oracle.AOParams(spec.n_ao, spec.ao_tmin, spec.ao_tmax, 0), 7, ref, W, H)
    assert _diff(acc.download(np.float32), ref) == 0


@pytest.mark.parametrize("n_ao", [1, 4, 16])
def test_ao_ray_counts_per_pixel(ctx, oracle, n_ao):
    sd = scenes.cornell_box()
    W, H = 32, 24
    gs = solhip.GpuScene(ctx, sd)
    spec = solhip.spec_default(solhip.PROG_AO)
    spec.n_ao = n_ao
    prog = ctx.program(solhip.PROG_AO, spec)
    img, acc, _, hits = _images(ctx, W, H, n_hits=W * H * (1 + n_ao))
    u = sd.uniforms(W, H, 2)
    ctx.trace(prog, gs.tlas, gs.tables, u, 2, img, W, H)
    ref = np.zeros((H, W, 4), np.float32)
    _, ph, ah = oracle.Scene(sd).render_ao(u, oracle.AOParams(n_ao, spec.ao_tmin, spec.ao_tmax, 0), 2, ref, W, H,
                                           want_hits=True)
    assert _diff(hits.download(HIT), np.concatenate([ph, ah])) == 0
    assert _diff(acc.download(np.float32), ref) == 0


def test_ao_rejects_bad_n_ao(ctx):
    spec = solhip.spec_default(solhip.PROG_AO)
    spec.n_ao = 6
    with pytest.raises(solhip.SolhipError) as e:
        ctx.program(solhip.PROG_AO, spec)
    assert e.value.code == -6


@pytest.mark.parametrize("name", ["cornell", "heightfield128"])
def test_ao_glsl_matches_oracle(ctx, oracle, name):
    f, W, H = AO_SCENES[name]
    sd = f()
    gs = solhip.GpuScene(ctx, sd, index_size=8)
    blue = scenes.blue_noise_synthetic()
    prog = ctx.program(solhip.PROG_AO_GLSL)
    img, acc, _, _ = _images(ctx, W, H, blue=blue)
    sc = oracle.Scene(sd)
    ref = np.zeros((H, W, 4), np.float32)
    for frame in range(2):
        u = sd.uniforms(W, H, frame)
        ctx.trace(prog, gs.tlas, gs.tables, u, 0, img, W, H)
        sc.render_ao_glsl(u, 0, ref, blue, W, H)
        assert _diff(acc.download(np.float32), ref) == 0, "frame %d" % frame


@pytest.mark.parametrize("name,sky", [("cornell", 0), ("instances", 1), ("cornell_sky", 1)])
def test_pathtrace_glsl_matches_oracle(ctx, oracle, name, sky):
    sd = scenes.random_instances() if name == "instances" else scenes.cornell_box()
    W, H = 40, 32
    gs = solhip.GpuScene(ctx, sd)
    spec = solhip.spec_default(solhip.PROG_PATHTRACE_GLSL)
    spec.enable_sky = sky
    prog = ctx.program(solhip.PROG_PATHTRACE_GLSL, spec)
    img, acc, out, _ = _images(ctx, W, H, rgba8=True)
    sc = oracle.Scene(sd)
    ref = np.zeros((H, W, 4), np.float32)
    ref8 = np.zeros((H, W, 4), np.uint8)
    for frame in range(2):
        u = sd.uniforms(W, H, frame)
        ctx.stats(reset=True)
        ctx.trace(prog, gs.tlas, gs.tables, u, 0, img, W, H)
        st = ctx.stats()
        nrays = sc.render_pathtrace_glsl(u, oracle.PTParams(sky, spec.spp_per_frame, spec.max_depth), 0, ref, ref8,
                                         W, H)
        assert _diff(acc.download(np.float32), ref) == 0, "frame %d" % frame
        g8 = out.download(np.uint8).reshape(H, W, 4).astype(int)
        assert np.abs(g8 - ref8.astype(int)).max() <= 1
        assert st.rays == nrays and st.overflow == 0


def test_pathtrace_index_size_8_equals_4(ctx):
    sd = scenes.cornell_box()
    W, H = 24, 24
    u = sd.uniforms(W, H, 3)
    res = []
    for isz in (4, 8):
        gs = solhip.GpuScene(ctx, sd, index_size=isz)
        prog = ctx.program(solhip.PROG_PATHTRACE_GLSL)
        img, acc, _, _ = _images(ctx, W, H)
        ctx.trace(prog, gs.tlas, gs.tables, u, 3, img, W, H)
        res.append(acc.download(np.float32))
    assert _diff(res[0], res[1]) == 0


def test_debug_matches_oracle(ctx, oracle):
    sd = scenes.cornell_box()
    W, H = 50, 30
    gs = solhip.GpuScene(ctx, sd)
    prog = ctx.program(solhip.PROG_DEBUG)
    img, acc, _, _ = _images(ctx, W, H)
    u = sd.uniforms(W, H, 0)
    ctx.trace(prog, gs.tlas, None, u, 0, img, W, H)
    ref = oracle.Scene(sd).render_debug(u, W, H)
    assert _diff(acc.download(np.float32), ref) == 0


def test_profile_counts_and_times(ctx):
    sd = scenes.heightfield(128)
    W, H = 64, 36
    gs = solhip.GpuScene(ctx, sd)
    prog = ctx.program(solhip.PROG_AO)
    img, acc, _, _ = _images(ctx, W, H)
    for k in ("ao_primary", "ao_rays"):
        ctx.kernel_stats(k, reset=True)
    ctx.profile(solhip.api.PROFILE_TIME | solhip.api.PROFILE_COUNT)
    try:
        ctx.trace(prog, gs.tlas, gs.tables, sd.uniforms(W, H, 0), 0, img, W, H)
        p = ctx.kernel_stats("ao_primary", reset=True)
        r = ctx.kernel_stats("ao_rays", reset=True)
    finally:
        ctx.profile(0)
    assert p.launches == 1 and r.launches == 1 and p.total_ms > 0 and r.total_ms > 0
    assert p.counters.rays == W * H and p.counters.node_fetches >= W * H
    assert r.counters.rays > 0 and r.counters.node_fetches >= r.counters.rays


def test_invalid_handles_are_rejected(ctx):
    with pytest.raises(solhip.SolhipError) as e:
        ctx.trace_rays(12345, np.zeros(1, solhip.records.RAY))
    assert e.value.code == -2
