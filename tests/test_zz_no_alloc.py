"""solhip_trace_noalloc (the AO program's scratch in the library's fixed
device array, no device allocation) against solhip_trace: the same AO images
and hit records bit for bit, on two calls in a row, the second on a smaller
frame (the fixed array keeps the first call's contents); frames beyond the
array's size and the wavefront tracer are refused before any launch."""
import numpy as np
import pytest

import solhip
from solhip import scenes
from solhip.records import HIT, MISS

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32).reshape(-1)


def _ao_frame(ctx, gs, prog, n_ao, u, W, H, noalloc):
    acc = ctx.upload(np.full((H, W, 4), 0.25, np.float32))
    hits = ctx.buffer(W * H * (1 + n_ao) * HIT.itemsize)
    img = solhip.Images(acc.h, 0, 0, 0, 0, hits.h)
    (ctx.trace_noalloc if noalloc else ctx.trace)(prog, gs.tlas, gs.tables, u, 0, img, W, H)
    ctx.sync()
    return acc.download(np.float32), hits.download(HIT)


@pytest.mark.parametrize("build", [0, solhip.BUILD_PREFER_FAST_TRACE], ids=["morton", "sah"])
def test_noalloc_equals_trace_on_two_calls(ctx, build):
    sd = scenes.heightfield(128)
    gs = solhip.GpuScene(ctx, sd, build_flags=build)
    spec = solhip.spec_default(solhip.PROG_AO)
    prog = ctx.program(solhip.PROG_AO, spec)
    for W, H, frame in ((96, 64, 3), (40, 24, 4)):   # the second call is smaller
        print("noalloc AO", W, H, "frame", frame, flush=True)
        u = sd.uniforms(W, H, frame)
        ctx.stats(reset=True)
        a_acc, a_hits = _ao_frame(ctx, gs, prog, spec.n_ao, u, W, H, noalloc=False)
        a_rays = ctx.stats(reset=True).rays
        b_acc, b_hits = _ao_frame(ctx, gs, prog, spec.n_ao, u, W, H, noalloc=True)
        st = ctx.stats()
        assert np.array_equal(_bits(a_acc), _bits(b_acc)), (W, H)
        assert np.array_equal(_bits(a_hits), _bits(b_hits)), (W, H)
        assert st.rays == a_rays and st.overflow == 0
        assert (b_hits["inst"][: W * H] != MISS).mean() > 0.5
    ctx.program_destroy(prog)


def test_noalloc_megakernel_program_equals_trace(ctx):
    sd = scenes.cornell_box()
    gs = solhip.GpuScene(ctx, sd)
    prog = ctx.program(solhip.PROG_PATHTRACE_GLSL)
    W, H = 32, 24
    out = []
    for noalloc in (False, True):
        acc = ctx.upload(np.zeros((H, W, 4), np.float32))
        img = solhip.Images(acc.h, 0, 0, 0, 0, 0)
        (ctx.trace_noalloc if noalloc else ctx.trace)(prog, gs.tlas, gs.tables, sd.uniforms(W, H, 1), 0, img, W, H)
        ctx.sync()
        out.append(acc.download(np.float32))
    assert np.array_equal(_bits(out[0]), _bits(out[1]))
    ctx.program_destroy(prog)


def test_noalloc_refuses_frames_beyond_the_fixed_array(ctx):
    sd = scenes.cornell_box()
    gs = solhip.GpuScene(ctx, sd)
    prog = ctx.program(solhip.PROG_AO)
    W, H = 4096, 1024   # W*H*(32 + 8) > 128 MiB
    assert W * H * 40 > solhip.api.NOALLOC_SCRATCH_BYTES
    acc = ctx.buffer(W * H * 16)
    img = solhip.Images(acc.h, 0, 0, 0, 0, 0)
    with pytest.raises(solhip.SolhipError) as e:
        ctx.trace_noalloc(prog, gs.tlas, gs.tables, sd.uniforms(W, H, 0), 0, img, W, H)
    assert e.value.code == -1 and "fixed scratch" in str(e.value)
    ctx.program_destroy(prog)


def test_noalloc_refuses_the_wavefront_tracer(ctx):
    sd = scenes.cornell_box()
    gs = solhip.GpuScene(ctx, sd)
    prog = ctx.program(solhip.PROG_PATHTRACE)
    W, H = 16, 16
    acc = ctx.upload(np.zeros((H, W, 4), np.float32))
    img = solhip.Images(acc.h, 0, 0, 0, 0, 0)
    with pytest.raises(solhip.SolhipError) as e:
        ctx.trace_noalloc(prog, gs.tlas, gs.tables, sd.uniforms(W, H, 0), 0, img, W, H)
    assert e.value.code == -6
    ctx.program_destroy(prog)
