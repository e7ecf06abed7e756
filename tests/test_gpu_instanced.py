"""The batched instanced workload (the 10,000-instance scene's generator at a
size the oracle affords): every BLAS and the TLAS rebuilt each frame, the
instances moved between frames through TLAS::regenerate's path
(solhip_tlas_rebuild plus the SceneInstance table, src/ray/mod.rs:162-193),
the wavefront tracer with the workload's settings (depth 8, roulette from
bounce 3, next-event estimation, gradient sky, graph replay, a random stream
per sample).  Bar: the accumulated image and RGBA8 output bit-identical with
the oracle rendering each frame from the moved transforms."""
import numpy as np
import pytest

import raygen
import solhip
from solhip import scenes

pytestmark = pytest.mark.gpu


def _moved(sd, f, rng):
    """frame f's transforms: every instance translated and turned a little"""
    out = []
    for m in scenes.instance_xforms(sd):
        a = 0.05 * f * rng.uniform(-1, 1)
        c, s = np.cos(a), np.sin(a)
        r = np.array([[c, 0, s, 0], [0, 1, 0, 0], [-s, 0, c, 0], [0, 0, 0, 1]])
        m2 = m @ r
        m2[:3, 3] += 0.2 * f * rng.uniform(-1, 1, size=3)
        out.append(m2)
    return out


@pytest.mark.parametrize("graph,per_pass", [(1, 0), (0, 3)])
def test_instanced_workload_with_moving_instances_matches_oracle(ctx, oracle, graph, per_pass):
    sd0 = scenes.instanced(200, seed=2, tube=(50, 100), spread=8.0)
    assert len(sd0.blas) == 5 and [b.tri_count for b in sd0.blas] == [20, 1280, 20480, 12, 10000]
    gs = solhip.GpuScene(ctx, sd0)
    W, H, F = 40, 32, 3
    spec = solhip.spec_default(solhip.PROG_PATHTRACE)
    spec.enable_sky, spec.spp_per_frame, spec.max_depth, spec.nee, spec.rr_start = 2, 4, 8, 1, 3
    spec.wf_graph = graph
    spec.sample_streams = 1
    spec.wf_samples_per_pass = per_pass
    prog = ctx.program(solhip.PROG_PATHTRACE, spec)
    acc = ctx.upload(np.zeros((H, W, 4), np.float32))
    out = ctx.upload(np.zeros((H, W, 4), np.uint8))
    img = solhip.Images(acc.h, out.h, 0, 0, 0, 0)
    ref = np.zeros((H, W, 4), np.float32)
    ref8 = np.zeros((H, W, 4), np.uint8)
    cfg = oracle.PTConfig(2, 4, 8, 3, 1, 1)
    rng = np.random.default_rng(5)
    for f in range(F):
        print("instanced frame", f, "graph", graph, flush=True)
        sd = scenes.with_transforms(sd0, _moved(sd0, f, rng)) if f else sd0
        for k in range(len(sd.blas)):
            gs.rebuild_blas(k)
        gs.move_instances(sd)
        u = sd.uniforms(W, H, f)
        ctx.stats(reset=True)
        ctx.trace(prog, gs.tlas, gs.tables, u, 0, img, W, H)
        ctx.sync()
        assert ctx.stats().overflow == 0
        n = oracle.Scene(sd).render_pathtrace(u, cfg, 0, ref, ref8, W, H)
        got = acc.download(np.float32).reshape(H, W, 4)
        bad = np.nonzero((got.view(np.uint32) != ref.view(np.uint32)).any(-1))
        assert len(bad[0]) == 0, (f, len(bad[0]), got[bad][:3], ref[bad][:3])
        assert np.array_equal(out.download(np.uint8).reshape(H, W, 4), ref8)
        assert n > W * H * 4
    prim = oracle.Scene(sd0).trace(raygen.camera_rays(sd0, W, H, oracle))
    assert (prim["inst"] != oracle.MISS).mean() > 0.2   # the instances cover a good part of the image
    ctx.program_destroy(prog)
