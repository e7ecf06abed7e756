"""The render-request server (solhip_server_*) on the device.

Jobs of mixed scene, image size, sample count and estimator settings share
one wavefront.  Each job's image must be bit-identical to the oracle's
rendering of the same estimator, spec and uniforms (sr_render_pathtrace), and
to a solhip_trace of SOLHIP_PROG_PATHTRACE, whatever else runs beside it and
whatever the admission order."""
import numpy as np
import pytest

import solhip
from solhip import scenes

pytestmark = pytest.mark.gpu


def _spec(sky=0, spp=3, depth=4, rr=0, nee=0, streams=0, lanes=0):
    s = solhip.spec_default(solhip.PROG_PATHTRACE)
    s.enable_sky, s.spp_per_frame, s.max_depth, s.rr_start, s.nee = sky, spp, depth, rr, nee
    s.sample_streams, s.wf_samples_per_pass = streams, lanes
    return s


def _eager(ctx, gs, sd, spec, W, H, frame):
    prog = ctx.program(solhip.PROG_PATHTRACE, spec)
    acc = ctx.upload(np.zeros((H, W, 4), np.float32))
    out = ctx.upload(np.zeros((H, W, 4), np.uint8))
    ctx.trace(prog, gs.tlas, gs.tables, sd.uniforms(W, H, frame), 0, solhip.Images(acc.h, out.h, 0, 0, 0, 0), W, H)
    a, o = acc.download(np.float32).reshape(H, W, 4), out.download(np.uint8).reshape(H, W, 4)
    ctx.program_destroy(prog)
    return a, o


JOBS = [  # (scene, W, H, spec kwargs, frame)
    ("cornell", 40, 32, dict(spp=3, depth=4, nee=1, rr=2), 0),
    ("instances", 48, 24, dict(sky=1, spp=2, depth=6), 1),
    ("cornell", 17, 13, dict(spp=5, depth=3, nee=1), 2),
    ("instances", 33, 20, dict(sky=2, spp=4, depth=8, nee=1, rr=3), 3),
    ("cornell", 64, 48, dict(spp=1, depth=5), 4),
    ("instances", 30, 22, dict(sky=2, spp=4, depth=6, nee=1, rr=2, streams=1), 5),
    # per-sample streams on several slots per pixel: 3 lanes for 5 samples
    # (lanes with 2, 2 and 1 samples), and the default lanes for 40 samples (3)
    ("cornell", 29, 21, dict(spp=5, depth=5, nee=1, rr=2, streams=1, lanes=3), 6),
    ("instances", 24, 20, dict(sky=2, spp=40, depth=4, nee=1, rr=3, streams=1), 7),
]
SLOT_BYTES = 220
LANES = [3 if kw.get("lanes") == 3 else (3 if kw.get("spp", 0) == 40 else 1) for _, _, _, kw, _ in JOBS]


def _run_jobs(ctx, budget, max_jobs=64):
    sds = {"cornell": scenes.cornell_box(), "instances": scenes.random_instances()}
    gss = {k: solhip.GpuScene(ctx, v) for k, v in sds.items()}
    srv = solhip.Server(ctx, budget, max_jobs)
    sid = {k: srv.scene(v) for k, v in gss.items()}
    jobs = []
    for name, W, H, kw, frame in JOBS:
        acc = ctx.upload(np.zeros((H, W, 4), np.float32))
        out = ctx.upload(np.zeros((H, W, 4), np.uint8))
        jid = srv.submit(sid[name], W, H, _spec(**kw), sds[name].uniforms(W, H, frame), acc, out)
        jobs.append((jid, name, W, H, kw, frame, acc, out))
    live = srv.run()
    assert live == 0
    return srv, sds, gss, jobs


def _oracle_render(sd, kw, W, H, frame):
    """the job's image by the oracle: sr_render_pathtrace with the job's
    estimator settings into a zeroed accumulator, accum_start 0"""
    import solref
    s = _spec(**kw)
    cfg = solref.PTConfig(s.enable_sky, s.spp_per_frame, s.max_depth, s.rr_start, s.nee, s.sample_streams)
    a = np.zeros((H, W, 4), np.float32)
    o = np.zeros((H, W, 4), np.uint8)
    solref.Scene(sd).render_pathtrace(sd.uniforms(W, H, frame), cfg, 0, a, o, W, H)
    return a, o


def _check(ctx, sds, gss, jobs):
    for jid, name, W, H, kw, frame, acc, out in jobs:
        a = acc.download(np.float32).reshape(H, W, 4)
        o = out.download(np.uint8).reshape(H, W, 4)
        a_or, o_or = _oracle_render(sds[name], kw, W, H, frame)
        assert (a.view(np.uint32) != a_or.view(np.uint32)).sum() == 0, ("oracle", jid, name, kw)
        assert np.array_equal(o, o_or), ("oracle", jid, name, kw)
        a_ref, o_ref = _eager(ctx, gss[name], sds[name], _spec(**kw), W, H, frame)
        assert (a.view(np.uint32) != a_ref.view(np.uint32)).sum() == 0, (jid, name, kw)
        assert np.array_equal(o, o_ref), (jid, name, kw)


def test_mixed_jobs_equal_single_renders(ctx):
    print("server: mixed jobs, large budget", flush=True)
    srv, sds, gss, jobs = _run_jobs(ctx, 64 << 20)
    for jid, *_ in jobs:
        st = srv.job(jid)
        assert st.state == solhip.JOB_DONE and st.submit_ms <= st.admit_ms <= st.done_ms
    _check(ctx, sds, gss, jobs)
    srv.close()


def test_budget_for_one_job_at_a_time(ctx):
    """the budget holds the largest job only: jobs run one after another in
    submission order and still give the same images"""
    print("server: tight budget", flush=True)
    largest = max(W * H * ln for (_, W, H, _, _), ln in zip(JOBS, LANES))
    srv, sds, gss, jobs = _run_jobs(ctx, largest * SLOT_BYTES + 64 * SLOT_BYTES)
    sts = [srv.job(j[0]) for j in jobs]
    assert all(s.state == solhip.JOB_DONE for s in sts)
    for a, b in zip(sts, sts[1:]):
        assert b.admit_ms >= a.admit_ms
    _check(ctx, sds, gss, jobs)
    srv.close()


def test_max_jobs_one(ctx):
    print("server: one job index", flush=True)
    srv, sds, gss, jobs = _run_jobs(ctx, 64 << 20, max_jobs=1)
    _check(ctx, sds, gss, jobs)
    srv.close()


def test_bad_submissions(ctx):
    sd = scenes.cornell_box()
    gs = solhip.GpuScene(ctx, sd)
    srv = solhip.Server(ctx, 1 << 20, 4)
    sid = srv.scene(gs)
    acc = ctx.upload(np.zeros((8, 8, 4), np.float32))
    with pytest.raises(solhip.SolhipError):
        srv.submit(sid + 7, 8, 8, _spec(), sd.uniforms(8, 8, 0), acc)
    with pytest.raises(solhip.SolhipError):
        srv.submit(sid, 16, 16, _spec(), sd.uniforms(16, 16, 0), acc)       # accum too small
    with pytest.raises(solhip.SolhipError):
        srv.submit(sid, 4096, 4096, _spec(), sd.uniforms(8, 8, 0), acc)     # beyond the slot budget
    with pytest.raises(solhip.SolhipError):
        srv.submit(sid, 8, 8, _spec(spp=0), sd.uniforms(8, 8, 0), acc)
    with pytest.raises(solhip.SolhipError):
        srv.submit(sid, 8, 8, _spec(streams=2), sd.uniforms(8, 8, 0), acc)
    with pytest.raises(solhip.SolhipError):
        solhip.Server(ctx, 1 << 20, 0)
    assert srv.run() == 0
    srv.close()
