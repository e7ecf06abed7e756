"""Poisson render-request load on the render server (solhip_server_*).

A request is (scene, resolution, spp, depth) drawn with a seeded generator:
scene uniform over the resident scenes, resolution uniform over
RESOLUTIONS, spp log-uniform in [1, 256], depth uniform in {4..8}, next-event
estimation on, Russian roulette from bounce 3, gradient sky.  Arrivals are a
Poisson process of the given rate on the host clock; between arrivals the
server runs wavefront steps.  Reported: latency percentiles (submit -> last
pixel done), jobs/s, Mrays/s and the makespan.
"""
import math
import time

import numpy as np

from . import api

RESOLUTIONS = [(256, 256), (512, 512), (1280, 720), (1920, 1080)]


def draw_requests(n, rate, seed=3, resolutions=RESOLUTIONS, spp_max=256):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    reqs = []
    for k in range(n):
        W, H = resolutions[rng.integers(len(resolutions))]
        spp = int(round(math.exp(rng.uniform(0.0, math.log(spp_max)))))
        reqs.append({"t": float(t[k]), "scene": int(rng.integers(1 << 30)), "W": W, "H": H,
                     "spp": max(1, min(spp_max, spp)), "depth": int(rng.integers(4, 9))})
    return reqs


def run_load(ctx, gpu_scenes, scene_data, reqs, hbm_budget=8 << 30, max_jobs=64, steps_per_poll=2, streams=0):
    """serves reqs (draw_requests) on a fresh server; returns a stats dict.
    streams = 1: per-sample random streams, so a job may run on several
    slots per pixel (the server picks how many)"""
    srv = api.Server(ctx, hbm_budget, max_jobs)
    sids = [srv.scene(g) for g in gpu_scenes]
    bufs = {}
    jobs = []
    ctx.stats(reset=True)
    t0 = time.perf_counter()
    k = 0
    live = 0
    while k < len(reqs) or live:
        now = time.perf_counter() - t0
        while k < len(reqs) and reqs[k]["t"] <= now:
            r = reqs[k]
            si = r["scene"] % len(sids)
            W, H = r["W"], r["H"]
            bufs[k] = ctx.buffer(W * H * 16)   # each job its own image
            spec = api.spec_default(api.PROG_PATHTRACE)
            spec.enable_sky, spec.spp_per_frame, spec.max_depth, spec.rr_start, spec.nee = 2, r["spp"], r["depth"], 3, 1
            spec.sample_streams = streams
            jid = srv.submit(sids[si], W, H, spec, scene_data[si].uniforms(W, H, k), bufs[k], accum_start=k)
            jobs.append((jid, r))
            k += 1
        live = srv.run(steps_per_poll)
        if not live and k < len(reqs):
            time.sleep(max(0.0, reqs[k]["t"] - (time.perf_counter() - t0)))
    makespan = time.perf_counter() - t0
    rays = ctx.stats().rays
    lat = np.array([srv.job(j).done_ms - srv.job(j).submit_ms for j, _ in jobs])
    wait = np.array([srv.job(j).admit_ms - srv.job(j).submit_ms for j, _ in jobs])
    srv.close()
    for b in bufs.values():
        b.destroy()
    return {"jobs": len(jobs), "makespan_s": makespan, "jobs_per_s": len(jobs) / makespan,
            "mrays_per_s": rays / makespan / 1e6, "rays": int(rays),
            "latency_ms_p50": float(np.percentile(lat, 50)), "latency_ms_p99": float(np.percentile(lat, 99)),
            "queue_ms_p50": float(np.percentile(wait, 50)), "queue_ms_p99": float(np.percentile(wait, 99)),
            "pixels": int(sum(r["W"] * r["H"] for _, r in jobs)),
            "samples": int(sum(r["W"] * r["H"] * r["spp"] for _, r in jobs))}
