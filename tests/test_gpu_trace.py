"""GPU parity of the BVH build + traversal (solhip_trace_rays) against the
oracle's brute-force closest hit.  Bar: hit records (t, u, v, instance,
primitive) bit-identical, including misses and the (t, inst, prim) tie rule."""
import numpy as np
import pytest

import raygen
import solhip
from solhip import scenes
from solhip.records import HIT

pytestmark = pytest.mark.gpu


def _check(ctx, oracle, sd, rays, brute=True, flags=0, build=0):
    gs = solhip.GpuScene(ctx, sd, build_flags=build)
    ctx.stats(reset=True)
    g = ctx.trace_rays(gs.tlas, rays, flags)
    st = ctx.stats()
    assert st.overflow == 0
    sc = oracle.Scene(sd)
    r = sc.trace(rays, use_bvh=not brute)
    if flags & solhip.TRACE_TERMINATE_ON_FIRST_HIT:
        assert np.array_equal(g["inst"] == 0xFFFFFFFF, r["inst"] == 0xFFFFFFFF)
        return g, r
    diff = np.nonzero((g.view(np.uint32).reshape(-1, 6) != r.view(np.uint32).reshape(-1, 6)).any(axis=1))[0]
    assert len(diff) == 0, "%d/%d hits differ, first %s: gpu %s ref %s" % (
        len(diff), len(rays), diff[:5], g[diff[:3]], r[diff[:3]])
    return g, r


SMALL_SCENES = {
    "triangle": scenes.single_triangle,
    "cornell": scenes.cornell_box,
    "soup": lambda: scenes.random_soup(3000, seed=5),
    "instances": scenes.random_instances,
    "heightfield64": lambda: scenes.heightfield(64),
}


BUILDS = pytest.mark.parametrize("build", [0, solhip.BUILD_PREFER_FAST_TRACE], ids=["morton", "sah"])


@BUILDS
@pytest.mark.parametrize("name", sorted(SMALL_SCENES))
def test_trace_random_rays_match_brute_force(ctx, oracle, name, build):
    sd = SMALL_SCENES[name]()
    rays = raygen.random_rays(4000, (0, 0, 0), 2.0 if name != "instances" else 6.0, seed=11)
    _check(ctx, oracle, sd, rays, build=build)


@BUILDS
@pytest.mark.parametrize("name", sorted(SMALL_SCENES))
def test_trace_aimed_rays_match_brute_force(ctx, oracle, name, build):
    sd = SMALL_SCENES[name]()
    rays = raygen.aimed_rays(sd, 3000, seed=12)
    g, r = _check(ctx, oracle, sd, rays, build=build)
    # rays aimed at vertices / shared edges of closed meshes must not leak
    if name in ("heightfield64",):
        assert (g["inst"] != 0xFFFFFFFF).mean() > 0.5


@pytest.mark.parametrize("name", ["cornell", "heightfield64", "instances"])
def test_trace_camera_rays_match_brute_force(ctx, oracle, name):
    sd = SMALL_SCENES[name]()
    rays = raygen.camera_rays(sd, 48, 32, oracle)
    _check(ctx, oracle, sd, rays)


def test_trace_axis_parallel_rays(ctx, oracle):
    sd = scenes.cornell_box()
    _check(ctx, oracle, sd, raygen.axis_rays(3000))


def test_trace_terminate_on_first_hit_agrees_on_occlusion(ctx, oracle):
    sd = scenes.random_soup(2000, seed=9)
    rays = raygen.random_rays(3000, (0, 0, 0), 1.5, seed=13, tmax=0.5)
    _check(ctx, oracle, sd, rays, flags=solhip.TRACE_TERMINATE_ON_FIRST_HIT)


def test_trace_tmin_tmax_window(ctx, oracle):
    sd = scenes.random_soup(1500, seed=4)
    rays = raygen.random_rays(3000, (0, 0, 0), 1.0, seed=14, tmin=0.3, tmax=0.9)
    _check(ctx, oracle, sd, rays)


def test_empty_tlas_misses(ctx):
    sd = scenes.empty_scene()
    gs = solhip.GpuScene(ctx, sd)
    rays = raygen.random_rays(100, (0, 0, 0), 1.0)
    g = ctx.trace_rays(gs.tlas, rays)
    assert (g["inst"] == 0xFFFFFFFF).all() and (g["t"] == -1.0).all()


def test_blas_build_rejects_unknown_flags(ctx):
    sd = scenes.single_triangle()
    b = sd.blas[0]
    vb = ctx.upload(b.vertices)
    ib = ctx.upload(b.indices.astype(np.uint32))
    with pytest.raises(solhip.SolhipError):
        ctx.blas_build(vb, 64, len(b.vertices), ib, len(b.indices), flags=0x80)


@BUILDS
def test_blas_rebuild_keeps_build_flags(ctx, oracle, build):
    """a rebuild in place uses the flags of the first build: same tree, same
    fetch counts"""
    sd = scenes.heightfield(64)
    gs = solhip.GpuScene(ctx, sd, build_flags=build)
    rays = raygen.camera_rays(sd, 48, 32, oracle)
    counts = []
    for rebuild in (False, True):
        if rebuild:
            b = sd.blas[0]
            ctx.blas_rebuild(gs.blas[0], gs.bufs[0], 64, len(b.vertices), gs.bufs[1], len(b.indices))
            gs.update_instances(sd.instance_descs)
        ctx.stats(reset=True)
        ctx.trace_rays(gs.tlas, rays, solhip.TRACE_COUNT)
        st = ctx.stats(reset=True)
        counts.append((st.node_fetches, st.tri_fetches))
    assert counts[0] == counts[1]


def test_bvh_info_and_counts(ctx, oracle):
    sd = scenes.heightfield(128)
    gs = solhip.GpuScene(ctx, sd)
    info = ctx.blas_info(gs.blas[0])
    assert info.prims == sd.tri_count
    assert 1 <= info.nodes < info.prims
    assert info.depth >= 2
    rays = raygen.camera_rays(sd, 64, 36, oracle)
    ctx.stats(reset=True)
    ctx.trace_rays(gs.tlas, rays, solhip.TRACE_COUNT)
    st = ctx.stats(reset=True)
    assert st.rays == len(rays) and st.overflow == 0
    assert st.node_fetches >= len(rays) and st.inst_fetches >= 1


@BUILDS
def test_heightfield_large_sampled(ctx, oracle, build):
    """512^2 heightfield (524k triangles) against the oracle's own BVH"""
    sd = scenes.heightfield(512)
    rays = np.concatenate([raygen.camera_rays(sd, 64, 36, oracle), raygen.aimed_rays(sd, 4000, seed=21, dist=0.5)])
    _check(ctx, oracle, sd, rays, brute=False, build=build)


@BUILDS
@pytest.mark.parametrize("name", ["soup", "instances", "heightfield128", "heightfield256"])
def test_gpu_tree_equals_host_replay(ctx, oracle, name, build):
    """the GPU build makes the tree the host replay of the same steps makes
    (tests/cpu_harness): same hits and the same node, triangle and instance
    fetch counts over a ray set (heightfield256: 131,072 triangles, so the
    SAH build's top nodes run on its 1024-thread workgroups)"""
    from cpu_harness import harness
    if name.startswith("heightfield"):
        sd = scenes.heightfield(int(name[len("heightfield"):]))
    else:
        sd = SMALL_SCENES[name]()
    rays = np.concatenate([raygen.camera_rays(sd, 64, 36, oracle), raygen.random_rays(3000, (0, 0, 0), 3.0, seed=4)])
    gs = solhip.GpuScene(ctx, sd, build_flags=build)
    ctx.stats(reset=True)
    g = ctx.trace_rays(gs.tlas, rays, solhip.TRACE_COUNT)
    st = ctx.stats(reset=True)
    h, c = harness.HostScene(sd, build_flags=build).trace(rays)
    assert np.array_equal(g.view(np.uint32), h.view(np.uint32))
    assert (st.node_fetches, st.tri_fetches, st.inst_fetches) == (int(c[0]), int(c[1]), int(c[2]))
