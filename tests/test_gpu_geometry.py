"""The BLAS boundary at the width of BLAS::new (src/ray/acceleration.rs:136-239):
several geometries per BLAS, non-indexed geometry, the geometry index of hits
(gl_GeometryIndexEXT), rejection of invalid geometry, and TLASes made stale by
a rebuilt or destroyed BLAS.  Parity bar: hit records bit-identical with the
oracle's brute force."""
import ctypes as C

import numpy as np
import pytest

import raygen
import solhip
from solhip import api, scenes

pytestmark = pytest.mark.gpu

BUILDS = pytest.mark.parametrize("build", [0, solhip.BUILD_PREFER_FAST_TRACE], ids=["morton", "sah"])


@BUILDS
def test_three_geometry_blas_matches_brute_force(ctx, oracle, build):
    """two instances of a BLAS over two indexed geometries and one
    non-indexed geometry, and a one-geometry BLAS: (t, u, v, instance,
    primitive within the geometry, geometry) bit for bit"""
    sd = scenes.multi_geometry_scene()
    print("multi-geometry trace, build", build, flush=True)
    gs = solhip.GpuScene(ctx, sd, build_flags=build)
    rays = np.concatenate([raygen.random_rays(4000, (0, 0, 0), 2.5, seed=31), raygen.aimed_rays(sd, 3000, seed=32),
                           raygen.axis_rays(300)])
    ctx.stats(reset=True)
    g = ctx.trace_rays(gs.tlas, rays)
    assert ctx.stats().overflow == 0
    r = oracle.Scene(sd).trace(rays, use_bvh=False)
    assert g.tobytes() == r.tobytes()
    hit = g["inst"] != solhip.records.MISS
    assert set(np.unique(g["geom"][hit & (g["inst"] < 2)]).tolist()) == {0, 1, 2}
    assert (g["geom"][~hit] == solhip.records.MISS).all()


def test_ao_hit_records_carry_the_geometry(ctx, oracle):
    """the AO program's hit records (images.hits_out) on the several-geometry
    scene equal the oracle's, geometry index included"""
    sd = scenes.multi_geometry_scene()
    W, H = 40, 32
    gs = solhip.GpuScene(ctx, sd)
    spec = solhip.spec_default(solhip.PROG_AO)
    prog = ctx.program(solhip.PROG_AO, spec)
    acc = ctx.upload(np.zeros((H, W, 4), np.float32))
    hits = ctx.buffer(W * H * (1 + spec.n_ao) * solhip.records.HIT.itemsize)
    u = sd.uniforms(W, H, 0)
    ctx.trace(prog, gs.tlas, gs.tables, u, 0, solhip.Images(acc.h, 0, 0, 0, 0, hits.h), W, H)
    ctx.sync()
    g = hits.download(solhip.records.HIT)
    ref = np.zeros((H, W, 4), np.float32)
    _, ph, ah = oracle.Scene(sd).render_ao(u, oracle.AOParams(spec.n_ao, spec.ao_tmin, spec.ao_tmax, 0), 0, ref, W,
                                           H, want_hits=True)
    assert g.tobytes() == np.concatenate([ph, ah]).tobytes()
    assert np.array_equal(acc.download(np.float32).reshape(H, W, 4), ref)
    ctx.program_destroy(prog)


def test_rebuild_with_other_geometries(ctx, oracle):
    """a rebuild may change the number of geometries; the TLAS built after it
    traces the new geometry"""
    sd = scenes.multi_geometry_scene()
    gs = solhip.GpuScene(ctx, sd)
    rays = raygen.random_rays(3000, (0, 0, 0), 2.5, seed=33)
    # BLAS 0 again from its last geometry only
    ctx.blas_rebuild_geometries(gs.blas[0], gs.blas_geoms[0][2:])
    gs.update_instances(sd.instance_descs)
    g = ctx.trace_rays(gs.tlas, rays)
    v2, _ = sd.blas[0].geometries[2]
    sd2 = scenes.assemble("one_of_three", [scenes.MultiBlas([(v2, None)]), sd.blas[1]], [0, 0, 1],
                          [np.vstack([d["transform"].reshape(3, 4), [0, 0, 0, 1]]) for d in sd.instance_descs],
                          [0, 1, 2], [sd.materials[k:k + 1] for k in range(3)], sd.camera, 3.0)
    r = oracle.Scene(sd2).trace(rays, use_bvh=False)
    assert g.tobytes() == r.tobytes()


def _geom(vaddr, stride, vcount, iaddr, icount, first=0, off=0):
    return api.Geometry(vaddr, stride, first, vcount, iaddr, off, icount, 1)


def _build_rc(ctx, geoms, rebuild_of=None):
    arr = (api.Geometry * len(geoms))(*geoms)
    if rebuild_of is not None:
        return api.lib().solhip_blas_rebuild(ctx.h, 0, rebuild_of, arr, len(geoms))
    h = C.c_uint64()
    rc = api.lib().solhip_blas_build(ctx.h, 0, arr, len(geoms), 0, C.byref(h))
    if rc == 0:
        ctx.blas_destroy(h.value)
    return rc


def test_invalid_geometry_is_rejected(ctx):
    """null addresses, bad strides, ragged index counts and indices at or
    beyond vertex_count fail with SOLHIP_ERROR_INVALID_ARGUMENT, for builds
    and rebuilds alike"""
    b = scenes.single_triangle().blas[0]
    vb = ctx.upload(b.vertices)
    ib = ctx.upload(b.indices)
    bad_idx = ctx.upload(np.array([0, 1, 3], np.uint32))   # vertex 3 of 3
    ok = _geom(vb.addr, 64, 3, ib.addr, 3)
    assert _build_rc(ctx, [ok]) == 0
    cases = {
        "null vertex address": _geom(0, 64, 3, ib.addr, 3),
        "stride < 12": _geom(vb.addr, 8, 3, ib.addr, 3),
        "stride not a multiple of 4": _geom(vb.addr, 66, 3, ib.addr, 3),
        "index_count % 3": _geom(vb.addr, 64, 3, ib.addr, 2),
        "indices without vertices": _geom(vb.addr, 64, 0, ib.addr, 3),
        "index_count without index buffer": _geom(vb.addr, 64, 3, 0, 3),
        "index >= vertex_count": _geom(vb.addr, 64, 3, bad_idx.addr, 3),
        "vertex_count below the indices": _geom(vb.addr, 64, 2, ib.addr, 3),
    }
    h = ctx.blas_build(vb, 64, 3, ib, 3)
    for name, g in cases.items():
        print("invalid geometry:", name, flush=True)
        assert _build_rc(ctx, [g]) == -1, name
        assert _build_rc(ctx, [ok, g]) == -1, name
        assert _build_rc(ctx, [g], rebuild_of=h) == -1, name
    assert _build_rc(ctx, []) == -1
    # the BLAS whose rebuild failed cannot be instanced until rebuilt
    descs = scenes.single_triangle().instance_descs.copy()
    descs["blas"] = h
    with pytest.raises(solhip.SolhipError) as e:
        ctx.tlas_build(descs)
    assert e.value.code == -8
    assert _build_rc(ctx, [ok], rebuild_of=h) == 0
    t = ctx.tlas_build(descs)
    rays = raygen.random_rays(64, (0, 0, 0), 1.0, seed=1)
    ctx.trace_rays(t, rays)
    ctx.tlas_destroy(t)
    ctx.blas_destroy(h)
    for buf in (vb, ib, bad_idx):
        buf.destroy()


def test_stale_tlas_is_refused(ctx, oracle):
    """a TLAS holds its BLASes' device pointers and boxes: after a BLAS rebuild
    or destroy, traces through it fail with SOLHIP_ERROR_STALE (no launch)
    until the TLAS is rebuilt, as Vulkan requires a TLAS build after its
    BLASes change"""
    sd = scenes.random_instances()
    gs = solhip.GpuScene(ctx, sd)
    rays = raygen.random_rays(500, (0, 0, 0), 6.0, seed=2)
    before = ctx.trace_rays(gs.tlas, rays)
    gs.rebuild_blas(1)
    with pytest.raises(solhip.SolhipError) as e:
        ctx.trace_rays(gs.tlas, rays)
    assert e.value.code == -8 and "rebuild the TLAS" in str(e.value)
    prog = ctx.program(solhip.PROG_AO)
    acc = ctx.upload(np.zeros((8, 8, 4), np.float32))
    with pytest.raises(solhip.SolhipError) as e:
        ctx.trace(prog, gs.tlas, gs.tables, sd.uniforms(8, 8, 0), 0, solhip.Images(acc.h, 0, 0, 0, 0, 0), 8, 8)
    assert e.value.code == -8
    gs.update_instances(sd.instance_descs)
    assert ctx.trace_rays(gs.tlas, rays).tobytes() == before.tobytes()
    # a server scene registered before a TLAS rebuild is stale too
    srv = solhip.Server(ctx, 64 << 20, 4)
    sid = srv.scene(gs)
    gs.update_instances(sd.instance_descs)
    spec = solhip.spec_default(solhip.PROG_PATHTRACE)
    spec.spp_per_frame = 1
    with pytest.raises(solhip.SolhipError) as e:
        srv.submit(sid, 8, 8, spec, sd.uniforms(8, 8, 0), acc)
    assert e.value.code == -8
    sid2 = srv.scene(gs)
    srv.submit(sid2, 8, 8, spec, sd.uniforms(8, 8, 0), acc)
    assert srv.run() == 0
    srv.close()
    # destroying a BLAS makes the TLAS stale as well
    ctx.blas_destroy(gs.blas[2])
    with pytest.raises(solhip.SolhipError) as e:
        ctx.trace_rays(gs.tlas, rays)
    assert e.value.code == -8 and "destroyed" in str(e.value)
    ctx.program_destroy(prog)
