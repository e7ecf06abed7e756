"""Host replay of the library's own build and traversal steps (bvh_logic.h,
traverse.h compiled for the CPU by tests/cpu_harness) against the oracle's
brute-force closest hit: the per-item code the GPU kernels run, checked
without a GPU.  The device-wide sort, refit and allocation are replaced by
sequential equivalents in the harness."""
import numpy as np
import pytest

import raygen
from cpu_harness import build as hbuild
from cpu_harness import harness
from solhip import scenes

SCENES = {
    "triangle": scenes.single_triangle,
    "cornell": scenes.cornell_box,
    "soup": lambda: scenes.random_soup(2000, seed=5),
    "instances": scenes.random_instances,
    "heightfield64": lambda: scenes.heightfield(64),
    "empty": scenes.empty_scene,
}


@pytest.fixture(scope="module", autouse=True)
def _built():
    import os
    if not os.path.exists(harness.LIB):
        hbuild.build()


@pytest.mark.parametrize("flags", [0, 1], ids=["morton", "sah"])
@pytest.mark.parametrize("name", sorted(SCENES))
def test_replay_matches_brute_force(oracle, name, flags):
    """both BLAS builds (Morton order, binned SAH) give the brute-force hits"""
    sd = SCENES[name]()
    hs = harness.HostScene(sd, build_flags=flags)
    sc = oracle.Scene(sd)
    rays = np.concatenate([raygen.random_rays(2000, (0, 0, 0), 2.0 if name != "instances" else 6.0, seed=11),
                           raygen.aimed_rays(sd, 1500, seed=12), raygen.axis_rays(500),
                           raygen.random_rays(500, (0, 0, 0), 1.0, seed=3, tmin=0.3, tmax=0.9)])
    g, counts = hs.trace(rays)
    r = sc.trace(rays, use_bvh=False)
    assert counts[3] == 0   # no stack overflow
    diff = np.nonzero((g.view(np.uint32).reshape(-1, 6) != r.view(np.uint32).reshape(-1, 6)).any(axis=1))[0]
    assert len(diff) == 0, (g[diff[:3]], r[diff[:3]])
    ga, _ = hs.trace(rays, any_hit=True)
    assert np.array_equal(ga["inst"] == 0xFFFFFFFF, r["inst"] == 0xFFFFFFFF)


def test_replay_tree_shape_heightfield():
    sd = scenes.heightfield(128)
    hs = harness.HostScene(sd)
    nodes, depth = hs.info(0)
    assert nodes < sd.tri_count // 2 and 3 <= depth <= 16


def test_replay_sah_build_needs_fewer_node_tests(oracle):
    """on a terrain the binned SAH build needs fewer node and triangle tests
    per camera ray than the Morton-order build, for the same hits"""
    sd = scenes.heightfield(128)
    rays = raygen.camera_rays(sd, 64, 36, oracle)
    (gm, cm), (gs, cs) = (harness.HostScene(sd, build_flags=f).trace(rays) for f in (0, 1))
    assert np.array_equal(gm.view(np.uint32), gs.view(np.uint32))
    assert cs[0] < 0.95 * cm[0] and cs[1] <= cm[1]


@pytest.mark.parametrize("name", sorted(SCENES))
def test_resumable_traversal_equals_trace_ray(name):
    """the resumable traversal (traverse_step.h) driven as 64 lanes that take
    a new ray as soon as theirs finishes gives trace_ray's closest and first
    hits bit for bit, with the same fetch counts and no stack overflow"""
    sd = SCENES[name]()
    hs = harness.HostScene(sd)
    rays = np.concatenate([raygen.random_rays(3000, (0, 0, 0), 2.0 if name != "instances" else 6.0, seed=21),
                           raygen.aimed_rays(sd, 1500, seed=22), raygen.axis_rays(500)])
    for any_hit in (False, True):
        a, ca = hs.trace(rays, any_hit=any_hit)
        b, cb = hs.trace(rays, any_hit=any_hit, steps=True)
        assert np.array_equal(a.view(np.uint32).reshape(-1), b.view(np.uint32).reshape(-1)), (name, any_hit)
        assert np.array_equal(ca[:3], cb[:3]) and cb[3] == 0


def _verts(p):
    return scenes.make_vertices(np.asarray(p, np.float32).reshape(-1, 3))


def test_build_vertex_fetch_over_several_geometries():
    """BLAS::new's geometries (src/ray/acceleration.rs:148-208) as the build
    kernels read them: an indexed geometry with a first_vertex offset, an
    indexed one, and a non-indexed one with a primitive offset (vertices
    3i..3i+2 from vertex_addr + primitive_offset + first_vertex * stride);
    triangles numbered geometry after geometry"""
    rng = np.random.default_rng(4)
    P0 = rng.normal(size=(12, 3))
    P1 = rng.normal(size=(7, 3))
    P2 = rng.normal(size=(11, 3))     # 3 triangles + 2 unused vertices
    i0 = rng.integers(0, 9, size=15).astype(np.uint32)   # 5 triangles over vertices 3..11
    i1 = rng.integers(0, 7, size=12).astype(np.uint32)
    v2 = _verts(P2)
    geoms = [dict(vertices=_verts(P0), indices=i0, first_vertex=3, vertex_count=9),
             dict(vertices=_verts(P1), indices=i1),
             dict(vertices=v2, indices=None, primitive_offset=64, first_vertex=1, vertex_count=9)]
    pos, geom, lprim, bad = harness.geometry_fetch(geoms)
    assert bad == 0 and len(pos) == 5 + 4 + 3
    assert geom.tolist() == [0] * 5 + [1] * 4 + [2] * 3
    assert lprim.tolist() == list(range(5)) + list(range(4)) + list(range(3))
    exp = np.concatenate([P0[3 + i0].reshape(-1, 3, 3), P1[i1].reshape(-1, 3, 3),
                          P2[2:11].reshape(-1, 3, 3)]).astype(np.float32)
    assert np.array_equal(pos, exp)


def test_build_counts_out_of_range_indices():
    """an index >= vertex_count is counted (the build then fails with
    SOLHIP_ERROR_INVALID_ARGUMENT) and read as the last vertex, never beyond"""
    P = np.arange(18, dtype=np.float32).reshape(6, 3)
    idx = np.array([0, 1, 2, 3, 6, 5, 9, 4, 100], np.uint32)
    pos, _, _, bad = harness.geometry_fetch([dict(vertices=_verts(P), indices=idx)])
    assert bad == 3
    assert np.array_equal(pos[1, 1], P[5]) and np.array_equal(pos[2, 0], P[5]) and np.array_equal(pos[2, 2], P[5])
    # a vertex_count below the indices used, and a geometry without vertices
    _, _, _, bad = harness.geometry_fetch([dict(vertices=_verts(P), indices=np.array([0, 1, 5], np.uint32),
                                                vertex_count=5)])
    assert bad == 1
    _, _, _, bad = harness.geometry_fetch([dict(vertices=_verts(P), indices=np.array([0, 0, 0], np.uint32),
                                                vertex_count=0)])
    assert bad == 3
