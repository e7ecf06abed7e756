"""Pins the CPU oracle (oracle/solref.c) against the golden vectors written by an
independent restatement, closed forms and brute force.  CPU only."""
import math
import os

import numpy as np
import pytest

import raygen
from solhip import scenes
from solhip.records import RAY

HERE = os.path.dirname(os.path.abspath(__file__))


# ---- RNG: sampling.glsl:18-47 -------------------------------------------------
def test_tea_and_pcg_match_golden_vectors(oracle):
    g = np.load(os.path.join(HERE, "golden", "rng_golden.npz"))
    for p, f, s, fl in zip(g["pixel"], g["frames"], g["seeds"], g["floats"]):
        assert oracle.tea(int(p), int(f)) == int(s)
        out, _ = oracle.next_rand_stream(int(s), fl.shape[0])
        assert out.view(np.uint32).tolist() == fl.view(np.uint32).tolist()
    for st, fl in zip(g["edge_states"], g["edge_floats"]):
        out, _ = oracle.next_rand_stream(int(st), fl.shape[0])
        assert out.view(np.uint32).tolist() == fl.view(np.uint32).tolist()


def _tea_cipher(v0, v1, key, cycles):
    """the Tiny Encryption Algorithm (Wheeler and Needham 1994) on Python
    integers modulo 2^32, with its key and cycle count as parameters"""
    M = 0xFFFFFFFF
    s = 0
    for _ in range(cycles):
        s = (s + 0x9E3779B9) & M
        v0 = (v0 + ((((v1 << 4) & M) + key[0]) & M ^ (v1 + s) & M ^ ((v1 >> 5) + key[1]) & M)) & M
        v1 = (v1 + ((((v0 << 4) & M) + key[2]) & M ^ (v0 + s) & M ^ ((v0 >> 5) + key[3]) & M)) & M
    return v0, v1


def test_tea_cipher_matches_the_published_test_vector():
    # the standard 32-cycle TEA with an all-zero key encrypts the zero block
    # to 41ea3a0a 94baa940: this pins the parametrised cipher above
    assert _tea_cipher(0, 0, (0, 0, 0, 0), 32) == (0x41EA3A0A, 0x94BAA940)


def test_tea_is_the_16_cycle_cipher_with_the_shader_key(oracle):
    """sampling.glsl:18-32 is TEA with 16 cycles and the key (0xa341316c,
    0xc8013ea4, 0xad90777d, 0x7e95761e), returning the first word: known
    answers from the cipher pinned by the published vector"""
    key = (0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E)
    rng = np.random.default_rng(11)
    cases = [(0, 0), (1, 0), (0, 1), (0xFFFFFFFF, 0xFFFFFFFF)] + [tuple(int(x) for x in rng.integers(0, 2 ** 32, 2))
                                                                   for _ in range(64)]
    for a, b in cases:
        assert oracle.tea(a, b) == _tea_cipher(a, b, key, 16)[0]


def test_next_rand_is_word_times_2_pow_minus_32(oracle):
    # the output can reach exactly 1.0 (word rounds to 2^32) and is never > 1
    out, _ = oracle.next_rand_stream(12345, 20000)
    assert out.min() >= 0.0 and out.max() <= 1.0
    assert abs(out.mean() - 0.5) < 0.01


def test_pcg_advance_equals_repeated_steps(oracle):
    for seed in (0, 1, 0xDEADBEEF, 2 ** 31):
        _, st = oracle.next_rand_stream(seed, 37)
        assert oracle.pcg_advance(seed, 37) == st
        assert oracle.pcg_advance(seed, 0) == seed


# ---- own sin/cos against libm (double) ------------------------------------------
def test_sincos_within_2_ulp_on_sampling_range(oracle):
    xs = np.linspace(0.0, 2 * math.pi, 20001, dtype=np.float32)
    for x in xs[::7]:
        s = oracle.sinf(float(x))
        c = oracle.cosf(float(x))
        rs, rc = math.sin(float(x)), math.cos(float(x))
        ulp = np.spacing(np.float32(1.0))
        assert abs(s - rs) <= 2 * ulp * max(1.0, abs(rs)) + 1e-7
        assert abs(c - rc) <= 2 * ulp * max(1.0, abs(rc)) + 1e-7


# ---- sampling closed forms --------------------------------------------------------
def test_onb_is_orthonormal(oracle):
    rng = np.random.default_rng(0)
    for _ in range(200):
        n = rng.normal(size=3)
        n = (n / np.linalg.norm(n)).astype(np.float32)
        u, v = oracle.onb(n)
        M = np.stack([u, v, n]).astype(np.float64)
        assert np.allclose(M @ M.T, np.eye(3), atol=2e-6)


def test_cosine_lobe_moments(oracle):
    # unit length; E[cos theta] = 2/3 and E[cos^2] = 1/2 for the cosine lobe
    rng = np.random.default_rng(1)
    n = np.array([0.3, -0.5, 0.8], np.float32)
    n /= np.linalg.norm(n)
    cs = []
    for xi in rng.random((20000, 2), dtype=np.float32):
        d = oracle.sample_cosine(n, xi)
        assert abs(np.linalg.norm(d) - 1.0) < 2e-6
        cs.append(float(d @ n))
    cs = np.array(cs)
    assert cs.min() >= -1e-6
    assert abs(cs.mean() - 2 / 3) < 0.01 and abs((cs ** 2).mean() - 0.5) < 0.01


def test_ggx_with_zero_roughness_returns_the_axis(oracle):
    rng = np.random.default_rng(2)
    n = np.array([0.0, 0.6, 0.8], np.float32)
    for xi in rng.random((100, 2), dtype=np.float32) * np.float32(0.999):
        assert np.array_equal(oracle.sample_ggx(n, xi, 0.0), n)


def test_ggx_alpha_one_is_cosine_squared_uniform(oracle):
    # alpha^2 = 1: cos^2(theta) = 1 - xi, uniform in [0, 1]
    rng = np.random.default_rng(3)
    n = np.array([0.0, 0.0, 1.0], np.float32)
    c2 = [float(oracle.sample_ggx(n, xi, 1.0)[2]) ** 2 for xi in rng.random((10000, 2), dtype=np.float32)]
    assert abs(np.mean(c2) - 0.5) < 0.01


def test_fresnel_normal_and_grazing(oracle):
    n = np.array([0.0, 0.0, 1.0], np.float32)
    assert abs(oracle.fresnel(n, n, 1.0 / 1.5) - 0.04) < 1e-6
    g = np.array([1.0, 0.0, 0.0], np.float32)
    assert abs(oracle.fresnel(n, g, 1.0 / 1.5) - 1.0) < 1e-6


# ---- camera --------------------------------------------------------------------------
def test_centre_ray_is_the_view_axis(oracle):
    cam = scenes.Camera((1.0, 2.0, 3.0), (0.0, 0.5, -1.0), up=(0.0, -1.0, 0.0), vfov_deg=40.0)
    u = cam.uniforms(64, 64, 0)
    o, d = oracle.camera_ray(u, 32, 32, 64, 64, 0.0, 0.0)
    axis = np.array([0.0, 0.5, -1.0]) - np.array([1.0, 2.0, 3.0])
    axis /= np.linalg.norm(axis)
    assert np.allclose(o, [1, 2, 3], atol=1e-6)
    assert np.allclose(d, axis, atol=1e-6)


def test_image_corners_span_the_vertical_fov(oracle):
    cam = scenes.Camera((0.0, 0.0, 5.0), (0.0, 0.0, 0.0), up=(0.0, -1.0, 0.0), vfov_deg=40.0)
    u = cam.uniforms(100, 100, 0)
    _, top = oracle.camera_ray(u, 50, 0, 100, 100, 0.0, 0.0)
    _, bot = oracle.camera_ray(u, 50, 99, 100, 100, 0.0, 1.0)
    ang = math.degrees(math.acos(float(np.clip(top @ bot, -1, 1))))
    assert abs(ang - 40.0) < 0.05
    assert top[1] > 0 > bot[1]   # row 0 is the top of the image (up = -Y, Vulkan clip space)


# ---- watertight triangle test ----------------------------------------------------------
def test_axis_aligned_quad_exact_t(oracle):
    v0, v1, v2 = [0, 0, 0], [1, 0, 0], [0, 1, 0]
    hit, t, u, v = oracle.intersect_triangle([0.25, 0.25, 2.0], [0, 0, -1], 0.0, 10.0, v0, v1, v2)
    assert hit and t == 2.0 and u == 0.25 and v == 0.25
    hit, *_ = oracle.intersect_triangle([0.25, 0.25, 2.0], [0, 0, -1], 0.0, 1.5, v0, v1, v2)
    assert not hit   # beyond tmax
    hit, t, *_ = oracle.intersect_triangle([0.25, 0.25, -2.0], [0, 0, 1], 0.0, 10.0, v0, v1, v2)
    assert hit and t == 2.0   # back face: no culling


def test_shared_edge_is_watertight(oracle):
    # two triangles sharing the edge (1,0)-(0,1); rays through points of the
    # edge hit at least one of them
    a = ([0, 0, 0], [1, 0, 0], [0, 1, 0])
    b = ([1, 0, 0], [1, 1, 0], [0, 1, 0])
    rng = np.random.default_rng(4)
    for s in rng.random(2000):
        p = np.array([1 - s, s, 0.0], np.float32)
        o = p + np.array([0.3, -0.2, 1.0], np.float32)
        d = p - o
        ha = oracle.intersect_triangle(o, d, 0.0, 10.0, *a)[0]
        hb = oracle.intersect_triangle(o, d, 0.0, 10.0, *b)[0]
        assert ha or hb


def test_instance_inverse_round_trip(oracle):
    rng = np.random.default_rng(5)
    for _ in range(50):
        m = np.eye(4)
        m[:3, :3] = rng.normal(size=(3, 3))
        m[:3, 3] = rng.normal(size=3)
        inv = oracle.instance_inverse(m[:3, :4].reshape(-1).astype(np.float32)).reshape(3, 4)
        M = np.vstack([inv, [0, 0, 0, 1]]) @ m
        assert np.allclose(M, np.eye(4), atol=1e-4)


# ---- closest hit: BVH == brute force ---------------------------------------------------
@pytest.mark.parametrize("make", [scenes.cornell_box, lambda: scenes.random_soup(1500, seed=1),
                                  scenes.random_instances, lambda: scenes.heightfield(32)])
def test_oracle_bvh_equals_brute_force(oracle, make):
    sd = make()
    sc = oracle.Scene(sd)
    rays = np.concatenate([raygen.random_rays(2000, (0, 0, 0), 3.0, seed=7), raygen.aimed_rays(sd, 1500, seed=8),
                           raygen.axis_rays(500)])
    a = sc.trace(rays, use_bvh=False)
    b = sc.trace(rays, use_bvh=True)
    assert a.tobytes() == b.tobytes()


def test_tie_rule_prefers_lower_instance_then_primitive(oracle):
    # two instances of the same triangle at the same place: equal t, the lower
    # instance index wins; duplicated triangles in one BLAS: the lower prim wins
    tri = [(-1.0, -1.0, 0.0), (1.0, -1.0, 0.0), (0.0, 1.0, 0.0)]
    b = scenes.soup_blas(tri + tri)
    sd = scenes.assemble("dup", [b], [0, 0], [np.eye(4), np.eye(4)], [0, 0], [scenes.material((1, 1, 1))],
                         scenes.Camera((0, 0, 3), (0, 0, 0)), 2.0)
    r = np.zeros(1, RAY)
    r["o"] = (0.0, 0.0, 2.0)
    r["d"] = (0.0, 0.0, -1.0)
    r["tmax"] = 10.0
    for bvh in (False, True):
        h = oracle.Scene(sd).trace(r, use_bvh=bvh)[0]
        assert (h["inst"], h["prim"], h["t"]) == (0, 0, 2.0)


# ---- programs: closed forms ---------------------------------------------------------------
def _plane(size=20.0, y=0.0):
    s = size
    return [(-s, y, -s), (s, y, -s), (s, y, s), (-s, y, -s), (s, y, s), (-s, y, s)]


def test_ao_of_an_open_plane_is_one(oracle):
    b = scenes.soup_blas(_plane())
    sd = scenes.assemble("plane", [b], [0], [np.eye(4)], [0], [scenes.material((1, 1, 1))],
                         scenes.Camera((0, 2, 2), (0, 0, 0)), 2.0)
    W = H = 24
    acc = np.zeros((H, W, 4), np.float32)
    oracle.Scene(sd).render_ao(sd.uniforms(W, H, 0), oracle.AOParams(8, 1e-4, 0.2, 0), 0, acc, W, H)
    assert (acc[..., :3] == 1.0).all()


def test_ao_under_a_parallel_ceiling_matches_closed_form(oracle):
    # cosine rays from the floor hit a ceiling at height h within range R
    # iff cos(theta) >= h / R: P(occluded) = 1 - (h / R)^2
    h, R = 0.1, 0.2
    floor = scenes.soup_blas(_plane(30.0, 0.0))
    ceil = scenes.soup_blas(_plane(30.0, h))
    sd = scenes.assemble("slab", [floor, ceil], [0, 1], [np.eye(4)] * 2, [0, 0], [scenes.material((1, 1, 1))],
                         scenes.Camera((0.0, 0.05, 0.0), (0.0, 0.0, -1.0)), 2.0)
    W, H = 64, 64
    u = sd.uniforms(W, H, 0)
    acc = np.zeros((H, W, 4), np.float32)
    n, ph, ah = oracle.Scene(sd).render_ao(u, oracle.AOParams(8, 1e-5, R, 0), 0, acc, W, H, want_hits=True)
    on_floor = ph["inst"] == 0
    assert on_floor.sum() > 500
    occl = 1.0 - acc[..., 0].reshape(-1)[on_floor]
    expect = 1.0 - (h / R) ** 2
    assert abs(occl.mean() - expect) < 0.02


def test_pathtrace_sky_closed_form(oracle):
    # empty scene, sky on: every path is one miss; looking straight up the
    # gradient is at its top colour (0.3, 0.4, 0.5), straight down at its
    # bottom colour (0.58, 0.45, 0.25)
    sd = scenes.empty_scene()
    sc = oracle.Scene(sd)
    for eye, col in (((0, 0, 0), (0.3, 0.4, 0.5)), ((0, 0, 0), (0.58, 0.45, 0.25))):
        target = (0.0, 1.0, 0.001) if col[0] == 0.3 else (0.0, -1.0, 0.001)
        sd.camera = scenes.Camera(eye, target, up=(0.0, 0.0, -1.0), vfov_deg=1.0)
        W = H = 4
        acc = np.zeros((H, W, 4), np.float32)
        sc.render_pathtrace_glsl(sd.uniforms(W, H, 0), oracle.PTParams(1, 8, 32), 0, acc, None, W, H)
        assert np.allclose(acc[..., :3], col, atol=1e-4)


def test_pathtrace_sees_emitter_radiance(oracle):
    # the camera looks straight into a large emissive quad: radiance 15
    quad = [(-5, -5, 0), (5, -5, 0), (5, 5, 0), (-5, -5, 0), (5, 5, 0), (-5, 5, 0)]
    b = scenes.soup_blas(quad)
    sd = scenes.assemble("light", [b], [0], [np.eye(4)], [0],
                         [scenes.material((1, 1, 1), emissive=(15, 15, 15))], scenes.Camera((0, 0, 2), (0, 0, 0)),
                         2.0)
    W = H = 8
    acc = np.zeros((H, W, 4), np.float32)
    out = np.zeros((H, W, 4), np.uint8)
    oracle.Scene(sd).render_pathtrace_glsl(sd.uniforms(W, H, 0), oracle.PTParams(0, 8, 32), 0, acc, out, W, H)
    assert (acc[..., :3] == 15.0).all() and (out[..., :3] == 255).all()


def test_accumulation_is_a_running_mean(oracle):
    # frames f = 0..3 with accum_start 0: value = mean of the four frame values
    sd = scenes.cornell_box()
    sc = oracle.Scene(sd)
    W = H = 8
    acc = np.zeros((H, W, 4), np.float32)
    singles = []
    for f in range(4):
        one = np.zeros((H, W, 4), np.float32)
        sc.render_ao(sd.uniforms(W, H, f), oracle.AOParams(8, 2e-4, 0.2, 0), f, one, W, H)
        singles.append(one[..., 0].astype(np.float64))
        sc.render_ao(sd.uniforms(W, H, f), oracle.AOParams(8, 2e-4, 0.2, 0), 0, acc, W, H)
    assert np.allclose(acc[..., 0], np.mean(singles, axis=0), atol=1e-6)


# ---------------------------------------------------------------------------
# the wavefront path tracer's estimator (next-event estimation, roulette)
# ---------------------------------------------------------------------------

def test_pathtrace_config_without_extensions_is_the_glsl_program(oracle):
    """nee = 0, rr_start = 0: the same draws and products as pathtrace.*"""
    for sd, sky in ((scenes.cornell_box(), 0), (scenes.random_instances(), 1), (scenes.cornell_box(), 1)):
        sc = oracle.Scene(sd)
        W, H = 20, 12
        a = np.zeros((H, W, 4), np.float32)
        b = np.zeros((H, W, 4), np.float32)
        o1 = np.zeros((H, W, 4), np.uint8)
        o2 = np.zeros((H, W, 4), np.uint8)
        for f in range(2):
            u = sd.uniforms(W, H, f)
            n1 = sc.render_pathtrace_glsl(u, oracle.PTParams(sky, 3, 6), 0, a, o1, W, H)
            n2 = sc.render_pathtrace(u, oracle.PTConfig(sky, 3, 6, 0, 0), 0, b, o2, W, H)
            assert n1 == n2
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)) and np.array_equal(o1, o2)


def test_cornell_light_table(oracle):
    """two triangles of the 0.49 x 0.49 quad at y = 0.999: total area 0.24,
    normals along y, emission 15, inclusive running sum"""
    sd = scenes.cornell_box()
    lights, total = oracle.Scene(sd).lights()
    assert len(lights) == 2
    assert abs(total - 0.24) < 1e-6 and abs(lights["cdf"][0] - 0.12) < 1e-6 and lights["cdf"][1] == np.float32(total)
    assert np.allclose(np.abs(lights["n"][:, 1]), 1.0) and np.allclose(lights["n"][:, [0, 2]], 0.0)
    assert (lights["e"] == 15.0).all()
    assert np.allclose(lights["p0"][:, 1], 0.999)
    # no emitter: empty table
    l2, t2 = oracle.Scene(scenes.random_instances()).lights()
    assert len(l2) == 0 and t2 == 0.0


def _lit_box():
    """closed box [-1,1]^3: white diffuse floor and walls (0.8), an emissive
    ceiling (radiance 1.5, an emitter since 1.5 >= 1) split in 2 x 2 quads, a
    small grey block on the floor; camera inside looking down at the floor.
    The light covers a large solid angle, so the plain estimator is not too
    noisy and a biased variant stands out."""
    L, R = -1.0, 1.0
    floor = scenes.quad_tris((L, L, L), (R, L, L), (R, L, R), (L, L, R))
    walls = (scenes.quad_tris((L, L, L), (L, R, L), (R, R, L), (R, L, L)) +
             scenes.quad_tris((L, L, R), (R, L, R), (R, R, R), (L, R, R)) +
             scenes.quad_tris((L, L, L), (L, L, R), (L, R, R), (L, R, L)) +
             scenes.quad_tris((R, L, L), (R, R, L), (R, R, R), (R, L, R)))
    ceil = []
    for x0 in (L, 0.0):
        for z0 in (L, 0.0):
            ceil += scenes.quad_tris((x0, R, z0), (x0, R, z0 + 1), (x0 + 1, R, z0 + 1), (x0 + 1, R, z0))
    block = scenes.box_tris((0.5, 0.5, 0.5), 30.0, (0.2, -1.0, 0.1))
    parts = [(floor + walls, scenes.material((0.8, 0.8, 0.8))), (block, scenes.material((0.5, 0.5, 0.5))),
             (ceil, scenes.material((1, 1, 1), emissive=(1.5, 1.5, 1.5)))]
    blas = [scenes.soup_blas(p) for p, _ in parts]
    cam = scenes.Camera((0.0, 0.3, 0.9), (0.0, -1.0, 0.0), up=(0.0, 0.0, -1.0), vfov_deg=60.0)
    return scenes.assemble("lit_box", blas, [0, 1, 2], [np.eye(4)] * 3, [0, 1, 2], [m for _, m in parts], cam, 2.0)


def _frame_means(oracle, sd, cfg, W, H, frames):
    """image mean of each independent frame (alpha = 1: accum_start = frame)"""
    sc = oracle.Scene(sd)
    out = []
    for f in range(frames):
        acc = np.zeros((H, W, 4), np.float32)
        sc.render_pathtrace(sd.uniforms(W, H, f), cfg, f, acc, None, W, H)
        out.append(acc[..., :3].astype(np.float64).mean(axis=(0, 1)))
    return np.array(out)


@pytest.mark.parametrize("nee,rr", [(1, 0), (0, 1), (1, 2)])
def test_nee_and_roulette_keep_the_expectation(oracle, nee, rr):
    """next-event estimation and Russian roulette are unbiased: the image
    mean agrees with the plain estimator within 5 standard errors (same
    scene, same depth limit, independent seeds per frame).  At this sample
    count a 10 % error in the next-event weight or a 10 % bias in the
    roulette weight is far outside that band."""
    sd = _lit_box()
    W, H, F = 16, 16, 24
    plain = _frame_means(oracle, sd, oracle.PTConfig(0, 32, 3, 0, 0), W, H, F)
    ext = _frame_means(oracle, sd, oracle.PTConfig(0, 32, 3, rr, nee), W, H, F)
    m1, m2 = plain.mean(0), ext.mean(0)
    se = np.sqrt(plain.var(0, ddof=1) / F + ext.var(0, ddof=1) / F)
    assert (np.abs(m1 - m2) < 5 * se).all(), (m1, m2, se)
    assert (se < 0.01 * m1).all()   # the band is tight


def test_sample_streams_keep_sample_zero_and_the_expectation(oracle):
    """sample_streams = 1 seeds sample i with tea(x + y*W + i*W*H, frame)
    and sums the samples in fixed point (2^20 units per 1.0): sample 0 is
    the sequential stream's, so one-sample images agree up to that rounding
    (half a unit, times the blend weight 1/4 of frame 3); more samples differ
    from the sequential stream but keep its expectation (image means within
    5 standard errors)"""
    sd = _lit_box()
    W, H = 16, 16
    sc = oracle.Scene(sd)
    a = np.zeros((H, W, 4), np.float32)
    b = np.zeros((H, W, 4), np.float32)
    sc.render_pathtrace(sd.uniforms(W, H, 3), oracle.PTConfig(0, 1, 3, 2, 1, 0), 0, a, None, W, H)
    sc.render_pathtrace(sd.uniforms(W, H, 3), oracle.PTConfig(0, 1, 3, 2, 1, 1), 0, b, None, W, H)
    d = np.abs(a[..., :3].astype(np.float64) - b[..., :3])
    assert (d <= 0.25 * 2.0 ** -21 + 1e-7 * np.abs(a[..., :3])).all(), d.max()
    assert (d > 0).any()   # the rounding is there
    sc.render_pathtrace(sd.uniforms(W, H, 3), oracle.PTConfig(0, 4, 3, 2, 1, 0), 0, a, None, W, H)
    sc.render_pathtrace(sd.uniforms(W, H, 3), oracle.PTConfig(0, 4, 3, 2, 1, 1), 0, b, None, W, H)
    assert not np.array_equal(a, b)
    F = 24
    seq = _frame_means(oracle, sd, oracle.PTConfig(0, 32, 3, 2, 1, 0), W, H, F)
    per = _frame_means(oracle, sd, oracle.PTConfig(0, 32, 3, 2, 1, 1), W, H, F)
    se = np.sqrt(seq.var(0, ddof=1) / F + per.var(0, ddof=1) / F)
    assert (np.abs(seq.mean(0) - per.mean(0)) < 5 * se).all(), (seq.mean(0), per.mean(0), se)


def test_nee_reduces_noise_for_a_small_light(oracle):
    """Cornell box (light 6 % of the ceiling), direct light only: sampling
    the emitter beats waiting for a cosine-sampled ray to hit it"""
    sd = scenes.cornell_box()
    W, H, F = 16, 16, 12
    plain = _frame_means(oracle, sd, oracle.PTConfig(0, 16, 1, 0, 0), W, H, F)
    nee = _frame_means(oracle, sd, oracle.PTConfig(0, 16, 1, 0, 1), W, H, F)
    assert (nee.var(0) < 0.5 * plain.var(0)).all()
    se = np.sqrt(plain.var(0, ddof=1) / F + nee.var(0, ddof=1) / F)
    assert (np.abs(plain.mean(0) - nee.mean(0)) < 5 * se).all()


def test_depth_zero_sees_only_emitters_and_sky(oracle):
    """max_bounces = 0: a path ends at its first surface unless that is an
    emitter, so next-event estimation adds nothing"""
    sd = scenes.cornell_box()
    W = H = 12
    a = np.zeros((H, W, 4), np.float32)
    b = np.zeros((H, W, 4), np.float32)
    oracle.Scene(sd).render_pathtrace(sd.uniforms(W, H, 0), oracle.PTConfig(0, 4, 0, 0, 0), 0, a, None, W, H)
    oracle.Scene(sd).render_pathtrace(sd.uniforms(W, H, 0), oracle.PTConfig(0, 4, 0, 0, 1), 0, b, None, W, H)
    assert np.array_equal(a, b)
    vals = np.unique(a[..., 0])
    assert set(np.round(vals * 4).astype(int)) <= {0, 15, 30, 45, 60}   # k of 4 samples saw the light


# ---------------------------------------------------------------------------
# the GLSL programs' own chains: ao.* (assets/glsl/ao.rgen:37-83,
# ao.rchit:45-88), pathtrace.rchit's lobe choice (pathtrace.rchit:91-112) and
# debug.* (debug.rgen:18-37, debug.rchit:9-13, debug.rmiss:6-9)
# ---------------------------------------------------------------------------
SKY_TOP = np.array([0.3, 0.4, 0.5])
SKY_BOTTOM = np.array([0.58, 0.45, 0.25])


def _sky(dirs):
    """pathtrace.rmiss:9-17 in float64 for unit directions (N, 3), without the
    sun (the tests aim away from it and assert so)"""
    dirs = np.asarray(dirs, np.float64)
    sun = np.array([0.0, 1.0, -0.25]) / np.linalg.norm([0.0, 1.0, -0.25])
    assert (dirs @ sun < 0.98).all()
    x = np.clip((0.5 * (dirs[:, 1] + 1.0) - 0.35) / 0.3, 0.0, 1.0)
    t = x * x * (3.0 - 2.0 * x)
    return SKY_BOTTOM * (1.0 - t[:, None]) + SKY_TOP * t[:, None]


def _camera_dirs(u, W, H, jx, jy):
    """camera rays of pathtrace.rgen:52-58 in float64 from the uniforms, for
    sub-pixel offsets (jx, jy): (H, W, 3) unit directions and the origin"""
    vi = np.asarray(u["view_inverse"][0], np.float64).reshape(4, 4).T
    pi = np.asarray(u["projection_inverse"][0], np.float64).reshape(4, 4).T
    xs, ys = np.meshgrid(np.arange(W) + jx, np.arange(H) + jy)
    ndc = np.stack([xs / W * 2.0 - 1.0, ys / H * 2.0 - 1.0, np.ones_like(xs), np.ones_like(xs)], -1)
    tgt = ndc @ pi.T
    t3 = tgt[..., :3] / np.linalg.norm(tgt[..., :3], axis=-1, keepdims=True)
    d = t3 @ vi[:3, :3].T
    return vi[:3, 3], d


def _blue_permutations(seed=5):
    """255 x 1 RGBA8 texture whose every channel is a permutation of
    1..255: a blue-noise draw at a uniform texel is uniform over k/255,
    k = 1..255, and never 0 (a cosine direction is then never tangent)"""
    rng = np.random.default_rng(seed)
    tex = np.zeros((1, 255, 4), np.uint8)
    for c in range(4):
        tex[0, :, c] = rng.permutation(255) + 1
    return tex


def test_ao_glsl_chain_on_an_open_plane_is_one(oracle):
    """ao.rchit adds 1 only from the second hit of a chain on (depth > 0,
    ao.rchit:84-86): over an open plane every chain is one hit and then an
    escape, so the image is exactly 1"""
    b = scenes.soup_blas(_plane(40.0))
    sd = scenes.assemble("plane", [b], [0], [np.eye(4)], [0], [scenes.material((1, 1, 1))],
                         scenes.Camera((0.0, 2.0, 2.0), (0.0, 0.0, 0.0)), 2.0)
    W = H = 24
    acc = np.zeros((H, W, 4), np.float32)
    oracle.Scene(sd).render_ao_glsl(sd.uniforms(W, H, 0), 0, acc, _blue_permutations(), W, H)
    assert (acc[..., :3] == 1.0).all()


def test_ao_glsl_chain_between_parallel_planes_matches_closed_form(oracle):
    """Floor at y = 0 and ceiling at y = h = 5, the camera between them
    looking down.  A chain segment leaves one plane with cos(theta) =
    sqrt(xi), xi = k/255 a uniform blue-noise draw (ao.rchit:45-51, 82;
    sampling.glsl:93-96), and reaches the other plane within the range 10
    (ao.rchit:83) iff cos(theta) >= h/10 = 0.5, i.e. k >= 64: p = 192/255.
    The chain stops after 5 hits (ao.rgen:71, max_samples = 4), so the hits
    after the first number K in 0..4 with P(K >= j) = p^j, and
    E[colour] = 1 - E[K]/4 (ao.rgen:74-76)."""
    h = 5.0
    floor = scenes.soup_blas(_plane(200.0, 0.0))
    ceil = scenes.soup_blas(_plane(200.0, h))
    sd = scenes.assemble("slab", [floor, ceil], [0, 1], [np.eye(4)] * 2, [0, 0], [scenes.material((1, 1, 1))],
                         scenes.Camera((0.0, 2.5, 0.0), (0.0, 0.0, 0.001), up=(0.0, 0.0, -1.0), vfov_deg=30.0),
                         2.0)
    W = H = 64
    acc = np.zeros((H, W, 4), np.float32)
    oracle.Scene(sd).render_ao_glsl(sd.uniforms(W, H, 0), 0, acc, _blue_permutations(), W, H)
    k = np.arange(1, 256)
    assert not np.any(np.abs(np.sqrt(k / 255.0) - 0.5) < 1e-4)   # no draw sits on the threshold
    p = np.count_nonzero(np.sqrt(k / 255.0) >= 0.5) / 255.0
    pk = np.array([p ** j * (1 - p) for j in range(4)] + [p ** 4])   # P(K = 0..4)
    ek = (np.arange(5) * pk).sum()
    vk = (np.arange(5) ** 2 * pk).sum() - ek ** 2
    col = acc[..., 0].astype(np.float64)
    # a pixel's colour is 1 - (sum of its 4 samples' K) / 16
    assert np.allclose(np.round((1.0 - col) * 16.0), (1.0 - col) * 16.0, atol=1e-5)
    se = math.sqrt(vk / (W * H * 4)) / 4.0
    assert abs(col.mean() - (1.0 - ek / 4.0)) < 5.0 * se, (col.mean(), 1.0 - ek / 4.0, se)


def _mirror_plane_scene(metallic, roughness, base, eye, target, vfov):
    b = scenes.soup_blas(_plane(500.0))
    return scenes.assemble("mirror", [b], [0], [np.eye(4)], [0],
                           [scenes.material(base, metallic=metallic, roughness=roughness)],
                           scenes.Camera(eye, target, up=(0.0, -1.0, 0.0), vfov_deg=vfov), 2.0)


def test_pathtrace_glsl_metal_mirror_reflects_the_sky(oracle):
    """metallic 1, roughness 0 (pathtrace.rchit:97-100): every path takes the
    metal lobe, GGX with alpha = 0 returns reflect(d, n) exactly
    (sampling.glsl:88-91), and the reflected ray escapes to the sky, so a
    pixel is base * vertex colour * sky(reflect(d, n)).  The camera looks
    almost horizontally over the plane, where the sky gradient is steep: each
    pixel lies between the closed form at its footprint's upper and lower
    edges."""
    base = (0.9, 0.6, 0.3)
    sd = _mirror_plane_scene(1.0, 0.0, base, (0.0, 1.0, 0.0), (0.0, 0.8, 1.0), 12.0)
    W = H = 48
    u = sd.uniforms(W, H, 0)
    acc = np.zeros((H, W, 4), np.float32)
    oracle.Scene(sd).render_pathtrace_glsl(u, oracle.PTParams(1, 8, 32), 0, acc, None, W, H)
    _, d0 = _camera_dirs(u, W, H, 0.0, 0.0)
    _, d1 = _camera_dirs(u, W, H, 1.0, 1.0)
    assert (d0[..., 1] < 0).all() and (d1[..., 1] < 0).all()   # every ray hits the plane
    refl = lambda d: d * np.array([1.0, -1.0, 1.0])
    s0 = _sky(refl(d0).reshape(-1, 3)).reshape(H, W, 3) * base
    s1 = _sky(refl(d1).reshape(-1, 3)).reshape(H, W, 3) * base
    lo, hi = np.minimum(s0, s1), np.maximum(s0, s1)
    img = acc[..., :3].astype(np.float64)
    assert ((img >= lo - 2e-6) & (img <= hi + 2e-6)).all()
    assert (hi - lo).max() < 0.02 and img[..., 0].std() > 0.01   # tight brackets, steep gradient


def test_pathtrace_glsl_dielectric_coat_reflects_with_fresnel_004(oracle):
    """metallic 0, roughness 0, black base (pathtrace.rchit:101-111): the
    microfacet normal is the surface normal, so the coat reflects with
    probability F(n, n, 1/1.5) = 0.04 (a mirror ray to the sky, value 1 *
    sky) and otherwise takes the diffuse lobe with value base = 0.  The
    camera looks steeply down, so every reflection sees the top sky colour:
    a pixel is k/8 of it (k of 8 samples reflected), and the mean of k/8 is
    0.04 within 5 standard errors."""
    sd = _mirror_plane_scene(0.0, 0.0, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.2, 0.6), 20.0)
    W = H = 64
    u = sd.uniforms(W, H, 0)
    _, d0 = _camera_dirs(u, W, H, 0.5, 0.5)
    assert (-d0[..., 1] > 0.35).all()   # reflections above the gradient band: sky = top colour
    acc = np.zeros((H, W, 4), np.float32)
    oracle.Scene(sd).render_pathtrace_glsl(u, oracle.PTParams(1, 8, 32), 0, acc, None, W, H)
    frac = acc[..., :3].astype(np.float64) / SKY_TOP
    k8 = frac * 8.0
    assert np.allclose(k8, np.round(k8), atol=1e-4)
    assert np.allclose(frac[..., 0], frac[..., 1], atol=1e-5) and np.allclose(frac[..., 0], frac[..., 2], atol=1e-5)
    n = W * H * 8
    se = math.sqrt(0.04 * 0.96 / n)
    assert abs(frac[..., 0].mean() - 0.04) < 5.0 * se, (frac[..., 0].mean(), se)


def test_debug_program_barycentrics_and_misses(oracle):
    """debug.rgen:18-37: one ray through each pixel centre, range
    [0.001, 1000]; a hit stores (1-u-v, u, v) (debug.rchit:11-12), a miss
    keeps the ray direction (debug.rmiss leaves the payload).  Closed form:
    the barycentrics of the ray/plane intersection solved in float64 from
    the triangle's vertices; misses equal the float64 camera ray; a copy of
    the triangle beyond the range is never seen."""
    v = np.array([(-1.2, -0.9, 0.0), (1.4, -1.0, 0.0), (-0.3, 1.3, 0.0)], np.float64)
    far = v + np.array([0.0, 0.0, -2000.0])
    near = scenes.soup_blas(v.astype(np.float32))
    beyond = scenes.soup_blas(far.astype(np.float32) * np.array([600.0, 600.0, 1.0], np.float32))
    sd = scenes.assemble("debug", [near, beyond], [0, 1], [np.eye(4)] * 2, [0, 0], [scenes.material((1, 1, 1))],
                         scenes.Camera((0.1, 0.2, 3.0), (0.0, 0.0, 0.0), vfov_deg=60.0), 2.0)
    W, H = 40, 30
    u = sd.uniforms(W, H, 0)
    out = oracle.Scene(sd).render_debug(u, W, H)
    o, d = _camera_dirs(u, W, H, 0.5, 0.5)
    t = -o[2] / d[..., 2]
    P = o + t[..., None] * d
    A = np.array([[v[0, 0], v[1, 0], v[2, 0]], [v[0, 1], v[1, 1], v[2, 1]], [1.0, 1.0, 1.0]])
    rhs = np.stack([P[..., 0], P[..., 1], np.ones_like(t)], -1)
    bary = np.linalg.solve(A, rhs.reshape(-1, 3).T).T.reshape(H, W, 3)
    inside = (bary > 1e-4).all(-1)
    outside = (bary < -1e-4).any(-1)
    assert inside.sum() > 200 and outside.sum() > 200
    assert np.allclose(out[inside][:, :3], bary[inside], atol=2e-5)
    assert np.allclose(out[inside][:, :3].sum(-1), 1.0, atol=1e-6)
    assert np.allclose(out[outside][:, :3], d[outside], atol=1e-6)
    assert (out[..., 3] == 0.0).all()


def test_oracle_hits_name_geometry_and_primitive(oracle):
    """a BLAS of three geometries (two indexed, one non-indexed): the BVH
    answer equals brute force, and every hit's (geometry, primitive within
    the geometry) names a triangle that the ray hits at the reported t
    (gl_GeometryIndexEXT / gl_PrimitiveID, src/ray/acceleration.rs:148-208)"""
    sd = scenes.multi_geometry_scene()
    sc = oracle.Scene(sd)
    rays = np.concatenate([raygen.random_rays(3000, (0, 0, 0), 2.5, seed=21), raygen.aimed_rays(sd, 2000, seed=22)])
    a = sc.trace(rays, use_bvh=False)
    b = sc.trace(rays, use_bvh=True)
    assert a.tobytes() == b.tobytes()
    hit = a["inst"] != oracle.MISS
    assert (a["geom"][~hit] == oracle.MISS).all()
    mb = sd.blas[0]
    geo_hits = a[hit & (sd.instance_descs["blas"][np.minimum(a["inst"], len(sd.instance_descs) - 1)] == 0)]
    assert set(np.unique(geo_hits["geom"]).tolist()) == {0, 1, 2}
    for h, r in zip(a[hit], rays[hit]):
        d = sd.instance_descs[h["inst"]]
        blas = sd.blas[int(d["blas"])]
        verts, inds = (blas.geometries[h["geom"]] if hasattr(blas, "geometries") else (blas.vertices, blas.indices))
        k = 3 * int(h["prim"])
        tri = verts["pos"][:, :3][inds[k:k + 3]] if inds is not None else verts["pos"][k:k + 3, :3]
        # the ray in object space with the oracle's fp32 inverse, evaluated
        # left to right as the traversal does
        m = oracle.instance_inverse(d["transform"]).astype(np.float32)
        ro, rd = r["o"].astype(np.float32), r["d"].astype(np.float32)
        o = np.array([((m[4 * a] * ro[0] + m[4 * a + 1] * ro[1]) + m[4 * a + 2] * ro[2]) + m[4 * a + 3]
                      for a in range(3)], np.float32)
        dd = np.array([(m[4 * a] * rd[0] + m[4 * a + 1] * rd[1]) + m[4 * a + 2] * rd[2] for a in range(3)], np.float32)
        ok, t, u, v = oracle.intersect_triangle(o, dd, float(r["tmin"]), float(r["tmax"]), *tri)
        assert ok and (np.float32(t), np.float32(u), np.float32(v)) == (h["t"], h["u"], h["v"])
