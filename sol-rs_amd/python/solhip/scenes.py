"""Seeded synthetic scenes and cameras.

A SceneData holds exactly what the sol-rs examples hand to the ray path:
  * blas            one triangle geometry per BLAS (ModelVertex array + u32
                    indices), as BLAS::new receives it (src/ray/acceleration.rs:136-239);
  * instance_descs  the 64-B TLAS records of TLAS::create_instances
                    (src/ray/acceleration.rs:315-342), blas = BLAS index;
  * scene_instances the 144-B SceneInstance table the hit shaders read
                    (src/ray/mod.rs:113-120);
  * objects         per SceneInstance.id: (vertices, indices) seen by the hit
                    shaders through the vertices[] / indices[] descriptor arrays;
  * materials / instance_material: materials[gl_InstanceID].
and a camera from which SceneUniforms are made (examples/*.rs:19-31).

Scenes:
  cornell_box()      36-triangle procedural Cornell box (Lambert walls, light)
  heightfield(n)     n x n-quad fBm terrain, 2*n*n triangles, one BLAS
  instanced(...)     TLAS of many instances drawn from 5 BLAS
  single_triangle(), quads() ... small cases for tests
"""
import math

import numpy as np

from .records import (GEOMETRY_INSTANCE_FORCE_OPAQUE, GEOMETRY_INSTANCE_TRIANGLE_FACING_CULL_DISABLE,
                      INSTANCE_DESC, MATERIAL, MODEL_VERTEX, SCENE_INSTANCE, UNIFORMS)


class Blas:
    def __init__(self, vertices, indices):
        self.vertices = np.ascontiguousarray(vertices, MODEL_VERTEX)
        self.indices = np.ascontiguousarray(indices, np.uint32).reshape(-1)

    @property
    def tri_count(self):
        return len(self.indices) // 3


class SceneData:
    def __init__(self, name, blas, instance_descs, scene_instances, objects, materials,
                 instance_material, camera, extent):
        self.name = name
        self.blas = blas
        self.instance_descs = instance_descs
        self.scene_instances = scene_instances
        self.objects = objects
        self.materials = materials
        self.instance_material = np.asarray(instance_material, np.int64)
        self.camera = camera
        self.extent = float(extent)

    @property
    def tri_count(self):
        return sum(b.tri_count for b in self.blas)

    def uniforms(self, W, H, frame):
        return self.camera.uniforms(W, H, frame)


# ----------------------------------------------------------------------------
# camera math (glam 0.20 look_at_rh / perspective_rh, src/scene/camera.rs:98-106)
# ----------------------------------------------------------------------------
def look_at_rh(eye, center, up):
    eye, center, up = (np.asarray(a, np.float64) for a in (eye, center, up))
    f = center - eye
    f = f / np.linalg.norm(f)
    s = np.cross(f, up)
    s = s / np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3] = s
    m[1, :3] = u
    m[2, :3] = -f
    m[0, 3] = -s @ eye
    m[1, 3] = -u @ eye
    m[2, 3] = f @ eye
    return m  # row-major math matrix


def perspective_rh(fov_y, aspect, z_near, z_far):
    h = 1.0 / math.tan(0.5 * fov_y)
    w = h / aspect
    r = z_far / (z_near - z_far)
    m = np.zeros((4, 4))
    m[0, 0] = w
    m[1, 1] = h
    m[2, 2] = r
    m[3, 2] = -1.0
    m[2, 3] = r * z_near
    return m


def colmajor(m):
    """math matrix (row-major numpy) -> 16 floats in glam's column-major order"""
    return np.asarray(m, np.float64).T.reshape(-1).astype(np.float32)


class Camera:
    """Camera::look_at + perspective (src/scene/camera.rs:54-113); vfov in degrees."""

    def __init__(self, eye, center, up=(0.0, -1.0, 0.0), vfov_deg=35.0, z_near=0.1, z_far=1000.0):
        self.eye, self.center, self.up = eye, center, up
        self.vfov = vfov_deg
        self.z_near, self.z_far = z_near, z_far

    def uniforms(self, W, H, frame):
        view = look_at_rh(self.eye, self.center, self.up)
        proj = perspective_rh(math.radians(self.vfov), W / H, self.z_near, self.z_far)
        u = np.zeros(1, UNIFORMS)
        u["model"][0] = colmajor(np.eye(4))
        u["view"][0] = colmajor(view)
        u["view_inverse"][0] = colmajor(np.linalg.inv(view))
        u["projection"][0] = colmajor(proj)
        u["projection_inverse"][0] = colmajor(np.linalg.inv(proj))
        u["model_view_projection"][0] = colmajor(proj @ view)
        u["frame"][0] = (W, H, frame)
        return u


# ----------------------------------------------------------------------------
# assembly helpers
# ----------------------------------------------------------------------------
def make_vertices(pos, normal=None, color=None):
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    v = np.zeros(len(pos), MODEL_VERTEX)
    v["pos"][:, :3] = pos
    v["pos"][:, 3] = 1.0
    if normal is None:
        v["normal"][:, 1] = 1.0
    else:
        v["normal"][:, :3] = np.asarray(normal, np.float32).reshape(-1, 3)
    v["normal"][:, 3] = 1.0  # src/scene/mod.rs:189 (w = 1)
    if color is None:
        v["color"][:] = 1.0
    else:
        c = np.asarray(color, np.float32).reshape(-1, 4)
        v["color"][:] = c
    return v


def material(base, emissive=(0, 0, 0), metallic=0.0, roughness=1.0):
    m = np.zeros(1, MATERIAL)
    m["base_color"][0, :3] = base
    m["base_color"][0, 3] = 1.0
    m["emissive"][0] = emissive
    m["metallic"] = metallic
    m["roughness"] = roughness
    return m


def affine_rows(m4):
    """4x4 math matrix -> the 3x4 row-major transform of InstanceDescriptor
    (TLAS::create_instances transposes glam's column-major matrix and keeps
    the first 12 floats, src/ray/acceleration.rs:330-331)."""
    return np.asarray(m4, np.float64)[:3, :4].reshape(-1).astype(np.float32)


def assemble(name, blas_list, inst_blas, inst_xforms, inst_material, materials, camera, extent,
             obj_of_inst=None, index_dtype=np.uint32):
    """Instances -> SceneData.  inst_xforms are 4x4 math matrices."""
    n = len(inst_blas)
    descs = np.zeros(n, INSTANCE_DESC)
    sis = np.zeros(n, SCENE_INSTANCE)
    flags = GEOMETRY_INSTANCE_FORCE_OPAQUE | GEOMETRY_INSTANCE_TRIANGLE_FACING_CULL_DISABLE
    for i in range(n):
        m = np.asarray(inst_xforms[i], np.float64)
        descs[i]["transform"] = affine_rows(m)
        descs[i]["id_mask"] = (i & 0xFFFFFF) | (0xFF << 24)
        descs[i]["offset_flags"] = (flags << 24)
        descs[i]["blas"] = inst_blas[i]
        sis[i]["id"] = inst_blas[i] if obj_of_inst is None else obj_of_inst[i]
        m32 = m.astype(np.float32).astype(np.float64)
        sis[i]["transform"] = colmajor(m32)
        sis[i]["transform_it"] = colmajor(np.linalg.inv(m32).T)
    objects = [(b.vertices, b.indices.astype(index_dtype)) for b in blas_list]
    mats = np.concatenate(materials) if len(materials) else np.zeros(0, MATERIAL)
    return SceneData(name, blas_list, descs, sis, objects, mats, inst_material, camera, extent)


def with_transforms(sd, inst_xforms):
    """the scene with its instances moved to new 4x4 object-to-world
    matrices: the TLAS records and the SceneInstance table (transform and
    transform_it, what SceneDescription::blas_transform / update write,
    src/ray/mod.rs:162-193) are recomputed as assemble() makes them"""
    descs = sd.instance_descs.copy()
    sis = sd.scene_instances.copy()
    for i, m in enumerate(inst_xforms):
        m = np.asarray(m, np.float64)
        descs[i]["transform"] = affine_rows(m)
        m32 = m.astype(np.float32).astype(np.float64)
        sis[i]["transform"] = colmajor(m32)
        sis[i]["transform_it"] = colmajor(np.linalg.inv(m32).T)
    return SceneData(sd.name, sd.blas, descs, sis, sd.objects, sd.materials, sd.instance_material, sd.camera,
                     sd.extent)


def instance_xforms(sd):
    """the instances' 4x4 object-to-world matrices (float64)"""
    out = []
    for d in sd.instance_descs:
        m = np.eye(4)
        m[:3, :4] = d["transform"].reshape(3, 4)
        out.append(m)
    return out


def quad_tris(p0, p1, p2, p3):
    """two triangles (p0,p1,p2), (p0,p2,p3)"""
    return [p0, p1, p2, p0, p2, p3]


def box_tris(size, rot_y_deg, offset):
    sx, sy, sz = size
    c = [(x, y, z) for x in (-sx / 2, sx / 2) for y in (0.0, sy) for z in (-sz / 2, sz / 2)]
    # corner index bits: x(4) y(2) z(1)
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    a = math.radians(rot_y_deg)
    ca, sa = math.cos(a), math.sin(a)
    pts = []
    for f in faces:
        q = [c[k] for k in f]
        pts += quad_tris(*q)
    out = []
    for (x, y, z) in pts:
        out.append((ca * x + sa * z + offset[0], y + offset[1], -sa * x + ca * z + offset[2]))
    return out


def flat_normals(pos):
    p = np.asarray(pos, np.float64).reshape(-1, 3, 3)
    n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return np.repeat(n, 3, axis=0)


def soup_blas(pos, color=None):
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    v = make_vertices(pos, flat_normals(pos), color)
    return Blas(v, np.arange(len(pos), dtype=np.uint32))


# ----------------------------------------------------------------------------
# scenes
# ----------------------------------------------------------------------------
def cornell_box():
    """Procedural Cornell box: room [-1,1]^3 open at +z; white floor, ceiling
    and back wall; red left, green right wall; a short box (0.6^3, -18 deg)
    and a tall box (0.6 x 1.2 x 0.6, +18 deg); an emissive quad light under
    the ceiling with area 0.06 of the ceiling and radiance 15.
    36 triangles in 6 BLAS/instances (one per material primitive, as
    SceneDescription::from_meshes makes them, src/ray/mod.rs:78-135)."""
    L, R = -1.0, 1.0
    floor = quad_tris((L, L, L), (R, L, L), (R, L, R), (L, L, R))
    ceil = quad_tris((L, R, L), (L, R, R), (R, R, R), (R, R, L))
    back = quad_tris((L, L, L), (L, R, L), (R, R, L), (R, L, L))
    left = quad_tris((L, L, L), (L, L, R), (L, R, R), (L, R, L))
    right = quad_tris((R, L, L), (R, R, L), (R, R, R), (R, L, R))
    short = box_tris((0.6, 0.6, 0.6), -18.0, (0.35, -1.0, 0.25))
    tall = box_tris((0.6, 1.2, 0.6), 18.0, (-0.35, -1.0, -0.3))
    h = math.sqrt(0.06 * 4.0) / 2.0
    y = 0.999
    light = quad_tris((-h, y, -h), (-h, y, h), (h, y, h), (h, y, -h))
    white, red, green = (0.73, 0.73, 0.73), (0.65, 0.05, 0.05), (0.12, 0.45, 0.15)
    parts = [(floor + ceil + back, white), (left, red), (right, green), (short, white),
             (tall, white), (light, (1.0, 1.0, 1.0))]
    blas = [soup_blas(p) for p, _ in parts]
    mats = [material(c) for _, c in parts[:-1]] + [material((1, 1, 1), emissive=(15, 15, 15))]
    n = len(blas)
    cam = Camera((0.0, 0.0, 3.4), (0.0, 0.0, 0.0), up=(0.0, -1.0, 0.0), vfov_deg=40.0)
    return assemble("cornell_box", blas, list(range(n)), [np.eye(4)] * n, list(range(n)), mats, cam, 2.0)


def _hash2(ix, iy, seed):
    h = (ix.astype(np.uint64) * np.uint64(0x8da6b343) + iy.astype(np.uint64) * np.uint64(0xd8163841)
         + np.uint64(seed) * np.uint64(0xcb1ab31f)) & np.uint64(0xFFFFFFFF)
    h = h.astype(np.uint32)
    h ^= h >> np.uint32(16)
    h *= np.uint32(0x7feb352d)
    h ^= h >> np.uint32(15)
    h *= np.uint32(0x846ca68b)
    h ^= h >> np.uint32(16)
    return h.astype(np.float64) / 4294967296.0 * 2.0 - 1.0


def value_noise(x, y, seed):
    ix, iy = np.floor(x), np.floor(y)
    fx, fy = x - ix, y - iy
    ix, iy = ix.astype(np.int64), iy.astype(np.int64)
    sx, sy = fx * fx * (3 - 2 * fx), fy * fy * (3 - 2 * fy)
    a = _hash2(ix, iy, seed)
    b = _hash2(ix + 1, iy, seed)
    c = _hash2(ix, iy + 1, seed)
    d = _hash2(ix + 1, iy + 1, seed)
    return (a * (1 - sx) + b * sx) * (1 - sy) + (c * (1 - sx) + d * sx) * sy


def fbm(x, y, octaves=6, lacunarity=2.0, gain=0.5, seed=1, base_freq=4.0):
    total, amp, freq, norm = 0.0, 1.0, base_freq, 0.0
    for o in range(octaves):
        total = total + amp * value_noise(x * freq, y * freq, seed + 101 * o)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return total / norm


def heightfield(n=1024, seed=1, amplitude=0.2, W=1920, H=1080):
    """n x n quads over x,z in [-1,1] (extent 2), height = amplitude * fBm
    (6 octaves, lacunarity 2, gain 0.5, value noise, seed); per-vertex
    normals by central differences; one BLAS, one instance.  Camera at 35
    degrees elevation looking at the centre."""
    xs = np.linspace(-1.0, 1.0, n + 1)
    X, Z = np.meshgrid(xs, xs)
    Y = amplitude * fbm(X, Z, seed=seed)
    step = 2.0 / n
    dYdx = np.gradient(Y, step, axis=1)
    dYdz = np.gradient(Y, step, axis=0)
    nrm = np.stack([-dYdx, np.ones_like(Y), -dYdz], axis=-1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    pos = np.stack([X, Y, Z], axis=-1).reshape(-1, 3)
    v = make_vertices(pos, nrm.reshape(-1, 3))
    i = np.arange(n)
    I, J = np.meshgrid(i, i)
    a = (J * (n + 1) + I).reshape(-1)
    b = a + 1
    c = a + (n + 1)
    d = c + 1
    idx = np.stack([a, c, b, b, c, d], axis=-1).reshape(-1).astype(np.uint32)
    blas = [Blas(v, idx)]
    el = math.radians(35.0)
    dist = 1.5
    cam = Camera((0.0, dist * math.sin(el), dist * math.cos(el)), (0.0, 0.0, 0.0), up=(0.0, -1.0, 0.0),
                 vfov_deg=40.0)
    mats = [material((0.8, 0.8, 0.8))]
    return assemble("heightfield_%d" % n, blas, [0], [np.eye(4)], [0], mats, cam, 2.0)


def icosphere(level):
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]
        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    P = np.array(verts)
    v = make_vertices(P, P)
    return Blas(v, np.array(faces, np.uint32).reshape(-1))


def box_blas():
    pos = box_tris((2.0, 2.0, 2.0), 0.0, (0.0, -1.0, 0.0))
    return soup_blas(pos)


def tube_blas(n_seg=500, n_around=1000, R=1.0, r=0.15):
    """swept tube along a (3,2) torus knot: 2*n_seg*n_around triangles"""
    s = np.linspace(0.0, 2 * math.pi, n_seg, endpoint=False)
    def knot(t):
        return np.stack([(R + 0.4 * np.cos(3 * t)) * np.cos(2 * t), (R + 0.4 * np.cos(3 * t)) * np.sin(2 * t),
                         0.4 * np.sin(3 * t)], axis=-1)
    C = knot(s)
    T = knot(s + 1e-4) - C
    T /= np.linalg.norm(T, axis=1, keepdims=True)
    up = np.array([0.0, 0.0, 1.0])
    Nn = np.cross(T, up)
    Nn /= np.linalg.norm(Nn, axis=1, keepdims=True)
    B = np.cross(T, Nn)
    a = np.linspace(0.0, 2 * math.pi, n_around, endpoint=False)
    ca, sa = np.cos(a)[None, :, None], np.sin(a)[None, :, None]
    nrm = ca * Nn[:, None, :] + sa * B[:, None, :]
    P = C[:, None, :] + r * nrm
    v = make_vertices(P.reshape(-1, 3), nrm.reshape(-1, 3))
    i = np.arange(n_seg)[:, None]
    j = np.arange(n_around)[None, :]
    a0 = i * n_around + j
    a1 = i * n_around + (j + 1) % n_around
    b0 = ((i + 1) % n_seg) * n_around + j
    b1 = ((i + 1) % n_seg) * n_around + (j + 1) % n_around
    idx = np.stack([a0, b0, a1, a1, b0, b1], axis=-1).reshape(-1).astype(np.uint32)
    return Blas(v, idx)


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def instanced(n_instances=10000, seed=2, tube=(500, 1000), spread=50.0):
    """TLAS of n_instances drawn uniformly from 5 BLAS (icospheres with 20,
    1280 and 20480 triangles, a 12-triangle box, a swept tube with
    2*tube[0]*tube[1] triangles), uniform random rotation, translation in
    [-spread, spread]^3, uniform scale in [0.5, 2]; per-instance Lambert
    albedo; lit by the gradient sky."""
    blas = [icosphere(0), icosphere(3), icosphere(5), box_blas(), tube_blas(*tube)]
    rng = np.random.default_rng(seed)
    which = rng.integers(0, 5, size=n_instances)
    xf = []
    for i in range(n_instances):
        m = np.eye(4)
        m[:3, :3] = random_rotation(rng) * rng.uniform(0.5, 2.0)
        m[:3, 3] = rng.uniform(-spread, spread, size=3)
        xf.append(m)
    albedo = rng.uniform(0.2, 0.9, size=(n_instances, 3))
    mats = [material(tuple(a)) for a in albedo]
    cam = Camera((0.0, 0.0, 2.6 * spread), (0.0, 0.0, 0.0), up=(0.0, -1.0, 0.0), vfov_deg=45.0,
                 z_far=10000.0)
    return assemble("instanced_%d" % n_instances, blas, list(which), xf, list(range(n_instances)), mats,
                    cam, 2.0 * spread)


def single_triangle():
    pos = [(-1.0, -1.0, 0.0), (1.0, -1.0, 0.0), (0.0, 1.0, 0.0)]
    b = soup_blas(pos)
    cam = Camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0))
    return assemble("triangle", [b], [0], [np.eye(4)], [0], [material((1, 1, 1))], cam, 2.0)


def empty_scene():
    cam = Camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0))
    return assemble("empty", [], [], [], [], [], cam, 1.0)


def random_soup(n_tris, seed=0, spread=1.0, size=0.2):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, size=(n_tris, 1, 3))
    p = c + rng.uniform(-size, size, size=(n_tris, 3, 3))
    b = soup_blas(p.reshape(-1, 3))
    cam = Camera((0.0, 0.0, 3.0 * spread + 1.0), (0.0, 0.0, 0.0))
    return assemble("soup_%d" % n_tris, [b], [0], [np.eye(4)], [0], [material((0.7, 0.7, 0.7))], cam,
                    2.0 * spread)


def random_instances(n_inst=50, seed=3, tris_per_blas=(30, 200, 5)):
    """several small random BLAS instanced with random affine transforms"""
    rng = np.random.default_rng(seed)
    blas = []
    for k, nt in enumerate(tris_per_blas):
        c = rng.uniform(-1, 1, size=(nt, 1, 3))
        p = c + rng.uniform(-0.3, 0.3, size=(nt, 3, 3))
        blas.append(soup_blas(p.reshape(-1, 3)))
    which = rng.integers(0, len(blas), size=n_inst)
    xf = []
    for i in range(n_inst):
        m = np.eye(4)
        m[:3, :3] = random_rotation(rng) * rng.uniform(0.3, 1.5, size=3)
        m[:3, 3] = rng.uniform(-4, 4, size=3)
        xf.append(m)
    mats = [material(tuple(rng.uniform(0.2, 0.9, size=3))) for _ in range(n_inst)]
    cam = Camera((0.0, 0.0, 12.0), (0.0, 0.0, 0.0))
    return assemble("random_instances", blas, list(which), xf, list(range(n_inst)), mats, cam, 8.0)


def blue_noise_synthetic(size=256, seed=7):
    """RGBA8 texture standing in for assets/textures/HDR_RGBA_0.png where the
    asset is not available (white noise, same shape and format)."""
    rng = np.random.default_rng(seed)
    return rng.integers(0, 256, size=(size, size, 4), dtype=np.uint8)


class MultiBlas:
    """A BLAS over several geometries (BLAS::new's Vec<GeometryInstance>,
    src/ray/acceleration.rs:136-208): each geometry is (ModelVertex array,
    u32 indices) or (ModelVertex array, None) for non-indexed triangles
    (vertices 3i..3i+2, vertex_count / 3 triangles).  Its triangles are
    numbered geometry after geometry; `vertices` / `indices` give the hit
    shaders' object view of that numbering (the geometries' vertices back to
    back, indices rebased onto them)."""

    def __init__(self, geometries):
        self.geometries = [(np.ascontiguousarray(v, MODEL_VERTEX),
                            None if i is None else np.ascontiguousarray(i, np.uint32).reshape(-1))
                           for v, i in geometries]
        verts, inds, base = [], [], 0
        for v, i in self.geometries:
            n = (len(v) // 3) * 3
            local = np.arange(n, dtype=np.uint32) if i is None else i
            verts.append(v)
            inds.append(local.astype(np.uint64) + base)
            base += len(v)
        self.vertices = np.concatenate(verts) if verts else np.zeros(0, MODEL_VERTEX)
        self.indices = (np.concatenate(inds) if inds else np.zeros(0, np.uint64)).astype(np.uint32)

    @property
    def tri_count(self):
        return len(self.indices) // 3

    @property
    def geo_begin(self):
        """first BLAS-wide triangle of each geometry"""
        out, t = [], 0
        for v, i in self.geometries:
            out.append(t)
            t += (len(v) // 3) if i is None else len(i) // 3
        return out


def multi_geometry_scene(seed=9):
    """two instances of a 3-geometry BLAS (two indexed soups, one
    non-indexed soup) plus an instance of a one-geometry BLAS, for the
    geometry index of hits (gl_GeometryIndexEXT)"""
    rng = np.random.default_rng(seed)

    def soup(n, c):
        p = c + rng.uniform(-0.6, 0.6, size=(n, 1, 3)) + rng.uniform(-0.25, 0.25, size=(n, 3, 3))
        return p.reshape(-1, 3)
    g0 = soup(40, np.array([-0.8, 0.0, 0.0]))
    g1 = soup(25, np.array([0.8, 0.2, 0.0]))
    g2 = soup(30, np.array([0.0, -0.6, 0.3]))
    # indexed geometries with shuffled, shared vertices; the non-indexed one in triangle order
    geoms = []
    for p in (g0, g1):
        uniq, inv = np.unique(np.round(p, 6), axis=0, return_inverse=True)
        perm = rng.permutation(len(uniq))
        v = make_vertices(uniq[perm])
        rank = np.empty(len(perm), np.int64)
        rank[perm] = np.arange(len(perm))
        geoms.append((v, rank[inv.reshape(-1)].astype(np.uint32)))
    geoms.append((make_vertices(g2, flat_normals(g2)), None))
    mb = MultiBlas(geoms)
    single = soup_blas(soup(20, np.array([0.0, 0.9, -0.4])))
    xf = [np.eye(4), np.eye(4), np.eye(4)]
    xf[1][:3, 3] = (0.3, -0.2, -1.5)
    xf[1][:3, :3] = random_rotation(rng)
    xf[2][:3, 3] = (0.0, 0.0, 0.2)
    cam = Camera((0.0, 0.0, 4.0), (0.0, 0.0, 0.0))
    return assemble("multi_geometry", [mb, single], [0, 0, 1], xf, [0, 1, 2],
                    [material((0.8, 0.7, 0.6)), material((0.5, 0.6, 0.7)), material((0.3, 0.8, 0.4))], cam, 3.0)
