"""The headless examples (sol-rs_amd/examples: 4-ray-ao and 5-pathtrace over
the C++ mirror of the sol crate) on the device, checked against the oracle.

Each program loads the Cornell glTF through the mirror (load_scene ->
SceneDescription::from_scene -> Pipeline -> ShaderBindingTable ->
cmd_trace_rays, examples/{4-ray-ao,5-pathtrace}.rs), traces a few
accumulated frames and dumps the exact arrays it handed to the ray path.  The
oracle renders the same frames from those arrays; the accumulation targets
must agree bit for bit."""
import json
import os
import subprocess

import numpy as np
import pytest

import solhip
from solhip import scenes
from solhip.records import INSTANCE_DESC, MATERIAL, MODEL_VERTEX, SCENE_INSTANCE, UNIFORMS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "sol-rs_amd", "bin")
CORNELL = os.path.join(ROOT, "tests", "golden", "assets", "cornell.gltf")

pytestmark = pytest.mark.gpu


def _run(prog, args):
    exe = os.path.join(BIN, prog)
    if not os.path.exists(exe):
        pytest.fail("%s not built (sol-rs_amd/build.py build_host)" % exe)
    print("example:", prog, " ".join(args), flush=True)
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def _scene_from_dump(d):
    """SceneData over the dumped arrays (objects[k] = BLAS k's section,
    materials[k] = its material)"""
    nblas = int(open(os.path.join(d, "manifest.txt")).read().split()[1])
    blas = []
    for k in range(nblas):
        v = np.fromfile(os.path.join(d, "blas_%d_vertices.bin" % k), MODEL_VERTEX)
        ix = np.fromfile(os.path.join(d, "blas_%d_indices.bin" % k), np.uint32)
        blas.append(scenes.Blas(v, ix))
    descs = np.fromfile(os.path.join(d, "instances.bin"), INSTANCE_DESC)
    sis = np.fromfile(os.path.join(d, "scene_instances.bin"), SCENE_INSTANCE)
    mats = np.fromfile(os.path.join(d, "materials.bin"), MATERIAL)
    objects = [(b.vertices, b.indices.astype(np.uint64)) for b in blas]
    return scenes.SceneData("dump", blas, descs, sis, objects, mats, np.arange(nblas), None, 1.0)


def _uniforms(d, f):
    return np.fromfile(os.path.join(d, "uniforms_%d.bin" % f), UNIFORMS)


@pytest.mark.parametrize("sky", [0, 1])
def test_pathtrace_example_matches_oracle(oracle, tmp_path, sky):
    W, H, F = 64, 48, 3
    d = str(tmp_path / "dump")
    acc_file = str(tmp_path / "accum.f32")
    args = ["--model", CORNELL, "--size", "%dx%d" % (W, H), "--frames", str(F), "--dump", d,
            "--accum", acc_file, "--out", str(tmp_path / "pt.ppm")]
    res = _run("pathtrace", args + (["--sky"] if sky else []))
    assert res["frames"] == F and res["overflow"] == 0
    sd = _scene_from_dump(d)
    assert sd.tri_count == 32 and len(sd.instance_descs) == 8
    spec = solhip.spec_default(solhip.PROG_PATHTRACE_GLSL)
    sc = oracle.Scene(sd)
    ref = np.zeros((H, W, 4), np.float32)
    ref8 = np.zeros((H, W, 4), np.uint8)
    rays = 0
    for f in range(F):
        rays += sc.render_pathtrace_glsl(_uniforms(d, f), oracle.PTParams(sky, spec.spp_per_frame, spec.max_depth),
                                         0, ref, ref8, W, H)
    gpu = np.fromfile(acc_file, np.float32).reshape(H, W, 4)
    assert np.count_nonzero(gpu.view(np.uint32) != ref.view(np.uint32)) == 0
    assert res["rays"] == rays
    assert ref[..., :3].max() > 0


def test_ray_ao_example_matches_oracle(oracle, tmp_path):
    W, H, F = 60, 40, 3
    d = str(tmp_path / "dump")
    acc_file = str(tmp_path / "ao.f32")
    res = _run("ray_ao", ["--model", CORNELL, "--size", "%dx%d" % (W, H), "--frames", str(F), "--dump", d,
                          "--accum", acc_file, "--out", str(tmp_path / "ao.ppm")])
    assert res["frames"] == F and res["overflow"] == 0
    sd = _scene_from_dump(d)
    blue = np.fromfile(os.path.join(d, "blue_noise.bin"), np.uint8).reshape(256, 256, 4)
    sc = oracle.Scene(sd)
    ref = np.zeros((H, W, 4), np.float32)
    for f in range(F):
        sc.render_ao_glsl(_uniforms(d, f), 0, ref, blue, W, H)
    gpu = np.fromfile(acc_file, np.float32).reshape(H, W, 4)
    assert np.count_nonzero(gpu.view(np.uint32) != ref.view(np.uint32)) == 0
    # the camera of examples/4-ray-ao.rs looks at the box from outside: some
    # pixels hit geometry (AO in [0, 1]) and some miss (cleared)
    assert 0.0 < ref[..., 0].max() <= 1.0


@pytest.mark.parametrize("streams", [0, 1])
def test_pathtrace_example_wavefront_nee_matches_oracle(oracle, tmp_path, streams):
    W, H, F = 48, 40, 2
    d = str(tmp_path / "dump")
    acc_file = str(tmp_path / "accum.f32")
    res = _run("pathtrace", ["--model", CORNELL, "--size", "%dx%d" % (W, H), "--frames", str(F), "--dump", d,
                             "--accum", acc_file, "--out", str(tmp_path / "pt.ppm"), "--wavefront", "--nee",
                             "--rr", "2", "--spp", "3", "--depth", "4"] + (["--streams"] if streams else []))
    assert res["frames"] == F and res["overflow"] == 0
    sd = _scene_from_dump(d)
    sc = oracle.Scene(sd)
    ref = np.zeros((H, W, 4), np.float32)
    rays = 0
    for f in range(F):
        rays += sc.render_pathtrace(_uniforms(d, f), oracle.PTConfig(0, 3, 4, 2, 1, streams), 0, ref, None, W, H)
    gpu = np.fromfile(acc_file, np.float32).reshape(H, W, 4)
    assert np.count_nonzero(gpu.view(np.uint32) != ref.view(np.uint32)) == 0
    assert res["rays"] == rays


def _write_tunnel_gltf(path):
    """a two-primitive mesh (the shape of assets/models/tunnel.gltf, whose
    mesh holds two primitives): an arched vault and a floor, with their own
    materials, and a camera node looking down the tunnel; one data-URI
    buffer.  SceneDescription::from_scene makes one BLAS per primitive
    (src/ray/mod.rs:59-156)."""
    import base64
    segs, rings, L = 16, 12, 3.0
    pos, nrm, idx = [], [], []
    for r in range(rings + 1):
        z = -L + 2 * L * r / rings
        for s in range(segs + 1):
            a = np.pi * s / segs
            pos.append((np.cos(a), np.sin(a), z))
            nrm.append((-np.cos(a), -np.sin(a), 0.0))
    for r in range(rings):
        for s in range(segs):
            i0 = r * (segs + 1) + s
            i1, i2, i3 = i0 + 1, i0 + segs + 1, i0 + segs + 2
            idx += [i0, i2, i1, i1, i2, i3]
    vault = (np.asarray(pos, np.float32), np.asarray(nrm, np.float32), np.asarray(idx, np.uint32))
    fp = np.asarray([(-1, 0, -L), (1, 0, -L), (-1, 0, L), (1, 0, L)], np.float32)
    floor = (fp, np.tile(np.asarray([0, 1, 0], np.float32), (4, 1)), np.asarray([0, 2, 1, 1, 2, 3], np.uint32))
    blob, views, accs, prims = b"", [], [], []
    for k, (p, n, ix) in enumerate((vault, floor)):
        attrs = {}
        for name, arr, typ in (("POSITION", p, "VEC3"), ("NORMAL", n, "VEC3"), (None, ix, "SCALAR")):
            views.append({"buffer": 0, "byteOffset": len(blob), "byteLength": arr.nbytes})
            acc = {"bufferView": len(views) - 1, "componentType": 5125 if name is None else 5126,
                   "count": len(arr), "type": typ}
            if name == "POSITION":
                acc["min"] = [float(v) for v in p.min(0)]
                acc["max"] = [float(v) for v in p.max(0)]
            accs.append(acc)
            blob += arr.tobytes()
            if name:
                attrs[name] = len(accs) - 1
        prims.append({"attributes": attrs, "indices": len(accs) - 1, "material": k})
    doc = {
        "asset": {"version": "2.0"},
        "buffers": [{"byteLength": len(blob),
                     "uri": "data:application/octet-stream;base64," + base64.b64encode(blob).decode()}],
        "bufferViews": views, "accessors": accs,
        "materials": [
            {"pbrMetallicRoughness": {"baseColorFactor": [0.8, 0.7, 0.6, 1.0], "metallicFactor": 0.0,
                                      "roughnessFactor": 1.0}},
            {"pbrMetallicRoughness": {"baseColorFactor": [0.3, 0.5, 0.8, 1.0], "metallicFactor": 0.5,
                                      "roughnessFactor": 0.3}}],
        "meshes": [{"name": "tunnel", "primitives": prims}],
        "cameras": [{"type": "perspective", "perspective": {"aspectRatio": 1.33, "yfov": 45.0, "znear": 0.01,
                                                           "zfar": 100.0}}],
        # the node matrix is taken as the view matrix (src/scene/camera.rs:73-94)
        "nodes": [{"mesh": 0}, {"camera": 0, "matrix": [-1, 0, 0, 0, 0, -1, 0, 0, 0, 0, 1, 0, 0, 0.5, -5.0, 1]}],
        "scenes": [{"nodes": [0, 1]}], "scene": 0,
    }
    with open(path, "w") as f:
        json.dump(doc, f)


def test_tunnel_two_primitives_pathtrace_example_matches_oracle(oracle, tmp_path):
    """the two-primitive mesh through load_scene -> from_scene (one BLAS and
    one instance per primitive) -> the pathtrace.* program with the sky on"""
    model = str(tmp_path / "tunnel.gltf")
    _write_tunnel_gltf(model)
    W, H, F = 64, 48, 2
    d = str(tmp_path / "dump")
    acc_file = str(tmp_path / "accum.f32")
    res = _run("pathtrace", ["--model", model, "--size", "%dx%d" % (W, H), "--frames", str(F), "--dump", d,
                             "--accum", acc_file, "--out", str(tmp_path / "pt.ppm"), "--sky"])
    assert res["frames"] == F and res["overflow"] == 0
    sd = _scene_from_dump(d)
    assert len(sd.blas) == 2 and [b.tri_count for b in sd.blas] == [16 * 12 * 2, 2]
    spec = solhip.spec_default(solhip.PROG_PATHTRACE_GLSL)
    sc = oracle.Scene(sd)
    ref = np.zeros((H, W, 4), np.float32)
    ref8 = np.zeros((H, W, 4), np.uint8)
    rays = 0
    for f in range(F):
        rays += sc.render_pathtrace_glsl(_uniforms(d, f), oracle.PTParams(1, spec.spp_per_frame, spec.max_depth),
                                         0, ref, ref8, W, H)
    gpu = np.fromfile(acc_file, np.float32).reshape(H, W, 4)
    assert np.count_nonzero(gpu.view(np.uint32) != ref.view(np.uint32)) == 0
    assert res["rays"] == rays
    # both primitives are seen: the pixel-centre rays of the dumped camera hit both BLAS
    from solhip.records import RAY
    u = _uniforms(d, 0)
    r = np.zeros(W * H, RAY)
    for k in range(W * H):
        r[k]["o"], r[k]["d"] = oracle.camera_ray(u, k % W, k // W, W, H, 0.5, 0.5)
    r["tmin"], r["tmax"] = 1e-3, 1e4
    hits = oracle.Scene(sd).trace(r)
    assert set(np.unique(hits["inst"])) >= {0, 1}


@pytest.mark.parametrize("model", ["cornell", "tunnel"])
def test_ray_debug_example_matches_oracle(oracle, tmp_path, model):
    """examples/3-ray-debug.rs over the mirror: debug.* program, the
    example's camera (Camera::new(size), look_at((5,5,5), 0, -Y))"""
    path = CORNELL
    if model == "tunnel":
        path = str(tmp_path / "tunnel.gltf")
        _write_tunnel_gltf(path)
    W, H = 50, 40
    d = str(tmp_path / "dump")
    tgt = str(tmp_path / "debug.f32")
    res = _run("ray_debug", ["--model", path, "--size", "%dx%d" % (W, H), "--frames", "1", "--dump", d,
                             "--target", tgt, "--out", str(tmp_path / "debug.ppm")])
    assert res["frames"] == 1 and res["overflow"] == 0
    sd = _scene_from_dump(d)
    ref = oracle.Scene(sd).render_debug(_uniforms(d, 0), W, H)
    gpu = np.fromfile(tgt, np.float32).reshape(H, W, 4)
    assert np.count_nonzero(gpu.view(np.uint32) != ref.view(np.uint32)) == 0
