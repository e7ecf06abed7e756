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
