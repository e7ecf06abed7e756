"""Host side of the C++ mirror of the sol crate (sol-rs_amd/host): glTF scene
decoding (src/scene/mod.rs:106-295) and PNG texture decoding
(src/texture.rs:487-496), checked through the gltf_info program against an
independent numpy decode of the same files.  No device is used."""
import json
import os
import subprocess

import numpy as np
import pytest

import gltf_py

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "sol-rs_amd", "bin")
GLTF_INFO = os.path.join(BIN, "gltf_info")
CORNELL = os.path.join(ROOT, "tests", "golden", "assets", "cornell.gltf")
REF_ASSETS = "/root/reference/assets"

pytestmark = pytest.mark.skipif(not os.path.exists(GLTF_INFO),
                                reason="host mirror not built (sol-rs_amd/build.py build_host)")


def _info(path, tmp_path):
    out = subprocess.run([GLTF_INFO, path, "--arrays", str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    return json.loads(out.stdout)


def _check_scene(path, tmp_path):
    info = _info(path, tmp_path)
    ref = gltf_py.decode(path)
    assert len(info["meshes"]) == len(ref["meshes"])
    tris = 0
    for m, (mi, mr) in enumerate(zip(info["meshes"], ref["meshes"])):
        v = np.fromfile(tmp_path / ("mesh_%d_vertices.bin" % m), np.float32).reshape(-1, 16)
        ix = np.fromfile(tmp_path / ("mesh_%d_indices.bin" % m), np.uint32)
        assert v.shape == mr["vertices"].shape
        # positions / normals / uvs are copied from f32 accessors: exact
        np.testing.assert_array_equal(v.astype(np.float64), mr["vertices"])
        np.testing.assert_array_equal(ix, mr["indices"])
        assert [tuple(s) for s in mi["sections"]] == [tuple(s) for s in mr["sections"]]
        t = np.array(mi["transform"]).reshape(4, 4).T
        np.testing.assert_allclose(t, mr["transform"], rtol=1e-6, atol=1e-6)
        tris += sum(s[3] for s in mr["sections"]) // 3
    assert info["triangles"] == tris
    mats = np.fromfile(tmp_path / "materials.bin", np.float32).reshape(-1, 12)
    assert len(mats) == len(ref["materials"])
    for row, (base, emi, met, rough) in zip(mats, ref["materials"]):
        np.testing.assert_array_equal(row[0:4], np.float32(base))
        np.testing.assert_array_equal(row[4:7], np.float32(emi))
        assert row[8] == np.float32(met) and row[9] == np.float32(rough)
    if ref["camera_view"] is None:
        assert info["camera"] is None
    else:
        np.testing.assert_allclose(np.array(info["camera"]["view"]).reshape(4, 4).T, ref["camera_view"],
                                   rtol=1e-6, atol=1e-6)
    return info


def test_cornell_decode(tmp_path):
    info = _check_scene(CORNELL, tmp_path)
    # 8 meshes (5 walls, light, 2 boxes), 32 triangles (scene description of
    # examples/5-pathtrace.rs's default model)
    assert len(info["meshes"]) == 8 and info["triangles"] == 32
    assert info["camera"] is not None


@pytest.mark.parametrize("name", ["Duck.gltf", "tunnel.gltf"])
def test_reference_models_decode(name, tmp_path):
    path = os.path.join(REF_ASSETS, "models", name)
    if not os.path.exists(path):
        pytest.skip("reference assets not present")
    _check_scene(path, tmp_path)


def _png(path, tmp_path):
    out_file = tmp_path / "texels.rgba8"
    r = subprocess.run([GLTF_INFO, "-", "--png", path, "--png-out", str(out_file)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    w, h = json.loads(r.stdout)["png"]
    return np.fromfile(out_file, np.uint8).reshape(h, w, 4)


def test_png_rgba8_matches_pil(tmp_path):
    """8-bit RGBA with all scanline filters: exact against PIL, flipped."""
    PIL = pytest.importorskip("PIL.Image")
    path = os.path.join(REF_ASSETS, "textures", "face.png")
    if not os.path.exists(path):
        pytest.skip("reference assets not present")
    ours = _png(path, tmp_path)
    ref = np.asarray(PIL.open(path).convert("RGBA"))[::-1]
    np.testing.assert_array_equal(ours, ref)


def _png16_numpy(path):
    """16-bit RGBA decode from the PNG specification (zlib stream of filtered
    scanlines, filter types 0-4), vectorised over the 8 bytes of a texel"""
    import zlib
    raw = open(path, "rb").read()
    off, idat, w, h = 8, b"", 0, 0
    while off < len(raw):
        n = int.from_bytes(raw[off:off + 4], "big")
        t = raw[off + 4:off + 8]
        if t == b"IHDR":
            w, h = int.from_bytes(raw[off + 8:off + 12], "big"), int.from_bytes(raw[off + 12:off + 16], "big")
            assert raw[off + 16] == 16 and raw[off + 17] == 6
        elif t == b"IDAT":
            idat += raw[off + 8:off + 8 + n]
        off += 12 + n
    data = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + w * 8)
    img = np.zeros((h, w, 8), np.int32)
    prev = np.zeros((w, 8), np.int32)
    for y in range(h):
        f, row = data[y, 0], data[y, 1:].reshape(w, 8).astype(np.int32)
        out = np.zeros((w, 8), np.int32)
        if f == 0:
            out = row
        elif f == 2:
            out = (row + prev) & 255
        else:
            a = np.zeros(8, np.int32)
            c = np.zeros(8, np.int32)
            for x in range(w):
                b = prev[x]
                if f == 1:
                    pr = a
                elif f == 3:
                    pr = (a + b) // 2
                else:
                    p = a + b - c
                    pa, pb, pc = np.abs(p - a), np.abs(p - b), np.abs(p - c)
                    pr = np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, c))
                out[x] = (row[x] + pr) & 255
                a, c = out[x], b
        img[y] = out
        prev = out
    return (img[:, :, 0::2] << 8) | img[:, :, 1::2]


def test_png_rgba16_blue_noise(tmp_path):
    """the 16-bit blue-noise texture: each channel becomes round(v / 257) of
    its 16-bit value v (image 0.24's u16 -> u8 conversion, v * 255 / 65535
    rounded), rows flipped.  The numpy decode is pinned to PIL through the
    high bytes."""
    PIL = pytest.importorskip("PIL.Image")
    path = os.path.join(REF_ASSETS, "textures", "HDR_RGBA_0.png")
    if not os.path.exists(path):
        pytest.skip("reference assets not present")
    ours = _png(path, tmp_path).astype(np.int64)
    v = _png16_numpy(path)[::-1].astype(np.int64)
    np.testing.assert_array_equal(v >> 8, np.asarray(PIL.open(path).convert("RGBA"))[::-1])
    np.testing.assert_array_equal(ours, np.floor(v / 257.0 + 0.5).astype(np.int64))


def test_png_rejects_garbage(tmp_path):
    bad = tmp_path / "bad.png"
    bad.write_bytes(b"\x89PNG\r\n\x1a\nnot really")
    r = subprocess.run([GLTF_INFO, "-", "--png", str(bad)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "error" in r.stderr


def test_gltf_missing_file_is_an_error(tmp_path):
    r = subprocess.run([GLTF_INFO, str(tmp_path / "none.gltf")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "load_scene" in r.stderr
