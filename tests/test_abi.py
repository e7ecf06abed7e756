"""The C-ABI library loads, exports every entry point include/solhip.h declares,
and the record layouts match the reference's byte sizes.  CPU only: no compute
call is made without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import solhip
from solhip import api, records

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "solhip.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int32_t|uint32_t|void)\s+(solhip_\w+)\s*\(", text)))


def test_every_declared_entry_point_is_exported():
    names = declared_functions()
    assert len(names) >= 25
    L = C.CDLL(api.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    assert set(declared_functions()) == set(api.SIGNATURES)


def test_record_sizes_match_the_reference():
    assert records.MODEL_VERTEX.itemsize == 64      # src/scene/mesh.rs:7-25
    assert records.MATERIAL.itemsize == 48          # src/scene/mod.rs:19-29
    assert records.SCENE_INSTANCE.itemsize == 144   # src/ray/mod.rs:16-24
    assert records.INSTANCE_DESC.itemsize == 64     # src/ray/acceleration.rs:407
    assert records.UNIFORMS.itemsize == 400         # examples/*.rs:9-17
    assert C.sizeof(api.Geometry) == 48
    assert C.sizeof(api.SceneTables) == 48
    assert C.sizeof(api.Spec) == 72
    # solhip_hit = {t, u, v, gl_InstanceID, gl_PrimitiveID, gl_GeometryIndexEXT}
    assert records.HIT.itemsize == 24 and records.HIT.names[-1] == "geom"
    # solhip_geometry: BuildRangeInfo::primitive_offset at byte 32, after index_addr
    assert api.Geometry.primitive_offset.offset == 32 and api.Geometry.index_count.offset == 40


def test_noalloc_scratch_size_matches_the_header():
    with open(os.path.join(ROOT, "include", "solhip.h")) as f:
        m = re.search(r"#define SOLHIP_NOALLOC_SCRATCH_BYTES \((\d+)ull << (\d+)\)", f.read())
    assert m and int(m.group(1)) << int(m.group(2)) == api.NOALLOC_SCRATCH_BYTES
    assert 1920 * 1080 * (32 + 32) + 256 + 8192 <= api.NOALLOC_SCRATCH_BYTES


def test_abi_version_and_spec_defaults():
    assert api.lib().solhip_abi_version() == 2
    assert api.ERRORS[-8] == "STALE"
    s = solhip.spec_default(solhip.PROG_PATHTRACE_GLSL)
    assert (s.spp_per_frame, s.max_depth, s.enable_sky) == (8, 32, 0)      # pathtrace.rgen:43-44, .rmiss:5
    s = solhip.spec_default(solhip.PROG_AO_GLSL)
    assert (s.spp_per_frame, s.max_depth) == (4, 4)                          # ao.rgen:42-43


def _has_gpu():
    try:
        c = solhip.Context(0)
        c.close()
        return True
    except solhip.SolhipError:
        return False


def test_no_cpu_fallback_without_a_device():
    if _has_gpu():
        pytest.skip("a device is present")
    with pytest.raises(solhip.SolhipError) as e:
        solhip.Context(0)
    assert e.value.code == -4 and "no CPU fallback" in str(e.value)


def test_invalid_context_handle_is_rejected():
    h = C.c_uint64()
    rc = api.lib().solhip_buffer_create(0xDEAD00, 16, C.byref(h))
    assert rc == -2 and "handle" in api.last_error()


def test_instance_descriptor_bitfields():
    sd = solhip.scenes.cornell_box()
    d = sd.instance_descs
    assert (d["id_mask"] >> 24 == 0xFF).all()                         # mask 0xFF
    assert ((d["id_mask"] & 0xFFFFFF) == np.arange(len(d))).all()      # custom id = i
    assert (d["offset_flags"] >> 24 == 0x5).all()                      # FORCE_OPAQUE | CULL_DISABLE
