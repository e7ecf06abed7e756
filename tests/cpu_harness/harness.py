"""ctypes wrapper of libbvhcpu.so (host replay of the library's BVH build and
traversal)."""
import ctypes as C
import os

import numpy as np

from solhip.records import HIT, RAY

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libbvhcpu.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from . import build as _b  # noqa
        L = C.CDLL(LIB)
        vp = C.c_void_p
        L.hc_scene_create.restype = vp
        L.hc_scene_create.argtypes = [vp, vp, C.c_uint32, vp, vp, vp, C.c_uint32, C.POINTER(C.c_int)]
        L.hc_scene_destroy.argtypes = [vp]
        L.hc_scene_info.argtypes = [vp, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.hc_trace.argtypes = [vp, vp, C.c_uint32, C.c_uint32, vp, vp]
        L.hc_trace_steps.argtypes = [vp, vp, C.c_uint32, C.c_uint32, vp, vp]
        L.hc_set_build_flags.argtypes = [C.c_uint32]
        L.hc_geometry_fetch.restype = C.c_uint32
        L.hc_geometry_fetch.argtypes = [vp, C.c_uint32, C.c_uint32, vp, vp, vp]
        _lib = L
    return _lib


class HostScene:
    def __init__(self, sd, build_flags=0):
        """build_flags: SOLHIP_BUILD_* of the BLAS builds (1: binned SAH)"""
        self._keep = []
        nb = len(sd.blas)
        ptrs = (C.c_void_p * max(1, nb))()
        counts = np.zeros(max(1, nb), np.uint32)
        for i, b in enumerate(sd.blas):
            t = np.ascontiguousarray(b.vertices["pos"][:, :3][b.indices.reshape(-1, 3)].reshape(-1), np.float32)
            self._keep.append(t)
            ptrs[i] = t.ctypes.data
            counts[i] = b.tri_count
        n = len(sd.instance_descs)
        xf = np.ascontiguousarray(sd.instance_descs["transform"].reshape(-1), np.float32) if n else np.zeros(12, np.float32)
        ib = np.ascontiguousarray(sd.instance_descs["blas"].astype(np.uint32)) if n else np.zeros(1, np.uint32)
        im = np.ascontiguousarray((sd.instance_descs["id_mask"] >> 24).astype(np.uint32)) if n else np.zeros(1, np.uint32)
        self._keep += [xf, ib, im, counts]
        err = C.c_int(0)
        lib().hc_set_build_flags(build_flags)
        self.h = lib().hc_scene_create(C.cast(ptrs, C.c_void_p), counts.ctypes.data, nb, xf.ctypes.data,
                                       ib.ctypes.data, im.ctypes.data, n, C.byref(err))
        if not self.h:
            raise RuntimeError("host BVH build failed: %d" % err.value)

    def info(self, blas):
        a, b = C.c_uint32(), C.c_uint32()
        lib().hc_scene_info(self.h, blas, C.byref(a), C.byref(b))
        return a.value, b.value

    def trace(self, rays, any_hit=False, steps=False):
        """trace_ray per ray, or (steps) the resumable traversal driven as
        64 refilling lanes"""
        rays = np.ascontiguousarray(rays, RAY)
        hits = np.zeros(len(rays), HIT)
        counts = np.zeros(4, np.uint64)
        fn = lib().hc_trace_steps if steps else lib().hc_trace
        fn(self.h, rays.ctypes.data, len(rays), 1 if any_hit else 0, hits.ctypes.data, counts.ctypes.data)
        return hits, counts

    def __del__(self):
        if getattr(self, "h", None):
            lib().hc_scene_destroy(self.h)
            self.h = None


GEOM_DESC = np.dtype([("vbase", "<u8"), ("idx", "<u8"), ("stride", "<u8"), ("first_vertex", "<u4"),
                      ("vertex_count", "<u4"), ("tri_begin", "<u4"), ("tri_count", "<u4")])


def geometry_fetch(geoms):
    """the BLAS build's vertex fetch over geometries given as dicts
    (vertices: ModelVertex array, indices: u32 array or None, first_vertex,
    vertex_count, primitive_offset in bytes): (positions (n, 3, 3), geometry
    index and in-geometry index per triangle, out-of-range references)"""
    keep = []
    d = np.zeros(len(geoms), GEOM_DESC)
    t = 0
    for k, g in enumerate(geoms):
        v = np.ascontiguousarray(g["vertices"])
        keep.append(v)
        vc = g.get("vertex_count", len(v))
        d[k]["stride"] = v.dtype.itemsize
        d[k]["first_vertex"] = g.get("first_vertex", 0)
        d[k]["vertex_count"] = vc
        d[k]["tri_begin"] = t
        if g.get("indices") is not None:
            i = np.ascontiguousarray(g["indices"], np.uint32)
            keep.append(i)
            d[k]["vbase"] = v.ctypes.data
            d[k]["idx"] = i.ctypes.data
            n = len(i) // 3
        else:
            d[k]["vbase"] = v.ctypes.data + g.get("primitive_offset", 0)
            d[k]["idx"] = 0
            n = vc // 3
        d[k]["tri_count"] = n
        t += n
    pos = np.zeros((t, 3, 3), np.float32)
    geom = np.zeros(t, np.uint32)
    lprim = np.zeros(t, np.uint32)
    bad = lib().hc_geometry_fetch(d.ctypes.data, len(geoms), t, pos.ctypes.data, geom.ctypes.data, lprim.ctypes.data)
    return pos, geom, lprim, int(bad)
