"""ctypes binding of libsolhip.so (include/solhip.h).

There is no fallback: if the library is missing or no gfx950 device is
present, every entry point raises.
"""
import ctypes as C
import os

import numpy as np

from .records import HIT, INSTANCE_DESC, MATERIAL, RAY, SCENE_INSTANCE, UNIFORMS

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SOLHIP_LIB") or os.path.normpath(os.path.join(_HERE, "..", "..", "lib", "libsolhip.so"))
HEADER_PATH = os.path.normpath(os.path.join(_HERE, "..", "..", "..", "include", "solhip.h"))

PROG_AO_GLSL, PROG_PATHTRACE_GLSL, PROG_DEBUG, PROG_AO, PROG_PATHTRACE = 1, 2, 3, 4, 5
TRACE_TERMINATE_ON_FIRST_HIT = 0x4
TRACE_COUNT = 0x100
BUILD_PREFER_FAST_TRACE = 0x1
# size of solhip_trace_noalloc's fixed device array (SOLHIP_NOALLOC_SCRATCH_BYTES)
NOALLOC_SCRATCH_BYTES = 128 << 20

ERRORS = {0: "OK", -1: "INVALID_ARGUMENT", -2: "INVALID_HANDLE", -3: "OUT_OF_MEMORY", -4: "NO_DEVICE",
          -5: "DEVICE", -6: "UNSUPPORTED", -7: "OVERFLOW", -8: "STALE"}


class SolhipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


class Geometry(C.Structure):
    _fields_ = [("vertex_addr", C.c_uint64), ("vertex_stride", C.c_uint64), ("first_vertex", C.c_uint32),
                ("vertex_count", C.c_uint32), ("index_addr", C.c_uint64), ("primitive_offset", C.c_uint64),
                ("index_count", C.c_uint32), ("opaque", C.c_uint32)]


class SceneTables(C.Structure):
    _fields_ = [("instances_addr", C.c_uint64), ("instance_count", C.c_uint32), ("index_size", C.c_uint32),
                ("vertex_table_addr", C.c_uint64), ("index_table_addr", C.c_uint64),
                ("material_table_addr", C.c_uint64), ("object_count", C.c_uint32), ("reserved", C.c_uint32)]


class Spec(C.Structure):
    _fields_ = [("enable_sky", C.c_uint32), ("spp_per_frame", C.c_uint32), ("max_depth", C.c_uint32),
                ("n_ao", C.c_uint32), ("ao_tmin", C.c_float), ("ao_tmax", C.c_float), ("ao_any_hit", C.c_uint32),
                ("rr_start", C.c_uint32), ("nee", C.c_uint32), ("wf_blocks_per_cu", C.c_uint32),
                ("wf_block", C.c_uint32), ("wf_graph", C.c_uint32), ("nrc", C.c_uint32),
                ("nrc_train_steps", C.c_uint32), ("nrc_learning_rate", C.c_float), ("wf_slices", C.c_uint32),
                ("sample_streams", C.c_uint32), ("wf_samples_per_pass", C.c_uint32)]


class Images(C.Structure):
    _fields_ = [("accum", C.c_uint64), ("out_rgba8", C.c_uint64), ("blue_noise", C.c_uint64),
                ("blue_w", C.c_uint32), ("blue_h", C.c_uint32), ("hits_out", C.c_uint64)]


class Counters(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("node_fetches", C.c_uint64), ("tri_fetches", C.c_uint64),
                ("inst_fetches", C.c_uint64), ("overflow", C.c_uint64)]


class KernelStats(C.Structure):
    _fields_ = [("launches", C.c_uint64), ("total_ms", C.c_double), ("counters", Counters)]


class JobDesc(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("W", C.c_uint32), ("H", C.c_uint32), ("spec", Spec),
                ("uniforms", C.c_uint8 * 400), ("accum_start_frame", C.c_int32), ("reserved", C.c_uint32),
                ("accum", C.c_uint64), ("out_rgba8", C.c_uint64)]


class JobStatus(C.Structure):
    _fields_ = [("state", C.c_uint32), ("reserved", C.c_uint32), ("submit_ms", C.c_double),
                ("admit_ms", C.c_double), ("done_ms", C.c_double)]


KERNELS = {"trace_rays": 0, "ao_primary": 1, "ao_rays": 2, "ao_glsl": 3, "pathtrace_glsl": 4, "debug": 5,
           "pt_generate": 6, "pt_extend": 7, "pt_shade": 8, "pt_shadow": 9, "pt_finalize": 10, "pt_lights": 11,
           "srv_admit": 12, "srv_extend": 13, "srv_shade": 14, "srv_shadow": 15, "srv_retire": 16}

JOB_QUEUED, JOB_RUNNING, JOB_DONE = 0, 1, 2
PROFILE_TIME, PROFILE_COUNT = 0x1, 0x2


class BvhInfo(C.Structure):
    _fields_ = [("nodes", C.c_uint32), ("prims", C.c_uint32), ("depth", C.c_uint32), ("leaf_max", C.c_uint32),
                ("bytes", C.c_uint64)]


# every entry point of include/solhip.h with its ctypes signature
_u32, _u64, _i32, _vp = C.c_uint32, C.c_uint64, C.c_int32, C.c_void_p
SIGNATURES = {
    "solhip_abi_version": (_u32, []),
    "solhip_last_error": (_i32, [C.c_char_p, C.c_size_t]),
    "solhip_context_create": (_i32, [_i32, _u32, C.POINTER(_u64)]),
    "solhip_context_destroy": (_i32, [_u64]),
    "solhip_stream_create": (_i32, [_u64, C.POINTER(_u64)]),
    "solhip_stream_destroy": (_i32, [_u64, _u64]),
    "solhip_stream_sync": (_i32, [_u64, _u64]),
    "solhip_buffer_create": (_i32, [_u64, _u64, C.POINTER(_u64)]),
    "solhip_buffer_destroy": (_i32, [_u64, _u64]),
    "solhip_buffer_upload": (_i32, [_u64, _u64, _u64, _vp, _u64]),
    "solhip_buffer_download": (_i32, [_u64, _u64, _u64, _vp, _u64]),
    "solhip_buffer_fill": (_i32, [_u64, _u64, _u64, _u32]),
    "solhip_buffer_address": (_i32, [_u64, _u64, C.POINTER(_u64)]),
    "solhip_blas_build": (_i32, [_u64, _u64, C.POINTER(Geometry), _u32, _u32, C.POINTER(_u64)]),
    "solhip_blas_rebuild": (_i32, [_u64, _u64, _u64, C.POINTER(Geometry), _u32]),
    "solhip_blas_destroy": (_i32, [_u64, _u64]),
    "solhip_tlas_build": (_i32, [_u64, _u64, _vp, _u32, C.POINTER(_u64)]),
    "solhip_tlas_rebuild": (_i32, [_u64, _u64, _u64, _vp, _u32]),
    "solhip_tlas_destroy": (_i32, [_u64, _u64]),
    "solhip_blas_info": (_i32, [_u64, _u64, C.POINTER(BvhInfo)]),
    "solhip_tlas_info": (_i32, [_u64, _u64, C.POINTER(BvhInfo)]),
    "solhip_trace_rays": (_i32, [_u64, _u64, _u64, _u64, _u64, _u32, _u32]),
    "solhip_spec_default": (None, [_u32, C.POINTER(Spec)]),
    "solhip_program_create": (_i32, [_u64, _u32, C.POINTER(Spec), C.POINTER(_u64)]),
    "solhip_program_destroy": (_i32, [_u64, _u64]),
    "solhip_trace": (_i32, [_u64, _u64, _u64, _u64, C.POINTER(SceneTables), _vp, _i32, C.POINTER(Images), _u32, _u32]),
    "solhip_trace_noalloc": (_i32, [_u64, _u64, _u64, _u64, C.POINTER(SceneTables), _vp, _i32, C.POINTER(Images), _u32,
                                    _u32]),
    "solhip_stats": (_i32, [_u64, C.POINTER(Counters), _u32]),
    "solhip_profile": (_i32, [_u64, _u32]),
    "solhip_kernel_stats_get": (_i32, [_u64, _u32, C.POINTER(KernelStats), _u32]),
    "solhip_pathtrace_autotune": (_i32, [_u64, _u64, _u64, C.POINTER(SceneTables), _vp, C.POINTER(Images), _u32, _u32,
                                         C.POINTER(Spec), _u32, C.c_char_p, C.POINTER(Spec),
                                         C.POINTER(C.c_float)]),
    "solhip_server_create": (_i32, [_u64, _u64, _u32, C.POINTER(_u64)]),
    "solhip_server_destroy": (_i32, [_u64, _u64]),
    "solhip_server_scene": (_i32, [_u64, _u64, _u64, C.POINTER(SceneTables), C.POINTER(_u32)]),
    "solhip_server_submit": (_i32, [_u64, _u64, C.POINTER(JobDesc), C.POINTER(_u64)]),
    "solhip_server_run": (_i32, [_u64, _u64, _u32, C.POINTER(_u32)]),
    "solhip_server_job": (_i32, [_u64, _u64, _u64, C.POINTER(JobStatus)]),
    "solhip_mlp_create": (_i32, [_u64, _u32, C.c_float, C.POINTER(_u64)]),
    "solhip_mlp_destroy": (_i32, [_u64, _u64]),
    "solhip_mlp_params": (_i32, [_u64, _u64, _vp, _vp]),
    "solhip_mlp_forward": (_i32, [_u64, _u64, _u64, _u64, _u32, _u64]),
    "solhip_mlp_train": (_i32, [_u64, _u64, _u64, _u64, _u64, _u32, _u32, C.POINTER(C.c_float)]),
    "solhip_mlp_gradient": (_i32, [_u64, _u64, _u64, _u64, _u64, _u32, _vp, C.POINTER(C.c_float)]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libsolhip.so not built (%s): run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    buf = C.create_string_buffer(1024)
    lib().solhip_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc):
    if rc != 0:
        raise SolhipError(rc, last_error())
    return rc


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def spec_default(kind):
    s = Spec()
    lib().solhip_spec_default(kind, C.byref(s))
    return s


class Buffer:
    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        h = C.c_uint64()
        check(lib().solhip_buffer_create(ctx.h, self.nbytes, C.byref(h)))
        self.h = h.value
        a = C.c_uint64()
        check(lib().solhip_buffer_address(ctx.h, self.h, C.byref(a)))
        self.addr = a.value

    def upload(self, arr, offset=0):
        arr = np.ascontiguousarray(arr)
        check(lib().solhip_buffer_upload(self.ctx.h, self.h, offset, _ptr(arr), arr.nbytes))

    def download(self, dtype, count=None, offset=0):
        dtype = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset) // dtype.itemsize
        out = np.empty(count, dtype)
        check(lib().solhip_buffer_download(self.ctx.h, self.h, offset, _ptr(out), out.nbytes))
        return out

    def fill(self, value=0, stream=0):
        check(lib().solhip_buffer_fill(self.ctx.h, stream, self.h, value))

    def destroy(self):
        if self.h:
            check(lib().solhip_buffer_destroy(self.ctx.h, self.h))
            self.h = 0


class Context:
    def __init__(self, device=0):
        h = C.c_uint64()
        check(lib().solhip_context_create(device, 0, C.byref(h)))
        self.h = h.value

    def close(self):
        if self.h:
            check(lib().solhip_context_destroy(self.h))
            self.h = 0

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def buffer(self, nbytes):
        return Buffer(self, nbytes)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        b = Buffer(self, max(arr.nbytes, 16))
        if arr.nbytes:
            b.upload(arr)
        return b

    def sync(self, stream=0):
        check(lib().solhip_stream_sync(self.h, stream))

    def stream(self):
        h = C.c_uint64()
        check(lib().solhip_stream_create(self.h, C.byref(h)))
        return h.value

    @staticmethod
    def geometry(vbuf, stride, vertex_count, ibuf=None, index_count=0, first_vertex=0, primitive_offset=0):
        """one solhip_geometry: indexed when ibuf is given, else non-indexed
        (vertex_count / 3 triangles)"""
        return Geometry(vbuf.addr, stride, first_vertex, vertex_count, ibuf.addr if ibuf is not None else 0,
                        primitive_offset, index_count if ibuf is not None else 0, 1)

    def blas_build_geometries(self, geoms, stream=0, flags=0):
        """BLAS over several geometries (BLAS::new's Vec<GeometryInstance>)"""
        arr = (Geometry * len(geoms))(*geoms)
        h = C.c_uint64()
        check(lib().solhip_blas_build(self.h, stream, arr, len(geoms), flags, C.byref(h)))
        return h.value

    def blas_rebuild_geometries(self, blas, geoms, stream=0):
        arr = (Geometry * len(geoms))(*geoms)
        check(lib().solhip_blas_rebuild(self.h, stream, blas, arr, len(geoms)))

    def blas_build(self, vbuf, stride, vertex_count, ibuf, index_count, first_vertex=0, primitive_offset=0,
                   stream=0, flags=0):
        """flags: BUILD_PREFER_FAST_TRACE (binned SAH) or 0 (Morton order)"""
        g = self.geometry(vbuf, stride, vertex_count, ibuf, index_count, first_vertex, primitive_offset)
        return self.blas_build_geometries([g], stream, flags)

    def blas_rebuild(self, blas, vbuf, stride, vertex_count, ibuf, index_count, stream=0):
        self.blas_rebuild_geometries(blas, [self.geometry(vbuf, stride, vertex_count, ibuf, index_count)], stream)

    def blas_destroy(self, blas):
        check(lib().solhip_blas_destroy(self.h, blas))

    def tlas_destroy(self, tlas):
        check(lib().solhip_tlas_destroy(self.h, tlas))

    def tlas_build(self, instances, stream=0):
        inst = np.ascontiguousarray(instances, INSTANCE_DESC)
        h = C.c_uint64()
        check(lib().solhip_tlas_build(self.h, stream, _ptr(inst) if len(inst) else None, len(inst), C.byref(h)))
        return h.value

    def tlas_rebuild(self, tlas, instances, stream=0):
        inst = np.ascontiguousarray(instances, INSTANCE_DESC)
        check(lib().solhip_tlas_rebuild(self.h, stream, tlas, _ptr(inst) if len(inst) else None, len(inst)))

    def blas_info(self, blas):
        i = BvhInfo()
        check(lib().solhip_blas_info(self.h, blas, C.byref(i)))
        return i

    def tlas_info(self, tlas):
        i = BvhInfo()
        check(lib().solhip_tlas_info(self.h, tlas, C.byref(i)))
        return i

    def trace_rays_dev(self, tlas, ray_buf, hit_buf, n, flags=0, stream=0):
        check(lib().solhip_trace_rays(self.h, stream, tlas, ray_buf.h, hit_buf.h, n, flags))

    def trace_rays(self, tlas, rays, flags=0):
        rays = np.ascontiguousarray(rays, RAY)
        rb = self.upload(rays)
        hb = self.buffer(max(16, len(rays) * HIT.itemsize))
        try:
            self.trace_rays_dev(tlas, rb, hb, len(rays), flags)
            self.sync()
            return hb.download(HIT, len(rays))
        finally:
            rb.destroy()
            hb.destroy()

    def program(self, kind, spec=None):
        h = C.c_uint64()
        check(lib().solhip_program_create(self.h, kind, C.byref(spec) if spec is not None else None, C.byref(h)))
        return h.value

    def program_destroy(self, p):
        check(lib().solhip_program_destroy(self.h, p))

    def trace(self, prog, tlas, tables, uniforms, accum_start, images, W, H, stream=0):
        u = np.ascontiguousarray(uniforms, UNIFORMS)
        check(lib().solhip_trace(self.h, stream, prog, tlas, C.byref(tables) if tables is not None else None,
                                 _ptr(u), accum_start, C.byref(images), W, H))

    def trace_noalloc(self, prog, tlas, tables, uniforms, accum_start, images, W, H, stream=0):
        """solhip_trace with the AO scratch in the library's fixed device array
        (no device allocation; one array per GPU, shared by all calls)"""
        u = np.ascontiguousarray(uniforms, UNIFORMS)
        check(lib().solhip_trace_noalloc(self.h, stream, prog, tlas, C.byref(tables) if tables is not None else None,
                                         _ptr(u), accum_start, C.byref(images), W, H))

    def stats(self, reset=False):
        c = Counters()
        check(lib().solhip_stats(self.h, C.byref(c), 1 if reset else 0))
        return c

    def profile(self, flags):
        check(lib().solhip_profile(self.h, flags))

    def autotune_pathtrace(self, tlas, tables, uniforms, images, W, H, spec, frames=3, cache_path=None, stream=0):
        """fastest wavefront launch shape for this workload: (spec, ms per frame)"""
        u = np.ascontiguousarray(uniforms, UNIFORMS)
        best = Spec()
        ms = C.c_float(0.0)
        path = cache_path.encode() if cache_path else None
        check(lib().solhip_pathtrace_autotune(self.h, stream, tlas, C.byref(tables), _ptr(u), C.byref(images), W, H,
                                              C.byref(spec), frames, path, C.byref(best), C.byref(ms)))
        return best, float(ms.value)

    def kernel_stats(self, kernel, reset=False):
        k = KERNELS[kernel] if isinstance(kernel, str) else kernel
        st = KernelStats()
        check(lib().solhip_kernel_stats_get(self.h, k, C.byref(st), 1 if reset else 0))
        return st


class GpuScene:
    """Uploads a SceneData (sol-rs_amd/python/solhip/scenes.py) and builds its
    BLASes and TLAS: what SceneDescription::from_scene does for the Vulkan
    path (src/ray/mod.rs:59-156).  build_flags: the BLAS build flags
    (BUILD_PREFER_FAST_TRACE: binned SAH, as BLAS::new asks of Vulkan; 0:
    Morton order, the fast build for geometry rebuilt every frame)."""

    def __init__(self, ctx, sd, index_size=4, build_flags=0):
        self.ctx = ctx
        self.sd = sd
        self.bufs = []
        self.blas = []
        self.blas_geoms = []   # per BLAS: its solhip_geometry records (kept for rebuilds)
        for b in sd.blas:
            # a scenes.Blas is one indexed geometry; a scenes.MultiBlas lists
            # (vertices, indices or None) per geometry
            geoms = []
            for verts, inds in (getattr(b, "geometries", None) or [(b.vertices, b.indices)]):
                vb = ctx.upload(verts)
                self.bufs.append(vb)
                if inds is None:
                    geoms.append(ctx.geometry(vb, 64, len(verts)))
                else:
                    ib = ctx.upload(np.asarray(inds).astype(np.uint32))
                    self.bufs.append(ib)
                    geoms.append(ctx.geometry(vb, 64, len(verts), ib, len(inds)))
            self.blas_geoms.append(geoms)
            self.blas.append(ctx.blas_build_geometries(geoms, flags=build_flags))
        descs = sd.instance_descs.copy()
        descs["blas"] = [self.blas[int(k)] for k in sd.instance_descs["blas"]] if len(descs) else []
        self.descs = descs
        self.tlas = ctx.tlas_build(descs)
        # hit-shader tables
        self.inst_buf = ctx.upload(np.ascontiguousarray(sd.scene_instances, SCENE_INSTANCE))
        vaddr, iaddr = [], []
        itype = np.uint64 if index_size == 8 else np.uint32
        for verts, inds in sd.objects:
            vb = ctx.upload(verts)
            ib = ctx.upload(np.asarray(inds).astype(itype))
            self.bufs += [vb, ib]
            vaddr.append(vb.addr)
            iaddr.append(ib.addr)
        self.vt = ctx.upload(np.array(vaddr or [0], np.uint64))
        self.it = ctx.upload(np.array(iaddr or [0], np.uint64))
        mats = np.ascontiguousarray(sd.materials, MATERIAL)
        self.mat_buf = ctx.upload(mats if len(mats) else np.zeros(1, MATERIAL))
        maddr = [self.mat_buf.addr + int(m) * MATERIAL.itemsize for m in sd.instance_material]
        self.mt = ctx.upload(np.array(maddr or [0], np.uint64))
        self.bufs += [self.inst_buf, self.vt, self.it, self.mat_buf, self.mt]
        self.tables = SceneTables(self.inst_buf.addr, len(sd.scene_instances), index_size, self.vt.addr, self.it.addr,
                                  self.mt.addr, len(sd.objects), 0)

    def rebuild_blas(self, k):
        """BLAS k again from its resident geometry (its TLAS then needs a rebuild)"""
        self.ctx.blas_rebuild_geometries(self.blas[k], self.blas_geoms[k])

    def move_instances(self, sd_moved):
        """TLAS::regenerate after SceneDescription::blas_transforms + update
        (src/ray/mod.rs:162-193): the TLAS rebuilt from the moved scene's
        instance records and its SceneInstance table uploaded in place"""
        self.ctx.sync()
        self.inst_buf.upload(np.ascontiguousarray(sd_moved.scene_instances, SCENE_INSTANCE))
        self.update_instances(sd_moved.instance_descs)

    def update_instances(self, instance_descs):
        descs = instance_descs.copy()
        descs["blas"] = [self.blas[int(k)] for k in instance_descs["blas"]]
        self.descs = descs
        self.ctx.tlas_rebuild(self.tlas, descs)


class Server:
    """render-request server: resident scenes, jobs of mixed scene / size /
    samples traced by one wavefront (solhip_server_*)"""

    def __init__(self, ctx, hbm_budget, max_jobs=64):
        self.ctx = ctx
        h = C.c_uint64()
        check(lib().solhip_server_create(ctx.h, hbm_budget, max_jobs, C.byref(h)))
        self.h = h.value

    def scene(self, gpu_scene):
        sid = C.c_uint32()
        check(lib().solhip_server_scene(self.ctx.h, self.h, gpu_scene.tlas, C.byref(gpu_scene.tables), C.byref(sid)))
        return sid.value

    def submit(self, scene, W, H, spec, uniforms, accum, out_rgba8=None, accum_start=0):
        d = JobDesc()
        d.scene, d.W, d.H, d.spec = scene, W, H, spec
        u = np.ascontiguousarray(uniforms, UNIFORMS).view(np.uint8).reshape(-1)
        C.memmove(d.uniforms, u.ctypes.data, 400)
        d.accum_start_frame = accum_start
        d.accum = accum.h
        d.out_rgba8 = out_rgba8.h if out_rgba8 is not None else 0
        jid = C.c_uint64()
        check(lib().solhip_server_submit(self.ctx.h, self.h, C.byref(d), C.byref(jid)))
        return jid.value

    def run(self, max_steps=1 << 30):
        live = C.c_uint32()
        check(lib().solhip_server_run(self.ctx.h, self.h, max_steps, C.byref(live)))
        return live.value

    def job(self, jid):
        st = JobStatus()
        check(lib().solhip_server_job(self.ctx.h, self.h, jid, C.byref(st)))
        return st

    def close(self):
        if self.h:
            check(lib().solhip_server_destroy(self.ctx.h, self.h))
            self.h = 0


MLP_PARAMS = 17408


class Mlp:
    """the radiance-cache network (solhip_mlp_*): 64 -> 4 x 64 ReLU -> 3"""

    def __init__(self, ctx, seed=1, lr=1e-2):
        self.ctx = ctx
        h = C.c_uint64()
        check(lib().solhip_mlp_create(ctx.h, seed, lr, C.byref(h)))
        self.h = h.value

    def params(self, set_to=None):
        out = np.zeros(MLP_PARAMS, np.float32)
        if set_to is not None:
            w = np.ascontiguousarray(set_to, np.float32)
            assert w.size == MLP_PARAMS
            check(lib().solhip_mlp_params(self.ctx.h, self.h, _ptr(out), _ptr(w)))
        else:
            check(lib().solhip_mlp_params(self.ctx.h, self.h, _ptr(out), None))
        return out

    def forward(self, x_buf, n, y_buf, stream=0):
        check(lib().solhip_mlp_forward(self.ctx.h, stream, self.h, x_buf.h, n, y_buf.h))

    def train(self, x_buf, t_buf, n, steps=1, stream=0):
        loss = C.c_float(0.0)
        check(lib().solhip_mlp_train(self.ctx.h, stream, self.h, x_buf.h, t_buf.h, n, steps, C.byref(loss)))
        return float(loss.value)

    def gradient(self, x_buf, t_buf, n, stream=0):
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_float(0.0)
        check(lib().solhip_mlp_gradient(self.ctx.h, stream, self.h, x_buf.h, t_buf.h, n, _ptr(g), C.byref(loss)))
        return g, float(loss.value)

    def close(self):
        if self.h:
            check(lib().solhip_mlp_destroy(self.ctx.h, self.h))
            self.h = 0

