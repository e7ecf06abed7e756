/*
 * solhip.h -- C ABI of the MI355X (gfx950) ray-tracing backend for sol-rs.
 *
 * Replaces the Vulkan hardware ray-tracing path of sol-rs: the
 * acceleration-structure and pipeline code of src/ray, the scene upload of
 * src/scene and the raygen / closest-hit / miss shaders of assets/glsl used by
 * examples/4-ray-ao.rs and examples/5-pathtrace.rs.  The Rust side binds these
 * symbols through a thin `extern "C"` crate (sol-hip-sys); the C++ mirror of
 * sol's ray:: / scene:: API (sol-rs_amd/host/) is written against this header.
 *
 * Conventions
 *   - every entry point returns int32_t: 0 = SOLHIP_OK, negative = error;
 *     solhip_last_error() returns the message of the last failure on the
 *     calling thread;
 *   - handles are opaque uint64_t; every *_create / *_build has a *_destroy;
 *   - a context is externally synchronized (like a VkQueue); work is enqueued
 *     on a solhip_stream (a HIP stream; it stands in for the
 *     vk::CommandBuffer the reference records into) and solhip_stream_sync
 *     waits for it;
 *   - device memory is owned by solhip_buffer objects; records passed by
 *     pointer (geometries, instances, uniforms, tables) are host memory and
 *     are consumed before the call returns;
 *   - there is no CPU fallback: without a usable gfx950 device,
 *     solhip_context_create fails with SOLHIP_ERROR_NO_DEVICE.
 */
#ifndef SOLHIP_H
#define SOLHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SOLHIP_ABI_VERSION 2u

typedef uint64_t solhip_ctx;
typedef uint64_t solhip_stream;
typedef uint64_t solhip_buffer;
typedef uint64_t solhip_blas;
typedef uint64_t solhip_tlas;
typedef uint64_t solhip_program;
typedef uint64_t solhip_server;
typedef uint64_t solhip_mlp;

enum {
    SOLHIP_OK = 0,
    SOLHIP_ERROR_INVALID_ARGUMENT = -1,
    SOLHIP_ERROR_INVALID_HANDLE = -2,
    SOLHIP_ERROR_OUT_OF_MEMORY = -3,
    SOLHIP_ERROR_NO_DEVICE = -4,
    SOLHIP_ERROR_DEVICE = -5,        /* HIP runtime / kernel failure */
    SOLHIP_ERROR_UNSUPPORTED = -6,
    SOLHIP_ERROR_OVERFLOW = -7,      /* a device-side capacity was exceeded */
    SOLHIP_ERROR_STALE = -8          /* a TLAS (or a server scene) refers to a BLAS / TLAS
                                        that was rebuilt or destroyed after it was built:
                                        rebuild it (the rule Vulkan has for its TLAS) */
};

/* ---- records with the reference's byte layouts ------------------------- */

/* src/scene/mesh.rs:7-25 ModelVertex (64 B) */
typedef struct { float pos[4], color[4], normal[4], uv[4]; } solhip_model_vertex;

/* src/scene/mod.rs:19-29 MaterialInfo (48 B) */
typedef struct {
    float base_color[4];
    float emissive[3];
    float padding0;
    float metallic, roughness, padding1, padding2;
} solhip_material;

/* src/ray/mod.rs:16-24 SceneInstance (144 B, column-major mat4s) */
typedef struct {
    uint32_t id, texture_offset;
    float padding[2];
    float transform[16];
    float transform_it[16];
} solhip_scene_instance;

/* src/ray/acceleration.rs:256-306 InstanceDescriptor (64 B):
 * transform = 3x4 row-major object-to-world, id_mask = custom id:24 | mask:8,
 * offset_flags = SBT offset:24 | VkGeometryInstanceFlagsKHR:8, and, where the
 * reference stores the BLAS device address, the solhip_blas handle. */
typedef struct {
    float transform[12];
    uint32_t id_mask;
    uint32_t offset_flags;
    uint64_t blas;
} solhip_instance;

/* examples/4-ray-ao.rs:9-17 / examples/5-pathtrace.rs:9-17 SceneUniforms
 * (std140, 400 B) */
typedef struct {
    float model[16], view[16], view_inverse[16];
    float projection[16], projection_inverse[16], model_view_projection[16];
    uint32_t frame[3];
    uint32_t pad;
} solhip_scene_uniforms;

/* BLAS geometry: replaces vk::AccelerationStructureGeometryTrianglesDataKHR +
 * BuildRangeInfo as BLAS::new fills them (src/ray/acceleration.rs:148-208).
 * Positions are R32G32B32_SFLOAT at byte offset 0 of each vertex.
 *   indexed (index_addr != 0): primitive_count = index_count / 3; the indices
 *     are UINT32 at index_addr + primitive_offset, and vertex k of a triangle
 *     is at vertex_addr + (first_vertex + index_k) * vertex_stride; every
 *     index must be < vertex_count (BLAS::new's max_vertex + 1);
 *   non-indexed (index_addr == 0, index_count == 0): primitive_count =
 *     vertex_count / 3 (acceleration.rs:185-188) and vertex k of triangle i is
 *     at vertex_addr + primitive_offset + (first_vertex + 3 i + k) * vertex_stride,
 *     VkAccelerationStructureBuildRangeInfoKHR's rule for non-indexed data. */
typedef struct {
    uint64_t vertex_addr;        /* device address (solhip_buffer_address) */
    uint64_t vertex_stride;      /* 64 for ModelVertex; >= 12, a multiple of 4 */
    uint32_t first_vertex;
    uint32_t vertex_count;
    uint64_t index_addr;         /* device address of a u32 index buffer, or 0 */
    uint64_t primitive_offset;   /* BuildRangeInfo::primitive_offset, bytes */
    uint32_t index_count;        /* indexed: a multiple of 3; non-indexed: 0 */
    uint32_t opaque;             /* GeometryFlagsKHR::OPAQUE; must be 1 */
} solhip_geometry;

/* ray / hit records of solhip_trace_rays (32 B / 24 B).  Closest hit is the
 * lexicographic minimum of (t, instance index, geometry index, primitive
 * index) over all triangles with t in [tmin, tmax]; on a miss t = -1 and
 * inst = prim = geom = 0xFFFFFFFF.  inst is the TLAS instance index
 * (gl_InstanceID), geom the geometry of the BLAS (gl_GeometryIndexEXT), prim
 * the triangle index within that geometry (gl_PrimitiveID), (u, v) the
 * barycentrics of vertices 1 and 2 (hitAttributeEXT.xy). */
typedef struct { float ox, oy, oz, tmin, dx, dy, dz, tmax; } solhip_ray;
typedef struct { float t, u, v; uint32_t inst, prim, geom; } solhip_hit;
#define SOLHIP_MISS 0xFFFFFFFFu

/* trace flags */
#define SOLHIP_TRACE_TERMINATE_ON_FIRST_HIT 0x4u  /* gl_RayFlagsTerminateOnFirstHitEXT */
#define SOLHIP_TRACE_COUNT                  0x100u /* count node/triangle/instance fetches */

/* hit-shader descriptor tables (set 1, bindings 3-6 of the examples):
 * scnDesc.i[] (SceneInstance), vertices[objId], indices[objId],
 * materials[gl_InstanceID].  The *_table_addr fields are device addresses of
 * arrays of uint64 device addresses (one per object / instance). */
typedef struct {
    uint64_t instances_addr;        /* solhip_scene_instance[instance_count] */
    uint32_t instance_count;
    uint32_t index_size;            /* 8 (the reference's u64 index storage) or 4 */
    uint64_t vertex_table_addr;     /* u64[objects] -> ModelVertex[] */
    uint64_t index_table_addr;      /* u64[objects] -> u32/u64 indices */
    uint64_t material_table_addr;   /* u64[instances] -> MaterialInfo, or 0 */
    uint32_t object_count;
    uint32_t reserved;
} solhip_scene_tables;

/* programs (ray::PipelineInfo shader sets, src/ray/pipeline.rs:5-52) */
enum {
    SOLHIP_PROG_AO_GLSL = 1,        /* ao.rgen/rchit/rmiss */
    SOLHIP_PROG_PATHTRACE_GLSL = 2, /* pathtrace.rgen/rchit/rmiss */
    SOLHIP_PROG_DEBUG = 3,          /* debug.rgen/rchit/rmiss */
    SOLHIP_PROG_AO = 4,             /* plain ray-traced AO (primary + n_ao rays) */
    SOLHIP_PROG_PATHTRACE = 5       /* wavefront path tracer: pathtrace.* estimator,
                                       optional next-event estimation and
                                       Russian roulette */
};

/* specialization + run configuration; solhip_spec_default gives the GLSL
 * constants (ao.rgen:42-43, pathtrace.rgen:43-44, pathtrace.rmiss:5). */
typedef struct {
    uint32_t enable_sky;      /* specialization constant 0 (ENABLE_SKYLIGHT) */
    uint32_t spp_per_frame;   /* pathtrace: 8 */
    uint32_t max_depth;       /* pathtrace: 32 bounces */
    uint32_t n_ao;            /* SOLHIP_PROG_AO: rays per primary hit (8) */
    float ao_tmin, ao_tmax;   /* SOLHIP_PROG_AO ray range */
    uint32_t ao_any_hit;      /* SOLHIP_PROG_AO: occlusion may stop at any hit */
    /* SOLHIP_PROG_PATHTRACE.  With rr_start = 0 and nee = 0 the image is
     * bit-identical to SOLHIP_PROG_PATHTRACE_GLSL's.  enable_sky = 2 gives the
     * sky gradient of pathtrace.rmiss without the sun disc. */
    uint32_t rr_start;        /* Russian roulette from this scattering vertex on; 0 = off */
    uint32_t nee;             /* next-event estimation to emissive triangles at diffuse vertices */
    uint32_t wf_blocks_per_cu;/* persistent grid of the extend / shadow kernels; 0 = default */
    uint32_t wf_block;        /* threads per block of the wavefront kernels (64..256); 0 = default */
    uint32_t wf_graph;        /* 1: replay the frame's passes as a captured hipGraph (recaptured
                                 when buffers, scene or launch shape change) */
    /* learned radiance cache (SOLHIP_PROG_PATHTRACE): paths end at their first
     * bounce vertex, where a network (solhip_mlp_*) trained online on the
     * tracer's own training paths predicts the radiance leaving it.  Biased;
     * the plain tracer (nrc = 0) is its reference.  Not captured in a graph
     * (training reads its batch size back). */
    uint32_t nrc;             /* 1: use the cache */
    uint32_t nrc_train_steps; /* Adam steps per frame (0 = 4) */
    float nrc_learning_rate;  /* Adam step size (0 = 1e-2) */
    /* SOLHIP_PROG_PATHTRACE: the image's 8x8 tiles split into this many
     * contiguous slices whose passes run on their own streams at the same
     * time, so one slice's slowest rays overlap the others' work; pixels are
     * independent, so the image does not depend on it.  0 = default (4
     * with sample_streams, else 1); always 1 with SOLHIP_PROFILE_TIME or nrc. */
    uint32_t wf_slices;
    /* SOLHIP_PROG_PATHTRACE random streams.  0: one stream per pixel runs
     * through all samples of the frame, as in pathtrace.rgen:47 (samples are
     * traced one after another).  1: sample s of pixel p starts its own
     * stream tea(p + s * W * H, frame) (sample 0 as before), so several
     * samples of a pixel can be in flight at once; the image is then
     * independent of wf_samples_per_pass. */
    uint32_t sample_streams;
    /* samples of every pixel traced together per pass (sample_streams = 1
     * only; 0 = enough for about 64 M paths in flight; always 1 with nrc) */
    uint32_t wf_samples_per_pass;
} solhip_spec;

/* per-dispatch images: W*H pixels, row-major, no padding */
typedef struct {
    solhip_buffer accum;        /* RGBA32F (accumulation / AO target), read-modify-write */
    solhip_buffer out_rgba8;    /* RGBA8 output (pathtrace), or 0 */
    solhip_buffer blue_noise;   /* RGBA8 texture (AO glsl), or 0 */
    uint32_t blue_w, blue_h;
    solhip_buffer hits_out;     /* optional: primary hits [W*H], then AO hits [W*H*n_ao] (SOLHIP_PROG_AO) */
} solhip_images;

typedef struct {
    uint64_t rays;             /* rays traced (all kinds) */
    uint64_t node_fetches;     /* with SOLHIP_TRACE_COUNT */
    uint64_t tri_fetches;
    uint64_t inst_fetches;
    uint64_t overflow;         /* traversal stack / queue overflows (must be 0) */
} solhip_counters;

/* ---- context / streams / errors ---------------------------------------- */
/* replaces SharedContext::new (src/context.rs:239-370) for the ray path */
int32_t solhip_context_create(int32_t device, uint32_t flags, solhip_ctx *out);
int32_t solhip_context_destroy(solhip_ctx ctx);
uint32_t solhip_abi_version(void);
int32_t solhip_last_error(char *buf, size_t len);
int32_t solhip_stream_create(solhip_ctx ctx, solhip_stream *out);
int32_t solhip_stream_destroy(solhip_ctx ctx, solhip_stream s);
int32_t solhip_stream_sync(solhip_ctx ctx, solhip_stream s);   /* stream 0 = default */

/* ---- buffers (src/buffer.rs: Buffer::new/from_data/update/get_device_address) */
int32_t solhip_buffer_create(solhip_ctx ctx, uint64_t bytes, solhip_buffer *out);
int32_t solhip_buffer_destroy(solhip_ctx ctx, solhip_buffer b);
int32_t solhip_buffer_upload(solhip_ctx ctx, solhip_buffer b, uint64_t offset, const void *src, uint64_t bytes);
int32_t solhip_buffer_download(solhip_ctx ctx, solhip_buffer b, uint64_t offset, void *dst, uint64_t bytes);
int32_t solhip_buffer_fill(solhip_ctx ctx, solhip_stream s, solhip_buffer b, uint32_t value);
int32_t solhip_buffer_address(solhip_ctx ctx, solhip_buffer b, uint64_t *addr);

/* ---- acceleration structures ------------------------------------------- */
/* BLAS::new (src/ray/acceleration.rs:136-239): GPU BVH build over the n
 * geometries g[0..n) (BLAS::new's Vec<GeometryInstance>), enqueued on s.
 * Their triangles are numbered geometry after geometry; the hit shaders of
 * the programs index an instance's index table with that BLAS-wide number,
 * which is gl_PrimitiveID for the one-geometry BLAS that
 * SceneDescription::from_meshes builds (src/ray/mod.rs:125-132).
 * Invalid geometry (null addresses with triangles, stride < 12 or not a
 * multiple of 4, an index_count that is not a multiple of 3, an index
 * >= vertex_count) fails with SOLHIP_ERROR_INVALID_ARGUMENT.  A rebuild or
 * destroy of a BLAS makes every TLAS built over it stale: traces through
 * such a TLAS fail with SOLHIP_ERROR_STALE until it is rebuilt, as Vulkan
 * requires a TLAS build after its BLASes change.  flags: 0 builds over Morton
 * order (fast build, for geometry rebuilt every frame);
 * SOLHIP_BUILD_PREFER_FAST_TRACE builds top down by binned SAH (slower build,
 * fewer node tests per ray), the flag BLAS::new passes to Vulkan
 * (VK_BUILD_ACCELERATION_STRUCTURE_PREFER_FAST_TRACE_BIT_KHR,
 * src/ray/acceleration.rs:213).  solhip_blas_rebuild keeps the flags. */
#define SOLHIP_BUILD_PREFER_FAST_TRACE 0x1u
int32_t solhip_blas_build(solhip_ctx ctx, solhip_stream s, const solhip_geometry *g, uint32_t n,
                          uint32_t flags, solhip_blas *out);
/* rebuild in place from the same (or updated) vertex data; n may differ
 * from the first build */
int32_t solhip_blas_rebuild(solhip_ctx ctx, solhip_stream s, solhip_blas b, const solhip_geometry *g, uint32_t n);
int32_t solhip_blas_destroy(solhip_ctx ctx, solhip_blas b);
/* TLAS::new / TLAS::regenerate (src/ray/acceleration.rs:344-467) */
int32_t solhip_tlas_build(solhip_ctx ctx, solhip_stream s, const solhip_instance *inst, uint32_t n,
                          solhip_tlas *out);
int32_t solhip_tlas_rebuild(solhip_ctx ctx, solhip_stream s, solhip_tlas t, const solhip_instance *inst, uint32_t n);
int32_t solhip_tlas_destroy(solhip_ctx ctx, solhip_tlas t);
/* BVH statistics: nodes, triangles, tree depth, bytes */
typedef struct { uint32_t nodes, prims, depth, leaf_max; uint64_t bytes; } solhip_bvh_info;
int32_t solhip_blas_info(solhip_ctx ctx, solhip_blas b, solhip_bvh_info *info);
int32_t solhip_tlas_info(solhip_ctx ctx, solhip_tlas t, solhip_bvh_info *info);

/* ---- tracing ------------------------------------------------------------ */
/* closest-hit (or first-hit) query for n rays in device buffers; the
 * behaviour of traceRayEXT with gl_RayFlagsOpaqueEXT and cull mask 0xFF. */
int32_t solhip_trace_rays(solhip_ctx ctx, solhip_stream s, solhip_tlas t, solhip_buffer rays,
                          solhip_buffer hits, uint32_t n, uint32_t flags);

/* ray::Pipeline::new + ShaderBindingTable (src/ray/pipeline.rs:61-122,
 * src/ray/sbt.rs:71-165) */
void solhip_spec_default(uint32_t kind, solhip_spec *spec);
int32_t solhip_program_create(solhip_ctx ctx, uint32_t kind, const solhip_spec *spec, solhip_program *out);
int32_t solhip_program_destroy(solhip_ctx ctx, solhip_program p);

/* ShaderBindingTable::cmd_trace_rays (src/ray/sbt.rs:167-180): one W x H
 * launch of the program with the examples' bindings (UBO set 0, set 1
 * tables/images, push constant accum_start_frame). */
int32_t solhip_trace(solhip_ctx ctx, solhip_stream s, solhip_program p, solhip_tlas t,
                     const solhip_scene_tables *tables, const solhip_scene_uniforms *uniforms,
                     int32_t accum_start_frame, const solhip_images *images, uint32_t W, uint32_t H);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that forbid
 * allocation): solhip_trace with the same arguments and the same results, but
 * the AO program's scratch (queue, fetch heads, per-ray occlusion) lives in
 * one fixed device array of SOLHIP_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends, and no device memory is allocated.  Needs
 * 256 + 8192 + W*H*(32 + n_ao) bytes (1920x1080 with up to 32 AO rays per
 * pixel); larger frames fail with SOLHIP_ERROR_INVALID_ARGUMENT.  The
 * megakernel programs (AO_GLSL, PATHTRACE_GLSL, DEBUG) use no scratch and
 * run as in solhip_trace; the wavefront PATHTRACE program, whose path state
 * is sized per image, is SOLHIP_ERROR_UNSUPPORTED.  The array keeps the last
 * call's contents, and every call on one GPU shares it: calls that overlap in
 * time on one GPU (two streams or host threads) race. */
#define SOLHIP_NOALLOC_SCRATCH_BYTES (128ull << 20)
int32_t solhip_trace_noalloc(solhip_ctx ctx, solhip_stream s, solhip_program p, solhip_tlas t,
                             const solhip_scene_tables *tables, const solhip_scene_uniforms *uniforms,
                             int32_t accum_start_frame, const solhip_images *images, uint32_t W, uint32_t H);

/* counters accumulated since the last reset (rays, fetches, overflows) */
int32_t solhip_stats(solhip_ctx ctx, solhip_counters *out, uint32_t reset);

/* ---- measurement (no counterpart in the reference, which only reads one
 * GPU timestamp pair per frame, src/renderer.rs:270-289) ------------------ */
enum {
    SOLHIP_KERNEL_TRACE_RAYS = 0,
    SOLHIP_KERNEL_AO_PRIMARY = 1,
    SOLHIP_KERNEL_AO_RAYS = 2,
    SOLHIP_KERNEL_AO_GLSL = 3,
    SOLHIP_KERNEL_PATHTRACE_GLSL = 4,
    SOLHIP_KERNEL_DEBUG = 5,
    SOLHIP_KERNEL_PT_GENERATE = 6,  /* wavefront path tracer: camera rays, per-pixel sums */
    SOLHIP_KERNEL_PT_EXTEND = 7,    /* closest hits of the extension queue */
    SOLHIP_KERNEL_PT_SHADE = 8,     /* materials, emission, scattering, queue compaction */
    SOLHIP_KERNEL_PT_SHADOW = 9,    /* any-hit shadow rays of next-event estimation */
    SOLHIP_KERNEL_PT_FINALIZE = 10, /* sample mean, accumulation, RGBA8 */
    SOLHIP_KERNEL_PT_LIGHTS = 11,   /* emitter table (next-event estimation) */
    SOLHIP_KERNEL_SRV_ADMIT = 12,   /* render server: first camera rays of an admitted job */
    SOLHIP_KERNEL_SRV_EXTEND = 13,  /* render server: closest hits of all live jobs' rays */
    SOLHIP_KERNEL_SRV_SHADE = 14,
    SOLHIP_KERNEL_SRV_SHADOW = 15,
    SOLHIP_KERNEL_SRV_RETIRE = 16,  /* render server: finished samples, next samples, finished pixels */
    SOLHIP_KERNEL_COUNT = 17
};
#define SOLHIP_PROFILE_TIME   0x1u  /* hipEvents around every launch, on the launch stream */
#define SOLHIP_PROFILE_COUNT  0x2u  /* fetch-counting kernel variants */
int32_t solhip_profile(solhip_ctx ctx, uint32_t flags);
typedef struct {
    uint64_t launches;
    double total_ms;           /* sum of event-measured launch durations */
    solhip_counters counters;  /* rays / fetches of this kernel */
} solhip_kernel_stats;
int32_t solhip_kernel_stats_get(solhip_ctx ctx, uint32_t kernel, solhip_kernel_stats *out, uint32_t reset);

/* ---- launch-shape autotuner for SOLHIP_PROG_PATHTRACE (no counterpart in
 * the reference, whose dispatch shape is the driver's) --------------------
 * Times `frames` frames of every candidate launch shape (wf_block in
 * {64, 128, 256} x three persistent grid sizes x wf_graph in {0, 1}, then,
 * with sample_streams, wf_slices in {1, 2, 4, 8} for the best of those) for
 * the estimator settings in *spec on this scene and image size, and returns
 * *spec with the fastest shape in *best (and its frame time).  Results persist in
 * cache_path (text, one line per workload: device, CUs, size, estimator
 * settings, instance and triangle counts); a cached workload is answered
 * without timing.  The timed frames render into images->accum / out_rgba8.
 * cache_path may be NULL (no persistence). */
int32_t solhip_pathtrace_autotune(solhip_ctx ctx, solhip_stream s, solhip_tlas t, const solhip_scene_tables *tables,
                                  const solhip_scene_uniforms *uniforms, const solhip_images *images, uint32_t W,
                                  uint32_t H, const solhip_spec *spec, uint32_t frames, const char *cache_path,
                                  solhip_spec *best, float *best_ms);

/* ---- render-request server (no counterpart in the reference, which
 * renders one image per command buffer) -----------------------------------
 * Jobs of mixed scene, size and sample count share one wavefront: every step
 * runs extend / shade / shadow / retire over the rays of all live jobs; a
 * path that ends starts its pixel's next sample in place, and a pixel whose
 * last sample ends is accumulated into its job's image.  Scenes (BVHs,
 * tables, emitter tables) stay resident; each job holds W*H path slots while
 * it runs.  Jobs are admitted first-come first-served while their path state
 * fits the HBM budget given at creation.  A job's image is bit-identical to
 * solhip_trace of SOLHIP_PROG_PATHTRACE with the same spec and uniforms. */
typedef struct {
    uint32_t scene;                  /* from solhip_server_scene */
    uint32_t W, H;
    solhip_spec spec;                /* SOLHIP_PROG_PATHTRACE settings (launch-shape fields ignored).
                                        With sample_streams, wf_samples_per_pass is the job's slots
                                        per pixel, each running every lanes-th sample (0 = about 16
                                        samples per slot within ~4 M slots); the image equals
                                        solhip_trace's for any value */
    solhip_scene_uniforms uniforms;
    int32_t accum_start_frame;
    uint32_t reserved;
    solhip_buffer accum;             /* RGBA32F W*H, read-modify-write */
    solhip_buffer out_rgba8;         /* RGBA8 W*H, or 0 */
} solhip_job_desc;

enum { SOLHIP_JOB_QUEUED = 0, SOLHIP_JOB_RUNNING = 1, SOLHIP_JOB_DONE = 2 };
typedef struct {
    uint32_t state;
    uint32_t reserved;
    double submit_ms, admit_ms, done_ms;   /* server clock (ms since creation) */
} solhip_job_status;

/* hbm_budget: bytes of path state and queues (about 190 B per pixel of the
 * running jobs); max_jobs <= 256 running at once */
int32_t solhip_server_create(solhip_ctx ctx, uint64_t hbm_budget, uint32_t max_jobs, solhip_server *out);
int32_t solhip_server_destroy(solhip_ctx ctx, solhip_server srv);
/* registers a resident scene: TLAS + hit-shader tables (+ its emitter table).
 * Rebuilding or destroying the TLAS (or one of its BLASes) afterwards makes
 * the scene stale: solhip_server_submit and solhip_server_run then fail with
 * SOLHIP_ERROR_STALE while it has jobs, and the scene must be registered
 * again under a new id. */
int32_t solhip_server_scene(solhip_ctx ctx, solhip_server srv, solhip_tlas t, const solhip_scene_tables *tables,
                            uint32_t *scene_id);
int32_t solhip_server_submit(solhip_ctx ctx, solhip_server srv, const solhip_job_desc *job, uint64_t *job_id);
/* runs up to max_steps wavefront steps (admitting and retiring jobs between
 * steps); returns early when nothing is queued or running; *live = jobs
 * queued or running afterwards */
int32_t solhip_server_run(solhip_ctx ctx, solhip_server srv, uint32_t max_steps, uint32_t *live);
int32_t solhip_server_job(solhip_ctx ctx, solhip_server srv, uint64_t job_id, solhip_job_status *out);

/* ---- radiance-cache network (no counterpart in the reference) ----------
 * A 64-input MLP, four hidden ReLU layers of 64, 3 linear outputs, no
 * biases; BF16 MFMA with FP32 accumulation; Adam (beta1 0.9, beta2 0.99,
 * eps 1e-8) on FP32 master weights; loss = relative L2,
 * mean over rows of sum_c (p_c - t_c)^2 / (p_c^2 + 0.01) with the
 * denominator held constant.  Parameters: 17408 floats, W_l[k][n] at
 * l*4096 + k*64 + n (l = 0..3), W_4[k][n] at 16384 + k*16 + n (n < 3; the
 * other columns are kept 0).  Inputs: n rows of 64 BF16; outputs and targets:
 * n float4 (x, y, z, unused). */
int32_t solhip_mlp_create(solhip_ctx ctx, uint32_t seed, float learning_rate, solhip_mlp *out);
int32_t solhip_mlp_destroy(solhip_ctx ctx, solhip_mlp m);
/* copies the FP32 parameters out (get, may be NULL) and/or in (set, may be NULL) */
int32_t solhip_mlp_params(solhip_ctx ctx, solhip_mlp m, float *get, const float *set);
int32_t solhip_mlp_forward(solhip_ctx ctx, solhip_stream s, solhip_mlp m, solhip_buffer in_bf16, uint32_t n,
                           solhip_buffer out_f32x4);
/* steps Adam steps on the full batch; *loss (may be NULL) = loss before the last step */
int32_t solhip_mlp_train(solhip_ctx ctx, solhip_stream s, solhip_mlp m, solhip_buffer in_bf16,
                         solhip_buffer target_f32x4, uint32_t n, uint32_t steps, float *loss);
/* the loss gradient (17408 floats) of one batch without updating */
int32_t solhip_mlp_gradient(solhip_ctx ctx, solhip_stream s, solhip_mlp m, solhip_buffer in_bf16,
                            solhip_buffer target_f32x4, uint32_t n, float *grad, float *loss);

#ifdef __cplusplus
}
#endif
#endif /* SOLHIP_H */
