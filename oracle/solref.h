/*
 * solref.h -- CPU oracle for the sol-rs hardware ray-tracing path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library (sol-rs_amd/)
 * links, loads or calls this code.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it, as the checker and as the reported
 * host baseline.
 *
 * It restates, line by line, the GLSL programs the reference hands to the
 * Vulkan ray-tracing pipeline plus the closest-hit query the Vulkan driver
 * answers behind traceRayEXT:
 *   assets/glsl/sampling.glsl, payload.glsl, postprocess.glsl (gamma),
 *   assets/glsl/ao.{rgen,rchit,rmiss}, pathtrace.{rgen,rchit,rmiss},
 *   assets/glsl/debug.{rgen,rchit,rmiss},
 *   src/ray/acceleration.rs (instance record layout, flags),
 *   src/ray/mod.rs (SceneInstance table), src/scene/{mod,mesh}.rs (records).
 *
 * Floating-point contract shared with the HIP kernels (both sides are built
 * with -ffp-contract=off): every expression is evaluated left to right in
 * fp32 with no fused multiply-add; division and sqrt are correctly rounded;
 * sin/cos come from sr_sinf/sr_cosf below (own polynomial, identical on both
 * sides).  Under this contract hit IDs, t, u, v are bit-identical between
 * this oracle and the GPU.
 *
 * Pinning of each function (what checks it besides the GPU) is listed in the
 * header comment of solref.c.
 */
#ifndef SOLREF_H
#define SOLREF_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- records shared with the reference (byte layouts are the contract) -- */

/* src/scene/mesh.rs:7-25  ModelVertex: 4 x vec4 = 64 B */
typedef struct { float pos[4], color[4], normal[4], uv[4]; } sr_model_vertex;

/* src/scene/mod.rs:19-29  MaterialInfo: 48 B */
typedef struct {
    float base_color[4];
    float emissive[3];
    float padding0;
    float metallic, roughness, padding1, padding2;
} sr_material;

/* src/ray/mod.rs:16-24  SceneInstance: 144 B, matrices column-major (glam) */
typedef struct {
    uint32_t id, texture_offset;
    float padding[2];
    float transform[16];
    float transform_it[16];
} sr_scene_instance;

/* src/ray/acceleration.rs:256-263  InstanceDescriptor: 64 B.  transform is
 * the 3x4 row-major object-to-world matrix; id_mask = id:24 | mask:8;
 * offset_flags = sbt offset:24 | flags:8; blas = BLAS index here (the
 * reference stores a device address). */
typedef struct {
    float transform[12];
    uint32_t id_mask;
    uint32_t offset_flags;
    uint64_t blas;
} sr_instance_desc;

/* examples/4-ray-ao.rs:9-17, examples/5-pathtrace.rs:9-17  SceneUniforms
 * (std140 == repr(C) here): 6 x mat4 + uvec3, padded to 400 B. */
typedef struct {
    float model[16], view[16], view_inverse[16];
    float projection[16], projection_inverse[16], model_view_projection[16];
    uint32_t frame[3];
    uint32_t pad;
} sr_uniforms;

/* ray: 32 B {o.xyz, tmin, d.xyz, tmax};  hit: 24 B {t,u,v,inst,prim,geom}.
 * Inside the renderers prim is the BLAS-wide triangle index (the geometries
 * of a BLAS numbered one after another) and geom is 0; the hits sr_trace and
 * sr_render_ao return carry gl_GeometryIndexEXT in geom and gl_PrimitiveID
 * (the index within that geometry) in prim. */
typedef struct { float ox, oy, oz, tmin, dx, dy, dz, tmax; } sr_ray;
typedef struct { float t, u, v; uint32_t inst, prim, geom; } sr_hit;
#define SR_MISS 0xFFFFFFFFu

/* ---- sampling.glsl / postprocess.glsl --------------------------------- */
uint32_t sr_tea(uint32_t val0, uint32_t val1);
float    sr_next_rand(uint32_t *rng);
float    sr_sinf(float x);
float    sr_cosf(float x);
float    sr_fresnel_dielectric(const float i[3], const float m[3], float eta);
void     sr_onb(const float n[3], float u[3], float v[3]);
void     sr_align_to_direction(const float n[3], float cos_theta, float phi, float out[3]);
void     sr_sample_ggx(const float n[3], const float xi[2], float alpha_sq, float out[3]);
void     sr_sample_cosine(const float n[3], const float xi[2], float out[3]);
uint32_t sr_pcg_advance(uint32_t rng, uint32_t steps);

/* camera ray of ao.rgen:49-54 / pathtrace.rgen:52-58 */
void sr_camera_ray(const sr_uniforms *u, uint32_t x, uint32_t y, uint32_t W, uint32_t H,
                   float jx, float jy, float origin[3], float dir[3]);

/* watertight ray/triangle test; returns 1 on a hit with t in [tmin, tmax] */
int sr_intersect_triangle(const float o[3], const float d[3], float tmin, float tmax,
                          const float v0[3], const float v1[3], const float v2[3],
                          float *t, float *u, float *v);

/* ---- scene (instances of triangle BLASes) ------------------------------ */
typedef struct {
    const void *vertices;       /* first vertex of the geometry */
    uint64_t vertex_stride;     /* bytes; positions are 3 floats at offset 0 */
    uint32_t vertex_count;
    const uint32_t *indices;    /* already offset to the geometry's first index;
                                   NULL: non-indexed, triangle i = vertices 3i..3i+2
                                   and vertex_count / 3 triangles
                                   (src/ray/acceleration.rs:169-188) */
    uint32_t index_count;
} sr_geometry;

typedef struct sr_scene sr_scene;
/* Builds the CPU acceleration structures, one geometry per BLAS.  Geometry
 * memory must outlive the scene. */
sr_scene *sr_scene_create(const sr_geometry *blas, uint32_t nblas,
                          const sr_instance_desc *inst, uint32_t ninst);
/* The same with geo_count[b] geometries in BLAS b (BLAS::new's
 * Vec<GeometryInstance>, src/ray/acceleration.rs:136-208), the geometries of
 * all BLASes back to back in geos. */
sr_scene *sr_scene_create_multi(const sr_geometry *geos, const uint32_t *geo_count, uint32_t nblas,
                                const sr_instance_desc *inst, uint32_t ninst);
void      sr_scene_destroy(sr_scene *s);

/* Closest hit for n rays: (t, instance, geometry, primitive) lexicographic
 * minimum over all hits with t in [tmin, tmax].  use_bvh = 0: brute force over every
 * triangle; 1: through the oracle's binary BVH (same answer).  cull_mask is
 * ANDed with the instance mask.  Threads: OpenMP, nthreads <= 0 = all. */
void sr_trace(const sr_scene *s, const sr_ray *rays, sr_hit *hits, uint64_t n,
              int use_bvh, uint32_t cull_mask, int nthreads);

/* ---- hit-shader inputs (the descriptor arrays the shaders index) ------- */
typedef struct {
    const sr_scene_instance *instances;     /* scnDesc.i[] */
    uint32_t instance_count;
    const sr_model_vertex *const *vertices; /* vertices[objId].v[] */
    const void *const *indices;             /* indices[objId].i[] */
    uint32_t index_size;                    /* 8 (reference: u64) or 4 */
    const sr_material *const *materials;    /* materials[gl_InstanceID].mat */
} sr_tables;

/* ---- programs ---------------------------------------------------------- */
/* ao.rgen/ao.rchit/ao.rmiss: accum is W*H float4 (read-modify-write),
 * blue is the RGBA8 blue-noise texture (bw x bh). */
void sr_render_ao_glsl(const sr_scene *s, const sr_tables *tb, const sr_uniforms *u,
                       int32_t accum_start_frame, float *accum,
                       const uint8_t *blue, uint32_t bw, uint32_t bh,
                       uint32_t W, uint32_t H, uint32_t y0, uint32_t y1, int nthreads);

/* Plain ray-traced AO (one jittered primary ray + n_ao cosine rays per
 * primary hit, range [ao_tmin, ao_tmax]); see solref.c for the estimator. */
typedef struct {
    uint32_t n_ao;
    float ao_tmin, ao_tmax;
    uint32_t any_hit;      /* 1: occlusion queries may stop at any hit */
} sr_ao_params;
void sr_render_ao(const sr_scene *s, const sr_tables *tb, const sr_uniforms *u,
                  const sr_ao_params *p, int32_t accum_start_frame, float *accum,
                  sr_hit *prim_hits, sr_hit *ao_hits, uint64_t *ray_count,
                  uint32_t W, uint32_t H, uint32_t y0, uint32_t y1, int nthreads);

/* pathtrace.rgen/rchit/rmiss; enable_sky = specialization constant 0 */
typedef struct {
    uint32_t enable_sky;
    uint32_t samples;     /* 8 in the shader */
    uint32_t max_bounces; /* 32 in the shader */
} sr_pt_params;
void sr_render_pathtrace_glsl(const sr_scene *s, const sr_tables *tb, const sr_uniforms *u,
                              const sr_pt_params *p, int32_t accum_start_frame,
                              float *accum, uint8_t *out_rgba8, uint64_t *ray_count,
                              uint32_t W, uint32_t H, uint32_t y0, uint32_t y1, int nthreads);

/* Wavefront path tracer (SOLHIP_PROG_PATHTRACE) restated as a per-pixel
 * loop.  With nee = 0 and rr_start = 0 it is exactly the pathtrace.* program
 * above (same RNG stream, same draws, same products).  Extensions, each
 * consuming its draws after the three of pathtrace.rchit:93-94:
 *  - next-event estimation of the diffuse lobe at every scattering vertex:
 *    one emissive triangle picked by area (draw 1), a uniform point on it
 *    (draws 2-3), an any-hit shadow ray, weighted by the probability
 *    1 - max(metallic, F) that pathtrace.rchit takes the diffuse lobe for
 *    that direction; when the diffuse lobe was sampled, the next vertex
 *    ignores emitters it hits (they were sampled);
 *  - Russian roulette from scattering vertex rr_start on: survival
 *    q = min(max(T), 0.95) (draw 4), T /= q.
 * A path dropped at max_bounces keeps what next-event estimation gave it.
 * enable_sky: 0 black, 1 pathtrace.rmiss sky with sun, 2 its gradient only. */
typedef struct {
    uint32_t enable_sky;
    uint32_t samples;
    uint32_t max_bounces;
    uint32_t rr_start;    /* 0 = no Russian roulette */
    uint32_t nee;         /* 1 = next-event estimation */
    uint32_t sample_streams; /* 1 = sample i starts its own stream tea(x + y*W + i*W*H, frame), and
                                the samples are summed in fixed point (2^20 units per 1.0, int64) */
} sr_pt_config;
void sr_render_pathtrace(const sr_scene *s, const sr_tables *tb, const sr_uniforms *u,
                         const sr_pt_config *p, int32_t accum_start_frame,
                         float *accum, uint8_t *out_rgba8, uint64_t *ray_count,
                         uint32_t W, uint32_t H, uint32_t y0, uint32_t y1, int nthreads);

/* light table of the estimator: every triangle of every instance whose
 * material is an emitter (some emissive channel >= 1, pathtrace.rchit:71), in
 * (instance, primitive) order, world-space vertices, unit geometric normal,
 * emission and the running (inclusive) area sum.  Returns the count (writes
 * at most cap records) and the total area. */
typedef struct {
    float p0[3], p1[3], p2[3], n[3], e[3];
    float cdf;
} sr_light;
uint32_t sr_lights(const sr_scene *s, const sr_tables *tb, sr_light *out, uint32_t cap, float *total_area);

/* debug.rgen/rchit/rmiss: barycentrics on hit, ray direction on miss */
void sr_render_debug(const sr_scene *s, const sr_uniforms *u, float *out_rgba32f,
                     uint32_t W, uint32_t H, int nthreads);

/* utilities */
void sr_instance_inverse(const float m[12], float inv[12]);
double sr_wall_seconds(void);

#ifdef __cplusplus
}
#endif
#endif
