// kernels.h -- launch interfaces of the trace / program kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "common.h"

namespace solhip {

// device copies of the hit-shader descriptor tables
struct DevTables {
    const solhip_scene_instance *instances;
    const uint64_t *vertex_table;   // -> solhip_model_vertex[]
    const uint64_t *index_table;    // -> u32[] / u64[]
    const uint64_t *material_table; // -> solhip_material (per instance), may be null
    uint32_t instance_count;
    uint32_t index_size;
};

// closest hit of a wavefront queue slot (20 B): the public solhip_hit
// without the geometry index, which the programs' hit shaders do not read
struct QHit {
    float t, u, v;
    uint32_t inst, prim;   // prim: BLAS-wide triangle index
};

struct Cam {
    float view_inverse[16];
    float projection_inverse[16];
};

// queue entry of the AO wavefront: a primary hit waiting for its AO rays
struct alignas(16) AoEntry {
    uint32_t pixel;
    float px, py, pz;
    float nx, ny, nz;
    uint32_t rng;
};

// launch timing hook (hipEvents on the launch stream), null when disabled
struct Timer {
    virtual void begin(hipStream_t s, int kernel) = 0;
    virtual void end(hipStream_t s, int kernel) = 0;
    virtual ~Timer() = default;
};

struct ProgArgs {
    SceneView sv;
    DevTables tb;
    Cam cam;
    uint32_t W, H, frame;
    float alpha;                    // 1 / (frame - accum_start + 1)
    float4 *accum;
    uint32_t *out_rgba8;
    const uint32_t *blue;           // RGBA8 texels
    uint32_t blue_w, blue_h;
    solhip_hit *hits_out;
    // AO
    uint32_t n_ao;
    float ao_tmin, ao_tmax;
    uint32_t ao_any;
    AoEntry *queue;
    uint32_t *qcount;       // [0] queue entries, [1] ray cursor of the persistent AO kernel
    uint8_t *occl;          // per AO ray: occluded (persistent AO kernel)
    uint32_t *aoheads;      // fetch heads of the persistent AO kernel's partitions, 128 B apart
    uint32_t ncu;
    // pathtrace
    uint32_t enable_sky, spp, max_depth;
    Counters *counters;   // per-kernel counter slots [SOLHIP_KERNEL_COUNT]
    bool count;           // fetch-counting kernel variants
};

// ---- wavefront path tracer (SOLHIP_PROG_PATHTRACE) ------------------------

// emitter triangle of next-event estimation (64 B): world-space vertices,
// unit geometric normal, emission, running (inclusive) area sum
struct alignas(16) LightTri {
    float p0[3], p1[3], p2[3], n[3], e[3];
    float cdf;
};

// extension / shadow queue entry (32 B): ray and the path it belongs to
// (kNoPath marks an empty slot of the first queue)
struct alignas(16) QRay {
    float ox, oy, oz;
    uint32_t path;
    float dx, dy, dz;
    float tmax;
};
constexpr uint32_t kNoPath = 0xFFFFFFFFu;

// per-frame constants of the wavefront kernels, in device memory so that a
// captured frame can be replayed with new uniforms
struct FrameConsts {
    Cam cam;
    uint32_t frame;
    float alpha;
    uint32_t pad[2];
};

// a radiance-cache query: path and its throughput at the query vertex
struct CacheQuery {
    uint32_t path;
    float T[3];
};

// radiance cache state of a wavefront frame (null pointers when off)
struct NrcArgs {
    uint32_t enabled;
    uint32_t ratio;          // one path in `ratio` is a training path
    uint32_t max_rec;        // training records per path (vertices 1 .. max_rec)
    uint32_t train_cap;      // training paths per pass
    uint32_t batch_cap;      // training rows per frame
    float lo[3], inv_ext[3]; // scene box -> [0, 1]^3
    uint16_t *qfeat;         // query features [nslots][64] BF16
    CacheQuery *qrec;
    uint32_t *counts;        // [0] queries, [1] training paths, [2] batch rows
    uint32_t *tidx;          // per pixel: training index or ~0
    uint32_t *tpath;         // per training index: pixel
    uint32_t *trec_n;        // per training index: records written
    uint16_t *tfeat;         // [train_cap * max_rec][64] BF16
    float4 *trec;            // [train_cap * max_rec][2]: (T, -), (L, -) on arrival
    uint16_t *bfeat;         // training batch features [batch_cap][64]
    float4 *btarget;         // training batch targets
};

struct WfArgs {
    SceneView sv;
    DevTables tb;
    const FrameConsts *fc;
    uint32_t W, H;
    float4 *accum;
    uint32_t *out_rgba8;
    uint32_t enable_sky, spp, max_depth, rr_start, nee;
    uint32_t streams;                 // 1: sample s of pixel p starts tea(p + s * npix, frame)
    uint32_t S;                       // samples per pass (> 1 only with streams)
    uint32_t npix;                    // W * H; path index = sample-in-pass * npix + pixel
    float rmin, rmax;                 // ray range of every segment (pathtrace.rgen:35)
    uint32_t nslots;                  // slots of the first queue (8x8 pixel tiles * 64)
    uint32_t slot_base;               // first slot of this image slice (0: the whole image)
    // path state, S * npix each: (T, rng), (L, depth | count_emission << 16);
    // per-pixel sum of the finished samples: float (one stream per pixel) or
    // order-free fixed point (per-sample streams, accum.h)
    float4 *st0, *st1, *psum;
    struct FixSum *fsum;
    QRay *q[2];                       // extension queues (ping-pong by bounce)
    QHit *hits;                       // closest hits, by slot of the current queue
    QRay *sq;                         // shadow queue
    float4 *scontrib;                 // shadow-ray contributions, by shadow slot
    uint32_t *counts;                 // [0], [1]: extension queue sizes, [2]: shadow queue size
    uint32_t *heads;                  // fetch heads, 2 per bounce, wf_head_words() words each
    const LightTri *lights;
    uint32_t nlights;
    NrcArgs nrc;
    uint32_t block;                   // threads per block
    uint32_t persist_blocks;          // grid of the extend / shadow kernels
    uint32_t shade_blocks;            // grid of the shade kernel
    Counters *counters;
    bool count;
};

int launch_trace_rays(hipStream_t s, const SceneView &sv, const solhip_ray *rays, solhip_hit *hits, uint32_t n,
                      uint32_t flags, Counters *counters, Timer *timer);
// rewrites the BLAS-wide triangle index of every hit (geom = 0) into the
// triangle index within its geometry and the geometry index, for instances
// whose BLAS has several geometries (refs: the TLAS's BlasRef per instance)
int launch_split_geometry(hipStream_t s, solhip_hit *hits, uint64_t n, const struct BlasRef *refs, uint32_t n_inst);
int launch_ao(hipStream_t s, const ProgArgs &a, Timer *timer);
// bytes of the persistent AO kernel's fetch heads
uint64_t ao_head_bytes();
int launch_ao_glsl(hipStream_t s, const ProgArgs &a, Timer *timer);
int launch_pathtrace_glsl(hipStream_t s, const ProgArgs &a, Timer *timer);
int launch_debug(hipStream_t s, const ProgArgs &a, Timer *timer);
int launch_fill(hipStream_t s, uint32_t *p, uint64_t nwords, uint32_t value);

// emitter table: per-instance triangle counts of emissive instances, their
// exclusive scan (n_inst + 1 entries, the last one the total), then the
// records and the running area sum.  lights_count needs no host round trip;
// the caller reads offsets[n_inst] before lights_fill.
int launch_lights_count(hipStream_t s, const DevTables &tb, const uint32_t *inst_tris, uint32_t n_inst,
                        uint32_t *offsets, Timer *timer);
int launch_lights_fill(hipStream_t s, const DevTables &tb, const uint32_t *offsets, uint32_t n_inst,
                       LightTri *lights, uint32_t nlights, Timer *timer);
// work the caller runs between the kernels of a pass (the cache queries
// after the shade that ends query paths)
struct PassHook {
    int (*after_query_shade)(void *user, hipStream_t s);
    void *user;
};
// one frame of the wavefront path tracer (kernels only: graph-capturable
// when the radiance cache is off)
int launch_pathtrace(hipStream_t s, const WfArgs &a, Timer *timer, const PassHook *hook = nullptr);
// the same in two parts: the passes over one image slice, and the final
// image write over the whole image (slot_base 0, all slots)
int launch_pathtrace_passes(hipStream_t s, const WfArgs &a, Timer *timer, const PassHook *hook = nullptr);
int launch_pathtrace_finalize(hipStream_t s, const WfArgs &a, Timer *timer);
// threads per CU that the traversal kernels' register budget keeps resident
uint32_t wf_threads_per_cu();
// words of WfArgs::heads per logical fetch head (its partitions' heads)
uint32_t wf_head_words();
// writes the frame constants through a kernel argument (no host buffer to keep alive)
int launch_set_frame(hipStream_t s, FrameConsts *dst, const FrameConsts &value);

}  // namespace solhip
