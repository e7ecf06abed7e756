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
    uint32_t *qcount;
    // pathtrace
    uint32_t enable_sky, spp, max_depth;
    Counters *counters;   // per-kernel counter slots [SOLHIP_KERNEL_COUNT]
    bool count;           // fetch-counting kernel variants
};

int launch_trace_rays(hipStream_t s, const SceneView &sv, const solhip_ray *rays, solhip_hit *hits, uint32_t n,
                      uint32_t flags, Counters *counters, Timer *timer);
int launch_ao(hipStream_t s, const ProgArgs &a, Timer *timer);
int launch_ao_glsl(hipStream_t s, const ProgArgs &a, Timer *timer);
int launch_pathtrace_glsl(hipStream_t s, const ProgArgs &a, Timer *timer);
int launch_debug(hipStream_t s, const ProgArgs &a, Timer *timer);
int launch_fill(hipStream_t s, uint32_t *p, uint64_t nwords, uint32_t value);

}  // namespace solhip
