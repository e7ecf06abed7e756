// objects.h -- the library's handle objects (what the opaque uint64
// handles of include/solhip.h refer to) and the registry helpers shared by
// the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "bvh.h"
#include "common.h"
#include "kernels.h"

namespace solhip {

// radiance-cache network (mlp.hip)
struct Mlp;
struct Context;
Mlp *mlp_new(Context *c, uint32_t seed, float lr);
void mlp_delete(Mlp *m);
int mlp_query(hipStream_t s, Mlp *m, const uint16_t *feat, const uint32_t *n_dev, uint32_t n_max,
              const CacheQuery *qrec, float4 *st1);
int mlp_train_dev(hipStream_t s, Mlp *m, const uint16_t *feat, const float4 *target, uint32_t n, uint32_t steps);

enum class Kind { Context, Stream, Buffer, Blas, Tlas, Program, Server, Mlp };

struct Obj {
    Kind kind;
    uint64_t ctx;
    virtual ~Obj() = default;
};

// hipEvent pairs around launches, resolved when statistics are read
struct EventTimer : Timer {
    struct Rec {
        int kernel;
        hipEvent_t a, b;
    };
    std::vector<Rec> pending;
    std::vector<Rec> open;
    void begin(hipStream_t s, int kernel) override
    {
        Rec r{kernel, nullptr, nullptr};
        if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) return;
        (void)hipEventRecord(r.a, s);
        open.push_back(r);
    }
    void end(hipStream_t s, int kernel) override
    {
        for (size_t i = open.size(); i-- > 0;) {
            if (open[i].kernel != kernel) continue;
            (void)hipEventRecord(open[i].b, s);
            pending.push_back(open[i]);
            open.erase(open.begin() + (long)i);
            return;
        }
    }
};

struct Context : Obj {
    int device = 0;
    uint32_t ncu = 256;   // compute units (persistent grids)
    hipStream_t stream = nullptr;
    Counters *counters = nullptr;  // [SOLHIP_KERNEL_COUNT]
    uint32_t profile = 0;
    EventTimer timer;
    uint64_t launches[SOLHIP_KERNEL_COUNT] = {};
    double total_ms[SOLHIP_KERNEL_COUNT] = {};
    Timer *timer_or_null() { return (profile & SOLHIP_PROFILE_TIME) ? &timer : nullptr; }
    void resolve_events()
    {
        for (auto &r : timer.pending) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess && r.kernel >= 0 && r.kernel < SOLHIP_KERNEL_COUNT) {
                launches[r.kernel] += 1;
                total_ms[r.kernel] += ms;
            }
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        timer.pending.clear();
    }
};

struct Stream : Obj {
    hipStream_t s = nullptr;
};

struct Buffer : Obj {
    void *ptr = nullptr;
    uint64_t bytes = 0;
};

// generation numbers: a BLAS or TLAS takes a new one at every (re)build, so
// that what was built over it can tell that it changed
uint64_t next_generation();

struct Blas : Obj {
    DeviceBvh bvh;
    DevScratch scratch;
    uint32_t build_flags = 0;   // SOLHIP_BUILD_*, kept for rebuilds
    uint32_t n_geo = 0;         // geometries of the last build
    DevBuf geo_dev;             // GeomDesc[n_geo], then u32 [n_geo, tri_begin[n_geo]]
    uint64_t gen = 0;
    // the u32 table BlasRef::geo points to (nullptr with one geometry)
    const uint32_t *geo_table() const
    {
        return n_geo > 1 ? (const uint32_t *)((const char *)geo_dev.ptr + (uint64_t)n_geo * sizeof(GeomDesc)) : nullptr;
    }
};

struct Tlas : Obj {
    DeviceBvh bvh;
    DevScratch scratch;
    DevBuf inst_dev;   // solhip_instance[]
    DevBuf refs_dev;   // BlasRef[]
    std::vector<uint64_t> blas_handles;   // distinct BLAS referenced by the instances
    std::vector<uint64_t> blas_gens;      // their generations when this TLAS was built
    uint32_t n_inst = 0;
    bool multi_geo = false;               // some instance's BLAS has several geometries
    bool built = false;                   // the last (re)build succeeded
    uint64_t gen = 0;
    DevBuf inst_tris_dev;                 // u32 triangle count of each instance's BLAS
};

struct Program : Obj {
    uint32_t kind_id = 0;
    solhip_spec spec{};
    DevBuf queue, qcount, occl, aoheads;
    // SOLHIP_PROG_PATHTRACE: path state, queues, emitter table, frame constants
    DevBuf st0, st1, psum, fsum, q0, q1, hits, sq, scontrib, counts, heads, lights, light_offsets, fc;
    // radiance cache (spec.nrc)
    Mlp *nrc_mlp = nullptr;
    DevBuf nq_feat, nq_rec, n_counts, n_tidx, n_tpath, n_trec_n, n_tfeat, n_trec, n_bfeat, n_btarget;
    uint64_t nrc_frames = 0, nrc_rows = 0;
    // captured frame (wf_graph): valid for the WfArgs it was captured with
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    WfArgs gkey{};
    bool gvalid = false;
    uint64_t graph_captures = 0;
    // image slices (spec.wf_slices): a stream each, fork and join events
    static constexpr uint32_t kMaxSlices = 16;
    hipStream_t slice_stream[kMaxSlices] = {};
    hipEvent_t fork_ev = nullptr, join_ev[kMaxSlices] = {};
    void drop_graph()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
        gvalid = false;
    }
    void release_all()
    {
        drop_graph();
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        cap_stream = nullptr;
        for (uint32_t j = 0; j < kMaxSlices; ++j) {
            if (slice_stream[j]) (void)hipStreamDestroy(slice_stream[j]);
            if (join_ev[j]) (void)hipEventDestroy(join_ev[j]);
            slice_stream[j] = nullptr;
            join_ev[j] = nullptr;
        }
        if (fork_ev) (void)hipEventDestroy(fork_ev);
        fork_ev = nullptr;
        for (DevBuf *b : {&queue, &qcount, &occl, &aoheads, &st0, &st1, &psum, &fsum, &q0, &q1, &hits, &sq, &scontrib, &counts, &heads,
                          &lights, &light_offsets, &fc, &nq_feat, &nq_rec, &n_counts, &n_tidx, &n_tpath, &n_trec_n,
                          &n_tfeat, &n_trec, &n_bfeat, &n_btarget})
            b->release();
        mlp_delete(nrc_mlp);
        nrc_mlp = nullptr;
    }
};

extern std::mutex g_mu;
extern std::unordered_map<uint64_t, Obj *> g_objs;

uint64_t reg(Obj *o);
Obj *unreg(uint64_t h, Kind k);

template <class T>
T *get(uint64_t h, Kind k)
{
    std::lock_guard<std::mutex> l(g_mu);
    auto it = g_objs.find(h);
    if (it == g_objs.end() || it->second->kind != k) return nullptr;
    return static_cast<T *>(it->second);
}

int bad_handle(const char *what);
int bad_arg(const std::string &what);
Context *ctx_of(solhip_ctx c);
hipStream_t stream_of(Context *cx, solhip_stream s, bool &ok);
uint32_t stack_depth_of(const Tlas *t);
// SOLHIP_OK, or SOLHIP_ERROR_STALE (with the message set) when a BLAS the
// TLAS was built over has been rebuilt or destroyed since
int check_fresh(const Tlas *t);
// device address of the fixed scratch of solhip_trace_noalloc on `device`
// (the current device), taken once per device
int noalloc_scratch_init(int device);
SceneView view_of(const Tlas *t);
// device scratch that solhip_trace_noalloc hands to the program instead of
// the program's own buffers
struct FixedScratch {
    void *ptr;
    uint64_t bytes;
};
// byte offsets of the AO wavefront's scratch within one allocation
struct AoScratchLayout {
    uint64_t qcount, heads, queue, occl, total;
};
AoScratchLayout ao_scratch_layout(uint64_t npix, uint32_t n_ao);
// solhip_trace, with the AO program's scratch in `fixed` when it is given
int trace_program(solhip_ctx ctx, solhip_stream st, solhip_program ph, solhip_tlas th,
                  const solhip_scene_tables *tables, const solhip_scene_uniforms *u, int32_t accum_start_frame,
                  const solhip_images *images, uint32_t W, uint32_t H, const FixedScratch *fixed);

// NEE emitter table of a scene into p's buffers; *nl = number of lights
int build_lights(hipStream_t s, const DevTables &tb, const Tlas *t, DevBuf &offsets, DevBuf &lights, uint32_t *nl,
                 Timer *timer);

}  // namespace solhip
