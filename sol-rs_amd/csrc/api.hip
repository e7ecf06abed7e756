// api.hip -- the C ABI of include/solhip.h.
//
// Objects live in a handle registry (opaque uint64 ids, type-checked on every
// call).  Device memory: hipMalloc; uploads/downloads block like
// Buffer::from_data / Buffer::update of the reference (src/buffer.rs:144-225).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "bvh.h"
#include "common.h"
#include "objects.h"
#include "kernels.h"
#include "accum.h"
#include "traverse.h"

namespace solhip {

thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

std::mutex g_mu;
std::unordered_map<uint64_t, Obj *> g_objs;
uint64_t g_next = 1;

uint64_t reg(Obj *o)
{
    std::lock_guard<std::mutex> l(g_mu);
    const uint64_t h = (g_next++ << 8) | (uint64_t)o->kind;
    g_objs[h] = o;
    return h;
}

Obj *unreg(uint64_t h, Kind k)
{
    std::lock_guard<std::mutex> l(g_mu);
    auto it = g_objs.find(h);
    if (it == g_objs.end() || it->second->kind != k) return nullptr;
    Obj *o = it->second;
    g_objs.erase(it);
    return o;
}

int bad_handle(const char *what)
{
    set_error(std::string("invalid ") + what + " handle");
    return SOLHIP_ERROR_INVALID_HANDLE;
}

int bad_arg(const std::string &what)
{
    set_error("invalid argument: " + what);
    return SOLHIP_ERROR_INVALID_ARGUMENT;
}

Context *ctx_of(solhip_ctx c)
{
    Context *cx = get<Context>(c, Kind::Context);
    if (cx) (void)hipSetDevice(cx->device);
    return cx;
}

hipStream_t stream_of(Context *cx, solhip_stream s, bool &ok)
{
    ok = true;
    if (s == 0) return cx->stream;
    Stream *st = get<Stream>(s, Kind::Stream);
    if (!st) {
        ok = false;
        return nullptr;
    }
    return st->s;
}

uint64_t next_generation()
{
    static std::atomic<uint64_t> g{0};
    return ++g;
}

int check_fresh(const Tlas *t)
{
    if (!t->built) {
        set_error("the TLAS's last build failed; rebuild it before tracing");
        return SOLHIP_ERROR_STALE;
    }
    for (size_t k = 0; k < t->blas_handles.size(); ++k) {
        const Blas *b = get<Blas>(t->blas_handles[k], Kind::Blas);
        if (!b || b->gen != t->blas_gens[k]) {
            set_error(std::string("a BLAS of this TLAS was ") + (b ? "rebuilt" : "destroyed") +
                      " after the TLAS was built; rebuild the TLAS (solhip_tlas_rebuild) before tracing");
            return SOLHIP_ERROR_STALE;
        }
    }
    return SOLHIP_OK;
}

// LDS stack entries a traversal of this TLAS needs: one per TLAS level plus
// one per level of the deepest referenced BLAS (plus one spare)
uint32_t stack_depth_of(const Tlas *t)
{
    uint32_t bd = 0;
    for (uint64_t h : t->blas_handles) {
        const Blas *b = get<Blas>(h, Kind::Blas);
        if (b) bd = std::max(bd, b->bvh.depth);
    }
    return t->bvh.depth + bd + 1;
}

SceneView view_of(const Tlas *t)
{
    SceneView v;
    v.stack_depth = std::min<uint32_t>(stack_depth_of(t), (uint32_t)kMaxStack);
    v.single = t->bvh.prim_count == 1 ? 1u : 0u;
    if (t->bvh.prim_count == 0) {
        v.tnodes = nullptr;
        v.insts = nullptr;
        v.inst_meta = nullptr;
    } else {
        v.tnodes = (const WideNode *)t->bvh.nodes.ptr;
        v.insts = (const InstRec *)t->bvh.leaves.ptr;
        v.inst_meta = (const uint32_t *)t->bvh.meta.ptr;
    }
    return v;
}

int tlas_build_into(Context *cx, hipStream_t s, Tlas *t, const solhip_instance *inst, uint32_t n)
{
    std::vector<BlasRef> refs(n);
    std::vector<uint32_t> tris(n ? n : 1, 0u);
    t->built = false;
    t->multi_geo = false;
    t->blas_handles.clear();
    t->blas_gens.clear();
    for (uint32_t i = 0; i < n; ++i) {
        Blas *b = get<Blas>(inst[i].blas, Kind::Blas);
        if (!b) {
            set_error("instance " + std::to_string(i) + " references an invalid BLAS handle");
            return SOLHIP_ERROR_INVALID_HANDLE;
        }
        if (b->gen == 0) {
            set_error("instance " + std::to_string(i) + " references a BLAS whose last build failed");
            return SOLHIP_ERROR_STALE;
        }
        if (std::find(t->blas_handles.begin(), t->blas_handles.end(), inst[i].blas) == t->blas_handles.end()) {
            t->blas_handles.push_back(inst[i].blas);
            t->blas_gens.push_back(b->gen);
        }
        if (b->bvh.prim_count == 0) {
            refs[i].nodes = nullptr;
            refs[i].tris = nullptr;
        } else {
            refs[i].nodes = (const WideNode *)b->bvh.nodes.ptr;
            refs[i].tris = (const TriRec *)b->bvh.leaves.ptr;
        }
        refs[i].root_box = (const float *)b->bvh.root_box.ptr;
        refs[i].geo = b->geo_table();
        if (refs[i].geo) t->multi_geo = true;
        tris[i] = b->bvh.prim_count;
    }
    const uint64_t nn = n ? n : 1;
    if (t->inst_dev.reserve(nn * sizeof(solhip_instance)) || t->refs_dev.reserve(nn * sizeof(BlasRef)) ||
        t->inst_tris_dev.reserve(nn * sizeof(uint32_t)))
        return SOLHIP_ERROR_OUT_OF_MEMORY;
    t->n_inst = n;
    if (n) {
        SOLHIP_HIP_CHECK(hipStreamSynchronize(s));
        SOLHIP_HIP_CHECK(hipMemcpy(t->inst_tris_dev.ptr, tris.data(), (uint64_t)n * sizeof(uint32_t), hipMemcpyHostToDevice));
        SOLHIP_HIP_CHECK(hipMemcpy(t->inst_dev.ptr, inst, (uint64_t)n * sizeof(solhip_instance), hipMemcpyHostToDevice));
        SOLHIP_HIP_CHECK(hipMemcpy(t->refs_dev.ptr, refs.data(), (uint64_t)n * sizeof(BlasRef), hipMemcpyHostToDevice));
    }
    (void)cx;
    const int rc = build_tlas(s, (const solhip_instance *)t->inst_dev.ptr, (const BlasRef *)t->refs_dev.ptr, n,
                              t->scratch, t->bvh);
    if (rc) return rc;
    t->built = true;
    t->gen = next_generation();
    return SOLHIP_OK;
}

// Checks the geometries of a BLAS build (BLAS::new's rules, see
// solhip_geometry) and fills the build kernels' descriptors; *tris = the
// BLAS-wide triangle count.
static int geometry_descs(const solhip_geometry *g, uint32_t n, std::vector<GeomDesc> &out, uint32_t *tris)
{
    if (!g || n == 0) return bad_arg("at least one geometry");
    if (n > (1u << 24)) return bad_arg("more than 2^24 geometries");
    out.resize(n);
    uint64_t total = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const solhip_geometry &G = g[k];
        const std::string at = "geometry " + std::to_string(k) + ": ";
        if (G.vertex_stride < 12 || G.vertex_stride % 4) return bad_arg(at + "vertex_stride must be >= 12 and a multiple of 4");
        uint64_t prims;
        GeomDesc &d = out[k];
        d.stride = G.vertex_stride;
        d.first_vertex = G.first_vertex;
        d.vertex_count = G.vertex_count;
        if (G.index_addr) {
            if (G.index_count % 3) return bad_arg(at + "index_count must be a multiple of 3");
            if (G.primitive_offset % 4) return bad_arg(at + "primitive_offset must be a multiple of 4 (UINT32 indices)");
            prims = G.index_count / 3;
            if (prims && (!G.vertex_addr || G.vertex_count == 0))
                return bad_arg(at + "indexed triangles need a vertex address and vertex_count > 0");
            d.vbase = reinterpret_cast<const uint8_t *>(G.vertex_addr);
            d.idx = reinterpret_cast<const uint32_t *>(G.index_addr + G.primitive_offset);
        } else {
            if (G.index_count) return bad_arg(at + "index_count > 0 without an index address");
            prims = G.vertex_count / 3;
            if (prims && !G.vertex_addr) return bad_arg(at + "null vertex address");
            d.vbase = reinterpret_cast<const uint8_t *>(G.vertex_addr + G.primitive_offset);
            d.idx = nullptr;
        }
        if ((uint64_t)G.first_vertex + G.vertex_count > 0xFFFFFFFFull) return bad_arg(at + "first_vertex + vertex_count >= 2^32");
        d.tri_begin = (uint32_t)total;
        d.tri_count = (uint32_t)prims;
        total += prims;
        if (total > 0x7FFFFFFFull) return bad_arg("more than 2^31 - 1 triangles in one BLAS");
    }
    *tris = (uint32_t)total;
    return SOLHIP_OK;
}

// (re)builds b over the geometries: descriptors to the device, then the BVH
static int blas_build_into(hipStream_t s, Blas *b, const solhip_geometry *g, uint32_t n)
{
    std::vector<GeomDesc> descs;
    uint32_t tris = 0;
    int rc = geometry_descs(g, n, descs, &tris);
    if (rc) return rc;
    b->gen = 0;   // stale until the build succeeds
    // descriptors, then the u32 table [n, tri_begin...] that TLAS records
    // point to when there are several geometries
    const uint64_t dbytes = (uint64_t)n * sizeof(GeomDesc);
    std::vector<uint8_t> blob(dbytes + 4ull * (n + 1));
    std::memcpy(blob.data(), descs.data(), dbytes);
    uint32_t *tab = reinterpret_cast<uint32_t *>(blob.data() + dbytes);
    tab[0] = n;
    for (uint32_t k = 0; k < n; ++k) tab[1 + k] = descs[k].tri_begin;
    if (b->geo_dev.reserve(blob.size())) return SOLHIP_ERROR_OUT_OF_MEMORY;
    // the previous build's kernels may still read the table
    SOLHIP_HIP_CHECK(hipStreamSynchronize(s));
    SOLHIP_HIP_CHECK(hipMemcpy(b->geo_dev.ptr, blob.data(), blob.size(), hipMemcpyHostToDevice));
    b->n_geo = n;
    rc = build_blas(s, (const GeomDesc *)b->geo_dev.ptr, n, tris, b->build_flags, b->scratch, b->bvh);
    if (rc) return rc;
    b->gen = next_generation();
    return SOLHIP_OK;
}

// emitter table: counts and offsets on the device, one read-back of the
// total to size the table
int build_lights(hipStream_t s, const DevTables &tb, const Tlas *t, DevBuf &offsets, DevBuf &lights, uint32_t *nl_out,
                 Timer *timer)
{
    const uint32_t n_inst = std::min(t->n_inst, tb.instance_count);
    if (offsets.reserve(((uint64_t)n_inst + 1) * 4)) return SOLHIP_ERROR_OUT_OF_MEMORY;
    uint32_t *offs = (uint32_t *)offsets.ptr;
    int rc = launch_lights_count(s, tb, (const uint32_t *)t->inst_tris_dev.ptr, n_inst, offs, timer);
    if (rc) return rc;
    uint32_t nl = 0;
    SOLHIP_HIP_CHECK(hipMemcpyAsync(&nl, offs + n_inst, 4, hipMemcpyDeviceToHost, s));
    SOLHIP_HIP_CHECK(hipStreamSynchronize(s));
    if (nl) {
        if (lights.reserve((uint64_t)nl * sizeof(LightTri))) return SOLHIP_ERROR_OUT_OF_MEMORY;
        rc = launch_lights_fill(s, tb, offs, n_inst, (LightTri *)lights.ptr, nl, timer);
        if (rc) return rc;
    }
    *nl_out = nl;
    return SOLHIP_OK;
}

// the radiance-cache frame: buffers, scene box, the passes with cache
// queries, then training on the frame's rows
struct NrcHookCtx {
    Program *p;
    const WfArgs *w;
};

static int nrc_query_hook(void *user, hipStream_t s)
{
    NrcHookCtx *h = (NrcHookCtx *)user;
    return mlp_query(s, h->p->nrc_mlp, h->w->nrc.qfeat, h->w->nrc.counts, h->w->nslots, h->w->nrc.qrec,
                     h->w->st1);
}

int trace_nrc(Context *c, hipStream_t s, Program *p, Tlas *t, WfArgs &w, Timer *timer)
{
    if (!p->nrc_mlp) {
        const float lr = p->spec.nrc_learning_rate > 0.0f ? p->spec.nrc_learning_rate : 1e-2f;
        p->nrc_mlp = mlp_new(c, 1u, lr);
        if (!p->nrc_mlp) return SOLHIP_ERROR_DEVICE;
    }
    NrcArgs &N = w.nrc;
    N.enabled = 1;
    N.max_rec = 4;
    const uint64_t npix = (uint64_t)w.W * w.H;
    N.batch_cap = 65536;
    N.ratio = (uint32_t)std::max<uint64_t>(1, npix * w.spp * N.max_rec / N.batch_cap);
    N.train_cap = (uint32_t)std::min<uint64_t>(npix, 2 * (npix / N.ratio) + 1024);
    const uint64_t ns = w.nslots, tc = N.train_cap, nr = tc * N.max_rec;
    if (p->nq_feat.reserve(ns * 128) || p->nq_rec.reserve(ns * sizeof(CacheQuery)) || p->n_counts.reserve(256) ||
        p->n_tidx.reserve(npix * 4) || p->n_tpath.reserve(tc * 4) || p->n_trec_n.reserve(tc * 4) ||
        p->n_tfeat.reserve(nr * 128) || p->n_trec.reserve(nr * 32) || p->n_bfeat.reserve((uint64_t)N.batch_cap * 128) ||
        p->n_btarget.reserve((uint64_t)N.batch_cap * 16))
        return SOLHIP_ERROR_OUT_OF_MEMORY;
    N.qfeat = (uint16_t *)p->nq_feat.ptr;
    N.qrec = (CacheQuery *)p->nq_rec.ptr;
    N.counts = (uint32_t *)p->n_counts.ptr;
    N.tidx = (uint32_t *)p->n_tidx.ptr;
    N.tpath = (uint32_t *)p->n_tpath.ptr;
    N.trec_n = (uint32_t *)p->n_trec_n.ptr;
    N.tfeat = (uint16_t *)p->n_tfeat.ptr;
    N.trec = (float4 *)p->n_trec.ptr;
    N.bfeat = (uint16_t *)p->n_bfeat.ptr;
    N.btarget = (float4 *)p->n_btarget.ptr;
    // scene box from the TLAS root (world space)
    float box[6] = {0, 0, 0, 1, 1, 1};
    if (t->bvh.prim_count) {
        SOLHIP_HIP_CHECK(hipMemcpyAsync(box, t->bvh.root_box.ptr, 24, hipMemcpyDeviceToHost, s));
        SOLHIP_HIP_CHECK(hipStreamSynchronize(s));
    }
    for (int a = 0; a < 3; ++a) {
        const float ext = box[3 + a] - box[a];
        N.lo[a] = box[a];
        N.inv_ext[a] = ext > 0.0f ? 1.0f / ext : 1.0f;
    }
    int rc = launch_fill(s, N.counts, 4, 0u);
    if (rc) return rc;
    NrcHookCtx hc{p, &w};
    PassHook hook{nrc_query_hook, &hc};
    rc = launch_pathtrace(s, w, timer, &hook);
    if (rc) return rc;
    uint32_t rows = 0;
    SOLHIP_HIP_CHECK(hipMemcpyAsync(&rows, N.counts + 2, 4, hipMemcpyDeviceToHost, s));
    SOLHIP_HIP_CHECK(hipStreamSynchronize(s));
    rows = std::min(rows, N.batch_cap);
    const uint32_t steps = p->spec.nrc_train_steps ? p->spec.nrc_train_steps : 4u;
    rc = mlp_train_dev(s, p->nrc_mlp, N.bfeat, N.btarget, rows, steps);
    p->nrc_frames++;
    p->nrc_rows += rows;
    return rc;
}

// SOLHIP_PROG_PATHTRACE: buffers, emitter table, the wavefront passes
// image slices of a frame when the spec leaves it open: with a stream per
// sample, passes are large and slices overlap one another's slowest rays;
// with one stream per pixel, passes are one sample each and slices only add
// launches
constexpr uint32_t kDefaultSlicesStreams = 4;
// paths in flight per pass that the default samples-per-pass aims for
constexpr uint64_t kPathsInFlight = 64ull << 20;

int trace_wavefront(Context *c, hipStream_t s, Program *p, Tlas *t, const ProgArgs &a)
{
    WfArgs w;
    std::memset(&w, 0, sizeof(w));
    w.sv = a.sv;
    w.tb = a.tb;
    w.W = a.W;
    w.H = a.H;
    w.accum = a.accum;
    w.out_rgba8 = a.out_rgba8;
    w.enable_sky = p->spec.enable_sky;
    w.spp = p->spec.spp_per_frame;
    w.max_depth = p->spec.max_depth;
    w.rr_start = p->spec.rr_start;
    w.nee = p->spec.nee;
    w.streams = p->spec.sample_streams ? 1u : 0u;
    w.npix = a.W * a.H;
    w.rmax = 10000.0f;   // pathtrace.rgen:35
    const uint64_t tiles = (uint64_t)((a.W + 7) / 8) * ((a.H + 7) / 8);
    if (tiles * 64 > 0xFFFFFFF0ull) return bad_arg("image too large");
    w.nslots = (uint32_t)(tiles * 64);
    w.block = p->spec.wf_block ? p->spec.wf_block : 64u;
    if ((uint64_t)(w.sv.stack_depth + 2) * sizeof(uint2) * w.block > 65536) {
        set_error("traversal stack of the wavefront kernels exceeds 64 KiB of LDS; use a smaller wf_block");
        return SOLHIP_ERROR_UNSUPPORTED;
    }
    // default: the waves per CU that the traversal kernels' registers allow
    const uint32_t per_cu = p->spec.wf_blocks_per_cu ? p->spec.wf_blocks_per_cu : std::max(1u, wf_threads_per_cu() / w.block);
    w.persist_blocks = c->ncu * per_cu;
    // samples per pass: one stream per pixel traces them one after another;
    // with a stream per sample, enough of them for about kPathsInFlight paths
    uint32_t S = 1;
    if (w.streams && !p->spec.nrc) {
        S = p->spec.wf_samples_per_pass ? p->spec.wf_samples_per_pass
                                         : (uint32_t)std::max<uint64_t>(1, kPathsInFlight / w.nslots);
        S = std::min(S, w.spp);
    }
    if ((uint64_t)S * w.nslots > 0xFFFFFFF0ull) return bad_arg("image too large for wf_samples_per_pass");
    w.S = S;
    w.shade_blocks = (uint32_t)std::min<uint64_t>(((uint64_t)S * w.nslots + w.block - 1) / w.block,
                                                  (uint64_t)c->ncu * 2048 / w.block);
    w.counters = c->counters;
    w.count = a.count;
    const uint64_t npix = (uint64_t)a.W * a.H, ns = (uint64_t)w.S * w.nslots, np = (uint64_t)w.S * npix;
    const uint64_t nheads = 2ull * (w.max_depth + 1);
    // image slices: profiling times every launch on one stream, and the
    // radiance cache works on the whole image between passes
    const uint32_t tiles32 = (uint32_t)tiles;
    uint32_t nslices = p->spec.wf_slices ? p->spec.wf_slices : (p->spec.sample_streams ? kDefaultSlicesStreams : 1u);
    if (p->spec.nrc || c->timer_or_null()) nslices = 1;
    nslices = std::max(1u, std::min(nslices, tiles32));
    if (p->st0.reserve(np * 16) || p->st1.reserve(np * 16) || p->psum.reserve(npix * 16) ||
        (w.streams && p->fsum.reserve(npix * 32)) ||
        p->q0.reserve(ns * sizeof(QRay)) || p->q1.reserve(ns * sizeof(QRay)) || p->hits.reserve(ns * sizeof(QHit)) ||
        p->sq.reserve(ns * sizeof(QRay)) || p->scontrib.reserve(ns * 16) || p->counts.reserve(256ull * nslices) ||
        p->heads.reserve(nheads * wf_head_words() * 4 * nslices) || p->fc.reserve(sizeof(FrameConsts)))
        return SOLHIP_ERROR_OUT_OF_MEMORY;
    FrameConsts fc;
    std::memset(&fc, 0, sizeof(fc));
    fc.cam = a.cam;
    fc.frame = a.frame;
    fc.alpha = a.alpha;
    w.fc = (const FrameConsts *)p->fc.ptr;
    w.st0 = (float4 *)p->st0.ptr;
    w.st1 = (float4 *)p->st1.ptr;
    w.psum = (float4 *)p->psum.ptr;
    w.fsum = w.streams ? (FixSum *)p->fsum.ptr : nullptr;
    w.q[0] = (QRay *)p->q0.ptr;
    w.q[1] = (QRay *)p->q1.ptr;
    w.hits = (QHit *)p->hits.ptr;
    w.sq = (QRay *)p->sq.ptr;
    w.scontrib = (float4 *)p->scontrib.ptr;
    w.counts = (uint32_t *)p->counts.ptr;
    w.heads = (uint32_t *)p->heads.ptr;
    Timer *timer = c->timer_or_null();
    int rc0 = launch_set_frame(s, (FrameConsts *)p->fc.ptr, fc);
    if (rc0) return rc0;
    if (w.nee) {
        uint32_t nl = 0;
        const int rc = build_lights(s, a.tb, t, p->light_offsets, p->lights, &nl, timer);
        if (rc) return rc;
        w.lights = (const LightTri *)p->lights.ptr;
        w.nlights = nl;
    }
    if (p->spec.nrc) return trace_nrc(c, s, p, t, w, timer);
    if (nslices == 1 && (!p->spec.wf_graph || timer)) return launch_pathtrace(s, w, timer);
    // the frame on `main`: fork to the slices' streams, their passes, join,
    // then the image write over the whole image
    auto frame = [&](hipStream_t main) -> int {
        if (nslices == 1) return launch_pathtrace(main, w, nullptr);
        for (uint32_t j = 0; j < nslices; ++j) {
            if (!p->slice_stream[j]) SOLHIP_HIP_CHECK(hipStreamCreateWithFlags(&p->slice_stream[j], hipStreamNonBlocking));
            if (!p->join_ev[j]) SOLHIP_HIP_CHECK(hipEventCreateWithFlags(&p->join_ev[j], hipEventDisableTiming));
        }
        if (!p->fork_ev) SOLHIP_HIP_CHECK(hipEventCreateWithFlags(&p->fork_ev, hipEventDisableTiming));
        SOLHIP_HIP_CHECK(hipEventRecord(p->fork_ev, main));
        for (uint32_t j = 0; j < nslices; ++j) {
            // slice j: tiles [tiles*j/K, tiles*(j+1)/K), its own counters and
            // fetch heads, its part of the queue arrays
            const uint32_t t0 = (uint32_t)((uint64_t)tiles32 * j / nslices);
            const uint32_t t1 = (uint32_t)((uint64_t)tiles32 * (j + 1) / nslices);
            WfArgs ws = w;
            ws.slot_base = t0 * 64u;
            ws.nslots = (t1 - t0) * 64u;
            const uint64_t qb = (uint64_t)w.S * ws.slot_base;   // the slice's part of the queues
            ws.q[0] = w.q[0] + qb;
            ws.q[1] = w.q[1] + qb;
            ws.hits = w.hits + qb;
            ws.sq = w.sq + qb;
            ws.scontrib = w.scontrib + qb;
            ws.counts = w.counts + 64u * j;
            ws.heads = w.heads + nheads * wf_head_words() * j;
            ws.shade_blocks = (uint32_t)std::min<uint64_t>(((uint64_t)w.S * ws.nslots + w.block - 1) / w.block,
                                                           w.shade_blocks);
            hipStream_t ss = p->slice_stream[j];
            SOLHIP_HIP_CHECK(hipStreamWaitEvent(ss, p->fork_ev, 0));
            const int rc = launch_pathtrace_passes(ss, ws, nullptr);
            SOLHIP_HIP_CHECK(hipEventRecord(p->join_ev[j], ss));
            SOLHIP_HIP_CHECK(hipStreamWaitEvent(main, p->join_ev[j], 0));
            if (rc) return rc;
        }
        return launch_pathtrace_finalize(main, w, nullptr);
    };
    if (!p->spec.wf_graph) return frame(s);
    // replay the captured frame; recapture when anything the kernels were
    // launched with differs (buffers, scene pointers, sizes, spec)
    if (!p->gvalid || std::memcmp(&p->gkey, &w, sizeof(w)) != 0) {
        p->drop_graph();
        if (!p->cap_stream) SOLHIP_HIP_CHECK(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
        SOLHIP_HIP_CHECK(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
        const int rc = frame(p->cap_stream);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(p->cap_stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess || !g) {
            set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
            return SOLHIP_ERROR_DEVICE;
        }
        p->graph = g;
        if (hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0) != hipSuccess) {
            p->drop_graph();
            set_error("hipGraphInstantiate failed");
            return SOLHIP_ERROR_DEVICE;
        }
        p->gkey = w;
        p->gvalid = true;
        p->graph_captures++;
    }
    SOLHIP_HIP_CHECK(hipGraphLaunch(p->gexec, s));
    return SOLHIP_OK;
}

}  // namespace solhip

using namespace solhip;

extern "C" {

uint32_t solhip_abi_version(void) { return SOLHIP_ABI_VERSION; }

int32_t solhip_last_error(char *buf, size_t len)
{
    if (!buf || len == 0) return SOLHIP_ERROR_INVALID_ARGUMENT;
    const size_t n = std::min(len - 1, g_last_error.size());
    std::memcpy(buf, g_last_error.data(), n);
    buf[n] = 0;
    return SOLHIP_OK;
}

int32_t solhip_context_create(int32_t device, uint32_t flags, solhip_ctx *out)
{
    (void)flags;
    if (!out) return bad_arg("out");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        set_error("no HIP device available (this library runs on gfx950 only; there is no CPU fallback)");
        return SOLHIP_ERROR_NO_DEVICE;
    }
    if (device < 0 || device >= count) return bad_arg("device index");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        set_error("hipGetDeviceProperties failed");
        return SOLHIP_ERROR_NO_DEVICE;
    }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error(std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950 only");
        return SOLHIP_ERROR_NO_DEVICE;
    }
    SOLHIP_HIP_CHECK(hipSetDevice(device));
    if (const int rc = noalloc_scratch_init(device)) return rc;
    Context *c = new Context();
    c->kind = Kind::Context;
    c->device = device;
    if (prop.multiProcessorCount > 0) c->ncu = (uint32_t)prop.multiProcessorCount;
    if (hipStreamCreate(&c->stream) != hipSuccess ||
        hipMalloc(&c->counters, sizeof(Counters) * SOLHIP_KERNEL_COUNT) != hipSuccess) {
        set_error("context allocation failed");
        delete c;
        return SOLHIP_ERROR_DEVICE;
    }
    (void)hipMemset(c->counters, 0, sizeof(Counters) * SOLHIP_KERNEL_COUNT);
    (void)hipDeviceSynchronize();
    const uint64_t h = reg(c);
    c->ctx = h;
    *out = h;
    return SOLHIP_OK;
}

int32_t solhip_context_destroy(solhip_ctx ctx)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    (void)hipDeviceSynchronize();
    // release every object created in this context
    std::vector<uint64_t> mine;
    {
        std::lock_guard<std::mutex> l(g_mu);
        for (auto &kv : g_objs)
            if (kv.second->ctx == ctx && kv.first != ctx) mine.push_back(kv.first);
    }
    for (uint64_t h : mine) {
        Obj *o = nullptr;
        {
            std::lock_guard<std::mutex> l(g_mu);
            auto it = g_objs.find(h);
            if (it == g_objs.end()) continue;
            o = it->second;
            g_objs.erase(it);
        }
        switch (o->kind) {
        case Kind::Stream: (void)hipStreamDestroy(static_cast<Stream *>(o)->s); break;
        case Kind::Buffer: (void)hipFree(static_cast<Buffer *>(o)->ptr); break;
        case Kind::Blas:
            static_cast<Blas *>(o)->bvh.release();
            static_cast<Blas *>(o)->scratch.release();
            static_cast<Blas *>(o)->geo_dev.release();
            break;
        case Kind::Tlas: {
            Tlas *t = static_cast<Tlas *>(o);
            t->bvh.release();
            t->scratch.release();
            t->inst_dev.release();
            t->refs_dev.release();
            break;
        }
        case Kind::Program: static_cast<Program *>(o)->queue.release(); static_cast<Program *>(o)->qcount.release(); break;
        default: break;
        }
        delete o;
    }
    unreg(ctx, Kind::Context);
    (void)hipStreamDestroy(c->stream);
    (void)hipFree(c->counters);
    delete c;
    return SOLHIP_OK;
}

int32_t solhip_stream_create(solhip_ctx ctx, solhip_stream *out)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    if (!out) return bad_arg("out");
    Stream *s = new Stream();
    s->kind = Kind::Stream;
    s->ctx = ctx;
    if (hipStreamCreate(&s->s) != hipSuccess) {
        delete s;
        set_error("hipStreamCreate failed");
        return SOLHIP_ERROR_DEVICE;
    }
    *out = reg(s);
    return SOLHIP_OK;
}

int32_t solhip_stream_destroy(solhip_ctx ctx, solhip_stream st)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    Stream *s = static_cast<Stream *>(unreg(st, Kind::Stream));
    if (!s) return bad_handle("stream");
    (void)hipStreamSynchronize(s->s);
    (void)hipStreamDestroy(s->s);
    delete s;
    return SOLHIP_OK;
}

int32_t solhip_stream_sync(solhip_ctx ctx, solhip_stream st)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    bool ok;
    hipStream_t s = stream_of(c, st, ok);
    if (!ok) return bad_handle("stream");
    SOLHIP_HIP_CHECK(hipStreamSynchronize(s));
    return SOLHIP_OK;
}

int32_t solhip_buffer_create(solhip_ctx ctx, uint64_t bytes, solhip_buffer *out)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    if (!out) return bad_arg("out");
    Buffer *b = new Buffer();
    b->kind = Kind::Buffer;
    b->ctx = ctx;
    b->bytes = bytes;
    if (hipMalloc(&b->ptr, bytes ? bytes : 16) != hipSuccess) {
        delete b;
        set_error("hipMalloc of " + std::to_string(bytes) + " bytes failed");
        return SOLHIP_ERROR_OUT_OF_MEMORY;
    }
    *out = reg(b);
    return SOLHIP_OK;
}

int32_t solhip_buffer_destroy(solhip_ctx ctx, solhip_buffer bh)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    Buffer *b = static_cast<Buffer *>(unreg(bh, Kind::Buffer));
    if (!b) return bad_handle("buffer");
    (void)hipDeviceSynchronize();
    (void)hipFree(b->ptr);
    delete b;
    return SOLHIP_OK;
}

int32_t solhip_buffer_upload(solhip_ctx ctx, solhip_buffer bh, uint64_t offset, const void *src, uint64_t bytes)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    Buffer *b = get<Buffer>(bh, Kind::Buffer);
    if (!b) return bad_handle("buffer");
    if (offset > b->bytes || bytes > b->bytes - offset) return bad_arg("upload range exceeds buffer");
    if (bytes == 0) return SOLHIP_OK;
    if (!src) return bad_arg("src");
    SOLHIP_HIP_CHECK(hipMemcpy((char *)b->ptr + offset, src, bytes, hipMemcpyHostToDevice));
    return SOLHIP_OK;
}

int32_t solhip_buffer_download(solhip_ctx ctx, solhip_buffer bh, uint64_t offset, void *dst, uint64_t bytes)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    Buffer *b = get<Buffer>(bh, Kind::Buffer);
    if (!b) return bad_handle("buffer");
    if (offset > b->bytes || bytes > b->bytes - offset) return bad_arg("download range exceeds buffer");
    if (bytes == 0) return SOLHIP_OK;
    if (!dst) return bad_arg("dst");
    SOLHIP_HIP_CHECK(hipMemcpy(dst, (const char *)b->ptr + offset, bytes, hipMemcpyDeviceToHost));
    return SOLHIP_OK;
}

int32_t solhip_buffer_fill(solhip_ctx ctx, solhip_stream st, solhip_buffer bh, uint32_t value)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    bool ok;
    hipStream_t s = stream_of(c, st, ok);
    if (!ok) return bad_handle("stream");
    Buffer *b = get<Buffer>(bh, Kind::Buffer);
    if (!b) return bad_handle("buffer");
    return launch_fill(s, (uint32_t *)b->ptr, b->bytes / 4, value);
}

int32_t solhip_buffer_address(solhip_ctx ctx, solhip_buffer bh, uint64_t *addr)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    Buffer *b = get<Buffer>(bh, Kind::Buffer);
    if (!b) return bad_handle("buffer");
    if (!addr) return bad_arg("addr");
    *addr = (uint64_t)b->ptr;
    return SOLHIP_OK;
}

int32_t solhip_blas_build(solhip_ctx ctx, solhip_stream st, const solhip_geometry *g, uint32_t n, uint32_t flags,
                          solhip_blas *out)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    bool ok;
    hipStream_t s = stream_of(c, st, ok);
    if (!ok) return bad_handle("stream");
    if (!g || !out) return bad_arg("geometry / out");
    if (flags & ~(uint32_t)SOLHIP_BUILD_PREFER_FAST_TRACE) return bad_arg("unknown build flags");
    Blas *b = new Blas();
    b->kind = Kind::Blas;
    b->ctx = ctx;
    b->build_flags = flags;
    int rc = blas_build_into(s, b, g, n);
    if (rc) {
        (void)hipStreamSynchronize(s);
        b->bvh.release();
        b->scratch.release();
        b->geo_dev.release();
        delete b;
        return rc;
    }
    *out = reg(b);
    return SOLHIP_OK;
}

int32_t solhip_blas_rebuild(solhip_ctx ctx, solhip_stream st, solhip_blas bh, const solhip_geometry *g, uint32_t n)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    bool ok;
    hipStream_t s = stream_of(c, st, ok);
    if (!ok) return bad_handle("stream");
    Blas *b = get<Blas>(bh, Kind::Blas);
    if (!b) return bad_handle("BLAS");
    return blas_build_into(s, b, g, n);
}

int32_t solhip_blas_destroy(solhip_ctx ctx, solhip_blas bh)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    Blas *b = static_cast<Blas *>(unreg(bh, Kind::Blas));
    if (!b) return bad_handle("BLAS");
    (void)hipDeviceSynchronize();
    b->bvh.release();
    b->scratch.release();
    b->geo_dev.release();
    delete b;
    return SOLHIP_OK;
}

int32_t solhip_tlas_build(solhip_ctx ctx, solhip_stream st, const solhip_instance *inst, uint32_t n, solhip_tlas *out)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    bool ok;
    hipStream_t s = stream_of(c, st, ok);
    if (!ok) return bad_handle("stream");
    if (!out || (n && !inst)) return bad_arg("instances / out");
    if (n > 0xFFFFFFu) return bad_arg("more than 2^24 instances");
    Tlas *t = new Tlas();
    t->kind = Kind::Tlas;
    t->ctx = ctx;
    int rc = tlas_build_into(c, s, t, inst, n);
    if (rc) {
        t->bvh.release();
        t->scratch.release();
        t->inst_dev.release();
        t->refs_dev.release();
        delete t;
        return rc;
    }
    *out = reg(t);
    return SOLHIP_OK;
}

int32_t solhip_tlas_rebuild(solhip_ctx ctx, solhip_stream st, solhip_tlas th, const solhip_instance *inst, uint32_t n)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    bool ok;
    hipStream_t s = stream_of(c, st, ok);
    if (!ok) return bad_handle("stream");
    Tlas *t = get<Tlas>(th, Kind::Tlas);
    if (!t) return bad_handle("TLAS");
    if (n && !inst) return bad_arg("instances");
    if (n > 0xFFFFFFu) return bad_arg("more than 2^24 instances");
    return tlas_build_into(c, s, t, inst, n);
}

int32_t solhip_tlas_destroy(solhip_ctx ctx, solhip_tlas th)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    Tlas *t = static_cast<Tlas *>(unreg(th, Kind::Tlas));
    if (!t) return bad_handle("TLAS");
    (void)hipDeviceSynchronize();
    t->bvh.release();
    t->scratch.release();
    t->inst_dev.release();
    t->refs_dev.release();
    delete t;
    return SOLHIP_OK;
}

static void fill_info(const DeviceBvh &b, solhip_bvh_info *info)
{
    info->nodes = b.node_count;
    info->prims = b.prim_count;
    info->depth = b.depth;
    info->leaf_max = b.leaf_max;
    info->bytes = (uint64_t)b.node_count * sizeof(WideNode) +
                  (uint64_t)b.prim_count * (b.kind == 0 ? sizeof(TriRec) : sizeof(InstRec) + 4);
}

int32_t solhip_blas_info(solhip_ctx ctx, solhip_blas bh, solhip_bvh_info *info)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    Blas *b = get<Blas>(bh, Kind::Blas);
    if (!b) return bad_handle("BLAS");
    if (!info) return bad_arg("info");
    fill_info(b->bvh, info);
    return SOLHIP_OK;
}

int32_t solhip_tlas_info(solhip_ctx ctx, solhip_tlas th, solhip_bvh_info *info)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    Tlas *t = get<Tlas>(th, Kind::Tlas);
    if (!t) return bad_handle("TLAS");
    if (!info) return bad_arg("info");
    fill_info(t->bvh, info);
    return SOLHIP_OK;
}

int32_t solhip_trace_rays(solhip_ctx ctx, solhip_stream st, solhip_tlas th, solhip_buffer rays, solhip_buffer hits,
                          uint32_t n, uint32_t flags)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    bool ok;
    hipStream_t s = stream_of(c, st, ok);
    if (!ok) return bad_handle("stream");
    Tlas *t = get<Tlas>(th, Kind::Tlas);
    if (!t) return bad_handle("TLAS");
    Buffer *rb = get<Buffer>(rays, Kind::Buffer);
    Buffer *hb = get<Buffer>(hits, Kind::Buffer);
    if (!rb || !hb) return bad_handle("buffer");
    if ((uint64_t)n * sizeof(solhip_ray) > rb->bytes) return bad_arg("ray buffer smaller than n rays");
    if ((uint64_t)n * sizeof(solhip_hit) > hb->bytes) return bad_arg("hit buffer smaller than n hits");
    if (const int rc = check_fresh(t)) return rc;
    if (stack_depth_of(t) > (uint32_t)kMaxStack) {
        set_error("acceleration structures too deep for the traversal stack");
        return SOLHIP_ERROR_UNSUPPORTED;
    }
    const int rc = launch_trace_rays(s, view_of(t), (const solhip_ray *)rb->ptr, (solhip_hit *)hb->ptr, n, flags,
                                     c->counters + SOLHIP_KERNEL_TRACE_RAYS, c->timer_or_null());
    if (rc || !t->multi_geo) return rc;
    return launch_split_geometry(s, (solhip_hit *)hb->ptr, n, (const BlasRef *)t->refs_dev.ptr, t->n_inst);
}

void solhip_spec_default(uint32_t kind, solhip_spec *spec)
{
    if (!spec) return;
    std::memset(spec, 0, sizeof(*spec));
    spec->enable_sky = 0;       // pathtrace.rmiss:5
    spec->spp_per_frame = kind == SOLHIP_PROG_AO_GLSL ? 4u : 8u;   // ao.rgen:43, pathtrace.rgen:44
    spec->max_depth = kind == SOLHIP_PROG_AO_GLSL ? 4u : 32u;      // ao.rgen:42, pathtrace.rgen:43
    spec->n_ao = 8;
    spec->ao_tmin = 2e-4f;
    spec->ao_tmax = 0.2f;
    spec->ao_any_hit = 0;
    spec->rr_start = 0;
    spec->nee = 0;
    spec->wf_blocks_per_cu = 0;
    spec->wf_block = 0;
}

int32_t solhip_program_create(solhip_ctx ctx, uint32_t kind, const solhip_spec *spec, solhip_program *out)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    if (!out) return bad_arg("out");
    if (kind < SOLHIP_PROG_AO_GLSL || kind > SOLHIP_PROG_PATHTRACE) return bad_arg("program kind");
    Program *p = new Program();
    p->kind = Kind::Program;
    p->ctx = ctx;
    p->kind_id = kind;
    if (spec)
        p->spec = *spec;
    else
        solhip_spec_default(kind, &p->spec);
    if (kind == SOLHIP_PROG_AO) {
        const uint32_t na = p->spec.n_ao;
        if (na == 0 || na > 64 || (na & (na - 1))) {
            delete p;
            set_error("n_ao must be a power of two in [1, 64]");
            return SOLHIP_ERROR_UNSUPPORTED;
        }
    }
    if ((kind == SOLHIP_PROG_PATHTRACE_GLSL || kind == SOLHIP_PROG_PATHTRACE) && (p->spec.spp_per_frame == 0)) {
        delete p;
        return bad_arg("spp_per_frame must be > 0");
    }
    if (kind == SOLHIP_PROG_PATHTRACE) {
        const uint32_t blk = p->spec.wf_block;
        if (p->spec.max_depth > 4096 || p->spec.enable_sky > 2 || (blk && (blk % 64 || blk > 256)) ||
            p->spec.wf_slices > Program::kMaxSlices || p->spec.sample_streams > 1 ||
            (p->spec.wf_samples_per_pass > 1 && !p->spec.sample_streams)) {
            delete p;
            set_error("pathtrace spec: max_depth <= 4096, enable_sky in {0,1,2}, wf_block a multiple of 64 <= 256, "
                      "wf_slices <= 16, sample_streams in {0,1}, wf_samples_per_pass > 1 only with sample_streams");
            return SOLHIP_ERROR_UNSUPPORTED;
        }
    }
    *out = reg(p);
    return SOLHIP_OK;
}

int32_t solhip_program_destroy(solhip_ctx ctx, solhip_program ph)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    Program *p = static_cast<Program *>(unreg(ph, Kind::Program));
    if (!p) return bad_handle("program");
    (void)hipDeviceSynchronize();
    p->release_all();
    delete p;
    return SOLHIP_OK;
}

int32_t solhip_trace(solhip_ctx ctx, solhip_stream st, solhip_program ph, solhip_tlas th,
                     const solhip_scene_tables *tables, const solhip_scene_uniforms *u, int32_t accum_start_frame,
                     const solhip_images *images, uint32_t W, uint32_t H)
{
    return trace_program(ctx, st, ph, th, tables, u, accum_start_frame, images, W, H, nullptr);
}

}  // extern "C"

namespace solhip {

int trace_program(solhip_ctx ctx, solhip_stream st, solhip_program ph, solhip_tlas th,
                  const solhip_scene_tables *tables, const solhip_scene_uniforms *u, int32_t accum_start_frame,
                  const solhip_images *images, uint32_t W, uint32_t H, const FixedScratch *fixed)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    bool ok;
    hipStream_t s = stream_of(c, st, ok);
    if (!ok) return bad_handle("stream");
    Program *p = get<Program>(ph, Kind::Program);
    if (!p) return bad_handle("program");
    Tlas *t = get<Tlas>(th, Kind::Tlas);
    if (!t) return bad_handle("TLAS");
    if (fixed && p->kind_id == SOLHIP_PROG_PATHTRACE) {
        set_error("the wavefront path tracer sizes its path state per image: no fixed-scratch variant");
        return SOLHIP_ERROR_UNSUPPORTED;
    }
    // the fixed-scratch variant records no timing events (they would be allocations)
    Timer *timer = fixed ? nullptr : c->timer_or_null();
    if (!u || !images) return bad_arg("uniforms / images");
    if (W == 0 || H == 0) return SOLHIP_OK;
    Buffer *acc = get<Buffer>(images->accum, Kind::Buffer);
    if (!acc) return bad_handle("accum buffer");
    const uint64_t npix = (uint64_t)W * H;
    if (acc->bytes < npix * 16) return bad_arg("accum buffer smaller than W*H*16 bytes");
    const bool needs_tables = p->kind_id != SOLHIP_PROG_DEBUG;
    if (needs_tables) {
        if (!tables || !tables->instances_addr || !tables->vertex_table_addr || !tables->index_table_addr)
            return bad_arg("scene tables");
        if (tables->index_size != 4 && tables->index_size != 8) return bad_arg("index_size must be 4 or 8");
        if ((p->kind_id == SOLHIP_PROG_PATHTRACE_GLSL || p->kind_id == SOLHIP_PROG_PATHTRACE) &&
            !tables->material_table_addr)
            return bad_arg("pathtrace needs the material table");
    }
    if (const int rc = check_fresh(t)) return rc;
    if (p->kind_id == SOLHIP_PROG_PATHTRACE && p->spec.sample_streams &&
        (uint64_t)p->spec.spp_per_frame * npix > (1ull << 32))
        return bad_arg("sample_streams: spp_per_frame * W * H must not exceed 2^32 (sample s of pixel p seeds "
                       "tea(p + s * W * H, frame) in 32 bits)");
    if (stack_depth_of(t) > (uint32_t)kMaxStack) {
        set_error("acceleration structures too deep for the traversal stack");
        return SOLHIP_ERROR_UNSUPPORTED;
    }
    ProgArgs a;
    std::memset(&a, 0, sizeof(a));
    a.sv = view_of(t);
    if (tables) {
        a.tb.instances = (const solhip_scene_instance *)tables->instances_addr;
        a.tb.vertex_table = (const uint64_t *)tables->vertex_table_addr;
        a.tb.index_table = (const uint64_t *)tables->index_table_addr;
        a.tb.material_table = (const uint64_t *)tables->material_table_addr;
        a.tb.instance_count = tables->instance_count;
        a.tb.index_size = tables->index_size;
    }
    std::memcpy(a.cam.view_inverse, u->view_inverse, 64);
    std::memcpy(a.cam.projection_inverse, u->projection_inverse, 64);
    a.W = W;
    a.H = H;
    a.frame = u->frame[2];
    // ao.rgen:78 / pathtrace.rgen:90: 1 / float(frame.z - accum_start_frame + 1), uint arithmetic
    const uint32_t denom = u->frame[2] - (uint32_t)accum_start_frame + 1u;
    a.alpha = 1.0f / (float)denom;
    a.accum = (float4 *)acc->ptr;
    if (images->out_rgba8) {
        Buffer *o = get<Buffer>(images->out_rgba8, Kind::Buffer);
        if (!o) return bad_handle("out_rgba8 buffer");
        if (o->bytes < npix * 4) return bad_arg("out_rgba8 smaller than W*H*4 bytes");
        a.out_rgba8 = (uint32_t *)o->ptr;
    }
    if (p->kind_id == SOLHIP_PROG_AO_GLSL) {
        Buffer *bn = get<Buffer>(images->blue_noise, Kind::Buffer);
        if (!bn) return bad_handle("blue_noise buffer");
        if (images->blue_w == 0 || images->blue_h == 0 ||
            bn->bytes < (uint64_t)images->blue_w * images->blue_h * 4)
            return bad_arg("blue-noise size");
        a.blue = (const uint32_t *)bn->ptr;
        a.blue_w = images->blue_w;
        a.blue_h = images->blue_h;
    }
    a.n_ao = p->spec.n_ao;
    a.ao_tmin = p->spec.ao_tmin;
    a.ao_tmax = p->spec.ao_tmax;
    a.ao_any = p->spec.ao_any_hit;
    a.enable_sky = p->spec.enable_sky;
    a.spp = p->spec.spp_per_frame;
    a.max_depth = p->spec.max_depth;
    a.counters = c->counters;
    a.count = (c->profile & SOLHIP_PROFILE_COUNT) != 0;
    if (images->hits_out) {
        Buffer *hb = get<Buffer>(images->hits_out, Kind::Buffer);
        if (!hb) return bad_handle("hits_out buffer");
        const uint64_t need = p->kind_id == SOLHIP_PROG_AO ? npix * (1 + a.n_ao) * sizeof(solhip_hit) : 0;
        if (hb->bytes < need) return bad_arg("hits_out too small");
        a.hits_out = (solhip_hit *)hb->ptr;
    }
    switch (p->kind_id) {
    case SOLHIP_PROG_AO: {
        const uint64_t hb = ao_head_bytes();
        if (fixed) {
            const AoScratchLayout L = ao_scratch_layout(npix, a.n_ao);
            if (L.total > fixed->bytes) {
                set_error("W*H*(32 + n_ao) + " + std::to_string(L.total - npix * (32 + a.n_ao)) + " = " +
                          std::to_string(L.total) + " bytes exceed the fixed scratch of " +
                          std::to_string(fixed->bytes) + " bytes");
                return SOLHIP_ERROR_INVALID_ARGUMENT;
            }
            uint8_t *base = (uint8_t *)fixed->ptr;
            a.qcount = (uint32_t *)(base + L.qcount);
            a.aoheads = hb ? (uint32_t *)(base + L.heads) : nullptr;
            a.queue = (AoEntry *)(base + L.queue);
            a.occl = base + L.occl;
        } else {
            if (p->queue.reserve(npix * sizeof(AoEntry)) || p->qcount.reserve(256) ||
                p->occl.reserve(npix * a.n_ao) || (hb && p->aoheads.reserve(hb)))
                return SOLHIP_ERROR_OUT_OF_MEMORY;
            a.aoheads = hb ? (uint32_t *)p->aoheads.ptr : nullptr;
            a.queue = (AoEntry *)p->queue.ptr;
            a.qcount = (uint32_t *)p->qcount.ptr;
            a.occl = (uint8_t *)p->occl.ptr;
        }
        a.ncu = c->ncu;
        int rc = launch_fill(s, a.qcount, 2, 0u);
        if (rc) return rc;
        rc = launch_ao(s, a, timer);
        if (rc || !a.hits_out || !t->multi_geo) return rc;
        return launch_split_geometry(s, a.hits_out, npix * (1 + a.n_ao), (const BlasRef *)t->refs_dev.ptr,
                                     t->n_inst);
    }
    case SOLHIP_PROG_AO_GLSL: return launch_ao_glsl(s, a, timer);
    case SOLHIP_PROG_PATHTRACE_GLSL: return launch_pathtrace_glsl(s, a, timer);
    case SOLHIP_PROG_DEBUG: return launch_debug(s, a, timer);
    case SOLHIP_PROG_PATHTRACE: return trace_wavefront(c, s, p, t, a);
    default: return bad_arg("program kind");
    }
}

AoScratchLayout ao_scratch_layout(uint64_t npix, uint32_t n_ao)
{
    auto up = [](uint64_t x) { return (x + 255) & ~255ull; };
    AoScratchLayout L;
    L.qcount = 0;
    L.heads = 256;
    L.queue = L.heads + up(ao_head_bytes());
    L.occl = L.queue + npix * sizeof(AoEntry);
    L.total = L.occl + npix * n_ao;
    return L;
}

}  // namespace solhip

extern "C" {

int32_t solhip_pathtrace_autotune(solhip_ctx ctx, solhip_stream st, solhip_tlas th, const solhip_scene_tables *tables,
                                  const solhip_scene_uniforms *u, const solhip_images *images, uint32_t W, uint32_t H,
                                  const solhip_spec *spec, uint32_t frames, const char *cache_path, solhip_spec *best,
                                  float *best_ms)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    bool ok;
    hipStream_t s = stream_of(c, st, ok);
    if (!ok) return bad_handle("stream");
    Tlas *t = get<Tlas>(th, Kind::Tlas);
    if (!t) return bad_handle("TLAS");
    if (!spec || !best || !u || !images || frames == 0) return bad_arg("spec / best / uniforms / images / frames");
    if (const int rc = check_fresh(t)) return rc;
    // workload key: device, CUs, image, estimator settings, scene size
    uint64_t tris = 0;
    for (uint64_t h : t->blas_handles) {
        const Blas *b = get<Blas>(h, Kind::Blas);
        if (b) tris += b->bvh.prim_count;
    }
    char key[256];
    std::snprintf(key, sizeof(key),
                  "gfx950 cu=%u %ux%u spp=%u depth=%u nee=%u rr=%u sky=%u streams=%u per_pass=%u inst=%u blas_tris=%llu",
                  c->ncu, W, H, spec->spp_per_frame, spec->max_depth, spec->nee, spec->rr_start, spec->enable_sky,
                  spec->sample_streams, spec->wf_samples_per_pass, t->n_inst, (unsigned long long)tris);
    if (cache_path) {
        if (FILE *f = std::fopen(cache_path, "r")) {
            char line[512];
            while (std::fgets(line, sizeof(line), f)) {
                const size_t kl = std::strlen(key);
                unsigned blk = 0, pcu = 0, gr = 0, sl = 0;
                float ms = 0.0f;
                if (std::strncmp(line, key, kl) == 0 && line[kl] == ' ' &&
                    std::sscanf(line + kl, " -> block=%u per_cu=%u graph=%u slices=%u ms=%f", &blk, &pcu, &gr, &sl,
                                &ms) == 5) {
                    *best = *spec;
                    best->wf_block = blk;
                    best->wf_blocks_per_cu = pcu;
                    best->wf_graph = gr;
                    best->wf_slices = sl;
                    if (best_ms) *best_ms = ms;
                    std::fclose(f);
                    return SOLHIP_OK;
                }
            }
            std::fclose(f);
        }
    }
    const uint32_t saved_profile = c->profile;
    c->profile = 0;   // per-launch events would defeat the graph
    const uint32_t depth = stack_depth_of(t);
    float best_t = 1e30f;
    solhip_spec winner = *spec;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SOLHIP_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = SOLHIP_ERROR_DEVICE;
    // one candidate: an untimed frame (allocation, capture), then the timed ones
    auto time_candidate = [&](const solhip_spec &cand) {
        solhip_program ph = 0;
        rc = solhip_program_create(ctx, SOLHIP_PROG_PATHTRACE, &cand, &ph);
        if (rc) return;
        solhip_scene_uniforms uf = *u;
        for (uint32_t f = 0; f <= frames && !rc; ++f) {
            if (f == 1) (void)hipEventRecord(e0, s);
            uf.frame[2] = u->frame[2] + f;
            rc = solhip_trace(ctx, st, ph, th, tables, &uf, (int32_t)uf.frame[2], images, W, H);
        }
        if (!rc) {
            (void)hipEventRecord(e1, s);
            (void)hipEventSynchronize(e1);
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            ms /= (float)frames;
            if (ms < best_t) {
                best_t = ms;
                winner = cand;
            }
        }
        (void)solhip_program_destroy(ctx, ph);
    };
    // launch shape, then (per-sample streams only) the image slices of the winner
    for (uint32_t blk : {64u, 128u, 256u}) {
        if ((uint64_t)(depth + 2) * 8 * blk > 65536) continue;
        const uint32_t base = std::max(1u, wf_threads_per_cu() / blk);
        for (uint32_t pcu : {base / 2 ? base / 2 : 1u, base, 2 * base})
            for (uint32_t gr : {0u, 1u}) {
                if (rc) break;
                solhip_spec cand = *spec;
                cand.wf_block = blk;
                cand.wf_blocks_per_cu = pcu;
                cand.wf_graph = gr;
                time_candidate(cand);
            }
    }
    if (spec->sample_streams && !spec->nrc) {
        const solhip_spec shape = winner;
        for (uint32_t sl : {1u, 2u, 4u, 8u}) {
            if (rc) break;
            solhip_spec cand = shape;
            cand.wf_slices = sl;
            time_candidate(cand);
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    c->profile = saved_profile;
    if (rc) return rc;
    *best = winner;
    if (best_ms) *best_ms = best_t;
    if (cache_path) {
        if (FILE *f = std::fopen(cache_path, "a")) {
            std::fprintf(f, "%s -> block=%u per_cu=%u graph=%u slices=%u ms=%.4f\n", key, winner.wf_block,
                         winner.wf_blocks_per_cu, winner.wf_graph, winner.wf_slices, best_t);
            std::fclose(f);
        }
    }
    return SOLHIP_OK;
}

int32_t solhip_stats(solhip_ctx ctx, solhip_counters *out, uint32_t reset)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    if (!out) return bad_arg("out");
    SOLHIP_HIP_CHECK(hipDeviceSynchronize());
    Counters h[SOLHIP_KERNEL_COUNT];
    SOLHIP_HIP_CHECK(hipMemcpy(h, c->counters, sizeof(h), hipMemcpyDeviceToHost));
    std::memset(out, 0, sizeof(*out));
    for (int k = 0; k < SOLHIP_KERNEL_COUNT; ++k) {
        out->rays += h[k].rays;
        out->node_fetches += h[k].nodes;
        out->tri_fetches += h[k].tris;
        out->inst_fetches += h[k].insts;
        out->overflow += h[k].overflow;
    }
    if (reset) {
        SOLHIP_HIP_CHECK(hipMemset(c->counters, 0, sizeof(h)));
        SOLHIP_HIP_CHECK(hipDeviceSynchronize());
    }
    return SOLHIP_OK;
}

int32_t solhip_profile(solhip_ctx ctx, uint32_t flags)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    if (flags & ~(SOLHIP_PROFILE_TIME | SOLHIP_PROFILE_COUNT)) return bad_arg("profile flags");
    c->profile = flags;
    return SOLHIP_OK;
}

int32_t solhip_kernel_stats_get(solhip_ctx ctx, uint32_t kernel, solhip_kernel_stats *out, uint32_t reset)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    if (!out || kernel >= SOLHIP_KERNEL_COUNT) return bad_arg("kernel / out");
    SOLHIP_HIP_CHECK(hipDeviceSynchronize());
    c->resolve_events();
    Counters h;
    SOLHIP_HIP_CHECK(hipMemcpy(&h, c->counters + kernel, sizeof(h), hipMemcpyDeviceToHost));
    out->launches = c->launches[kernel];
    out->total_ms = c->total_ms[kernel];
    out->counters.rays = h.rays;
    out->counters.node_fetches = h.nodes;
    out->counters.tri_fetches = h.tris;
    out->counters.inst_fetches = h.insts;
    out->counters.overflow = h.overflow;
    if (reset) {
        c->launches[kernel] = 0;
        c->total_ms[kernel] = 0.0;
        SOLHIP_HIP_CHECK(hipMemset(c->counters + kernel, 0, sizeof(h)));
        SOLHIP_HIP_CHECK(hipDeviceSynchronize());
    }
    return SOLHIP_OK;
}

}  // extern "C"
