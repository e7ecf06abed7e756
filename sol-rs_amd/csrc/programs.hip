// programs.hip -- the ray-tracing programs as kernels.
//
//   SOLHIP_PROG_AO            plain ray-traced AO as a wavefront:
//                             k_ao_primary  generate + extend + shade of the
//                                           primary rays; hits are compacted
//                                           into the AO queue with one atomic
//                                           per wave (ballot + mbcnt)
//                             k_ao_rays     n_ao rays per queue entry, the
//                                           lanes of one pixel reduce their
//                                           occlusion with a ballot
//   SOLHIP_PROG_AO_GLSL       ao.rgen / ao.rchit / ao.rmiss
//   SOLHIP_PROG_PATHTRACE_GLSL pathtrace.rgen / pathtrace.rchit / pathtrace.rmiss
//   SOLHIP_PROG_DEBUG         debug.rgen / debug.rchit / debug.rmiss
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "shading.h"
#include "traverse.h"
#include "traverse_step.h"
#include "fetch.h"

namespace solhip {

// minimum waves per SIMD requested for the AO ray kernel: it is VALU-issue
// bound, so the register budget is traded against latency hiding
#ifndef SOLHIP_AO_WAVES
#define SOLHIP_AO_WAVES 5
#endif
#define SOLHIP_AO_RAYS_OCCUPANCY __attribute__((amdgpu_waves_per_eu(SOLHIP_AO_WAVES)))
// persistent AO rays: a wave takes new rays for its idle lanes once at least
// this many are idle (0: one ray per lane, no refill)
#ifndef SOLHIP_AO_REFILL
#define SOLHIP_AO_REFILL 16
#endif
// fetch partitions of the persistent AO kernel (fetch.h)
#ifndef SOLHIP_AO_PARTS
#define SOLHIP_AO_PARTS 64
#endif

// ---------------------------------------------------------------------------
// plain ray-traced AO (wavefront)
// ---------------------------------------------------------------------------
// the primary hit of pixel pix (ray direction d, RNG after the jitter): a hit
// becomes an AO queue entry (returns true); a miss has nothing to occlude and
// takes the value 1 at once
__device__ __forceinline__ bool ao_primary_result(const ProgArgs &A, uint32_t pix, uint32_t rng, V3 d, const Hit &h,
                                                  AoEntry &e)
{
    if (A.hits_out) store_hit(A.hits_out + pix, h);
    if (h.inst != SOLHIP_MISS) {
        const Surf sf = fetch_surface(A.tb, h.inst, h.prim, h.u, h.v);
        const float sg = sgn(dot(neg(d), sf.nrm));
        e.pixel = pix;
        e.px = sf.pos.x;
        e.py = sf.pos.y;
        e.pz = sf.pos.z;
        e.nx = sf.nrm.x * sg;
        e.ny = sf.nrm.y * sg;
        e.nz = sf.nrm.z * sg;
        e.rng = rng;
        return true;
    }
    float4 *px = A.accum + pix;
    const float4 old = *px;
    *px = make_float4(mixf(old.x, 1.0f, A.alpha), mixf(old.y, 1.0f, A.alpha), mixf(old.z, 1.0f, A.alpha), 1.0f);
    if (A.hits_out) {
        solhip_hit m;
        m.t = -1.0f;
        m.u = 0.0f;
        m.v = 0.0f;
        m.inst = SOLHIP_MISS;
        m.prim = SOLHIP_MISS;
        m.geom = SOLHIP_MISS;
        solhip_hit *dst = A.hits_out + (uint64_t)A.W * A.H + (uint64_t)pix * A.n_ao;
        for (uint32_t k = 0; k < A.n_ao; ++k) dst[k] = m;
    }
    return false;
}

// the wave's hits into the AO queue with one atomic; every lane must call it
__device__ __forceinline__ void ao_queue_append(const ProgArgs &A, bool hit, const AoEntry &e)
{
    const uint64_t mask = __ballot(hit);
    if (mask) {
        const uint32_t lane = threadIdx.x & 63;
        uint32_t base = 0;
        if (lane == __builtin_ctzll(mask)) base = atomicAdd(A.qcount, (uint32_t)__popcll(mask));
        base = __shfl(base, __builtin_ctzll(mask));
        if (hit) {
            const uint32_t slot = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
            A.queue[slot] = e;
        }
    }
}

// camera ray of pixel (x, y) with the jitter of ao.rgen:49-54; rng is the
// pixel's stream after the two jitter numbers
__device__ __forceinline__ void ao_camera(const ProgArgs &A, uint32_t x, uint32_t y, uint32_t &rng, V3 &o, V3 &d)
{
    rng = tea(x + y * A.W, A.frame);
    const float jx = next_rand(rng);
    const float jy = next_rand(rng);
    camera_ray(A.cam.view_inverse, A.cam.projection_inverse, x, y, A.W, A.H, jx, jy, o, d);
}

template <bool COUNT>
__global__ void __launch_bounds__(kProgBlock) k_ao_primary(ProgArgs A)
{
    extern __shared__ uint2 stack[];
    uint32_t x, y;
    const bool active = tile_pixel(A.W, A.H, x, y);
    TravCounts tc{0, 0, 0, false};
    bool hit = false;
    AoEntry e;
    if (active) {
        uint32_t rng;
        V3 o, d;
        ao_camera(A, x, y, rng, o, d);
        const float tmin = fmaxf(1.0f, length3(o)) * 1e-3f;
        // the step traversal run to the end: the same visit order, hits and
        // counts as trace_ray, in fewer registers (5 waves per SIMD instead of 4)
        TravState st;
        trav_init(st, A.sv, o, d, tmin, 10000.0f);
        while (!trav_step<false, COUNT>(st, A.sv, 0xFFu, stack + threadIdx.x, kProgBlock, tc)) {
        }
        const Hit h = trav_result(st);
        hit = ao_primary_result(A, y * A.W + x, rng, d, h, e);
    }
    ao_queue_append(A, hit, e);
    count_rays<COUNT>(A.counters + SOLHIP_KERNEL_AO_PRIMARY, active ? 1u : 0u, tc);
}

// one lane per (queue entry, AO ray); n_ao is a power of two <= 64, so the
// rays of one pixel are adjacent lanes of one wave
template <bool COUNT>
__global__ void __launch_bounds__(kProgBlock) SOLHIP_AO_RAYS_OCCUPANCY k_ao_rays(ProgArgs A)
{
    extern __shared__ uint2 stack[];
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
    const uint64_t t = (uint64_t)lb * kProgBlock + threadIdx.x;
    const uint32_t count = *A.qcount;
    const uint32_t shift = __builtin_ctz(A.n_ao);
    const uint64_t q = t >> shift;
    const uint32_t k = (uint32_t)(t & (A.n_ao - 1));
    const bool active = q < count;
    TravCounts tc{0, 0, 0, false};
    bool occl = false;
    uint32_t pix = 0;
    if (active) {
        const AoEntry e = A.queue[q];
        pix = e.pixel;
        uint32_t rng = pcg_advance(e.rng, 2u * k);
        const float xi0 = next_rand(rng);
        const float xi1 = next_rand(rng);
        const V3 dir = sample_cosine(v3(e.nx, e.ny, e.nz), xi0, xi1);
        const V3 o = v3(e.px, e.py, e.pz);
        Hit h;
        if (A.ao_any)
            h = trace_ray<true, COUNT>(A.sv, o, dir, A.ao_tmin, A.ao_tmax, 0xFFu, stack + threadIdx.x, kProgBlock, tc);
        else
            h = trace_ray<false, COUNT>(A.sv, o, dir, A.ao_tmin, A.ao_tmax, 0xFFu, stack + threadIdx.x, kProgBlock, tc);
        occl = h.inst != SOLHIP_MISS;
        if (A.hits_out) store_hit(A.hits_out + (uint64_t)A.W * A.H + (uint64_t)pix * A.n_ao + k, h);
    }
    const uint64_t m = __ballot(occl);
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t group = lane & ~(A.n_ao - 1);
    const uint64_t gmask = (A.n_ao == 64) ? ~0ull : (((1ull << A.n_ao) - 1ull) << group);
    if (active && k == 0) {
        const uint32_t occ = __popcll(m & gmask);
        const float value = 1.0f - (float)occ / (float)A.n_ao;
        float4 *px = A.accum + pix;
        const float4 old = *px;
        *px = make_float4(mixf(old.x, value, A.alpha), mixf(old.y, value, A.alpha), mixf(old.z, value, A.alpha), 1.0f);
    }
    count_rays<COUNT>(A.counters + SOLHIP_KERNEL_AO_RAYS, active ? 1u : 0u, tc);
}

// Persistent AO rays: each lane runs the resumable traversal and, when its
// ray is done, records the occlusion and (once enough lanes of the wave are
// idle) takes the next ray from a global cursor, so incoherent rays do not
// leave most of a wave waiting for its slowest lane.  Ray r = (queue entry
// r / n_ao, AO ray r % n_ao), with the same direction as k_ao_rays.
template <bool ANY, bool COUNT>
__global__ void __launch_bounds__(kProgBlock) SOLHIP_AO_RAYS_OCCUPANCY k_ao_rays_persist(ProgArgs A)
{
    extern __shared__ uint2 stack[];
    const uint32_t count = *A.qcount;
    const uint32_t total = count * A.n_ao;
    const uint32_t shift = __builtin_ctz(A.n_ao);
    const uint32_t lane = threadIdx.x & 63u;
    TravCounts tc{0, 0, 0, false};
    uint32_t rays = 0;
    TravState st;
    bool has = false;
    uint32_t rid = 0, pix = 0;
    PartFetch pf;
    pf.init(A.aoheads, SOLHIP_AO_PARTS, total);
    while (true) {
        const uint64_t idle = __ballot(!has);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (!pf.exhausted && nidle >= (uint32_t)SOLHIP_AO_REFILL) {
            uint32_t end;
            const uint32_t base = pf.claim(nidle, end);
            if (!has) {
                const uint32_t my = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                if (my < end) {
                    const AoEntry e = A.queue[my >> shift];
                    uint32_t rng = pcg_advance(e.rng, 2u * (my & (A.n_ao - 1u)));
                    const float xi0 = next_rand(rng);
                    const float xi1 = next_rand(rng);
                    const V3 dir = sample_cosine(v3(e.nx, e.ny, e.nz), xi0, xi1);
                    trav_init(st, A.sv, v3(e.px, e.py, e.pz), dir, A.ao_tmin, A.ao_tmax);
                    pix = e.pixel;
                    rid = my;
                    has = true;
                }
            }
        }
        if (__ballot(has) == 0 && pf.exhausted) break;
        if (has && trav_step<ANY, COUNT>(st, A.sv, 0xFFu, stack + threadIdx.x, kProgBlock, tc)) {
            const Hit h = trav_result(st);
            A.occl[rid] = h.inst != SOLHIP_MISS ? 1u : 0u;
            if (A.hits_out) store_hit(A.hits_out + (uint64_t)A.W * A.H + (uint64_t)pix * A.n_ao + (rid & (A.n_ao - 1u)), h);
            ++rays;
            has = false;
        }
    }
    count_rays<COUNT>(A.counters + SOLHIP_KERNEL_AO_RAYS, rays, tc);
}

// one thread per queue entry: the AO value from its rays' occlusion bytes
__global__ void k_ao_resolve(ProgArgs A)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= *A.qcount) return;
    const uint8_t *o = A.occl + (uint64_t)q * A.n_ao;
    uint32_t occ = 0;
    for (uint32_t k = 0; k < A.n_ao; ++k) occ += o[k];
    const float value = 1.0f - (float)occ / (float)A.n_ao;
    float4 *px = A.accum + A.queue[q].pixel;
    const float4 old = *px;
    *px = make_float4(mixf(old.x, value, A.alpha), mixf(old.y, value, A.alpha), mixf(old.z, value, A.alpha), 1.0f);
}

// dynamic LDS: the traversal stack, sv.stack_depth 8-byte entries per lane
static size_t stack_bytes(const ProgArgs &a) { return (size_t)a.sv.stack_depth * sizeof(uint2) * kProgBlock; }

static int launch_status(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string(what) + " launch: " + hipGetErrorString(e));
        return SOLHIP_ERROR_DEVICE;
    }
    return 0;
}

#define SOLHIP_LAUNCH(kernel, id, grid, a, s, timer)                                            \
    do {                                                                                      \
        if (timer) (timer)->begin(s, id);                                                     \
        if ((a).count)                                                                        \
            hipLaunchKernelGGL(kernel<true>, dim3(grid), dim3(kProgBlock), stack_bytes(a), s, a); \
        else                                                                                  \
            hipLaunchKernelGGL(kernel<false>, dim3(grid), dim3(kProgBlock), stack_bytes(a), s, a); \
        if (timer) (timer)->end(s, id);                                                       \
    } while (0)

uint64_t ao_head_bytes() { return SOLHIP_AO_REFILL ? (uint64_t)SOLHIP_AO_PARTS * kHeadStride * 4 : 0; }

int launch_ao(hipStream_t s, const ProgArgs &a, Timer *timer)
{
    // the step traversal (primary and persistent AO kernels) may hold two
    // more stack entries than trace_ray
    ProgArgs p = a;
    p.sv.stack_depth = a.sv.stack_depth + 2;
    const size_t lds = (size_t)p.sv.stack_depth * sizeof(uint2) * kProgBlock;
    if (SOLHIP_AO_REFILL == 0 || !a.occl || !a.aoheads) {
        SOLHIP_LAUNCH(k_ao_primary, SOLHIP_KERNEL_AO_PRIMARY, tile_count(a.W, a.H), p, s, timer);
        const uint64_t lanes = (uint64_t)a.W * a.H * a.n_ao;
        const uint32_t blocks = (uint32_t)((lanes + kProgBlock - 1) / kProgBlock);
        SOLHIP_LAUNCH(k_ao_rays, SOLHIP_KERNEL_AO_RAYS, blocks, a, s, timer);
        return launch_status("AO");
    }
    // persistent grid: as many waves as fit (SOLHIP_AO_WAVES per SIMD, 4 SIMDs per CU)
    const uint32_t blocks = a.ncu * 4 * SOLHIP_AO_WAVES;
    // the fetch heads start at zero (a fill kernel: no memset node in a captured graph)
    int rc = launch_fill(s, a.aoheads, (uint64_t)SOLHIP_AO_PARTS * kHeadStride, 0u);
    if (rc) return rc;
    SOLHIP_LAUNCH(k_ao_primary, SOLHIP_KERNEL_AO_PRIMARY, tile_count(a.W, a.H), p, s, timer);
    if (timer) timer->begin(s, SOLHIP_KERNEL_AO_RAYS);
    if (a.ao_any) {
        if (a.count) hipLaunchKernelGGL((k_ao_rays_persist<true, true>), dim3(blocks), dim3(kProgBlock), lds, s, p);
        else hipLaunchKernelGGL((k_ao_rays_persist<true, false>), dim3(blocks), dim3(kProgBlock), lds, s, p);
    } else {
        if (a.count) hipLaunchKernelGGL((k_ao_rays_persist<false, true>), dim3(blocks), dim3(kProgBlock), lds, s, p);
        else hipLaunchKernelGGL((k_ao_rays_persist<false, false>), dim3(blocks), dim3(kProgBlock), lds, s, p);
    }
    if (timer) timer->end(s, SOLHIP_KERNEL_AO_RAYS);
    const uint32_t rblocks = (a.W * a.H + kProgBlock - 1) / kProgBlock;
    hipLaunchKernelGGL(k_ao_resolve, dim3(rblocks), dim3(kProgBlock), 0, s, p);
    return launch_status("AO");
}

// ---------------------------------------------------------------------------
// ao.rgen / ao.rchit / ao.rmiss
// ---------------------------------------------------------------------------
template <bool COUNT>
__global__ void __launch_bounds__(kProgBlock) k_ao_glsl(ProgArgs A)
{
    extern __shared__ uint2 stack[];
    uint32_t x, y;
    const bool active = tile_pixel(A.W, A.H, x, y);
    TravCounts tc{0, 0, 0, false};
    uint32_t rays = 0;
    if (active) {
        const int max_samples = 4, sample_count = 4;  // ao.rgen:42-43
        float ao = 0.0f;
        uint32_t rng = tea(x + y * A.W, A.frame);     // ao.rgen:45
        for (int i = 0; i < sample_count; i++) {
            const float jx = next_rand(rng);
            const float jy = next_rand(rng);          // ao.rgen:49
            V3 o, d;
            camera_ray(A.cam.view_inverse, A.cam.projection_inverse, x, y, A.W, A.H, jx, jy, o, d);
            float hitValue = 0.0f;                    // preparePayload, ao.rgen:26-35
            uint32_t depth = 0, done = 0;
            const uint32_t sampleId = (uint32_t)i;
            float rmin = fmaxf(1.0f, length3(o)) * 1e-3f, rmax = 10000.0f;
            for (;;) {
                const Hit h = trace_ray<false, COUNT>(A.sv, o, d, rmin, rmax, 0xFFu, stack + threadIdx.x,
                                                      kProgBlock, tc);
                ++rays;
                if (h.inst == SOLHIP_MISS) {
                    done = 1;                         // ao.rmiss:9
                } else {
                    const Surf sf = fetch_surface(A.tb, h.inst, h.prim, h.u, h.v);
                    const V3 no = v3(sf.pos.x + 0.00001f * d.x, sf.pos.y + 0.00001f * d.y,
                                     sf.pos.z + 0.00001f * d.z);   // ao.rchit:79
                    const uint32_t bi = depth + depth * sampleId;   // ao.rchit:80
                    const float r0 = next_rand(rng), r1 = next_rand(rng);
                    const float fw = (float)A.blue_w, fh = (float)A.blue_h;
                    const float fx = (float)x + r0 * fw, fy = (float)y + r1 * fh;
                    const float mx = fx - fw * floorf(fx / fw), my = fy - fh * floorf(fy / fh);
                    int cx = (int)mx, cy = (int)my;
                    if (cx >= (int)A.blue_w) cx -= (int)A.blue_w;
                    if (cy >= (int)A.blue_h) cy -= (int)A.blue_h;
                    if (cx < 0) cx += (int)A.blue_w;
                    if (cy < 0) cy += (int)A.blue_h;
                    const uint32_t texel = A.blue[(uint64_t)cy * A.blue_w + cx];
                    const float xi0 = (float)((texel >> (8 * (bi % 4u))) & 0xFFu) / 255.0f;
                    const float xi1 = (float)((texel >> (8 * ((bi + 1u) % 4u))) & 0xFFu) / 255.0f;
                    const float sg = sgn(dot(neg(d), sf.nrm));
                    const V3 hn = scale(sf.nrm, sg);                 // ao.rchit:81
                    d = sample_cosine(hn, xi0, xi1);                 // ao.rchit:82
                    o = no;
                    rmin = 0.001f;
                    rmax = 10.0f;                                    // ao.rchit:83
                    if (depth > 0) hitValue += 1.0f;                 // ao.rchit:84-86
                    depth++;
                }
                if (done == 1 || depth > (uint32_t)max_samples) break;   // ao.rgen:71
            }
            ao += 1.0f / (float)max_samples * hitValue;             // ao.rgen:74
        }
        const float color = 1.0f - ao / (float)sample_count;         // ao.rgen:76
        float4 *px = A.accum + (uint64_t)y * A.W + x;
        const float4 old = *px;
        *px = make_float4(mixf(old.x, color, A.alpha), mixf(old.y, color, A.alpha), mixf(old.z, color, A.alpha), 1.0f);
    }
    count_rays<COUNT>(A.counters + SOLHIP_KERNEL_AO_GLSL, rays, tc);
}

int launch_ao_glsl(hipStream_t s, const ProgArgs &a, Timer *timer)
{
    const uint32_t tiles = tile_count(a.W, a.H);
    SOLHIP_LAUNCH(k_ao_glsl, SOLHIP_KERNEL_AO_GLSL, tiles, a, s, timer);
    return launch_status("AO-glsl");
}

// ---------------------------------------------------------------------------
// pathtrace.rgen / pathtrace.rchit / pathtrace.rmiss
// ---------------------------------------------------------------------------
// pathtrace.rmiss:8-21
template <bool COUNT>
__global__ void __launch_bounds__(kProgBlock) k_pathtrace_glsl(ProgArgs A)
{
    extern __shared__ uint2 stack[];
    uint32_t x, y;
    const bool active = tile_pixel(A.W, A.H, x, y);
    TravCounts tc{0, 0, 0, false};
    uint32_t rays = 0;
    if (active) {
        const int maxBounces = (int)A.max_depth, sampleCount = (int)A.spp;
        float pc[3] = {0.0f, 0.0f, 0.0f};
        uint32_t rng = tea(x + y * A.W, A.frame);   // pathtrace.rgen:47
        for (int i = 0; i < sampleCount; i++) {
            const float jx = next_rand(rng);
            const float jy = next_rand(rng);        // :52
            V3 o, d;
            camera_ray(A.cam.view_inverse, A.cam.projection_inverse, x, y, A.W, A.H, jx, jy, o, d);
            const float rmin = fmaxf(1.0f, length3(o)) * 1e-3f, rmax = 10000.0f;   // :35 (never reset)
            uint32_t depth = 0, done = 0;
            float acc[3] = {1.0f, 1.0f, 1.0f};
            for (;;) {
                const Hit h = trace_ray<false, COUNT>(A.sv, o, d, rmin, rmax, 0xFFu, stack + threadIdx.x,
                                                      kProgBlock, tc);
                ++rays;
                V3 hv;
                if (h.inst == SOLHIP_MISS) {
                    hv = pt_miss(A.enable_sky, true, d);
                    done = 1;
                } else {
                    const float *mat = reinterpret_cast<const float *>(A.tb.material_table[h.inst]);   // :69
                    const float e0 = mat[4], e1 = mat[5], e2 = mat[6];
                    if (e0 >= 1.0f || e1 >= 1.0f || e2 >= 1.0f) {       // :71-76
                        hv = v3(e0, e1, e2);
                        done = 1;
                        depth++;
                    } else {
                        const Surf sf = fetch_surface(A.tb, h.inst, h.prim, h.u, h.v);
                        const V3 wI = normalize(d);                      // :90
                        const float sg = sgn(dot(sf.nrm, neg(wI)));
                        const V3 nO = scale(sf.nrm, sg);
                        const float metallic = mat[8], roughness = mat[9];
                        const float alphaSquared = roughness * roughness;
                        const float xi0 = next_rand(rng);
                        const float xi1 = next_rand(rng);                // :93
                        const float rnd = next_rand(rng);                // :94
                        o = v3(sf.pos.x + 0.0001f * nO.x, sf.pos.y + 0.0001f * nO.y, sf.pos.z + 0.0001f * nO.z);
                        const V3 base = v3(mat[0] * sf.col.x, mat[1] * sf.col.y, mat[2] * sf.col.z);
                        V3 nd;
                        if (rnd < metallic) {                            // :97-100
                            nd = sample_ggx(reflect(d, nO), xi0, xi1, alphaSquared);
                            hv = base;
                        } else {
                            const V3 m = sample_ggx(nO, xi0, xi1, alphaSquared);   // :102
                            if (rnd < fresnel_dielectric(nO, m, 1.0f / 1.5f)) {
                                nd = reflect(d, m);
                                hv = v3(1.0f, 1.0f, 1.0f);
                            } else {
                                nd = sample_cosine(nO, xi0, xi1);
                                hv = base;
                            }
                        }
                        d = nd;
                        depth++;                                         // :112
                    }
                }
                acc[0] = acc[0] * hv.x;                                  // pathtrace.rgen:77
                acc[1] = acc[1] * hv.y;
                acc[2] = acc[2] * hv.z;
                if (done == 1) break;
                if ((int)depth > maxBounces) {
                    acc[0] = acc[1] = acc[2] = 0.0f;
                    break;
                }
            }
            pc[0] = pc[0] + acc[0];
            pc[1] = pc[1] + acc[1];
            pc[2] = pc[2] + acc[2];
        }
        const float inv = 1.0f / (float)sampleCount;                     // :88
        pc[0] = inv * pc[0];
        pc[1] = inv * pc[1];
        pc[2] = inv * pc[2];
        float4 *px = A.accum + (uint64_t)y * A.W + x;
        const float4 old = *px;
        float m0 = mixf(old.x, pc[0], A.alpha), m1 = mixf(old.y, pc[1], A.alpha), m2 = mixf(old.z, pc[2], A.alpha);
        const bool bad = isnan(m0) || isnan(m1) || isnan(m2) || isinf(m0) || isinf(m1) || isinf(m2);
        if (bad) {                                                       // :94-99
            m0 = old.x;
            m1 = old.y;
            m2 = old.z;
        }
        *px = make_float4(m0, m1, m2, 1.0f);
        if (A.out_rgba8) {
            const float g = 1.0f / 2.2f;                                 // :102
            const uint32_t r8 = to_unorm8(powf(m0, g)), g8 = to_unorm8(powf(m1, g)), b8 = to_unorm8(powf(m2, g));
            A.out_rgba8[(uint64_t)y * A.W + x] = r8 | (g8 << 8) | (b8 << 16) | (255u << 24);
        }
    }
    count_rays<COUNT>(A.counters + SOLHIP_KERNEL_PATHTRACE_GLSL, rays, tc);
}

int launch_pathtrace_glsl(hipStream_t s, const ProgArgs &a, Timer *timer)
{
    const uint32_t tiles = tile_count(a.W, a.H);
    SOLHIP_LAUNCH(k_pathtrace_glsl, SOLHIP_KERNEL_PATHTRACE_GLSL, tiles, a, s, timer);
    return launch_status("pathtrace");
}

// ---------------------------------------------------------------------------
// debug.rgen / debug.rchit / debug.rmiss
// ---------------------------------------------------------------------------
template <bool COUNT>
__global__ void __launch_bounds__(kProgBlock) k_debug(ProgArgs A)
{
    extern __shared__ uint2 stack[];
    uint32_t x, y;
    const bool active = tile_pixel(A.W, A.H, x, y);
    TravCounts tc{0, 0, 0, false};
    if (active) {
        V3 o, d;
        camera_ray(A.cam.view_inverse, A.cam.projection_inverse, x, y, A.W, A.H, 0.5f, 0.5f, o, d);  // debug.rgen:20-26
        const Hit h = trace_ray<false, COUNT>(A.sv, o, d, 0.001f, 1000.0f, 0xFFu, stack + threadIdx.x, kProgBlock, tc);
        float4 r;
        if (h.inst == SOLHIP_MISS)
            r = make_float4(d.x, d.y, d.z, 0.0f);                       // debug.rgen:28
        else
            r = make_float4(1.0f - h.u - h.v, h.u, h.v, 0.0f);          // debug.rchit:11-12
        A.accum[(uint64_t)y * A.W + x] = r;
    }
    count_rays<COUNT>(A.counters + SOLHIP_KERNEL_DEBUG, active ? 1u : 0u, tc);
}

int launch_debug(hipStream_t s, const ProgArgs &a, Timer *timer)
{
    const uint32_t tiles = tile_count(a.W, a.H);
    SOLHIP_LAUNCH(k_debug, SOLHIP_KERNEL_DEBUG, tiles, a, s, timer);
    return launch_status("debug");
}

}  // namespace solhip
