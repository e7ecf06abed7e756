// wavefront.hip -- SOLHIP_PROG_PATHTRACE: the path tracer as a wavefront of
// queue-connected kernels, one sample per pixel per pass:
//
//   generate  camera rays of all pixels into extension queue 0 (8x8 pixel
//             tiles, one per 64 slots); folds the previous sample's radiance
//             into the per-pixel sum
//   extend    closest hit of every queued ray (persistent grid, waves fetch
//             64 slots at a time from an atomic head)
//   shade     pathtrace.rchit's material model; emission / sky into the
//             path's radiance, the next ray and the next-event shadow ray
//             appended to their queues with one atomic per wave (ballot +
//             popcount + mbcnt)
//   shadow    any-hit test of the shadow queue; unoccluded contributions are
//             added to the path's radiance
//   finalize  sample mean, accumulation (pathtrace.rgen:88-99), RGBA8 (:102)
//
// Every path owns its state (throughput, RNG, radiance, depth), so the
// result does not depend on queue order; with nee = 0 and rr_start = 0 the
// arithmetic and the random draws are exactly those of pathtrace.rgen /
// pathtrace.rchit, bit for bit.  Per path the additions to its radiance come
// in the order next-event(1), next-event(2), ..., emission, as in a
// sequential loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "kernels.h"
#include "pathshade.h"
#include "shading.h"
#include "traverse.h"
#include "traverse_step.h"
#include "fetch.h"
#include "accum.h"

namespace solhip {

namespace {

// slot of the first queue -> pixel
__device__ __forceinline__ bool slot_pixel(uint32_t slot, uint32_t W, uint32_t H, uint32_t &x, uint32_t &y)
{
    const uint32_t tiles_x = (W + 7) >> 3, tile = slot >> 6, lane = slot & 63u;
    x = (tile % tiles_x) * 8 + (lane & 7u);
    y = (tile / tiles_x) * 8 + (lane >> 3);
    return x < W && y < H;
}

// wave-aggregated append; every lane of the wave must call it
__device__ __forceinline__ uint32_t wave_append(uint32_t *counter, bool pred)
{
    const uint64_t m = __ballot(pred);
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t base = 0;
    if (m) {
        const int leader = __ffsll((unsigned long long)m) - 1;
        if ((int)lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
        base = __shfl(base, leader);
    }
    return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

// pathtrace.rgen:35: tmin from the camera origin (the same for every ray)
__device__ __forceinline__ float ray_tmin(const WfArgs &A)
{
    float o[4];
    mat4_mul(A.fc->cam.view_inverse, 0.0f, 0.0f, 0.0f, 1.0f, o);
    return fmaxf(1.0f, sqrtf(o[0] * o[0] + o[1] * o[1] + o[2] * o[2])) * 1e-3f;
}

// fetch partitions of the extend and shadow kernels (fetch.h)
#ifndef SOLHIP_WF_PARTS
#define SOLHIP_WF_PARTS 64
#endif
constexpr uint32_t kWfParts = SOLHIP_WF_PARTS;

// samples traced in pass `pass` (the last pass may hold fewer than S)
__device__ __forceinline__ uint32_t pass_samples(const WfArgs &A, uint32_t pass)
{
    return min(A.S, A.spp - pass * A.S);
}

// first queue of pass `pass`: entry e = sample-in-pass * nslots + slot, path
// index = sample-in-pass * npix + pixel
__global__ void k_pt_generate(WfArgs A, uint32_t pass)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = pass_samples(A, pass) * A.nslots;
    if (e == 0) {
        A.counts[0] = n;
        A.counts[1] = 0;
        A.counts[2] = 0;
    }
    if (e < 2 * (A.max_depth + 1) * kWfParts) A.heads[e * kHeadStride] = 0;
    if (e >= n) return;
    const uint32_t sl = e / A.nslots, slot = e - sl * A.nslots;
    uint32_t x, y;
    QRay r;
    if (!slot_pixel(A.slot_base + slot, A.W, A.H, x, y)) {
        r.ox = r.oy = r.oz = r.dx = r.dy = r.dz = r.tmax = 0.0f;
        r.path = kNoPath;
        A.q[0][e] = r;
        return;
    }
    const uint32_t pix = y * A.W + x, p = sl * A.npix + pix;
    uint32_t rng;
    if (A.streams) {
        // sample s of the pixel starts its own stream (s = 0: pathtrace.rgen:47)
        rng = tea(pix + (pass * A.S + sl) * A.npix, A.fc->frame);
    } else if (pass == 0) {
        rng = tea(x + y * A.W, A.fc->frame);   // pathtrace.rgen:47
    } else {
        // one sample per pass: the RNG runs on from the previous sample
        rng = as_u(A.st0[p].w);
    }
    if (!A.streams && pass > 0) {
        // the previous sample's radiance joins the sum
        const float4 s1 = A.st1[p];
        const float4 ps = pass == 1 ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : A.psum[pix];
        A.psum[pix] = make_float4(ps.x + s1.x, ps.y + s1.y, ps.z + s1.z, 0.0f);
    }
    const float jx = next_rand(rng);
    const float jy = next_rand(rng);   // :52
    V3 o, d;
    camera_ray(A.fc->cam.view_inverse, A.fc->cam.projection_inverse, x, y, A.W, A.H, jx, jy, o, d);
    A.st0[p] = make_float4(1.0f, 1.0f, 1.0f, as_f(rng));
    A.st1[p] = make_float4(0.0f, 0.0f, 0.0f, as_f(1u << 16));   // depth 0, emission counts
    r.ox = o.x;
    r.oy = o.y;
    r.oz = o.z;
    r.path = p;
    r.dx = d.x;
    r.dy = d.y;
    r.dz = d.z;
    r.tmax = A.rmax;
    A.q[0][e] = r;
}

// per-sample streams: the radiance of pass `pass`'s samples joins each
// pixel's order-free sum (accum.h), before the next pass's generate
// overwrites the path state
__device__ __forceinline__ void fold_pass(const WfArgs &A, uint32_t pass, uint32_t pix, FixSum &fs)
{
    const uint32_t sk = pass_samples(A, pass);
    for (uint32_t sl = 0; sl < sk; ++sl) {
        const float4 s1 = A.st1[sl * A.npix + pix];
        fix_add(fs, s1.x, s1.y, s1.z);
    }
}

__global__ void k_pt_fold(WfArgs A, uint32_t pass)
{
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t x, y;
    if (slot >= A.nslots || !slot_pixel(A.slot_base + slot, A.W, A.H, x, y)) return;
    const uint32_t pix = y * A.W + x;
    FixSum fs;
    if (pass == 0) fix_clear(fs);
    else fs = A.fsum[pix];
    fold_pass(A, pass, pix, fs);
    A.fsum[pix] = fs;
}

// Register budget of the traversal kernels (waves per SIMD they must fit).
// BLAS exits restore the world ray's inverse direction from the ray state
// (trav_step's WINV), and the traversal kernels ask for 4 waves per SIMD so
// that those registers fit: on 10,000 instances (3.9 BLAS exits per ray)
// 1279 -> 1356 Mrays/s; the inverse at 5 waves spilled (1275), 4 waves
// without it gave 1295
#ifndef SOLHIP_WF_WINV
#define SOLHIP_WF_WINV true
#endif
#ifndef SOLHIP_WF_WAVES
#define SOLHIP_WF_WAVES 4
#endif
#define SOLHIP_WF_OCCUPANCY __attribute__((amdgpu_waves_per_eu(SOLHIP_WF_WAVES)))

// lanes of a wave take new queue slots once at least this many are idle
// (0: one slot per lane per 64-slot fetch, trace_ray to completion)
#ifndef SOLHIP_WF_REFILL
#define SOLHIP_WF_REFILL 16
#endif

template <bool ANY>
__device__ __forceinline__ void queue_result(const WfArgs &A, uint32_t i, uint32_t path, const Hit &h)
{
    if (!ANY) {
        QHit out;
        out.t = h.t;
        out.u = h.u;
        out.v = h.v;
        out.inst = h.inst;
        out.prim = h.prim;
        A.hits[i] = out;
    } else if (h.inst == SOLHIP_MISS) {
        const float4 c = A.scontrib[i];
        float4 s1 = A.st1[path];
        s1.x = s1.x + c.x;
        s1.y = s1.y + c.y;
        s1.z = s1.z + c.z;
        A.st1[path] = s1;
    }
}

// Extension (closest hit -> hits[i]) or shadow (any hit; unoccluded ->
// contribution into the path radiance) for queue slots [0, cnt), taken from
// the atomic head.  The resumable traversal lets a lane start its next slot
// while the rest of the wave is still traversing.  sv.stack_depth includes
// the two extra entries the step traversal may hold.
template <bool ANY, bool COUNT>
__device__ __forceinline__ void trace_queue(const WfArgs &A, const QRay *q, uint32_t cnt, uint32_t *heads, int kernel)
{
    extern __shared__ uint2 stack[];
    const uint32_t lane = threadIdx.x & 63u;
    const float tmin = ray_tmin(A);
    TravCounts tc{0, 0, 0, false};
    uint32_t rays = 0;
#if SOLHIP_WF_REFILL
    TravState st;
    bool has = false;
    uint32_t idx = 0, path = 0;
    PartFetch pf;
    pf.init(heads, kWfParts, cnt);
    while (true) {
        const uint64_t idle = __ballot(!has);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (!pf.exhausted && nidle >= (uint32_t)SOLHIP_WF_REFILL) {
            uint32_t end;
            const uint32_t base = pf.claim(nidle, end);
            if (!has) {
                const uint32_t my = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                if (my < end) {
                    const QRay r = q[my];
                    if (r.path == kNoPath) {
                        if (!ANY) {
                            Hit miss;
                            miss.t = -1.0f;
                            miss.u = miss.v = 0.0f;
                            miss.inst = miss.prim = SOLHIP_MISS;
                            queue_result<false>(A, my, r.path, miss);
                        }
                    } else {
                        trav_init(st, A.sv, v3(r.ox, r.oy, r.oz), v3(r.dx, r.dy, r.dz), tmin, r.tmax);
                        idx = my;
                        path = r.path;
                        has = true;
                    }
                }
            }
        }
        if (__ballot(has) == 0 && pf.exhausted) break;
        if (has && trav_step<ANY, COUNT, SOLHIP_WF_WINV>(st, A.sv, 0xFFu, stack + threadIdx.x, A.block, tc)) {
            queue_result<ANY>(A, idx, path, trav_result(st));
            ++rays;
            has = false;
        }
    }
#else
    for (;;) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(heads, 64u);
        base = __shfl(base, 0);
        if (base >= cnt) break;
        const uint32_t i = base + lane;
        if (i < cnt) {
            const QRay r = q[i];
            Hit h;
            h.t = -1.0f;
            h.u = h.v = 0.0f;
            h.inst = h.prim = SOLHIP_MISS;
            if (r.path != kNoPath) {
                h = trace_ray<ANY, COUNT>(A.sv, v3(r.ox, r.oy, r.oz), v3(r.dx, r.dy, r.dz), tmin, r.tmax, 0xFFu,
                                          stack + threadIdx.x, A.block, tc);
                ++rays;
            }
            if (!ANY || r.path != kNoPath) queue_result<ANY>(A, i, r.path, h);
        }
    }
#endif
    count_rays<COUNT>(A.counters + kernel, rays, tc);
}

template <bool COUNT>
__global__ void SOLHIP_WF_OCCUPANCY k_pt_extend(WfArgs A, uint32_t b)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // the queues this bounce's shade fills were last read a bounce ago
        A.counts[(b + 1) & 1] = 0;
        A.counts[2] = 0;
    }
    trace_queue<false, COUNT>(A, A.q[b & 1], A.counts[b & 1], A.heads + 2 * b * kWfParts * kHeadStride,
                              SOLHIP_KERNEL_PT_EXTEND);
}

template <bool COUNT>
__global__ void SOLHIP_WF_OCCUPANCY k_pt_shadow(WfArgs A, uint32_t b)
{
    trace_queue<true, COUNT>(A, A.sq, A.counts[2], A.heads + (2 * b + 1) * kWfParts * kHeadStride,
                             SOLHIP_KERNEL_PT_SHADOW);
}

// The shade kernel runs blocks of WAVES waves and appends to its queues once
// per block.  With shadow rays (two appends per ray) 8-wave blocks: the
// per-wave atomics on the two counters limited the Cornell box (3599 ->
// 4223 Mrays/s); without them 2-wave blocks: larger blocks wait longer at the
// append barriers on scattered scenes (10k instances: 1259 with 1 wave, 1277
// with 2, 1225 with 8).
// waves per SIMD the shade kernel is compiled for (0: the compiler's choice,
// 115 VGPRs and 4 waves; 5 fits in 90 VGPRs without spills: Cornell box
// 4176 -> 4287 Mrays/s, 10k instances 1252 -> 1258; 6 spills and is slower)
#ifndef SOLHIP_WF_SHADE_WAVES
#define SOLHIP_WF_SHADE_WAVES 5
#endif
#if SOLHIP_WF_SHADE_WAVES
#define SOLHIP_WF_SHADE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(SOLHIP_WF_SHADE_WAVES)))
#else
#define SOLHIP_WF_SHADE_OCCUPANCY
#endif

template <uint32_t WAVES>
__global__ void __launch_bounds__(64 * WAVES) SOLHIP_WF_SHADE_OCCUPANCY k_pt_shade(WfArgs A, uint32_t b)
{
    __shared__ uint32_t lds[WAVES + 1];
    const uint32_t cnt = A.counts[b & 1];
    const QRay *qin = A.q[b & 1];
    QRay *qout = A.q[(b + 1) & 1];
    uint32_t *cout = A.counts + ((b + 1) & 1);
    const uint32_t nthreads = gridDim.x * blockDim.x;
    ShadeParams P;
    P.tb = A.tb;
    P.enable_sky = A.enable_sky;
    P.max_depth = A.max_depth;
    P.rr_start = A.rr_start;
    P.lights = A.lights;
    P.nlights = A.nlights;
    P.total = A.nlights ? A.lights[A.nlights - 1].cdf : 0.0f;
    P.nee = A.nee && A.nlights > 0 && P.total > 0.0f;
    P.rmax = A.rmax;
    for (uint32_t base = blockIdx.x * blockDim.x; base < cnt; base += nthreads) {
        const uint32_t i = base + threadIdx.x;
        ShadeOut o;
        if (i < cnt) {
            const QRay r = qin[i];
            if (r.path != kNoPath) {
                const uint32_t p = r.path;
                float4 s0 = A.st0[p], s1 = A.st1[p];
                shade_vertex(P, p, r, A.hits[i], s0, s1, o);
                A.st0[p] = s0;
                A.st1[p] = s1;
            }
        }
        const uint32_t s_next = block_append<WAVES>(cout, o.cont, lds);
        if (o.cont) qout[s_next] = o.nr;
        const uint32_t s_sh = block_append<WAVES>(A.counts + 2, o.shadow, lds);
        if (o.shadow) {
            A.sq[s_sh] = o.sr;
            A.scontrib[s_sh] = o.sc;
        }
    }
}

// ---- radiance cache (NrcArgs) ---------------------------------------------
//
// Query paths end at their first bounce vertex (depth 1), where the network
// predicts the radiance leaving the vertex towards the previous one; their
// throughput times the prediction joins the path radiance.  One path in
// `ratio` (per pass, by a hash of pixel and sample) is a training path: it
// runs unbiased to the end and records, at each non-emitter vertex 1 ..
// max_rec, the features with the throughput and radiance on arrival; after
// the pass the radiance gathered from a vertex on, divided by the
// throughput into it, is that vertex's training target.

__device__ __forceinline__ uint16_t bf16_bits(float f)
{
    const __bf16 h = (__bf16)f;
    return __builtin_bit_cast(uint16_t, h);
}

// 64 features: frequency encoding of the normalised position (6 octaves x
// sin/cos x 3 = 36), shading normal, direction to the previous vertex,
// albedo, roughness, metallic, a constant 1, zeros
__device__ __forceinline__ void nrc_features(const NrcArgs &N, V3 pos, V3 nO, V3 d, V3 albedo, float roughness,
                                             float metallic, uint16_t *row)
{
    float f[64];
    const float pn[3] = {(pos.x - N.lo[0]) * N.inv_ext[0], (pos.y - N.lo[1]) * N.inv_ext[1],
                         (pos.z - N.lo[2]) * N.inv_ext[2]};
    int k = 0;
    for (int a = 0; a < 3; ++a) {
        float scale = kPi;
        for (int o = 0; o < 6; ++o) {
            float sn, cs;
            own_sincos(pn[a] * scale, sn, cs);
            f[k++] = sn;
            f[k++] = cs;
            scale = scale * 2.0f;
        }
    }
    const V3 wo = neg(normalize(d));
    f[k++] = nO.x;
    f[k++] = nO.y;
    f[k++] = nO.z;
    f[k++] = wo.x;
    f[k++] = wo.y;
    f[k++] = wo.z;
    f[k++] = albedo.x;
    f[k++] = albedo.y;
    f[k++] = albedo.z;
    f[k++] = roughness;
    f[k++] = metallic;
    f[k++] = 1.0f;
    while (k < 64) f[k++] = 0.0f;
    for (int c = 0; c < 8; ++c) {
        uint4 v;
        v.x = bf16_bits(f[8 * c + 0]) | ((uint32_t)bf16_bits(f[8 * c + 1]) << 16);
        v.y = bf16_bits(f[8 * c + 2]) | ((uint32_t)bf16_bits(f[8 * c + 3]) << 16);
        v.z = bf16_bits(f[8 * c + 4]) | ((uint32_t)bf16_bits(f[8 * c + 5]) << 16);
        v.w = bf16_bits(f[8 * c + 6]) | ((uint32_t)bf16_bits(f[8 * c + 7]) << 16);
        reinterpret_cast<uint4 *>(row)[c] = v;
    }
}

// training paths of this pass
__global__ void k_nrc_select(WfArgs A, uint32_t pass)
{
    const NrcArgs &N = A.nrc;
    const uint32_t npix = A.W * A.H, lane = threadIdx.x & 63u, nthreads = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < npix; base += nthreads) {
        const uint32_t p = base + lane;
        const bool sel = p < npix && tea(p, A.fc->frame * A.spp + pass) % N.ratio == 0u;
        const uint32_t idx = wave_append(N.counts + 1, sel);
        if (p < npix) {
            const bool ok = sel && idx < N.train_cap;
            N.tidx[p] = ok ? idx : 0xFFFFFFFFu;
            if (ok) {
                N.tpath[idx] = p;
                N.trec_n[idx] = 0;
            }
        }
    }
}

__global__ void k_pt_shade_nrc(WfArgs A, uint32_t b)
{
    const uint32_t cnt = A.counts[b & 1];
    const QRay *qin = A.q[b & 1];
    QRay *qout = A.q[(b + 1) & 1];
    uint32_t *cout = A.counts + ((b + 1) & 1);
    const NrcArgs &N = A.nrc;
    const uint32_t lane = threadIdx.x & 63u, nthreads = gridDim.x * blockDim.x;
    ShadeParams P;
    P.tb = A.tb;
    P.enable_sky = A.enable_sky;
    P.max_depth = A.max_depth;
    P.rr_start = A.rr_start;
    P.lights = A.lights;
    P.nlights = A.nlights;
    P.total = A.nlights ? A.lights[A.nlights - 1].cdf : 0.0f;
    P.nee = A.nee && A.nlights > 0 && P.total > 0.0f;
    P.rmax = A.rmax;
    for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < cnt; base += nthreads) {
        const uint32_t i = base + lane;
        ShadeOut o;
        bool query = false;
        CacheQuery qr;
        uint16_t qf[64];
        if (i < cnt) {
            const QRay r = qin[i];
            if (r.path != kNoPath) {
                const uint32_t p = r.path;
                const QHit h = A.hits[i];
                float4 s0 = A.st0[p], s1 = A.st1[p];
                const uint32_t depth = as_u(s1.w) & 0xFFFFu;
                bool shade = true;
                if (h.inst != SOLHIP_MISS && depth >= 1) {
                    const float *mat = reinterpret_cast<const float *>(A.tb.material_table[h.inst]);
                    const uint32_t ti = N.tidx[p];
                    const bool record = ti != 0xFFFFFFFFu && depth <= N.max_rec;
                    const bool ask = ti == 0xFFFFFFFFu && depth == 1;
                    if (!is_emitter(mat) && (record || ask)) {
                        const Surf sf = fetch_surface(A.tb, h.inst, h.prim, h.u, h.v);
                        const V3 d = v3(r.dx, r.dy, r.dz);
                        const V3 wI = normalize(d);
                        const V3 nO = scale(sf.nrm, sgn(dot(sf.nrm, neg(wI))));
                        const V3 alb = v3(mat[0] * sf.col.x, mat[1] * sf.col.y, mat[2] * sf.col.z);
                        if (ask) {
                            nrc_features(N, sf.pos, nO, d, alb, mat[9], mat[8], qf);
                            qr.path = p;
                            qr.T[0] = s0.x;
                            qr.T[1] = s0.y;
                            qr.T[2] = s0.z;
                            query = true;
                            shade = false;   // the cache answers for the rest of the path
                        } else {
                            const uint32_t rec = ti * N.max_rec + (depth - 1);
                            nrc_features(N, sf.pos, nO, d, alb, mat[9], mat[8], N.tfeat + (uint64_t)rec * 64);
                            N.trec[2ull * rec] = make_float4(s0.x, s0.y, s0.z, 0.0f);
                            N.trec[2ull * rec + 1] = make_float4(s1.x, s1.y, s1.z, 0.0f);
                            N.trec_n[ti] = depth;
                        }
                    }
                }
                if (shade) {
                    shade_vertex(P, p, r, h, s0, s1, o);
                    A.st0[p] = s0;
                    A.st1[p] = s1;
                }
            }
        }
        const uint32_t s_next = wave_append(cout, o.cont);
        if (o.cont) qout[s_next] = o.nr;
        const uint32_t s_sh = wave_append(A.counts + 2, o.shadow);
        if (o.shadow) {
            A.sq[s_sh] = o.sr;
            A.scontrib[s_sh] = o.sc;
        }
        const uint32_t s_q = wave_append(N.counts + 0, query);
        if (query) {
            N.qrec[s_q] = qr;
            uint4 *dst = reinterpret_cast<uint4 *>(N.qfeat + (uint64_t)s_q * 64);
            for (int c = 0; c < 8; ++c) dst[c] = reinterpret_cast<const uint4 *>(qf)[c];
        }
    }
}

// training rows of this pass: (features, radiance from the vertex on / throughput into it)
__global__ void k_nrc_targets(WfArgs A)
{
    const NrcArgs &N = A.nrc;
    const uint32_t ntrain = min(N.counts[1], N.train_cap);
    const uint32_t nthreads = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ntrain; i += nthreads) {
        const float4 Lf = A.st1[N.tpath[i]];
        const uint32_t nrec = N.trec_n[i];
        for (uint32_t k = 0; k < nrec; ++k) {
            const uint64_t rec = (uint64_t)i * N.max_rec + k;
            const float4 T = N.trec[2 * rec], L = N.trec[2 * rec + 1];
            if (!(T.x > 0.0f && T.y > 0.0f && T.z > 0.0f)) continue;
            const float4 tg = make_float4((Lf.x - L.x) / T.x, (Lf.y - L.y) / T.y, (Lf.z - L.z) / T.z, 0.0f);
            if (!(isfinite(tg.x) && isfinite(tg.y) && isfinite(tg.z))) continue;
            const uint32_t slot = atomicAdd(N.counts + 2, 1u);
            if (slot >= N.batch_cap) continue;
            const uint4 *src = reinterpret_cast<const uint4 *>(N.tfeat + rec * 64);
            uint4 *dst = reinterpret_cast<uint4 *>(N.bfeat + (uint64_t)slot * 64);
            for (int c = 0; c < 8; ++c) dst[c] = src[c];
            N.btarget[slot] = tg;
        }
    }
}

__global__ void k_pt_finalize(WfArgs A)
{
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t x, y;
    if (slot >= A.nslots || !slot_pixel(A.slot_base + slot, A.W, A.H, x, y)) return;
    const uint32_t p = y * A.W + x;
    const uint32_t passes = (A.spp + A.S - 1) / A.S;
    float pc0, pc1, pc2;
    bool bad_sum = false;
    if (A.streams) {
        // the last pass's samples join the order-free sum of the earlier ones
        FixSum fs;
        if (passes == 1) fix_clear(fs);
        else fs = A.fsum[p];
        fold_pass(A, passes - 1, p, fs);
        pc0 = fix_value(fs.c[0]);
        pc1 = fix_value(fs.c[1]);
        pc2 = fix_value(fs.c[2]);
        bad_sum = fs.bad != 0;
    } else {
        // one sample per pass: the last one joins the float sum
        const float4 s1 = A.st1[p];
        const float4 ps = passes == 1 ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : A.psum[p];
        pc0 = ps.x + s1.x;
        pc1 = ps.y + s1.y;
        pc2 = ps.z + s1.z;
    }
    const float inv = 1.0f / (float)A.spp;   // pathtrace.rgen:88
    pc0 = inv * pc0;
    pc1 = inv * pc1;
    pc2 = inv * pc2;
    float4 *px = A.accum + p;
    const float4 old = *px;
    const float alpha = A.fc->alpha;
    float m0 = mixf(old.x, pc0, alpha), m1 = mixf(old.y, pc1, alpha), m2 = mixf(old.z, pc2, alpha);
    const bool bad = bad_sum || isnan(m0) || isnan(m1) || isnan(m2) || isinf(m0) || isinf(m1) || isinf(m2);
    if (bad) {   // :94-99
        m0 = old.x;
        m1 = old.y;
        m2 = old.z;
    }
    *px = make_float4(m0, m1, m2, 1.0f);
    if (A.out_rgba8) {
        const float g = 1.0f / 2.2f;   // :102
        const uint32_t r8 = to_unorm8(powf(m0, g)), g8 = to_unorm8(powf(m1, g)), b8 = to_unorm8(powf(m2, g));
        A.out_rgba8[p] = r8 | (g8 << 8) | (b8 << 16) | (255u << 24);
    }
}

__global__ void k_set_frame(FrameConsts *dst, FrameConsts v)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *dst = v;
}

// ---- emitter table ----------------------------------------------------------

__global__ void k_lights_count(DevTables tb, const uint32_t *inst_tris, uint32_t n, uint32_t *cnt)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *mat = reinterpret_cast<const float *>(tb.material_table[i]);
    cnt[i] = is_emitter(mat) ? inst_tris[i] : 0u;
}

// exclusive scan of n counts in place, offsets[n] = total (one block of 1024)
__global__ void __launch_bounds__(1024) k_scan_counts(uint32_t *offsets, uint32_t n)
{
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x, chunk = (n + 1023) / 1024;
    const uint32_t lo = min(n, t * chunk), hi = min(n, lo + chunk);
    uint32_t s = 0;
    for (uint32_t i = lo; i < hi; ++i) s += offsets[i];
    part[t] = s;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - s;   // exclusive prefix of this chunk
    for (uint32_t i = lo; i < hi; ++i) {
        const uint32_t c = offsets[i];
        offsets[i] = run;
        run += c;
    }
    if (t == 1023) offsets[n] = part[1023];
}

__global__ void k_lights_fill(DevTables tb, const uint32_t *offsets, uint32_t n_inst, LightTri *lights,
                              uint32_t nlights)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nlights) return;
    // instance: the last j with offsets[j] <= i
    uint32_t lo = 0, hi = n_inst;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (offsets[mid] <= i) lo = mid;
        else hi = mid;
    }
    const uint32_t inst = lo, prim = i - offsets[inst];
    const solhip_scene_instance *si = tb.instances + inst;
    const uint32_t obj = si->id;
    uint32_t ix[3];
    for (int k = 0; k < 3; ++k) {
        if (tb.index_size == 8)
            ix[k] = (uint32_t)(reinterpret_cast<const uint64_t *>(tb.index_table[obj]))[3ull * prim + k];
        else
            ix[k] = (reinterpret_cast<const uint32_t *>(tb.index_table[obj]))[3ull * prim + k];
    }
    const float4 *vb = reinterpret_cast<const float4 *>(tb.vertex_table[obj]);
    float w[3][4];
    for (int k = 0; k < 3; ++k) {
        const float4 pv = vb[4ull * ix[k]];
        mat4_mul(si->transform, pv.x, pv.y, pv.z, 1.0f, w[k]);
    }
    const float e1x = w[1][0] - w[0][0], e1y = w[1][1] - w[0][1], e1z = w[1][2] - w[0][2];
    const float e2x = w[2][0] - w[0][0], e2y = w[2][1] - w[0][1], e2z = w[2][2] - w[0][2];
    const float c[3] = {e1y * e2z - e1z * e2y, e1z * e2x - e1x * e2z, e1x * e2y - e1y * e2x};
    const float len = sqrtf(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    const float *mat = reinterpret_cast<const float *>(tb.material_table[inst]);
    LightTri L;
    for (int a = 0; a < 3; ++a) {
        L.p0[a] = w[0][a];
        L.p1[a] = w[1][a];
        L.p2[a] = w[2][a];
        L.n[a] = len > 0.0f ? c[a] / len : 0.0f;
        L.e[a] = mat[4 + a];
    }
    L.cdf = 0.5f * len;
    lights[i] = L;
}

// running area sum in light order (sequential: the float sums must follow
// the host restatement's order exactly)
__global__ void k_lights_cdf(LightTri *lights, uint32_t n)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    float run = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        run = run + lights[i].cdf;
        lights[i].cdf = run;
    }
}

int launch_status(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string(what) + " launch: " + hipGetErrorString(e));
        return SOLHIP_ERROR_DEVICE;
    }
    return 0;
}

}  // namespace

uint32_t wf_threads_per_cu() { return 4u * SOLHIP_WF_WAVES * 64u; }

uint32_t wf_head_words() { return kWfParts * kHeadStride; }

int launch_lights_count(hipStream_t s, const DevTables &tb, const uint32_t *inst_tris, uint32_t n_inst,
                        uint32_t *offsets, Timer *timer)
{
    if (timer) timer->begin(s, SOLHIP_KERNEL_PT_LIGHTS);
    if (n_inst) hipLaunchKernelGGL(k_lights_count, dim3((n_inst + 255) / 256), dim3(256), 0, s, tb, inst_tris, n_inst, offsets);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, s, offsets, n_inst);
    if (timer) timer->end(s, SOLHIP_KERNEL_PT_LIGHTS);
    return launch_status("light count");
}

int launch_lights_fill(hipStream_t s, const DevTables &tb, const uint32_t *offsets, uint32_t n_inst,
                       LightTri *lights, uint32_t nlights, Timer *timer)
{
    if (!nlights) return 0;
    if (timer) timer->begin(s, SOLHIP_KERNEL_PT_LIGHTS);
    hipLaunchKernelGGL(k_lights_fill, dim3((nlights + 255) / 256), dim3(256), 0, s, tb, offsets, n_inst, lights,
                       nlights);
    hipLaunchKernelGGL(k_lights_cdf, dim3(1), dim3(64), 0, s, lights, nlights);
    if (timer) timer->end(s, SOLHIP_KERNEL_PT_LIGHTS);
    return launch_status("light table");
}

int launch_set_frame(hipStream_t s, FrameConsts *dst, const FrameConsts &value)
{
    hipLaunchKernelGGL(k_set_frame, dim3(1), dim3(64), 0, s, dst, value);
    return launch_status("frame constants");
}

int launch_pathtrace(hipStream_t s, const WfArgs &a, Timer *timer, const PassHook *hook)
{
    const int rc = launch_pathtrace_passes(s, a, timer, hook);
    return rc ? rc : launch_pathtrace_finalize(s, a, timer);
}

int launch_pathtrace_finalize(hipStream_t s, const WfArgs &a, Timer *timer)
{
    const uint32_t blk = a.block;
    const uint32_t fin_blocks = (a.nslots + blk - 1) / blk;
    if (timer) timer->begin(s, SOLHIP_KERNEL_PT_FINALIZE);
    hipLaunchKernelGGL(k_pt_finalize, dim3(fin_blocks), dim3(blk), 0, s, a);
    if (timer) timer->end(s, SOLHIP_KERNEL_PT_FINALIZE);
    return launch_status("pathtrace finalize");
}

int launch_pathtrace_passes(hipStream_t s, const WfArgs &a0, Timer *timer, const PassHook *hook)
{
    // the step traversal may hold two more stack entries than trace_ray
    WfArgs a = a0;
    if (SOLHIP_WF_REFILL) a.sv.stack_depth += 2;
    const uint32_t blk = a.block;
    const size_t lds = (size_t)a.sv.stack_depth * sizeof(uint2) * blk;
    const uint32_t nheads = 2 * (a.max_depth + 1);
    const uint32_t gen_blocks = (std::max(a.S * a.nslots, nheads * kWfParts) + blk - 1) / blk;
    const uint32_t fold_blocks = (a.nslots + blk - 1) / blk;
    const uint32_t passes = (a.spp + a.S - 1) / a.S;
    auto shade_grid = [&](uint32_t waves) {
        return std::max(1u, (uint32_t)(((uint64_t)a.shade_blocks * blk + 64 * waves - 1) / (64 * waves)));
    };
    const bool shadows = a.nee && a.nlights > 0;
    auto timed = [&](int id, auto &&launch) {
        if (timer) timer->begin(s, id);
        launch();
        if (timer) timer->end(s, id);
    };
    const bool nrc = a.nrc.enabled != 0;
    for (uint32_t pass = 0; pass < passes; ++pass) {
        if (a.streams && pass > 0)
            timed(SOLHIP_KERNEL_PT_GENERATE,
                  [&] { hipLaunchKernelGGL(k_pt_fold, dim3(fold_blocks), dim3(blk), 0, s, a, pass - 1); });
        timed(SOLHIP_KERNEL_PT_GENERATE,
              [&] { hipLaunchKernelGGL(k_pt_generate, dim3(gen_blocks), dim3(blk), 0, s, a, pass); });
        if (nrc) {
            int rc = launch_fill(s, a.nrc.counts, 2, 0u);   // queries, training paths
            if (rc) return rc;
            hipLaunchKernelGGL(k_nrc_select, dim3(a.shade_blocks), dim3(blk), 0, s, a, pass);
        }
        for (uint32_t b = 0; b <= a.max_depth; ++b) {
            timed(SOLHIP_KERNEL_PT_EXTEND, [&] {
                if (a.count) hipLaunchKernelGGL(k_pt_extend<true>, dim3(a.persist_blocks), dim3(blk), lds, s, a, b);
                else hipLaunchKernelGGL(k_pt_extend<false>, dim3(a.persist_blocks), dim3(blk), lds, s, a, b);
            });
            timed(SOLHIP_KERNEL_PT_SHADE, [&] {
                if (nrc) hipLaunchKernelGGL(k_pt_shade_nrc, dim3(a.shade_blocks), dim3(blk), 0, s, a, b);
                else if (shadows) hipLaunchKernelGGL(k_pt_shade<8>, dim3(shade_grid(8)), dim3(512), 0, s, a, b);
                else hipLaunchKernelGGL(k_pt_shade<2>, dim3(shade_grid(2)), dim3(128), 0, s, a, b);
            });
            if (nrc && b == 1 && hook) {
                int rc = hook->after_query_shade(hook->user, s);
                if (rc) return rc;
            }
            if (shadows && b < a.max_depth) {
                timed(SOLHIP_KERNEL_PT_SHADOW, [&] {
                    if (a.count) hipLaunchKernelGGL(k_pt_shadow<true>, dim3(a.persist_blocks), dim3(blk), lds, s, a, b);
                    else hipLaunchKernelGGL(k_pt_shadow<false>, dim3(a.persist_blocks), dim3(blk), lds, s, a, b);
                });
            }
        }
        if (nrc) hipLaunchKernelGGL(k_nrc_targets, dim3(a.shade_blocks), dim3(blk), 0, s, a);
        int rc = launch_status("pathtrace pass");
        if (rc) return rc;
    }
    return SOLHIP_OK;
}

}  // namespace solhip
