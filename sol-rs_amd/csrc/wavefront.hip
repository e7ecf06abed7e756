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

__global__ void k_pt_generate(WfArgs A, uint32_t pass)
{
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot == 0) {
        A.counts[0] = A.nslots;
        A.counts[1] = 0;
        A.counts[2] = 0;
    }
    if (slot < 2 * (A.max_depth + 1)) A.heads[slot] = 0;
    if (slot >= A.nslots) return;
    uint32_t x, y;
    QRay r;
    if (!slot_pixel(slot, A.W, A.H, x, y)) {
        r.ox = r.oy = r.oz = r.dx = r.dy = r.dz = r.tmax = 0.0f;
        r.path = kNoPath;
        A.q[0][slot] = r;
        return;
    }
    const uint32_t p = y * A.W + x;
    uint32_t rng;
    if (pass == 0) {
        rng = tea(x + y * A.W, A.fc->frame);   // pathtrace.rgen:47
    } else {
        // the RNG runs on from the previous sample; its radiance joins the sum
        const float4 s0 = A.st0[p], s1 = A.st1[p];
        rng = as_u(s0.w);
        const float4 ps = pass == 1 ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : A.psum[p];
        A.psum[p] = make_float4(ps.x + s1.x, ps.y + s1.y, ps.z + s1.z, 0.0f);
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
    A.q[0][slot] = r;
}

template <bool ANY, bool COUNT>
__device__ __forceinline__ void trace_queue(const WfArgs &A, const QRay *q, uint32_t cnt, uint32_t *head, int kernel)
{
    extern __shared__ uint2 stack[];
    const uint32_t lane = threadIdx.x & 63u;
    const float tmin = ray_tmin(A);
    TravCounts tc{0, 0, 0, false};
    uint32_t rays = 0;
    for (;;) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(head, 64u);
        base = __shfl(base, 0);
        if (base >= cnt) break;
        const uint32_t i = base + lane;
        if (i < cnt) {
            const QRay r = q[i];
            if (!ANY) {
                solhip_hit out;
                out.t = -1.0f;
                out.u = out.v = 0.0f;
                out.inst = out.prim = SOLHIP_MISS;
                if (r.path != kNoPath) {
                    const Hit h = trace_ray<false, COUNT>(A.sv, v3(r.ox, r.oy, r.oz), v3(r.dx, r.dy, r.dz), tmin,
                                                          r.tmax, 0xFFu, stack + threadIdx.x, A.block, tc);
                    ++rays;
                    out.t = h.t;
                    out.u = h.u;
                    out.v = h.v;
                    out.inst = h.inst;
                    out.prim = h.prim;
                }
                A.hits[i] = out;
            } else {
                const Hit h = trace_ray<true, COUNT>(A.sv, v3(r.ox, r.oy, r.oz), v3(r.dx, r.dy, r.dz), tmin, r.tmax,
                                                     0xFFu, stack + threadIdx.x, A.block, tc);
                ++rays;
                if (h.inst == SOLHIP_MISS) {
                    const float4 c = A.scontrib[i];
                    float4 s1 = A.st1[r.path];
                    s1.x = s1.x + c.x;
                    s1.y = s1.y + c.y;
                    s1.z = s1.z + c.z;
                    A.st1[r.path] = s1;
                }
            }
        }
    }
    count_rays<COUNT>(A.counters + kernel, rays, tc);
}

template <bool COUNT>
__global__ void k_pt_extend(WfArgs A, uint32_t b)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // the queues this bounce's shade fills were last read a bounce ago
        A.counts[(b + 1) & 1] = 0;
        A.counts[2] = 0;
    }
    trace_queue<false, COUNT>(A, A.q[b & 1], A.counts[b & 1], A.heads + 2 * b, SOLHIP_KERNEL_PT_EXTEND);
}

template <bool COUNT>
__global__ void k_pt_shadow(WfArgs A, uint32_t b)
{
    trace_queue<true, COUNT>(A, A.sq, A.counts[2], A.heads + 2 * b + 1, SOLHIP_KERNEL_PT_SHADOW);
}

__global__ void k_pt_shade(WfArgs A, uint32_t b)
{
    const uint32_t cnt = A.counts[b & 1];
    const QRay *qin = A.q[b & 1];
    QRay *qout = A.q[(b + 1) & 1];
    uint32_t *cout = A.counts + ((b + 1) & 1);
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
        const uint32_t s_next = wave_append(cout, o.cont);
        if (o.cont) qout[s_next] = o.nr;
        const uint32_t s_sh = wave_append(A.counts + 2, o.shadow);
        if (o.shadow) {
            A.sq[s_sh] = o.sr;
            A.scontrib[s_sh] = o.sc;
        }
    }
}

__global__ void k_pt_finalize(WfArgs A)
{
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t x, y;
    if (slot >= A.nslots || !slot_pixel(slot, A.W, A.H, x, y)) return;
    const uint32_t p = y * A.W + x;
    const float4 s1 = A.st1[p];
    const float4 ps = A.spp == 1 ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : A.psum[p];
    float pc0 = ps.x + s1.x, pc1 = ps.y + s1.y, pc2 = ps.z + s1.z;
    const float inv = 1.0f / (float)A.spp;   // pathtrace.rgen:88
    pc0 = inv * pc0;
    pc1 = inv * pc1;
    pc2 = inv * pc2;
    float4 *px = A.accum + p;
    const float4 old = *px;
    const float alpha = A.fc->alpha;
    float m0 = mixf(old.x, pc0, alpha), m1 = mixf(old.y, pc1, alpha), m2 = mixf(old.z, pc2, alpha);
    const bool bad = isnan(m0) || isnan(m1) || isnan(m2) || isinf(m0) || isinf(m1) || isinf(m2);
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

int launch_pathtrace(hipStream_t s, const WfArgs &a, Timer *timer)
{
    const uint32_t blk = a.block;
    const size_t lds = (size_t)a.sv.stack_depth * sizeof(uint2) * blk;
    const uint32_t nheads = 2 * (a.max_depth + 1);
    const uint32_t gen_blocks = (std::max(a.nslots, nheads) + blk - 1) / blk;
    const uint32_t fin_blocks = (a.nslots + blk - 1) / blk;
    const bool shadows = a.nee && a.nlights > 0;
    auto timed = [&](int id, auto &&launch) {
        if (timer) timer->begin(s, id);
        launch();
        if (timer) timer->end(s, id);
    };
    for (uint32_t pass = 0; pass < a.spp; ++pass) {
        timed(SOLHIP_KERNEL_PT_GENERATE,
              [&] { hipLaunchKernelGGL(k_pt_generate, dim3(gen_blocks), dim3(blk), 0, s, a, pass); });
        for (uint32_t b = 0; b <= a.max_depth; ++b) {
            timed(SOLHIP_KERNEL_PT_EXTEND, [&] {
                if (a.count) hipLaunchKernelGGL(k_pt_extend<true>, dim3(a.persist_blocks), dim3(blk), lds, s, a, b);
                else hipLaunchKernelGGL(k_pt_extend<false>, dim3(a.persist_blocks), dim3(blk), lds, s, a, b);
            });
            timed(SOLHIP_KERNEL_PT_SHADE,
                  [&] { hipLaunchKernelGGL(k_pt_shade, dim3(a.shade_blocks), dim3(blk), 0, s, a, b); });
            if (shadows && b < a.max_depth) {
                timed(SOLHIP_KERNEL_PT_SHADOW, [&] {
                    if (a.count) hipLaunchKernelGGL(k_pt_shadow<true>, dim3(a.persist_blocks), dim3(blk), lds, s, a, b);
                    else hipLaunchKernelGGL(k_pt_shadow<false>, dim3(a.persist_blocks), dim3(blk), lds, s, a, b);
                });
            }
        }
        int rc = launch_status("pathtrace pass");
        if (rc) return rc;
    }
    timed(SOLHIP_KERNEL_PT_FINALIZE, [&] { hipLaunchKernelGGL(k_pt_finalize, dim3(fin_blocks), dim3(blk), 0, s, a); });
    return launch_status("pathtrace finalize");
}

}  // namespace solhip
