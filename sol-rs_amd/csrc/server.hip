// server.hip -- the render-request server: jobs of mixed scene, image size
// and sample count traced by one wavefront.
//
// Device state is a pool of path slots (one per pixel of every running job),
// two extension queues, a shadow queue and a retire list.  One step is
//
//   extend  closest hits of every queued ray, each against its job's scene
//   shade   shade_vertex (the same code as the per-frame wavefront): next
//           ray, shadow ray, or the path's sample ends (-> retire list)
//   shadow  any-hit tests; unoccluded contributions join the path radiance
//   retire  the ended sample joins the pixel sum; the pixel's next sample
//           starts in the same slot (its RNG runs on, as in pathtrace.rgen's
//           sample loop) or, after the last sample, the pixel is accumulated
//           into the job's image and counted done
//
// so per pixel the arithmetic and random draws are those of the per-frame
// wavefront, in the same order: a job's image is bit-identical to a
// solhip_trace of the same program.  Admission is first-come first-served
// into free slot ranges; the host polls per-job done counters between steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <string>

#include "objects.h"
#include "pathshade.h"
#include "shading.h"
#include "traverse.h"
#include "traverse_step.h"
#include "fetch.h"
#include "accum.h"

namespace solhip {

namespace {

constexpr uint32_t kSrvBlock = 64;

// register budget of the traversal kernels: waves per SIMD they must fit
// BLAS exits restore the world ray's inverse direction from the ray state
// (trav_step's WINV), with a budget of 4 waves per SIMD for its registers:
// C3 capacity 12.04 -> 12.09 jobs/s, p50 at half load 130 -> 122 ms
#ifndef SOLHIP_SRV_WINV
#define SOLHIP_SRV_WINV true
#endif
#ifndef SOLHIP_SRV_WAVES
#define SOLHIP_SRV_WAVES 4
#endif
#define SOLHIP_SRV_OCCUPANCY __attribute__((amdgpu_waves_per_eu(SOLHIP_SRV_WAVES)))
// fetch partitions of the extend and shadow kernels (fetch.h)
constexpr uint32_t kSrvParts = 64;
constexpr uint32_t kMaxJobs = 256;   // job index lives in 8 bits of the slot meta word

// per-job constants (device array, one entry per running-job index)
struct JobDev {
    SceneView sv;
    DevTables tb;
    Cam cam;
    const LightTri *lights;
    float4 *accum;
    uint32_t *out_rgba8;
    uint32_t nlights;
    float total;
    uint32_t W, H, frame, spp, max_depth, rr_start, enable_sky, nee;
    float alpha, rmax;
    uint32_t base;      // first slot
    uint32_t streams;   // spec.sample_streams: sample s starts tea(p + s * W * H, frame)
    uint32_t lanes;     // slots per pixel (per-sample streams): slot base + j * npix + p runs samples j, j + lanes, ...
    uint32_t npix;      // W * H
};

struct SrvArgs {
    const JobDev *jobs;
    float4 *st0, *st1, *psum;
    FixSum *fsum;         // per-sample streams: per-slot order-free sums (lane 0's count: lanes done)
    uint32_t *pmeta;      // job | sample << 8
    QRay *q[2];
    QHit *hits;
    QRay *sq;
    float4 *scontrib;
    uint32_t *retire;
    uint32_t *counts;     // [0], [1] extension queues, [2] shadow queue, [3] retire list
    uint32_t *heads;      // partitioned fetch heads: extend, then shadow (kSrvParts each)
    uint32_t *done;       // finished pixels per job index
    Counters *counters;
};

__device__ __forceinline__ uint32_t srv_wave_append(uint32_t *counter, bool pred)
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

// shade and retire run blocks of kSrvShadeWaves waves and append to the
// queues once per block: three appends per pass over one counter each for
// all rays of every job made the per-wave atomics on those few addresses the
// largest cost of a step
constexpr uint32_t kSrvShadeWaves = 8;
constexpr uint32_t kSrvShadeBlock = 64 * kSrvShadeWaves;

// pathtrace.rgen:35 for this job's camera
__device__ __forceinline__ float job_tmin(const JobDev &J)
{
    float o[4];
    mat4_mul(J.cam.view_inverse, 0.0f, 0.0f, 0.0f, 1.0f, o);
    return fmaxf(1.0f, sqrtf(o[0] * o[0] + o[1] * o[1] + o[2] * o[2])) * 1e-3f;
}

// a new sample of pixel (x, y): jittered camera ray from the running RNG
__device__ __forceinline__ QRay start_sample(const JobDev &J, uint32_t slot, uint32_t x, uint32_t y, uint32_t &rng,
                                             float4 &s0, float4 &s1)
{
    const float jx = next_rand(rng);
    const float jy = next_rand(rng);
    V3 o, d;
    camera_ray(J.cam.view_inverse, J.cam.projection_inverse, x, y, J.W, J.H, jx, jy, o, d);
    s0 = make_float4(1.0f, 1.0f, 1.0f, as_f(rng));
    s1 = make_float4(0.0f, 0.0f, 0.0f, as_f(1u << 16));
    QRay r;
    r.ox = o.x;
    r.oy = o.y;
    r.oz = o.z;
    r.path = slot;
    r.dx = d.x;
    r.dy = d.y;
    r.dz = d.z;
    r.tmax = J.rmax;
    return r;
}

__global__ void k_srv_set_job(JobDev *jobs, uint32_t *done, uint32_t j, JobDev v)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        jobs[j] = v;
        done[j] = 0;
    }
}

// first sample of every pixel of job j, appended to queue cur (8x8 tiles)
// the first sample of every slot of job j (lane-major, 8x8 tiles within a
// lane), appended to queue cur; nslots = lanes * tile slots
__global__ void k_srv_admit(SrvArgs A, uint32_t j, uint32_t cur, uint32_t nslots)
{
    const JobDev &J = A.jobs[j];
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t tslots = nslots / J.lanes, ln = t / tslots, tt = t - ln * tslots;
    const uint32_t tiles_x = (J.W + 7) >> 3, tile = tt >> 6, lane = tt & 63u;
    const uint32_t x = (tile % tiles_x) * 8 + (lane & 7u), y = (tile / tiles_x) * 8 + (lane >> 3);
    const bool valid = t < nslots && x < J.W && y < J.H;
    QRay r;
    if (valid) {
        const uint32_t p = y * J.W + x, slot = J.base + ln * J.npix + p;
        // pathtrace.rgen:47; lane ln starts sample ln of its own stream
        uint32_t rng = tea(p + ln * J.npix, J.frame);
        float4 s0, s1;
        r = start_sample(J, slot, x, y, rng, s0, s1);
        A.st0[slot] = s0;
        A.st1[slot] = s1;
        if (J.streams) {
            FixSum fs;
            fix_clear(fs);
            A.fsum[slot] = fs;
        } else {
            A.psum[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        A.pmeta[slot] = j | (ln << 8);
    }
    const uint32_t at = srv_wave_append(A.counts + cur, valid);
    if (valid) A.q[cur][at] = r;
}

// extension / shadow rays of all running jobs, each against its job's scene
// through the resumable traversal; a lane takes new slots once at least 16
// lanes of its wave are idle
template <bool ANY>
__device__ __forceinline__ void srv_trace(const SrvArgs &A, const QRay *q, uint32_t cnt, uint32_t *heads, int kernel)
{
    extern __shared__ uint2 stack[];
    const uint32_t lane = threadIdx.x & 63u;
    TravCounts tc{0, 0, 0, false};
    uint32_t rays = 0;
    TravState st;
    SceneView sv;
    bool has = false;
    uint32_t idx = 0, path = 0;
    PartFetch pf;
    pf.init(heads, kSrvParts, cnt);
    while (true) {
        const uint64_t idle = __ballot(!has);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (!pf.exhausted && nidle >= 16u) {
            uint32_t end;
            const uint32_t base = pf.claim(nidle, end);
            if (!has) {
                const uint32_t my = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                if (my < end) {
                    const QRay r = q[my];
                    const JobDev &J = A.jobs[A.pmeta[r.path] & 0xFFu];
                    sv = J.sv;
                    sv.stack_depth += 2;   // the step traversal's two extra entries
                    trav_init(st, sv, v3(r.ox, r.oy, r.oz), v3(r.dx, r.dy, r.dz), job_tmin(J), r.tmax);
                    idx = my;
                    path = r.path;
                    has = true;
                }
            }
        }
        if (__ballot(has) == 0 && pf.exhausted) break;
        if (has && trav_step<ANY, false, SOLHIP_SRV_WINV>(st, sv, 0xFFu, stack + threadIdx.x, kSrvBlock, tc)) {
            const Hit h = trav_result(st);
            ++rays;
            if (!ANY) {
                QHit out;
                out.t = h.t;
                out.u = h.u;
                out.v = h.v;
                out.inst = h.inst;
                out.prim = h.prim;
                A.hits[idx] = out;
            } else if (h.inst == SOLHIP_MISS) {
                const float4 c = A.scontrib[idx];
                float4 s1 = A.st1[path];
                s1.x = s1.x + c.x;
                s1.y = s1.y + c.y;
                s1.z = s1.z + c.z;
                A.st1[path] = s1;
            }
            has = false;
        }
    }
    count_rays<false>(A.counters + kernel, rays, tc);
}

__global__ void __launch_bounds__(kSrvBlock) SOLHIP_SRV_OCCUPANCY k_srv_extend(SrvArgs A,
                                                                                                  uint32_t cur)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // filled by this step's shade / retire; last read a step ago
        A.counts[cur ^ 1u] = 0;
        A.counts[2] = 0;
        A.counts[3] = 0;
    }
    srv_trace<false>(A, A.q[cur], A.counts[cur], A.heads, SOLHIP_KERNEL_SRV_EXTEND);
}

__global__ void __launch_bounds__(kSrvBlock) SOLHIP_SRV_OCCUPANCY k_srv_shadow(SrvArgs A)
{
    srv_trace<true>(A, A.sq, A.counts[2], A.heads + kSrvParts * kHeadStride, SOLHIP_KERNEL_SRV_SHADOW);
}

// waves per SIMD the shade kernel is compiled for (0: the compiler's choice,
// 126 VGPRs and 4 waves; 5 fits in 96 VGPRs, but the 40-request load is
// level with it: p50 1352-1354 vs 1356 ms)
#ifndef SOLHIP_SRV_SHADE_WAVES
#define SOLHIP_SRV_SHADE_WAVES 0
#endif
#if SOLHIP_SRV_SHADE_WAVES
#define SOLHIP_SRV_SHADE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(SOLHIP_SRV_SHADE_WAVES)))
#else
#define SOLHIP_SRV_SHADE_OCCUPANCY
#endif

__global__ void __launch_bounds__(kSrvShadeBlock) SOLHIP_SRV_SHADE_OCCUPANCY k_srv_shade(SrvArgs A, uint32_t cur)
{
    __shared__ uint32_t lds[kSrvShadeWaves + 1];
    const uint32_t cnt = A.counts[cur];
    const uint32_t nthreads = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < cnt; base += nthreads) {
        const uint32_t i = base + threadIdx.x;
        ShadeOut o;
        bool ended = false;
        uint32_t slot = 0;
        if (i < cnt) {
            const QRay r = A.q[cur][i];
            slot = r.path;
            const JobDev &J = A.jobs[A.pmeta[slot] & 0xFFu];
            ShadeParams P;
            P.tb = J.tb;
            P.enable_sky = J.enable_sky;
            P.max_depth = J.max_depth;
            P.rr_start = J.rr_start;
            P.lights = J.lights;
            P.nlights = J.nlights;
            P.total = J.total;
            P.nee = J.nee && J.nlights > 0 && J.total > 0.0f;
            P.rmax = J.rmax;
            float4 s0 = A.st0[slot], s1 = A.st1[slot];
            shade_vertex(P, slot, r, A.hits[i], s0, s1, o);
            A.st0[slot] = s0;
            A.st1[slot] = s1;
            ended = !o.cont;
        }
        const uint32_t s_next = block_append<kSrvShadeWaves>(A.counts + (cur ^ 1u), o.cont, lds);
        if (o.cont) A.q[cur ^ 1u][s_next] = o.nr;
        const uint32_t s_sh = block_append<kSrvShadeWaves>(A.counts + 2, o.shadow, lds);
        if (o.shadow) {
            A.sq[s_sh] = o.sr;
            A.scontrib[s_sh] = o.sc;
        }
        const uint32_t s_rt = block_append<kSrvShadeWaves>(A.counts + 3, ended, lds);
        if (ended) A.retire[s_rt] = slot;
    }
}

__global__ void __launch_bounds__(kSrvShadeBlock) k_srv_retire(SrvArgs A, uint32_t cur)
{
    __shared__ uint32_t lds[kSrvShadeWaves + 1];
    if (blockIdx.x == 0) {
        // this step's extend and shadow are complete
        for (uint32_t k = threadIdx.x; k < 2 * kSrvParts; k += blockDim.x) A.heads[k * kHeadStride] = 0;
    }
    const uint32_t cnt = A.counts[3];
    const uint32_t nthreads = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < cnt; base += nthreads) {
        const uint32_t i = base + threadIdx.x;
        bool again = false;
        QRay r;
        if (i < cnt) {
            const uint32_t slot = A.retire[i];
            const uint32_t meta = A.pmeta[slot], job = meta & 0xFFu;
            const JobDev &J = A.jobs[job];
            const uint32_t sample = (meta >> 8) + J.lanes;   // this slot's next sample
            const uint32_t k = slot - J.base, ln = k / J.npix, p = k - ln * J.npix, x = p % J.W, y = p / J.W;
            const float4 s1 = A.st1[slot];
            float pc0 = 0.0f, pc1 = 0.0f, pc2 = 0.0f;
            FixSum fs;
            fix_clear(fs);
            bool last = true;
            if (J.streams) {
                fs = A.fsum[slot];
                fix_add(fs, s1.x, s1.y, s1.z);
            } else {
                const float4 ps = A.psum[slot];
                pc0 = ps.x + s1.x;
                pc1 = ps.y + s1.y;
                pc2 = ps.z + s1.z;
            }
            if (sample < J.spp) {
                if (J.streams) {
                    A.fsum[slot].c[0] = fs.c[0];
                    A.fsum[slot].c[1] = fs.c[1];
                    A.fsum[slot].c[2] = fs.c[2];
                    A.fsum[slot].bad = fs.bad;
                } else {
                    A.psum[slot] = make_float4(pc0, pc1, pc2, 0.0f);
                }
                uint32_t rng = J.streams ? tea(p + sample * J.npix, J.frame) : as_u(A.st0[slot].w);
                float4 s0n, s1n;
                r = start_sample(J, slot, x, y, rng, s0n, s1n);
                A.st0[slot] = s0n;
                A.st1[slot] = s1n;
                A.pmeta[slot] = job | (sample << 8);
                again = true;
                last = false;
            } else if (J.streams) {
                // this lane of the pixel is done: publish its sum; the pixel's
                // last lane (counted on lane 0's record) adds all of them up.
                // The count is only ever changed atomically, so the partial
                // sum is stored field by field.
                A.fsum[slot].c[0] = fs.c[0];
                A.fsum[slot].c[1] = fs.c[1];
                A.fsum[slot].c[2] = fs.c[2];
                A.fsum[slot].bad = fs.bad;
                if (J.lanes > 1) {
                    __threadfence();
                    last = atomicAdd(&A.fsum[J.base + p].count, 1u) == J.lanes - 1u;
                    if (last) {
                        __threadfence();
                        for (uint32_t l = 0; l < J.lanes; ++l)
                            if (l != ln) fix_merge(fs, A.fsum[J.base + l * J.npix + p]);
                    }
                }
                pc0 = fix_value(fs.c[0]);
                pc1 = fix_value(fs.c[1]);
                pc2 = fix_value(fs.c[2]);
            }
            if (!again && last) {
                // pathtrace.rgen:88-102
                const float inv = 1.0f / (float)J.spp;
                const float c0 = inv * pc0, c1 = inv * pc1, c2 = inv * pc2;
                float4 *px = J.accum + p;
                const float4 old = *px;
                float m0 = mixf(old.x, c0, J.alpha), m1 = mixf(old.y, c1, J.alpha), m2 = mixf(old.z, c2, J.alpha);
                const bool bad = (J.streams && fs.bad) || isnan(m0) || isnan(m1) || isnan(m2) || isinf(m0) ||
                                 isinf(m1) || isinf(m2);
                if (bad) {
                    m0 = old.x;
                    m1 = old.y;
                    m2 = old.z;
                }
                *px = make_float4(m0, m1, m2, 1.0f);
                if (J.out_rgba8) {
                    const float g = 1.0f / 2.2f;
                    const uint32_t r8 = to_unorm8(powf(m0, g)), g8 = to_unorm8(powf(m1, g)),
                                   b8 = to_unorm8(powf(m2, g));
                    J.out_rgba8[p] = r8 | (g8 << 8) | (b8 << 16) | (255u << 24);
                }
                atomicAdd(A.done + job, 1u);
            }
        }
        const uint32_t at = block_append<kSrvShadeWaves>(A.counts + (cur ^ 1u), again, lds);
        if (again) A.q[cur ^ 1u][at] = r;
    }
}

int srv_status(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string(what) + " launch: " + hipGetErrorString(e));
        return SOLHIP_ERROR_DEVICE;
    }
    return 0;
}

double now_ms(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

// bytes of device state per path slot: st0, st1, psum (48), order-free sum
// (32), meta (4), two extension queues (64), hits (20), shadow queue +
// contribution (48), retire (4)
constexpr uint64_t kSlotBytes = 48 + 32 + 4 + 64 + 20 + 48 + 4;
// slots per pixel of a job with per-sample streams when its spec leaves it
// open: enough lanes that a lane runs about kLaneSamples samples, within
// about kJobSlots slots for the job
constexpr uint32_t kLaneSamples = 16;
constexpr uint64_t kJobSlots = 4ull << 20;
constexpr uint32_t kLaneShare = 4;

struct SrvScene {
    SceneView sv;
    uint64_t tlas = 0, tlas_gen = 0;   // the TLAS and its generation at registration
    DevTables tb;
    DevBuf light_offsets, lights;
    uint32_t nlights = 0;
    float total = 0.0f;
};

struct SrvJob {
    solhip_job_desc desc{};
    uint32_t state = SOLHIP_JOB_QUEUED;
    double t_submit = 0, t_admit = 0, t_done = 0;
    uint32_t base = 0, npix = 0, jidx = 0;
    uint32_t lanes_max = 1;          // slots per pixel the job may use (per-sample streams)
    uint32_t lanes = 1, slots = 0;   // slots per pixel, slots of the job (lanes * npix), set at admission
};

struct Server : Obj {
    Context *cx = nullptr;
    hipStream_t stream = nullptr;
    uint32_t max_slots = 0, max_jobs = 0, max_stack = 1;
    DevBuf jobs, st0, st1, psum, fsum, pmeta, q0, q1, hits, sq, scontrib, retire, counts, heads, done;
    uint32_t *done_host = nullptr;   // pinned copy of the done counters
    std::vector<SrvScene *> scenes;
    std::map<uint64_t, SrvJob> all;
    std::deque<uint64_t> waiting;
    std::vector<uint64_t> running_by_idx;   // job id per job index (0 = free)
    std::map<uint32_t, uint32_t> free_ranges;   // start -> length
    uint64_t next_id = 1;
    uint32_t cur = 0;
    std::chrono::steady_clock::time_point t0;

    SrvArgs args() const
    {
        SrvArgs a;
        std::memset(&a, 0, sizeof(a));
        a.jobs = (const JobDev *)jobs.ptr;
        a.st0 = (float4 *)st0.ptr;
        a.st1 = (float4 *)st1.ptr;
        a.psum = (float4 *)psum.ptr;
        a.fsum = (FixSum *)fsum.ptr;
        a.pmeta = (uint32_t *)pmeta.ptr;
        a.q[0] = (QRay *)q0.ptr;
        a.q[1] = (QRay *)q1.ptr;
        a.hits = (QHit *)hits.ptr;
        a.sq = (QRay *)sq.ptr;
        a.scontrib = (float4 *)scontrib.ptr;
        a.retire = (uint32_t *)retire.ptr;
        a.counts = (uint32_t *)counts.ptr;
        a.heads = (uint32_t *)heads.ptr;
        a.done = (uint32_t *)done.ptr;
        a.counters = cx->counters;
        return a;
    }
    bool alloc_range(uint32_t n, uint32_t &base)
    {
        for (auto it = free_ranges.begin(); it != free_ranges.end(); ++it) {
            if (it->second < n) continue;
            base = it->first;
            const uint32_t rest = it->second - n;
            free_ranges.erase(it);
            if (rest) free_ranges[base + n] = rest;
            return true;
        }
        return false;
    }
    void free_range(uint32_t base, uint32_t n)
    {
        auto it = free_ranges.emplace(base, n).first;
        auto nx = std::next(it);
        if (nx != free_ranges.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_ranges.erase(nx);
        }
        if (it != free_ranges.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_ranges.erase(it);
            }
        }
    }
    void release_all()
    {
        for (DevBuf *b : {&jobs, &st0, &st1, &psum, &fsum, &pmeta, &q0, &q1, &hits, &sq, &scontrib, &retire, &counts, &heads,
                          &done})
            b->release();
        for (SrvScene *sc : scenes) {
            sc->light_offsets.release();
            sc->lights.release();
            delete sc;
        }
        scenes.clear();
        if (done_host) (void)hipHostFree(done_host);
        done_host = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }

    // admit waiting jobs (first come, first served) while slots and job
    // indices are free
    int admit()
    {
        while (!waiting.empty()) {
            SrvJob &jb = all[waiting.front()];
            uint32_t jidx = 0;
            while (jidx < max_jobs && running_by_idx[jidx]) ++jidx;
            if (jidx == max_jobs) return 0;
            // slots per pixel: what the job asks for, but extra lanes take at
            // most 1/kLaneShare of the largest free range, so that jobs
            // behind it are not kept waiting; one lane if that is all that fits
            uint64_t lanes = jb.lanes_max, largest = 0;
            for (const auto &r : free_ranges) largest = std::max<uint64_t>(largest, r.second);
            lanes = std::max<uint64_t>(1, std::min<uint64_t>(lanes, largest / ((uint64_t)kLaneShare * jb.npix)));
            uint32_t base = 0;
            if (!alloc_range((uint32_t)(lanes * jb.npix), base)) {
                if (lanes == 1 || !alloc_range(jb.npix, base)) return 0;
                lanes = 1;
            }
            jb.lanes = (uint32_t)lanes;
            jb.slots = jb.lanes * jb.npix;
            const SrvScene *sc = scenes[jb.desc.scene];
            Buffer *acc = get<Buffer>(jb.desc.accum, Kind::Buffer);
            Buffer *out = jb.desc.out_rgba8 ? get<Buffer>(jb.desc.out_rgba8, Kind::Buffer) : nullptr;
            if (!acc) {
                free_range(base, jb.slots);
                return bad_handle("job accum buffer");
            }
            JobDev v;
            std::memset(&v, 0, sizeof(v));
            v.sv = sc->sv;
            v.tb = sc->tb;
            std::memcpy(v.cam.view_inverse, jb.desc.uniforms.view_inverse, 64);
            std::memcpy(v.cam.projection_inverse, jb.desc.uniforms.projection_inverse, 64);
            v.lights = (const LightTri *)sc->lights.ptr;
            v.nlights = sc->nlights;
            v.total = sc->total;
            v.accum = (float4 *)acc->ptr;
            v.out_rgba8 = out ? (uint32_t *)out->ptr : nullptr;
            v.W = jb.desc.W;
            v.H = jb.desc.H;
            v.frame = jb.desc.uniforms.frame[2];
            v.spp = jb.desc.spec.spp_per_frame;
            v.max_depth = jb.desc.spec.max_depth;
            v.rr_start = jb.desc.spec.rr_start;
            v.enable_sky = jb.desc.spec.enable_sky;
            v.nee = jb.desc.spec.nee;
            v.streams = jb.desc.spec.sample_streams;
            const uint32_t denom = jb.desc.uniforms.frame[2] - (uint32_t)jb.desc.accum_start_frame + 1u;
            v.alpha = 1.0f / (float)denom;   // pathtrace.rgen:90
            v.rmax = 10000.0f;
            v.base = base;
            v.lanes = jb.lanes;
            v.npix = jb.npix;
            hipLaunchKernelGGL(k_srv_set_job, dim3(1), dim3(64), 0, stream, (JobDev *)jobs.ptr, (uint32_t *)done.ptr,
                               jidx, v);
            const uint32_t tiles = ((v.W + 7) / 8) * ((v.H + 7) / 8), nslots = tiles * 64 * jb.lanes;
            Timer *timer = cx->timer_or_null();
            if (timer) timer->begin(stream, SOLHIP_KERNEL_SRV_ADMIT);
            hipLaunchKernelGGL(k_srv_admit, dim3((nslots + kSrvBlock - 1) / kSrvBlock), dim3(kSrvBlock), 0, stream,
                               args(), jidx, cur, nslots);
            if (timer) timer->end(stream, SOLHIP_KERNEL_SRV_ADMIT);
            int rc = srv_status("server admit");
            if (rc) return rc;
            jb.state = SOLHIP_JOB_RUNNING;
            jb.base = base;
            jb.jidx = jidx;
            jb.t_admit = now_ms(t0);
            running_by_idx[jidx] = waiting.front();
            waiting.pop_front();
        }
        return 0;
    }

    int step()
    {
        const SrvArgs a = args();
        const uint32_t persist = cx->ncu * 4 * SOLHIP_SRV_WAVES;
        const uint32_t shade_blocks =
            std::min<uint32_t>((max_slots + kSrvShadeBlock - 1) / kSrvShadeBlock, cx->ncu * 4);
        const size_t lds = (size_t)(max_stack + 2) * sizeof(uint2) * kSrvBlock;
        Timer *timer = cx->timer_or_null();
        auto timed = [&](int id, auto &&launch) {
            if (timer) timer->begin(stream, id);
            launch();
            if (timer) timer->end(stream, id);
        };
        timed(SOLHIP_KERNEL_SRV_EXTEND,
              [&] { hipLaunchKernelGGL(k_srv_extend, dim3(persist), dim3(kSrvBlock), lds, stream, a, cur); });
        timed(SOLHIP_KERNEL_SRV_SHADE,
              [&] { hipLaunchKernelGGL(k_srv_shade, dim3(shade_blocks), dim3(kSrvShadeBlock), 0, stream, a, cur); });
        timed(SOLHIP_KERNEL_SRV_SHADOW,
              [&] { hipLaunchKernelGGL(k_srv_shadow, dim3(persist), dim3(kSrvBlock), lds, stream, a); });
        timed(SOLHIP_KERNEL_SRV_RETIRE,
              [&] { hipLaunchKernelGGL(k_srv_retire, dim3(shade_blocks), dim3(kSrvShadeBlock), 0, stream, a, cur); });
        cur ^= 1u;
        return srv_status("server step");
    }

    // reads the done counters; finished jobs give back their slots
    int poll()
    {
        SOLHIP_HIP_CHECK(hipMemcpyAsync(done_host, done.ptr, sizeof(uint32_t) * max_jobs, hipMemcpyDeviceToHost,
                                        stream));
        SOLHIP_HIP_CHECK(hipStreamSynchronize(stream));
        const double t = now_ms(t0);
        for (uint32_t j = 0; j < max_jobs; ++j) {
            const uint64_t id = running_by_idx[j];
            if (!id) continue;
            SrvJob &jb = all[id];
            if (done_host[j] >= jb.npix) {
                jb.state = SOLHIP_JOB_DONE;
                jb.t_done = t;
                free_range(jb.base, jb.slots);
                running_by_idx[j] = 0;
            }
        }
        return 0;
    }
    uint32_t live() const
    {
        uint32_t n = (uint32_t)waiting.size();
        for (uint64_t id : running_by_idx) n += id ? 1u : 0u;
        return n;
    }
};

}  // namespace solhip

using namespace solhip;

extern "C" {

int32_t solhip_server_create(solhip_ctx ctx, uint64_t hbm_budget, uint32_t max_jobs, solhip_server *out)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    if (!out) return bad_arg("out");
    if (max_jobs == 0 || max_jobs > kMaxJobs) return bad_arg("max_jobs must be in [1, 256]");
    const uint64_t slots = hbm_budget / kSlotBytes;
    if (slots < 64 || slots > 0xFFFFFF00ull) return bad_arg("HBM budget out of range");
    Server *sv = new Server();
    sv->kind = Kind::Server;
    sv->ctx = ctx;
    sv->cx = c;
    sv->max_slots = (uint32_t)slots;
    sv->max_jobs = max_jobs;
    sv->running_by_idx.assign(max_jobs, 0);
    sv->free_ranges[0] = sv->max_slots;
    sv->t0 = std::chrono::steady_clock::now();
    const uint64_t n = slots;
    bool fail = sv->jobs.reserve((uint64_t)max_jobs * sizeof(JobDev)) || sv->st0.reserve(n * 16) ||
                sv->st1.reserve(n * 16) || sv->psum.reserve(n * 16) || sv->fsum.reserve(n * sizeof(FixSum)) ||
                sv->pmeta.reserve(n * 4) ||
                sv->q0.reserve(n * sizeof(QRay)) || sv->q1.reserve(n * sizeof(QRay)) ||
                sv->hits.reserve(n * sizeof(QHit)) || sv->sq.reserve(n * sizeof(QRay)) ||
                sv->scontrib.reserve(n * 16) || sv->retire.reserve(n * 4) || sv->counts.reserve(256) ||
                sv->heads.reserve(2ull * kSrvParts * kHeadStride * 4) || sv->done.reserve((uint64_t)max_jobs * 4);
    if (!fail) fail = hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking) != hipSuccess ||
                      hipHostMalloc((void **)&sv->done_host, sizeof(uint32_t) * max_jobs, hipHostMallocDefault) !=
                          hipSuccess;
    if (fail) {
        sv->release_all();
        delete sv;
        set_error("render server allocation failed");
        return SOLHIP_ERROR_OUT_OF_MEMORY;
    }
    if (launch_fill(sv->stream, (uint32_t *)sv->counts.ptr, 64, 0u) ||
        launch_fill(sv->stream, (uint32_t *)sv->heads.ptr, 2ull * kSrvParts * kHeadStride, 0u) ||
        launch_fill(sv->stream, (uint32_t *)sv->done.ptr, max_jobs, 0u) || hipStreamSynchronize(sv->stream)) {
        sv->release_all();
        delete sv;
        return SOLHIP_ERROR_DEVICE;
    }
    *out = reg(sv);
    return SOLHIP_OK;
}

int32_t solhip_server_destroy(solhip_ctx ctx, solhip_server h)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    Server *sv = static_cast<Server *>(unreg(h, Kind::Server));
    if (!sv) return bad_handle("server");
    (void)hipStreamSynchronize(sv->stream);
    sv->release_all();
    delete sv;
    return SOLHIP_OK;
}

// the scene's TLAS still exists with the generation it had at registration,
// and its BLASes are those it was built over
static int scene_fresh(const SrvScene *sc)
{
    const Tlas *t = get<Tlas>(sc->tlas, Kind::Tlas);
    if (!t || t->gen != sc->tlas_gen) {
        set_error(std::string("the scene's TLAS was ") + (t ? "rebuilt" : "destroyed") +
                  " after solhip_server_scene registered it; register the scene again");
        return SOLHIP_ERROR_STALE;
    }
    return check_fresh(t);
}

int32_t solhip_server_scene(solhip_ctx ctx, solhip_server h, solhip_tlas th, const solhip_scene_tables *tables,
                            uint32_t *scene_id)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    Server *sv = get<Server>(h, Kind::Server);
    if (!sv) return bad_handle("server");
    Tlas *t = get<Tlas>(th, Kind::Tlas);
    if (!t) return bad_handle("TLAS");
    if (!tables || !tables->instances_addr || !tables->vertex_table_addr || !tables->index_table_addr ||
        !tables->material_table_addr || (tables->index_size != 4 && tables->index_size != 8) || !scene_id)
        return bad_arg("scene tables");
    if (const int rc = check_fresh(t)) return rc;
    if (stack_depth_of(t) > (uint32_t)kMaxStack) {
        set_error("acceleration structures too deep for the traversal stack");
        return SOLHIP_ERROR_UNSUPPORTED;
    }
    SrvScene *sc = new SrvScene();
    sc->sv = view_of(t);
    sc->tlas = th;
    sc->tlas_gen = t->gen;
    sc->tb.instances = (const solhip_scene_instance *)tables->instances_addr;
    sc->tb.vertex_table = (const uint64_t *)tables->vertex_table_addr;
    sc->tb.index_table = (const uint64_t *)tables->index_table_addr;
    sc->tb.material_table = (const uint64_t *)tables->material_table_addr;
    sc->tb.instance_count = tables->instance_count;
    sc->tb.index_size = tables->index_size;
    uint32_t nl = 0;
    int rc = build_lights(sv->stream, sc->tb, t, sc->light_offsets, sc->lights, &nl, nullptr);
    if (rc) {
        sc->light_offsets.release();
        sc->lights.release();
        delete sc;
        return rc;
    }
    sc->nlights = nl;
    if (nl) {
        LightTri last;
        SOLHIP_HIP_CHECK(hipMemcpyAsync(&last, (const LightTri *)sc->lights.ptr + (nl - 1), sizeof(LightTri),
                                        hipMemcpyDeviceToHost, sv->stream));
        SOLHIP_HIP_CHECK(hipStreamSynchronize(sv->stream));
        sc->total = last.cdf;
    }
    sv->max_stack = std::max(sv->max_stack, sc->sv.stack_depth);
    sv->scenes.push_back(sc);
    *scene_id = (uint32_t)(sv->scenes.size() - 1);
    return SOLHIP_OK;
}

int32_t solhip_server_submit(solhip_ctx ctx, solhip_server h, const solhip_job_desc *job, uint64_t *job_id)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    Server *sv = get<Server>(h, Kind::Server);
    if (!sv) return bad_handle("server");
    if (!job || !job_id) return bad_arg("job");
    if (job->scene >= sv->scenes.size()) return bad_arg("scene id");
    if (const int rc = scene_fresh(sv->scenes[job->scene])) return rc;
    if (job->W == 0 || job->H == 0) return bad_arg("image size");
    const uint64_t npix = (uint64_t)job->W * job->H;
    if (job->spec.sample_streams && (uint64_t)job->spec.spp_per_frame * npix > (1ull << 32))
        return bad_arg("sample_streams: spp_per_frame * W * H must not exceed 2^32");
    if (npix > sv->max_slots) return bad_arg("image larger than the server's slot budget");
    if (job->spec.spp_per_frame == 0 || job->spec.spp_per_frame > 0xFFFFFFu || job->spec.max_depth > 4096 ||
        job->spec.enable_sky > 2 || job->spec.sample_streams > 1)
        return bad_arg("job spec");
    Buffer *acc = get<Buffer>(job->accum, Kind::Buffer);
    if (!acc || acc->bytes < npix * 16) return bad_arg("job accum buffer");
    if (job->out_rgba8) {
        Buffer *o = get<Buffer>(job->out_rgba8, Kind::Buffer);
        if (!o || o->bytes < npix * 4) return bad_arg("job out_rgba8 buffer");
    }
    SrvJob jb;
    jb.desc = *job;
    jb.npix = (uint32_t)npix;
    if (job->spec.sample_streams) {
        const uint32_t spp = job->spec.spp_per_frame;
        uint64_t lanes = job->spec.wf_samples_per_pass;
        if (lanes == 0)
            lanes = std::min<uint64_t>((spp + kLaneSamples - 1) / kLaneSamples, std::max<uint64_t>(1, kJobSlots / npix));
        lanes = std::max<uint64_t>(1, std::min<uint64_t>({lanes, spp, sv->max_slots / npix}));
        jb.lanes_max = (uint32_t)lanes;
    }
    jb.t_submit = now_ms(sv->t0);
    const uint64_t id = sv->next_id++;
    sv->all[id] = jb;
    sv->waiting.push_back(id);
    *job_id = id;
    return SOLHIP_OK;
}

int32_t solhip_server_run(solhip_ctx ctx, solhip_server h, uint32_t max_steps, uint32_t *live)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    Server *sv = get<Server>(h, Kind::Server);
    if (!sv) return bad_handle("server");
    // a scene rebuilt since its registration would leave jobs reading freed
    // or changed acceleration structures: refuse to run them
    for (const auto &kv : sv->all) {
        const SrvJob &jb = kv.second;
        if (jb.state == SOLHIP_JOB_DONE) continue;
        if (const int rc = scene_fresh(sv->scenes[jb.desc.scene])) return rc;
    }
    for (uint32_t k = 0; k < max_steps; ++k) {
        int rc = sv->admit();
        if (rc) return rc;
        bool running = false;
        for (uint64_t id : sv->running_by_idx) running = running || id != 0;
        if (!running) break;
        rc = sv->step();
        if (rc) return rc;
        rc = sv->poll();
        if (rc) return rc;
    }
    if (live) *live = sv->live();
    return SOLHIP_OK;
}

int32_t solhip_server_job(solhip_ctx ctx, solhip_server h, uint64_t job_id, solhip_job_status *out)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    Server *sv = get<Server>(h, Kind::Server);
    if (!sv) return bad_handle("server");
    auto it = sv->all.find(job_id);
    if (it == sv->all.end() || !out) return bad_arg("job id");
    std::memset(out, 0, sizeof(*out));
    out->state = it->second.state;
    out->submit_ms = it->second.t_submit;
    out->admit_ms = it->second.t_admit;
    out->done_ms = it->second.t_done;
    return SOLHIP_OK;
}

}  // extern "C"
