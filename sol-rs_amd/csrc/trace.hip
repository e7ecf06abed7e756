// trace.hip -- raw ray queries (solhip_trace_rays) and small utility kernels.
#include <hip/hip_runtime.h>

#include "bvh.h"
#include "kernels.h"
#include "traverse.h"
#include "traverse_step.h"

namespace solhip {

constexpr int kTraceBlock = 64;

// XCD-aware block order: workgroups are dispatched round-robin over the 8
// XCDs, so physical block b runs on XCD b % 8.  Mapping each XCD to a
// contiguous range of logical blocks keeps neighbouring rays (and the BVH
// nodes they touch) in one XCD's L2.
__device__ __forceinline__ uint32_t xcd_logical_block(uint32_t b, uint32_t nb)
{
    const uint32_t x = b & 7u, q = nb >> 3, r = nb & 7u;
    return x * q + min(x, r) + (b >> 3);
}

__device__ __forceinline__ void add_counts(Counters *c, const TravCounts &t, bool active)
{
    // wave-level sums, one atomic per wave and counter
    unsigned long long nodes = active ? t.nodes : 0, tris = active ? t.tris : 0, insts = active ? t.insts : 0;
    for (int off = 32; off > 0; off >>= 1) {
        nodes += __shfl_xor(nodes, off);
        tris += __shfl_xor(tris, off);
        insts += __shfl_xor(insts, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&c->nodes, nodes);
        atomicAdd(&c->tris, tris);
        atomicAdd(&c->insts, insts);
    }
}

template <bool ANY, bool COUNT>
__global__ void __launch_bounds__(kTraceBlock) k_trace_rays(SceneView sv, const solhip_ray *rays, solhip_hit *hits,
                                                            uint32_t n, Counters *counters)
{
    extern __shared__ uint2 stack[];
    const uint32_t lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const uint32_t i = lb * kTraceBlock + threadIdx.x;
    const bool active = i < n;
    TravCounts tc{0, 0, 0, false};
    if (active) {
        const float4 *rp = reinterpret_cast<const float4 *>(rays + i);
        const float4 a = rp[0], b = rp[1];
        // the step traversal run to the end (trace_ray's visit order, hits
        // and counts, in fewer registers); sv.stack_depth includes its two
        // extra entries
        TravState st;
        trav_init(st, sv, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), a.w, b.w);
        while (!trav_step<ANY, COUNT>(st, sv, 0xFFu, stack + threadIdx.x, kTraceBlock, tc)) {
        }
        const Hit h = trav_result(st);
        solhip_hit o;
        o.t = h.t;
        o.u = h.u;
        o.v = h.v;
        o.inst = h.inst;
        o.prim = h.prim;
        o.geom = h.inst == SOLHIP_MISS ? SOLHIP_MISS : 0u;
        hits[i] = o;
    }
    const uint64_t ov = __ballot(active && tc.overflow);
    const uint32_t na = __popcll(__ballot(active));
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&counters->rays, (unsigned long long)na);
        if (ov) atomicAdd(&counters->overflow, (unsigned long long)__popcll(ov));
    }
    if (COUNT) add_counts(counters, tc, active);
}

int launch_trace_rays(hipStream_t s, const SceneView &sv, const solhip_ray *rays, solhip_hit *hits, uint32_t n,
                      uint32_t flags, Counters *counters, Timer *timer)
{
    if (n == 0) return 0;
    if (timer) timer->begin(s, SOLHIP_KERNEL_TRACE_RAYS);
    const dim3 grid((n + kTraceBlock - 1) / kTraceBlock), block(kTraceBlock);
    SceneView v = sv;
    v.stack_depth += 2;
    const size_t lds = (size_t)v.stack_depth * sizeof(uint2) * kTraceBlock;
    const bool any = flags & SOLHIP_TRACE_TERMINATE_ON_FIRST_HIT, count = flags & SOLHIP_TRACE_COUNT;
    if (any && count)
        hipLaunchKernelGGL((k_trace_rays<true, true>), grid, block, lds, s, v, rays, hits, n, counters);
    else if (any)
        hipLaunchKernelGGL((k_trace_rays<true, false>), grid, block, lds, s, v, rays, hits, n, counters);
    else if (count)
        hipLaunchKernelGGL((k_trace_rays<false, true>), grid, block, lds, s, v, rays, hits, n, counters);
    else
        hipLaunchKernelGGL((k_trace_rays<false, false>), grid, block, lds, s, v, rays, hits, n, counters);
    if (timer) timer->end(s, SOLHIP_KERNEL_TRACE_RAYS);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string("trace launch: ") + hipGetErrorString(e));
        return SOLHIP_ERROR_DEVICE;
    }
    return 0;
}

// (geometry, triangle in geometry) of hits on several-geometry BLASes: the
// geometry is the last one whose first BLAS-wide triangle is <= prim
__global__ void k_split_geometry(solhip_hit *hits, uint64_t n, const BlasRef *refs, uint32_t n_inst)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    solhip_hit h = hits[i];
    if (h.inst >= n_inst) return;   // a miss
    const uint32_t *geo = refs[h.inst].geo;
    if (!geo) return;
    uint32_t lo = 0, hi = geo[0];
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (geo[1 + mid] <= h.prim) lo = mid;
        else hi = mid;
    }
    h.prim -= geo[1 + lo];
    h.geom = lo;
    hits[i] = h;
}

int launch_split_geometry(hipStream_t s, solhip_hit *hits, uint64_t n, const BlasRef *refs, uint32_t n_inst)
{
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_split_geometry, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, hits, n, refs, n_inst);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string("split-geometry launch: ") + hipGetErrorString(e));
        return SOLHIP_ERROR_DEVICE;
    }
    return 0;
}

// fill kernel (used instead of hipMemsetAsync so that callers may capture
// the work into a HIP graph)
__global__ void k_fill(uint32_t *p, uint64_t n, uint32_t v)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = i; k < n; k += stride) p[k] = v;
}

int launch_fill(hipStream_t s, uint32_t *p, uint64_t nwords, uint32_t value)
{
    if (nwords == 0) return 0;
    uint64_t blocks = (nwords + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill, dim3((uint32_t)blocks), dim3(256), 0, s, p, nwords, value);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string("fill launch: ") + hipGetErrorString(e));
        return SOLHIP_ERROR_DEVICE;
    }
    return 0;
}

}  // namespace solhip
