// bvh_build.hip -- GPU LBVH build for BLAS (triangles) and TLAS (instances).
//
// Replaces the driver's acceleration-structure build behind
// vkCmdBuildAccelerationStructuresKHR (src/ray/acceleration.rs:44-126 via
// BLAS::new 136-239 and TLAS::new/regenerate 344-467).
//
// Pipeline (all on the stream the caller passes):
//   1. primitive boxes + centroid bounds   (triangles: 3 vertex fetches at the
//      caller's stride; instances: 8 transformed corners of the BLAS root box)
//   2. 30-bit Morton codes of the centroids (10 bits per axis)
//   3. stable LSD radix sort of (code, index), 8-bit digits: per-tile LDS
//      histograms, one exclusive scan, per-tile stable scatter whose local
//      ranks come from wave64 ballots (match on the 8 digit bits)
//   4. Karras 2012 binary radix tree (index tie-break for equal codes)
//   5. bottom-up box refit in two phases: inside each 256-leaf block with
//      LDS flags, then across blocks from the frontier nodes with one
//      agent-scope acq_rel flag per node
//   6. top-down collapse, one launch per level, to 8-wide nodes: a node
//      repeatedly opens its largest-area internal child until it has 8
//      children; children of <= leaf_max primitives become leaves, the others
//      new nodes; slots are assigned by the octant heuristic; child boxes are
//      quantised to 8 bits, rounded outward and padded by one step
//   7. leaf records in node order (48-B triangles / 64-B instances)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "bvh.h"
#include "bvh_logic.h"
#include "common.h"

namespace solhip {

namespace {

constexpr int kBlock = 256;
constexpr int kSortItems = 8;                 // keys per thread per tile
constexpr int kSortTile = kBlock * kSortItems;  // 2048 keys

__device__ __forceinline__ uint32_t f2ord(float f)
{
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u)
{
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}
// block-wide min/max of centroids, then one atomic per block and component
__device__ void reduce_bounds(float c[3], bool valid, uint32_t *bounds)
{
    __shared__ uint32_t smin[3][kBlock / 64], smax[3][kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        uint32_t mn = valid ? f2ord(c[a]) : 0xFFFFFFFFu;
        uint32_t mx = valid ? f2ord(c[a]) : 0u;
        for (int off = 32; off > 0; off >>= 1) {
            mn = min(mn, (uint32_t)__shfl_xor((int)mn, off));
            mx = max(mx, (uint32_t)__shfl_xor((int)mx, off));
        }
        if (lane == 0) {
            smin[a][wave] = mn;
            smax[a][wave] = mx;
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int a = threadIdx.x;
        uint32_t mn = 0xFFFFFFFFu, mx = 0u;
        for (int w = 0; w < kBlock / 64; ++w) {
            mn = min(mn, smin[a][w]);
            mx = max(mx, smax[a][w]);
        }
        atomicMin(&bounds[a], mn);
        atomicMax(&bounds[3 + a], mx);
    }
}

__device__ __forceinline__ const GeomDesc &geom_of(const GeomDesc *geo, uint32_t n_geo, uint32_t t)
{
    return geo[geom_index(geo, n_geo, t)];
}

// vertex k of BLAS-wide triangle t of geometry G; an index value >=
// vertex_count is counted in *bad (the build then fails with an error)
__device__ __forceinline__ const float *tri_vertex(const GeomDesc &G, uint32_t t, int k, uint32_t *bad)
{
    uint32_t iv;
    if (!tri_vertex_index(G, t, k, iv)) atomicAdd(bad, 1u);
    return reinterpret_cast<const float *>(G.vbase + (uint64_t)(G.first_vertex + iv) * G.stride);
}

__global__ void __launch_bounds__(kBlock) k_tri_boxes(const GeomDesc *geo, uint32_t n_geo, uint32_t n, float4 *boxes,
                                                      uint32_t *bounds, uint32_t *bad)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    float c[3] = {0.0f, 0.0f, 0.0f};
    const bool valid = i < n;
    if (valid) {
        const GeomDesc &G = geom_of(geo, n_geo, i);
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int k = 0; k < 3; ++k) {
            const float *p = tri_vertex(G, i, k, bad);
            for (int a = 0; a < 3; ++a) {
                lo[a] = fminf(lo[a], p[a]);
                hi[a] = fmaxf(hi[a], p[a]);
            }
        }
        boxes[2ull * i] = make_float4(lo[0], lo[1], lo[2], 0.0f);
        boxes[2ull * i + 1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
        for (int a = 0; a < 3; ++a) c[a] = 0.5f * lo[a] + 0.5f * hi[a];
    }
    reduce_bounds(c, valid, bounds);
}

// instance world boxes: the 8 corners of the BLAS root box through the 3x4
// object-to-world transform, padded relatively
__global__ void __launch_bounds__(kBlock) k_inst_boxes(const solhip_instance *inst, const BlasRef *refs, uint32_t n,
                                                       float4 *boxes, uint32_t *bounds)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    float c[3] = {0.0f, 0.0f, 0.0f};
    const bool valid = i < n;
    if (valid) {
        float lo[3], hi[3];
        instance_box(inst[i].transform, refs[i].nodes != nullptr, refs[i].root_box, lo, hi);
        boxes[2ull * i] = make_float4(lo[0], lo[1], lo[2], 0.0f);
        boxes[2ull * i + 1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
        for (int a = 0; a < 3; ++a) c[a] = 0.5f * lo[a] + 0.5f * hi[a];
    }
    reduce_bounds(c, valid, bounds);
}

__global__ void __launch_bounds__(kBlock) k_morton(const float4 *boxes, uint32_t n, const uint32_t *bounds,
                                                   uint32_t *keys, uint32_t *vals)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 lo = boxes[2ull * i], hi = boxes[2ull * i + 1];
    const float l3[3] = {lo.x, lo.y, lo.z}, h3[3] = {hi.x, hi.y, hi.z};
    const float blo[3] = {ord2f(bounds[0]), ord2f(bounds[1]), ord2f(bounds[2])};
    const float bhi[3] = {ord2f(bounds[3]), ord2f(bounds[4]), ord2f(bounds[5])};
    keys[i] = morton30(l3, h3, blo, bhi);
    vals[i] = i;
}

// ---- radix sort ---------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_hist(const uint32_t *keys, uint32_t n, int shift, uint32_t *hist,
                                                 uint32_t ntiles)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kSortTile;
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const uint64_t i = base + (uint64_t)r * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// scan of up to 1024 * kScanItems values per block; block sums out
constexpr int kScanItems = 4;
__global__ void __launch_bounds__(kBlock) k_scan_block(uint32_t *data, uint32_t n, uint32_t *block_sums)
{
    __shared__ uint32_t s[kBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kBlock * kScanItems;
    uint32_t v[kScanItems];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t i = base + (uint64_t)threadIdx.x * kScanItems + k;
        v[k] = i < n ? data[i] : 0u;
        sum += v[k];
    }
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        const uint32_t t = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t run = s[threadIdx.x] - sum;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t i = base + (uint64_t)threadIdx.x * kScanItems + k;
        if (i < n) data[i] = run;
        run += v[k];
    }
    if (threadIdx.x == kBlock - 1 && block_sums) block_sums[blockIdx.x] = s[kBlock - 1];
}

__global__ void __launch_bounds__(kBlock) k_scan_add(uint32_t *data, uint32_t n, const uint32_t *block_offs)
{
    const uint64_t base = (uint64_t)blockIdx.x * kBlock * kScanItems;
    const uint32_t add = block_offs[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t i = base + (uint64_t)threadIdx.x * kScanItems + k;
        if (i < n) data[i] += add;
    }
}

__global__ void __launch_bounds__(kBlock) k_scatter(const uint32_t *keys_in, const uint32_t *vals_in, uint32_t n,
                                                    int shift, const uint32_t *offs, uint32_t ntiles,
                                                    uint32_t *keys_out, uint32_t *vals_out, uint32_t *err)
{
    __shared__ uint32_t base_off[256];   // global offset of this tile's digits, advanced per round
    __shared__ uint32_t wave_cnt[kBlock / 64][256];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    base_off[threadIdx.x] = offs[(uint64_t)threadIdx.x * ntiles + blockIdx.x];
    const uint64_t tile = (uint64_t)blockIdx.x * kSortTile;
    const uint64_t lanes_below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int r = 0; r < kSortItems; ++r) {
        for (int w = 0; w < kBlock / 64; ++w) wave_cnt[w][threadIdx.x] = 0;
        __syncthreads();
        const uint64_t i = tile + (uint64_t)r * kBlock + threadIdx.x;
        const bool valid = i < n;
        const uint32_t key = valid ? keys_in[i] : 0u;
        const uint32_t val = valid ? vals_in[i] : 0u;
        const uint32_t digit = (key >> shift) & 0xFFu;
        // lanes of this wave holding the same digit (8 ballots)
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t m = __ballot((digit >> b) & 1u);
            peers &= ((digit >> b) & 1u) ? m : ~m;
        }
        const uint32_t rank_in_wave = __popcll(peers & lanes_below);
        if (valid && rank_in_wave == 0) wave_cnt[wave][digit] = __popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = base_off[digit] + rank_in_wave;
            for (int w = 0; w < wave; ++w) pos += wave_cnt[w][digit];
            if (pos < n) {
                keys_out[pos] = key;
                vals_out[pos] = val;
            } else {
                atomicAdd(err, 1u);
            }
        }
        __syncthreads();
        uint32_t tot = 0;
        for (int w = 0; w < kBlock / 64; ++w) tot += wave_cnt[w][threadIdx.x];
        base_off[threadIdx.x] += tot;
        __syncthreads();
    }
}

// ---- Karras 2012 radix tree ---------------------------------------------
__global__ void __launch_bounds__(kBlock) k_karras(const uint32_t *keys, int n, uint32_t *left, uint32_t *right,
                                                   uint32_t *parent, uint32_t *rfirst, uint32_t *ncount)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n - 1) return;
    karras_node(keys, n, i, left, right, parent, rfirst, ncount);
    if (i == 0) parent[0] = 0xFFFFFFFFu;
}

// bottom-up boxes; nbox[2*node] = lo, [2*node+1] = hi (ncount of internal
// nodes comes from k_karras: the size of the node's leaf range).
//
// Two phases.  k_refit_local: block b owns sorted leaves [256b, 256b+256);
// an internal node whose leaf range lies inside that span has its index in
// it too (one end of a radix-tree node's range is the node itself), so the
// block builds those nodes with workgroup-scope LDS flags and LDS boxes (the
// second child to arrive merges).  A node whose parent's range leaves the
// span is a frontier node: k_refit_cross walks up from the frontier nodes
// with agent-scope flags, over the few nodes that span blocks.
constexpr uint32_t kRefitSpan = kBlock;

__device__ __forceinline__ void box_union(float4 &lo, float4 &hi, float4 blo, float4 bhi)
{
    lo = make_float4(fminf(lo.x, blo.x), fminf(lo.y, blo.y), fminf(lo.z, blo.z), 0.0f);
    hi = make_float4(fmaxf(hi.x, bhi.x), fmaxf(hi.y, bhi.y), fmaxf(hi.z, bhi.z), 0.0f);
}

__global__ void __launch_bounds__(kBlock) k_refit_local(const float4 *pbox, const uint32_t *vals, int n,
                                                        const uint32_t *left, const uint32_t *right,
                                                        const uint32_t *parent, const uint32_t *rfirst,
                                                        uint32_t *ncount, float4 *nbox, uint32_t *frontier,
                                                        uint32_t *nfrontier, uint32_t *err)
{
    __shared__ uint32_t lflag[kRefitSpan];
    __shared__ float4 lbox[kRefitSpan][2];
    const uint32_t L0 = blockIdx.x * kRefitSpan, L1 = min((uint32_t)n, L0 + kRefitSpan);
    lflag[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t k = L0 + threadIdx.x;
    if (k >= L1) return;
    const uint32_t leaf_base = (uint32_t)(n - 1);
    uint32_t node = leaf_base + k;
    const uint32_t pk = vals[k];
    float4 lo = pbox[2ull * pk], hi = pbox[2ull * pk + 1];
    nbox[2ull * node] = lo;
    nbox[2ull * node + 1] = hi;
    ncount[node] = 1;
    if (n == 1) return;
    while (true) {
        const uint32_t par = parent[node];
        if (par == 0xFFFFFFFFu) break;
        if (par >= leaf_base) {
            atomicAdd(err, 1u);
            break;
        }
        const uint32_t plo = rfirst[par], phi = plo + ncount[par] - 1u;
        if (plo < L0 || phi >= L1) {
            frontier[atomicAdd(nfrontier, 1u)] = node;
            break;
        }
        const uint32_t old = __hip_atomic_fetch_add(&lflag[par - L0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (old == 0) break;
        const uint32_t l = left[par], r = right[par];
        const uint32_t sib = l == node ? r : l;
        float4 slo, shi;
        if (sib >= leaf_base) {
            if (sib - leaf_base >= (uint32_t)n) {
                atomicAdd(err, 1u);
                break;
            }
            const uint32_t ps = vals[sib - leaf_base];
            slo = pbox[2ull * ps];
            shi = pbox[2ull * ps + 1];
        } else {
            if (sib < L0 || sib >= L1) {
                atomicAdd(err, 1u);
                break;
            }
            slo = lbox[sib - L0][0];
            shi = lbox[sib - L0][1];
        }
        box_union(lo, hi, slo, shi);
        lbox[par - L0][0] = lo;
        lbox[par - L0][1] = hi;
        nbox[2ull * par] = lo;
        nbox[2ull * par + 1] = hi;
        node = par;
    }
}

__global__ void __launch_bounds__(kBlock) k_refit_cross(int n, const uint32_t *left, const uint32_t *right,
                                                        const uint32_t *parent, uint32_t *flags, float4 *nbox,
                                                        const uint32_t *frontier, const uint32_t *nfrontier,
                                                        uint32_t *err)
{
    const uint32_t count = min(*nfrontier, (uint32_t)n);
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < count; f += stride) {
        uint32_t node = frontier[f];
        float4 lo = nbox[2ull * node], hi = nbox[2ull * node + 1];
        while (true) {
            const uint32_t par = parent[node];
            if (par == 0xFFFFFFFFu) break;
            if (par >= (uint32_t)(n - 1)) {
                atomicAdd(err, 1u);
                break;
            }
            const uint32_t old = __hip_atomic_fetch_add(&flags[par], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (old == 0) break;
            const uint32_t l = left[par], r = right[par];
            if (l >= 2u * (uint32_t)n - 1u || r >= 2u * (uint32_t)n - 1u) {
                atomicAdd(err, 1u);
                break;
            }
            const uint32_t sib = l == node ? r : l;
            box_union(lo, hi, nbox[2ull * sib], nbox[2ull * sib + 1]);
            nbox[2ull * par] = lo;
            nbox[2ull * par + 1] = hi;
            node = par;
        }
    }
}

// ---- binned SAH binary tree (fast-trace BLAS builds) ---------------------
// Top down, one launch per tree level, one workgroup per node of the level
// (a task: first position of the node's range in the primitive order, the
// binary node).  The workgroup finds the node's centroid range and box, bins
// its primitives (sah_choose), and partitions the range stably into the
// other order buffer.  A side with one primitive is a leaf, placed at its
// final position (leaf node n-1+position); a larger side gets the next
// internal node index and becomes a task of the next level.  Node indices
// follow the order in which workgroups finish, but the tree's shape does not
// depend on it: it equals the host replay's.  Nodes of at least big_min
// primitives go to a second task list, run by 1024-thread workgroups (the
// top levels, where a few workgroups work through millions of primitives).
constexpr uint32_t kSahBigMin = 65536;

struct SahArgs {
    const float4 *boxes;   // 2 per primitive
    const uint32_t *ids_in;
    uint32_t *ids_out;
    uint32_t *vals;        // primitive of each leaf position
    const uint2 *tasks_in;     // this launch's tasks (small or big list)
    uint2 *tasks_out;          // next level, nodes of fewer than big_min primitives
    uint2 *big_out;            // next level, nodes of big_min or more
    uint32_t *ntasks_out, *nbig_out;
    uint32_t *node_alloc;
    uint32_t *left, *right, *rfirst, *ncount;
    float4 *nbox;
    uint32_t n, task_cap, big_cap, big_min, level;
    uint32_t *err;
};

struct SahLdsBins {
    const uint32_t (*cnt)[kSahBins];
    const uint32_t (*lo_)[kSahBins][3];
    const uint32_t (*hi_)[kSahBins][3];
    __device__ uint32_t count(int a, int b) const { return cnt[a][b]; }
    __device__ float lo(int a, int b, int k) const { return ord2f(lo_[a][b][k]); }
    __device__ float hi(int a, int b, int k) const { return ord2f(hi_[a][b][k]); }
};

__device__ __forceinline__ void prim_centroid(const float4 *boxes, uint32_t p, float4 &l, float4 &h, float c[3])
{
    l = boxes[2ull * p];
    h = boxes[2ull * p + 1];
    c[0] = 0.5f * l.x + 0.5f * h.x;
    c[1] = 0.5f * l.y + 0.5f * h.y;
    c[2] = 0.5f * l.z + 0.5f * h.z;
}

template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_sah_level(SahArgs A)
{
    __shared__ uint32_t s_cnt[3][kSahBins];
    __shared__ uint32_t s_lo[3][kSahBins][3], s_hi[3][kSahBins][3];
    __shared__ uint32_t s_red[BLOCK / 64][12];
    __shared__ uint32_t s_bounds[12];
    __shared__ SahSplit s_axis_best[3];
    __shared__ float s_axis_cost[3];
    __shared__ int s_split[3];
    __shared__ uint32_t s_wave[BLOCK / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint2 t = A.tasks_in[blockIdx.x];
    const uint32_t lo = t.x, node = t.y, n = A.n;
    // block-uniform checks: every thread of the block takes the same branch
    if (node >= n - 1u) {
        if (tid == 0) atomicAdd(A.err, 1u);
        return;
    }
    const uint32_t cnt = A.ncount[node];
    if (cnt < 2u || cnt > n || lo > n - cnt) {
        if (tid == 0) atomicAdd(A.err, 1u);
        return;
    }
    const uint32_t hi = lo + cnt;
    // pass 1: centroid range and node box (ordered-integer min / max)
    uint32_t r[12];
    for (int k = 0; k < 3; ++k) {
        r[k] = 0xFFFFFFFFu;      // centroid min
        r[3 + k] = 0u;           // centroid max
        r[6 + k] = 0xFFFFFFFFu;  // box min
        r[9 + k] = 0u;           // box max
    }
    for (uint32_t i = lo + tid; i < hi; i += BLOCK) {
        const uint32_t p = min(A.ids_in[i], n - 1u);
        float4 l, h;
        float c[3];
        prim_centroid(A.boxes, p, l, h, c);
        const float lv[3] = {l.x, l.y, l.z}, hv[3] = {h.x, h.y, h.z};
        for (int k = 0; k < 3; ++k) {
            r[k] = min(r[k], f2ord(c[k]));
            r[3 + k] = max(r[3 + k], f2ord(c[k]));
            r[6 + k] = min(r[6 + k], f2ord(lv[k]));
            r[9 + k] = max(r[9 + k], f2ord(hv[k]));
        }
    }
    for (int k = 0; k < 12; ++k) {
        const bool is_min = (k % 6) < 3;
        for (int off = 32; off > 0; off >>= 1) {
            const uint32_t o = (uint32_t)__shfl_xor((int)r[k], off);
            r[k] = is_min ? min(r[k], o) : max(r[k], o);
        }
    }
    if (lane == 0)
        for (int k = 0; k < 12; ++k) s_red[wave][k] = r[k];
    // bins start empty
    for (uint32_t j = tid; j < 3u * kSahBins; j += BLOCK) {
        (&s_cnt[0][0])[j] = 0u;
        for (int k = 0; k < 3; ++k) {
            (&s_lo[0][0][0])[3 * j + k] = f2ord(INFINITY);
            (&s_hi[0][0][0])[3 * j + k] = f2ord(-INFINITY);
        }
    }
    __syncthreads();
    if (tid < 12) {
        const bool is_min = (tid % 6) < 3;
        uint32_t v = s_red[0][tid];
        for (int w = 1; w < BLOCK / 64; ++w) v = is_min ? min(v, s_red[w][tid]) : max(v, s_red[w][tid]);
        s_bounds[tid] = v;
    }
    __syncthreads();
    float cmin[3], scale[3];
    for (int k = 0; k < 3; ++k) {
        cmin[k] = ord2f(s_bounds[k]);
        scale[k] = sah_scale(ord2f(s_bounds[3 + k]) - cmin[k]);
    }
    if (tid == 0) {
        A.nbox[2ull * node] = make_float4(ord2f(s_bounds[6]), ord2f(s_bounds[7]), ord2f(s_bounds[8]), 0.0f);
        A.nbox[2ull * node + 1] = make_float4(ord2f(s_bounds[9]), ord2f(s_bounds[10]), ord2f(s_bounds[11]), 0.0f);
    }
    // pass 2: bins.  Neighbouring primitives mostly share a bin, so a wave
    // whose valid lanes all fall in one bin reduces them with shuffles and
    // makes one update; otherwise each lane updates its bin.
    for (uint32_t base = lo; base < hi; base += BLOCK) {
        const uint32_t i = base + tid;
        const bool valid = i < hi;
        float4 l = make_float4(INFINITY, INFINITY, INFINITY, 0.0f), h = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
        float c[3] = {0.0f, 0.0f, 0.0f};
        if (valid) prim_centroid(A.boxes, min(A.ids_in[i], n - 1u), l, h, c);
        const uint64_t vmask = __ballot(valid);
        if (vmask == 0) continue;   // wave-uniform
        const float lv[3] = {l.x, l.y, l.z}, hv[3] = {h.x, h.y, h.z};
        for (int a = 0; a < 3; ++a) {
            if (!(scale[a] > 0.0f)) continue;   // block-uniform
            const int b = sah_bin(c[a], cmin[a], scale[a]);
            const int b0 = __shfl(b, 0);         // lane 0 is valid when any lane is
            if (__ballot(valid && b != b0) == 0) {
                uint32_t v[6];
                for (int k = 0; k < 3; ++k) {
                    v[k] = f2ord(lv[k]);
                    v[3 + k] = f2ord(hv[k]);
                }
                for (int off = 32; off > 0; off >>= 1)
                    for (int k = 0; k < 6; ++k) {
                        const uint32_t o = (uint32_t)__shfl_xor((int)v[k], off);
                        v[k] = k < 3 ? min(v[k], o) : max(v[k], o);
                    }
                if (lane == 0) {
                    atomicAdd(&s_cnt[a][b0], (uint32_t)__popcll(vmask));
                    for (int k = 0; k < 3; ++k) {
                        atomicMin(&s_lo[a][b0][k], v[k]);
                        atomicMax(&s_hi[a][b0][k], v[3 + k]);
                    }
                }
            } else if (valid) {
                atomicAdd(&s_cnt[a][b], 1u);
                for (int k = 0; k < 3; ++k) {
                    atomicMin(&s_lo[a][b][k], f2ord(lv[k]));
                    atomicMax(&s_hi[a][b][k], f2ord(hv[k]));
                }
            }
        }
    }
    __syncthreads();
    // split: one thread per axis, then the first strict minimum in axis order
    // (sah_choose's result)
    if (tid < 3) {
        SahSplit sp{-1, 0, cnt / 2};
        float cost = INFINITY;
        if (A.level < kSahMedianLevel && scale[tid] > 0.0f) sah_axis(SahLdsBins{s_cnt, s_lo, s_hi}, (int)tid, cost, sp);
        s_axis_best[tid] = sp;
        s_axis_cost[tid] = cost;
    }
    __syncthreads();
    if (tid == 0) {
        SahSplit sp{-1, 0, cnt / 2};
        float cost = INFINITY;
        for (int a = 0; a < 3; ++a)
            if (s_axis_cost[a] < cost) {
                cost = s_axis_cost[a];
                sp = s_axis_best[a];
            }
        s_split[0] = sp.axis;
        s_split[1] = sp.bin;
        s_split[2] = (int)sp.nleft;
    }
    __syncthreads();
    const int axis = s_split[0], sbin = s_split[1];
    const uint32_t nleft = (uint32_t)s_split[2], nright = cnt - nleft;
    // pass 3: stable partition into ids_out, chunk by chunk
    uint32_t lbase = 0, rbase = 0;
    for (uint32_t base = lo; base < hi; base += BLOCK) {
        const uint32_t i = base + tid;
        const bool valid = i < hi;
        uint32_t p = 0;
        bool left_side = false;
        float4 l, h;
        if (valid) {
            p = min(A.ids_in[i], n - 1u);
            float c[3];
            prim_centroid(A.boxes, p, l, h, c);
            left_side = axis >= 0 ? sah_bin(c[axis], cmin[axis], scale[axis]) <= sbin : (i - lo) < nleft;
        }
        const uint64_t m = __ballot(left_side);
        if (lane == 0) s_wave[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t woff = 0, total = 0;
        for (uint32_t w = 0; w < BLOCK / 64; ++w) {
            woff += w < wave ? s_wave[w] : 0u;
            total += s_wave[w];
        }
        const uint32_t lrank = woff + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (valid) {
            const uint32_t rrank = (i - base) - lrank;
            const uint32_t pos = left_side ? lo + lbase + lrank : lo + nleft + rbase + rrank;
            if (pos < hi && (left_side ? lbase + lrank < nleft : rbase + rrank < nright)) {
                A.ids_out[pos] = p;
                if ((left_side ? nleft : nright) == 1u) {
                    // a one-primitive side: leaf n-1+pos at its final position
                    A.vals[pos] = p;
                    A.nbox[2ull * (n - 1u + pos)] = l;
                    A.nbox[2ull * (n - 1u + pos) + 1] = h;
                    A.ncount[n - 1u + pos] = 1u;
                }
            } else {
                atomicAdd(A.err, 1u);
            }
        }
        lbase += total;
        rbase += min((uint32_t)BLOCK, hi - base) - total;
        __syncthreads();   // s_wave is rewritten by the next chunk
    }
    if (tid == 0) {
        if (lbase != nleft || nleft == 0u || nleft >= cnt) atomicAdd(A.err, 1u);
        for (int side = 0; side < 2; ++side) {
            const uint32_t start = side ? lo + nleft : lo, m = side ? nright : nleft;
            uint32_t child;
            if (m == 1u) {
                child = n - 1u + start;
            } else {
                child = atomicAdd(A.node_alloc, 1u);
                const bool big = m >= A.big_min;
                const uint32_t slot = atomicAdd(big ? A.nbig_out : A.ntasks_out, 1u);
                if (child >= n - 1u || slot >= (big ? A.big_cap : A.task_cap) || m == 0u) {
                    atomicAdd(A.err, 1u);
                    child = 0u;
                } else {
                    A.rfirst[child] = start;
                    A.ncount[child] = m;
                    (big ? A.big_out : A.tasks_out)[slot] = make_uint2(start, child);
                }
            }
            if (side) A.right[node] = child;
            else A.left[node] = child;
        }
    }
}

// the root task, in the big list (big) or the small one
__global__ void k_sah_init(uint32_t n, bool big, uint2 *small0, uint2 *big0, uint32_t *ncount, uint32_t *rfirst,
                           uint32_t *counters)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        (big ? big0 : small0)[0] = make_uint2(0u, 0u);
        ncount[0] = n;
        rfirst[0] = 0u;
        counters[6] = 0u;   // small tasks of the next level
        counters[7] = 1u;   // internal nodes allocated (the root)
        counters[8] = 0u;   // big tasks of the next level
    }
}

__global__ void __launch_bounds__(kBlock) k_iota(uint32_t *v, uint32_t n)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) v[i] = i;
}

// ---- collapse to 8-wide quantised nodes ---------------------------------
struct DevAlloc {
    uint32_t *counters;
    __device__ uint32_t operator()(int which, uint32_t amount) const { return atomicAdd(&counters[which], amount); }
};

__global__ void __launch_bounds__(64) k_collapse(CollapseArgs A)
{
    const uint32_t w = A.begin + blockIdx.x * 64 + threadIdx.x;
    if (w >= A.end) return;
    if (collapse_node(A, w, DevAlloc{A.counters})) atomicAdd(&A.counters[2], 1u);
}

__global__ void __launch_bounds__(kBlock) k_write_tris(const uint32_t *leaf_ids, uint32_t n, const GeomDesc *geo,
                                                       uint32_t n_geo, TriRec *out, uint32_t *bad)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t p = leaf_ids[i];
    const GeomDesc &G = geom_of(geo, n_geo, p);
    TriRec r;
    for (int k = 0; k < 3; ++k) {
        const float *q = tri_vertex(G, p, k, bad);
        r.v[3 * k] = q[0];
        r.v[3 * k + 1] = q[1];
        r.v[3 * k + 2] = q[2];
    }
    r.prim = p;
    r.geom = (uint32_t)(&G - geo);
    r.lprim = p - G.tri_begin;
    out[i] = r;
}

__global__ void __launch_bounds__(kBlock) k_write_insts(const uint32_t *leaf_ids, uint32_t n,
                                                        const solhip_instance *inst, const BlasRef *refs,
                                                        InstRec *out, uint32_t *meta)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t p = leaf_ids[i];
    InstRec r;
    inverse3x4(inst[p].transform, r.inv);
    r.nodes = refs[p].nodes;
    r.tris = refs[p].tris;
    out[i] = r;
    meta[i] = (p & 0xFFFFFFu) | (inst[p].id_mask & 0xFF000000u);
}

// structural check of the collapsed tree before any traversal touches it
__global__ void __launch_bounds__(kBlock) k_validate(const WideNode *wn, uint32_t nodes, uint32_t prims, uint32_t *bad)
{
    const uint32_t w = blockIdx.x * kBlock + threadIdx.x;
    if (w >= nodes) return;
    if (!node_valid(wn[w], w, nodes, prims)) atomicAdd(bad, 1u);
}

__global__ void k_root_box(const float4 *nbox, uint32_t root, float *out)
{
    if (threadIdx.x == 0) {
        const float4 l = nbox[2ull * root], h = nbox[2ull * root + 1];
        out[0] = l.x;
        out[1] = l.y;
        out[2] = l.z;
        out[3] = h.x;
        out[4] = h.y;
        out[5] = h.z;
    }
}

__global__ void k_fill_u32(uint32_t *p, uint32_t n, uint32_t v)
{
    if (threadIdx.x < n) p[threadIdx.x] = v;
}

__global__ void __launch_bounds__(kBlock) k_fill_u32_n(uint32_t *p, uint32_t n, uint32_t v)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

int exclusive_scan(hipStream_t s, uint32_t *data, uint32_t n, DevScratch &scratch, int level)
{
    const uint32_t per_block = kBlock * kScanItems;
    const uint32_t nb = cdiv(n, per_block);
    if (nb <= 1) {
        hipLaunchKernelGGL(k_scan_block, dim3(1), dim3(kBlock), 0, s, data, n, (uint32_t *)nullptr);
        return 0;
    }
    uint32_t *sums = scratch.scan_level(level, nb);
    if (!sums) return SOLHIP_ERROR_OUT_OF_MEMORY;
    hipLaunchKernelGGL(k_scan_block, dim3(nb), dim3(kBlock), 0, s, data, n, sums);
    int rc = exclusive_scan(s, sums, nb, scratch, level + 1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kBlock), 0, s, data, n, (const uint32_t *)sums);
    return 0;
}

}  // namespace

uint32_t *DevScratch::scan_level(int level, uint32_t count)
{
    if (level >= (int)scan_bufs.size()) scan_bufs.resize(level + 1);
    DevBuf &b = scan_bufs[level];
    if (b.bytes < (uint64_t)count * 4) {
        if (b.reserve((uint64_t)count * 4) != 0) return nullptr;
    }
    return (uint32_t *)b.ptr;
}

int DevBuf::reserve(uint64_t bytes_needed)
{
    if (bytes >= bytes_needed && ptr) return 0;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    uint64_t want = std::max<uint64_t>(bytes_needed, 256);
    if (hipMalloc(&ptr, want) != hipSuccess) {
        ptr = nullptr;
        set_error("hipMalloc failed in BVH build (" + std::to_string(want) + " bytes)");
        return SOLHIP_ERROR_OUT_OF_MEMORY;
    }
    bytes = want;
    return 0;
}

void DevBuf::release()
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

static int build_binary_lbvh(hipStream_t s, uint32_t n, DevScratch &sc, const uint32_t *&vals_out);

// Binary tree by binned SAH (k_sah_level), one or two launches (small and big
// nodes) and one read-back of the next level's task counts per level.  Order
// buffers: vals[0] / vals[1]; leaf order: keys[0]; small tasks: flags /
// frontier (at most n / 2 per level); big tasks: the two halves of hist.
static int build_binary_sah(hipStream_t s, uint32_t n, DevScratch &sc, const uint32_t *&vals_out)
{
    uint32_t *counters = (uint32_t *)sc.counters.ptr;
    uint32_t *pinned = sc.pinned;
    uint32_t *ids[2] = {(uint32_t *)sc.vals[0].ptr, (uint32_t *)sc.vals[1].ptr};
    uint32_t *vals = (uint32_t *)sc.keys[0].ptr;
    uint2 *tasks[2] = {(uint2 *)sc.flags.ptr, (uint2 *)sc.frontier.ptr};
    const uint32_t big_cap = (uint32_t)std::min<uint64_t>(n / kSahBigMin + 1, sc.hist.bytes / 16);
    uint2 *bigs[2] = {(uint2 *)sc.hist.ptr, (uint2 *)sc.hist.ptr + big_cap};
    uint32_t *ncount = (uint32_t *)sc.ncount.ptr, *rfirst = (uint32_t *)sc.rfirst.ptr;
    const uint32_t cap = n / 2;
    if (big_cap < 1) {
        set_error("SAH BVH build: no room for the task lists");
        return SOLHIP_ERROR_DEVICE;
    }
    const bool root_big = n >= kSahBigMin;
    hipLaunchKernelGGL(k_iota, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, s, ids[0], n);
    hipLaunchKernelGGL(k_sah_init, dim3(1), dim3(64), 0, s, n, root_big, tasks[0], bigs[0], ncount, rfirst, counters);
    uint32_t nsmall = root_big ? 0u : 1u, nbig = root_big ? 1u : 0u, level = 0;
    int cur = 0;
    while (nsmall || nbig) {
        SahArgs a;
        a.boxes = (const float4 *)sc.boxes.ptr;
        a.ids_in = ids[cur];
        a.ids_out = ids[cur ^ 1];
        a.vals = vals;
        a.tasks_out = tasks[cur ^ 1];
        a.big_out = bigs[cur ^ 1];
        a.ntasks_out = counters + 6;
        a.nbig_out = counters + 8;
        a.node_alloc = counters + 7;
        a.left = (uint32_t *)sc.left.ptr;
        a.right = (uint32_t *)sc.right.ptr;
        a.rfirst = rfirst;
        a.ncount = ncount;
        a.nbox = (float4 *)sc.nbox.ptr;
        a.n = n;
        a.task_cap = cap;
        a.big_cap = big_cap;
        a.big_min = kSahBigMin;
        a.level = level;
        a.err = counters + 3;
        if (nbig) {
            a.tasks_in = bigs[cur];
            hipLaunchKernelGGL(k_sah_level<1024>, dim3(nbig), dim3(1024), 0, s, a);
        }
        if (nsmall) {
            a.tasks_in = tasks[cur];
            hipLaunchKernelGGL(k_sah_level<kBlock>, dim3(nsmall), dim3(kBlock), 0, s, a);
        }
        SOLHIP_HIP_CHECK(hipMemcpyAsync(pinned, counters, 36, hipMemcpyDeviceToHost, s));
        SOLHIP_HIP_CHECK(hipStreamSynchronize(s));
        if (pinned[3]) {
            set_error("SAH BVH build range check failed at level " + std::to_string(level));
            return SOLHIP_ERROR_DEVICE;
        }
        nsmall = pinned[6];
        nbig = pinned[8];
        ++level;
        if (nsmall > cap || nbig > big_cap || level > kSahMedianLevel + 40) {
            set_error("SAH BVH build did not converge");
            return SOLHIP_ERROR_DEVICE;
        }
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, counters + 6, 1u, 0u);
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, counters + 8, 1u, 0u);
        cur ^= 1;
    }
    if (pinned[7] != n - 1) {
        set_error("SAH BVH build made " + std::to_string(pinned[7]) + " internal nodes for " + std::to_string(n) +
                  " primitives");
        return SOLHIP_ERROR_DEVICE;
    }
    vals_out = vals;
    return 0;
}

// Generic build over primitive boxes already in scratch.boxes (2 float4 per
// primitive) with centroid bounds in scratch.bounds: a binary tree (Morton
// order and radix tree, or binned SAH when `sah`), then the 8-wide collapse.
static int build_common(hipStream_t s, uint32_t n, uint32_t leaf_max, bool sah, DevScratch &sc, DeviceBvh &out)
{
    // counters (set by reset_build_counters before the box kernel): [0] wide
    // nodes allocated, [1] leaf records allocated, [2] collapse range errors,
    // [3] sort / refit / SAH range errors, [4] invalid nodes, [5] refit
    // frontier size, [6] SAH tasks of the next level, [7] SAH internal nodes
    // allocated, [8] SAH big tasks, [9] vertex references out of range
    uint32_t *counters = (uint32_t *)sc.counters.ptr;
    uint32_t *pinned = sc.pinned;
    uint32_t *left = (uint32_t *)sc.left.ptr, *right = (uint32_t *)sc.right.ptr;
    uint32_t *rfirst = (uint32_t *)sc.rfirst.ptr;
    float4 *nbox = (float4 *)sc.nbox.ptr;
    uint32_t *ncount = (uint32_t *)sc.ncount.ptr;
    const uint32_t *vals = nullptr;
    if (sah && n > 1) {
        int rc = build_binary_sah(s, n, sc, vals);
        if (rc) return rc;
    } else {
        int rc = build_binary_lbvh(s, n, sc, vals);
        if (rc) return rc;
    }
    // binary root: internal node 0, or leaf 0 (node index n-1 = 0) when n == 1
    const uint32_t root = 0u;
    hipLaunchKernelGGL(k_root_box, dim3(1), dim3(64), 0, s, (const float4 *)nbox, root, (float *)out.root_box.ptr);
    // collapse, level by level
    WideNode *wn = (WideNode *)out.nodes.ptr;
    uint32_t *wsrc = (uint32_t *)sc.wsrc.ptr;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, wsrc, 1u, root);
    uint32_t begin = 0, end = 1, depth = 0;
    while (begin < end) {
        CollapseArgs a;
        a.n = (int)n;
        a.leaf_max = leaf_max;
        a.node_cap = n;
        a.left = left;
        a.right = right;
        a.rfirst = rfirst;
        a.ncount = ncount;
        a.nbox = nbox;
        a.vals = vals;
        a.wnodes = wn;
        a.wsrc = wsrc;
        a.counters = counters;
        a.leaf_ids = (uint32_t *)sc.leaf_ids.ptr;
        a.begin = begin;
        a.end = end;
        hipLaunchKernelGGL(k_collapse, dim3(cdiv(end - begin, 64)), dim3(64), 0, s, a);
        SOLHIP_HIP_CHECK(hipMemcpyAsync(pinned, counters, 40, hipMemcpyDeviceToHost, s));
        SOLHIP_HIP_CHECK(hipStreamSynchronize(s));
        if (pinned[9]) {
            set_error("BLAS geometry holds " + std::to_string(pinned[9]) +
                      " vertex references at or beyond its vertex_count");
            return SOLHIP_ERROR_INVALID_ARGUMENT;
        }
        if (pinned[2] || pinned[3]) {
            set_error("BVH build range check failed (collapse " + std::to_string(pinned[2]) + ", sort/refit " +
                      std::to_string(pinned[3]) + ")");
            return SOLHIP_ERROR_DEVICE;
        }
        begin = end;
        end = pinned[0];
        ++depth;
        if (end > n || depth > 4096) {
            set_error("BVH collapse did not converge");
            return SOLHIP_ERROR_DEVICE;
        }
    }
    out.node_count = end;
    out.depth = depth;
    out.prim_count = n;
    if (pinned[1] != n) {
        set_error("BVH collapse placed " + std::to_string(pinned[1]) + " of " + std::to_string(n) + " primitives");
        return SOLHIP_ERROR_DEVICE;
    }
    hipLaunchKernelGGL(k_validate, dim3(cdiv(end, kBlock)), dim3(kBlock), 0, s, (const WideNode *)wn, end, n,
                       counters + 4);
    SOLHIP_HIP_CHECK(hipMemcpyAsync(pinned + 4, counters + 4, 4, hipMemcpyDeviceToHost, s));
    SOLHIP_HIP_CHECK(hipStreamSynchronize(s));
    if (pinned[4] != 0) {
        set_error("BVH validation failed on " + std::to_string(pinned[4]) + " nodes");
        return SOLHIP_ERROR_DEVICE;
    }
    return 0;
}

// Binary tree over Morton order: codes, LSD radix sort, Karras radix tree,
// two-phase refit.  Leaf order: the sorted values.
static int build_binary_lbvh(hipStream_t s, uint32_t n, DevScratch &sc, const uint32_t *&vals_out)
{
    const uint32_t g = cdiv(n, kBlock);
    uint32_t *counters = (uint32_t *)sc.counters.ptr;
    // Morton codes
    hipLaunchKernelGGL(k_morton, dim3(g), dim3(kBlock), 0, s, (const float4 *)sc.boxes.ptr, n,
                       (const uint32_t *)sc.bounds.ptr, (uint32_t *)sc.keys[0].ptr, (uint32_t *)sc.vals[0].ptr);
    // radix sort, 4 passes of 8 bits (codes have 30 bits)
    const uint32_t ntiles = cdiv(n, kSortTile);
    int cur = 0;
    for (int pass = 0; pass < 4; ++pass) {
        const int shift = pass * 8;
        hipLaunchKernelGGL(k_hist, dim3(ntiles), dim3(kBlock), 0, s, (const uint32_t *)sc.keys[cur].ptr, n, shift,
                           (uint32_t *)sc.hist.ptr, ntiles);
        int rc = exclusive_scan(s, (uint32_t *)sc.hist.ptr, 256u * ntiles, sc, 0);
        if (rc) return rc;
        hipLaunchKernelGGL(k_scatter, dim3(ntiles), dim3(kBlock), 0, s, (const uint32_t *)sc.keys[cur].ptr,
                           (const uint32_t *)sc.vals[cur].ptr, n, shift, (const uint32_t *)sc.hist.ptr, ntiles,
                           (uint32_t *)sc.keys[cur ^ 1].ptr, (uint32_t *)sc.vals[cur ^ 1].ptr, counters + 3);
        cur ^= 1;
    }
    const uint32_t *keys = (const uint32_t *)sc.keys[cur].ptr;
    const uint32_t *vals = (const uint32_t *)sc.vals[cur].ptr;
    uint32_t *left = (uint32_t *)sc.left.ptr, *right = (uint32_t *)sc.right.ptr;
    uint32_t *parent = (uint32_t *)sc.parent.ptr, *rfirst = (uint32_t *)sc.rfirst.ptr;
    uint32_t *flags = (uint32_t *)sc.flags.ptr;
    float4 *nbox = (float4 *)sc.nbox.ptr;
    uint32_t *ncount = (uint32_t *)sc.ncount.ptr;
    if (n > 1) {
        hipLaunchKernelGGL(k_karras, dim3(cdiv(n - 1, kBlock)), dim3(kBlock), 0, s, keys, (int)n, left, right, parent,
                           rfirst, ncount);
        hipLaunchKernelGGL(k_fill_u32_n, dim3(cdiv(n - 1, kBlock)), dim3(kBlock), 0, s, flags, n - 1, 0u);
    }
    uint32_t *frontier = (uint32_t *)sc.frontier.ptr;
    hipLaunchKernelGGL(k_refit_local, dim3(cdiv(n, kRefitSpan)), dim3(kBlock), 0, s, (const float4 *)sc.boxes.ptr,
                       vals, (int)n, left, right, parent, rfirst, ncount, nbox, frontier, counters + 5, counters + 3);
    if (n > 1)
        hipLaunchKernelGGL(k_refit_cross, dim3(std::min<uint32_t>(cdiv(n, kBlock), 2048)), dim3(kBlock), 0, s, (int)n,
                           left, right, parent, flags, nbox, frontier, counters + 5, counters + 3);
    vals_out = vals;
    return 0;
}

int DevScratch::ensure(uint32_t n)
{
    const uint64_t N = std::max<uint32_t>(n, 1);
    const uint64_t ntiles = (N + kSortTile - 1) / kSortTile;
    int rc = 0;
    rc |= boxes.reserve(N * 32);
    rc |= bounds.reserve(64);
    rc |= keys[0].reserve(N * 4);
    rc |= keys[1].reserve(N * 4);
    rc |= vals[0].reserve(N * 4);
    rc |= vals[1].reserve(N * 4);
    rc |= hist.reserve(256 * ntiles * 4);
    rc |= left.reserve(N * 4);
    rc |= right.reserve(N * 4);
    rc |= parent.reserve(2 * N * 4);
    rc |= rfirst.reserve(N * 4);
    rc |= flags.reserve(N * 4);
    rc |= nbox.reserve(2 * N * 32);
    rc |= ncount.reserve(2 * N * 4);
    rc |= counters.reserve(256);
    rc |= wsrc.reserve(N * 4);
    rc |= leaf_ids.reserve(N * 4);
    rc |= frontier.reserve(N * 4);
    if (!pinned) {
        if (hipHostMalloc((void **)&pinned, 256, hipHostMallocDefault) != hipSuccess) {
            pinned = nullptr;
            set_error("hipHostMalloc failed");
            return SOLHIP_ERROR_OUT_OF_MEMORY;
        }
    }
    return rc ? SOLHIP_ERROR_OUT_OF_MEMORY : 0;
}

void DevScratch::release()
{
    for (DevBuf *b : {&boxes, &bounds, &keys[0], &keys[1], &vals[0], &vals[1], &hist, &left, &right, &parent, &rfirst,
                      &flags, &nbox, &ncount, &counters, &wsrc, &leaf_ids, &frontier})
        b->release();
    for (auto &b : scan_bufs) b.release();
    scan_bufs.clear();
    if (pinned) (void)hipHostFree(pinned);
    pinned = nullptr;
}

static const uint32_t kOrdInit[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};

// build counters before the first kernel of a build: [0] wide nodes
// allocated (the root), [1]-[5] 0, [6]-[8] SAH task counts (0), [9] vertex
// references out of range (0); see build_common
static int reset_build_counters(hipStream_t s, DevScratch &sc)
{
    uint32_t *pinned = sc.pinned;
    pinned[16] = 1u;
    for (int k = 17; k < 26; ++k) pinned[k] = 0u;
    SOLHIP_HIP_CHECK(hipMemcpyAsync(sc.counters.ptr, pinned + 16, 40, hipMemcpyHostToDevice, s));
    return 0;
}

int build_blas(hipStream_t s, const GeomDesc *geo_dev, uint32_t n_geo, uint32_t n, uint32_t flags, DevScratch &sc,
               DeviceBvh &out)
{
    int rc = sc.ensure(n);
    if (rc) return rc;
    const uint64_t nn = std::max<uint32_t>(n, 1);
    rc = out.nodes.reserve(nn * sizeof(WideNode));
    rc |= out.leaves.reserve(nn * sizeof(TriRec));
    rc |= out.root_box.reserve(64);
    if (rc) return SOLHIP_ERROR_OUT_OF_MEMORY;
    out.kind = 0;
    out.leaf_max = kLeafMaxTri;
    if (n == 0) {
        out.node_count = 0;
        out.prim_count = 0;
        out.depth = 0;
        return 0;
    }
    rc = reset_build_counters(s, sc);
    if (rc) return rc;
    uint32_t *bad = (uint32_t *)sc.counters.ptr + 9;
    SOLHIP_HIP_CHECK(hipMemcpyAsync(sc.bounds.ptr, kOrdInit, 24, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_tri_boxes, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, s, geo_dev, n_geo, n,
                       (float4 *)sc.boxes.ptr, (uint32_t *)sc.bounds.ptr, bad);
    rc = build_common(s, n, kLeafMaxTri, (flags & SOLHIP_BUILD_PREFER_FAST_TRACE) != 0, sc, out);
    if (rc) return rc;
    hipLaunchKernelGGL(k_write_tris, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, s, (const uint32_t *)sc.leaf_ids.ptr, n,
                       geo_dev, n_geo, (TriRec *)out.leaves.ptr, bad);
    SOLHIP_HIP_CHECK(hipGetLastError());
    return 0;
}

int build_tlas(hipStream_t s, const solhip_instance *inst_dev, const BlasRef *refs_dev, uint32_t n, DevScratch &sc,
               DeviceBvh &out)
{
    int rc = sc.ensure(n);
    if (rc) return rc;
    const uint64_t nn = std::max<uint32_t>(n, 1);
    rc = out.nodes.reserve(nn * sizeof(WideNode));
    rc |= out.leaves.reserve(nn * sizeof(InstRec));
    rc |= out.meta.reserve(nn * 4);
    rc |= out.root_box.reserve(64);
    if (rc) return SOLHIP_ERROR_OUT_OF_MEMORY;
    out.kind = 1;
    out.leaf_max = kLeafMaxInst;
    if (n == 0) {
        out.node_count = 0;
        out.prim_count = 0;
        out.depth = 0;
        return 0;
    }
    rc = reset_build_counters(s, sc);
    if (rc) return rc;
    SOLHIP_HIP_CHECK(hipMemcpyAsync(sc.bounds.ptr, kOrdInit, 24, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_inst_boxes, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, s, inst_dev, refs_dev, n,
                       (float4 *)sc.boxes.ptr, (uint32_t *)sc.bounds.ptr);
    rc = build_common(s, n, kLeafMaxInst, false, sc, out);
    if (rc) return rc;
    hipLaunchKernelGGL(k_write_insts, dim3(cdiv(n, kBlock)), dim3(kBlock), 0, s, (const uint32_t *)sc.leaf_ids.ptr, n,
                       inst_dev, refs_dev, (InstRec *)out.leaves.ptr, (uint32_t *)out.meta.ptr);
    SOLHIP_HIP_CHECK(hipGetLastError());
    return 0;
}

}  // namespace solhip
