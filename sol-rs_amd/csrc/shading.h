// shading.h -- pieces shared by the program kernels: pixel tiling, the
// hit-shader surface fetch, per-wave ray counting, the miss shader.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.h"
#include "traverse.h"

namespace solhip {

#ifndef SOLHIP_PROG_BLOCK
#define SOLHIP_PROG_BLOCK 64
#endif
constexpr int kProgBlock = SOLHIP_PROG_BLOCK;  // a multiple of 64
constexpr int kTileH = kProgBlock / 8;          // pixel tiles are 8 x kTileH

__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb)
{
    const uint32_t x = b & 7u, q = nb >> 3, r = nb & 7u;
    return x * q + min(x, r) + (b >> 3);
}

// pixel of this lane: 8 x kTileH tiles (one wave = 8x8 pixels), tiles in
// row-major order, XCD-contiguous
__device__ __forceinline__ bool tile_pixel(uint32_t W, uint32_t H, uint32_t &x, uint32_t &y)
{
    const uint32_t tiles_x = (W + 7) / 8;
    const uint32_t t = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t tx = t % tiles_x, ty = t / tiles_x;
    x = tx * 8 + (threadIdx.x & 7);
    y = ty * kTileH + (threadIdx.x >> 3);
    return x < W && y < H;
}

__host__ inline uint32_t tile_count(uint32_t W, uint32_t H) { return ((W + 7) / 8) * ((H + kTileH - 1) / kTileH); }

struct Surf {
    V3 pos, nrm, col;
};

// ao.rchit:56-77 / pathtrace.rchit:59-88: world-space hit point, normal
// (transform_it, normalised) and interpolated vertex colour
__device__ __forceinline__ Surf fetch_surface(const DevTables &tb, uint32_t inst, uint32_t prim, float u, float v)
{
    const solhip_scene_instance *si = tb.instances + inst;
    const uint32_t obj = si->id;
    uint32_t i0, i1, i2;
    if (tb.index_size == 8) {
        const uint64_t *ib = reinterpret_cast<const uint64_t *>(tb.index_table[obj]) + 3ull * prim;
        i0 = (uint32_t)ib[0];
        i1 = (uint32_t)ib[1];
        i2 = (uint32_t)ib[2];
    } else {
        const uint32_t *ib = reinterpret_cast<const uint32_t *>(tb.index_table[obj]) + 3ull * prim;
        i0 = ib[0];
        i1 = ib[1];
        i2 = ib[2];
    }
    const float4 *vb = reinterpret_cast<const float4 *>(tb.vertex_table[obj]);
    const float4 p0 = vb[4ull * i0], c0 = vb[4ull * i0 + 1], n0 = vb[4ull * i0 + 2];
    const float4 p1 = vb[4ull * i1], c1 = vb[4ull * i1 + 1], n1 = vb[4ull * i1 + 2];
    const float4 p2 = vb[4ull * i2], c2 = vb[4ull * i2 + 1], n2 = vb[4ull * i2 + 2];
    const float bx = 1.0f - u - v, by = u, bz = v;
    const V3 n = v3(n0.x * bx + n1.x * by + n2.x * bz, n0.y * bx + n1.y * by + n2.y * bz,
                    n0.z * bx + n1.z * by + n2.z * bz);
    const V3 p = v3(p0.x * bx + p1.x * by + p2.x * bz, p0.y * bx + p1.y * by + p2.y * bz,
                    p0.z * bx + p1.z * by + p2.z * bz);
    Surf s;
    s.col = v3(c0.x * bx + c1.x * by + c2.x * bz, c0.y * bx + c1.y * by + c2.y * bz,
               c0.z * bx + c1.z * by + c2.z * bz);
    float nw[4], pw[4];
    mat4_mul(si->transform_it, n.x, n.y, n.z, 0.0f, nw);
    s.nrm = normalize(v3(nw[0], nw[1], nw[2]));
    mat4_mul(si->transform, p.x, p.y, p.z, 1.0f, pw);
    s.pos = v3(pw[0], pw[1], pw[2]);
    return s;
}

// a traversal result as a public hit record: geometry 0 (launch_split_geometry
// splits the BLAS-wide triangle index of several-geometry BLASes afterwards)
__device__ __forceinline__ void store_hit(solhip_hit *dst, const Hit &h)
{
    solhip_hit o;
    o.t = h.t;
    o.u = h.u;
    o.v = h.v;
    o.inst = h.inst;
    o.prim = h.prim;
    o.geom = h.inst == SOLHIP_MISS ? SOLHIP_MISS : 0u;
    *dst = o;
}

// per-wave sums of rays (and fetches when COUNT), one atomic per wave
template <bool COUNT>
__device__ __forceinline__ void count_rays(Counters *c, uint32_t rays_this_lane, const TravCounts &tc)
{
    unsigned long long r = rays_this_lane;
    unsigned long long nd = COUNT ? tc.nodes : 0, tr = COUNT ? tc.tris : 0, in = COUNT ? tc.insts : 0;
    for (int off = 32; off > 0; off >>= 1) {
        r += __shfl_xor(r, off);
        if (COUNT) {
            nd += __shfl_xor(nd, off);
            tr += __shfl_xor(tr, off);
            in += __shfl_xor(in, off);
        }
    }
    const uint64_t ov = __ballot(tc.overflow);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&c->rays, r);
        if (ov) atomicAdd(&c->overflow, (unsigned long long)__popcll(ov));
        if (COUNT) {
            atomicAdd(&c->nodes, nd);
            atomicAdd(&c->tris, tr);
            atomicAdd(&c->insts, in);
        }
    }
}

__device__ __forceinline__ float length3(V3 a) { return sqrtf(dot(a, a)); }

// pathtrace.rmiss:8-21; sun = false drops the sun disc (gradient sky only)
__device__ __forceinline__ V3 pt_miss(uint32_t enable_sky, bool sun, V3 d)
{
    if (enable_sky) {
        const V3 wI = normalize(d);
        const float e = 0.5f * (wI.y + 1.0f);
        const float tt = saturate((e - 0.35f) / (0.65f - 0.35f));
        const float t = tt * tt * (3.0f - 2.0f * tt);
        const V3 sd = normalize(v3(0.0f, 1.0f, -0.25f));
        const float isSun = (sun && dot(wI, sd) > 0.99f) ? 1.0f : 0.0f;
        return v3(mixf(mixf(0.58f, 0.3f, t), 120.0f, isSun), mixf(mixf(0.45f, 0.4f, t), 100.0f, isSun),
                  mixf(mixf(0.25f, 0.5f, t), 50.0f, isSun));
    }
    return v3(0.0f, 0.0f, 0.0f);
}

}  // namespace solhip
