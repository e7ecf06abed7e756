// fetch.h -- work distribution of the persistent traversal kernels, and the
// block-wide queue append of the kernels that fill their queues.
//
// The items [0, total) of one launch are split into P contiguous partitions,
// each with its own fetch head; heads sit kHeadStride words (128 B) apart so
// that no two share a cache line.  A wave starts on partition wave % P, claims
// items from its head with one atomic per refill, and moves on to the next
// partition when its own is used up, until it has seen all P.  Refills then
// spread their atomics over P addresses instead of one, which is what lets a
// wave refill its idle lanes often.  Every item is claimed exactly once (a
// partition is left only after an atomic showed it used up), and a wave that
// has seen all partitions is done.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace solhip {

constexpr uint32_t kHeadStride = 32;   // words between two fetch heads

struct PartFetch {
    uint32_t *heads;   // P heads, kHeadStride words apart, zero at launch
    uint32_t total, P, part, moved, lo, hi;
    bool exhausted;

    __device__ __forceinline__ void bounds()
    {
        lo = (uint32_t)((uint64_t)total * part / P);
        hi = (uint32_t)((uint64_t)total * (part + 1) / P);
    }

    __device__ __forceinline__ void init(uint32_t *h, uint32_t nparts, uint32_t n_items)
    {
        heads = h;
        total = n_items;
        P = nparts;
        part = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) % P;
        moved = 0;
        exhausted = false;
        bounds();
    }

    // Wave-uniform: claims n items of the current partition.  Returns the
    // first; items base .. min(base + n, end) - 1 are the caller's.
    __device__ __forceinline__ uint32_t claim(uint32_t n, uint32_t &end)
    {
        uint32_t base = 0;
        if ((threadIdx.x & 63u) == 0) base = atomicAdd(heads + part * kHeadStride, n);
        base = lo + __shfl(base, 0);
        end = hi;
        if (base + n >= hi) {
            // this partition is used up: the next one, until all were seen
            if (++moved >= P) {
                exhausted = true;
            } else {
                part = part + 1 == P ? 0 : part + 1;
                bounds();
            }
        }
        return base;
    }
};

// Block-wide append of one item per thread that has one: every thread of the
// block calls it with block-uniform control flow; one atomic per block (a
// producer kernel that appended once per wave put every wave of the GPU on
// one counter address, which limited it).  lds: WAVES + 1 words.
template <uint32_t WAVES>
__device__ __forceinline__ uint32_t block_append(uint32_t *counter, bool pred, uint32_t *lds)
{
    const uint64_t m = __ballot(pred);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (uint32_t w = 0; w < WAVES; ++w) {
            const uint32_t c = lds[w];
            lds[w] = total;
            total += c;
        }
        lds[WAVES] = total ? atomicAdd(counter, total) : 0u;
    }
    __syncthreads();
    const uint32_t base = lds[WAVES] + lds[wave];
    __syncthreads();   // lds is reused by the next call
    return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

}  // namespace solhip
