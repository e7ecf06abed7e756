// fetch.h -- work distribution of the persistent traversal kernels.
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

}  // namespace solhip
