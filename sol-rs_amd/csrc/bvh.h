// bvh.h -- host-side objects of the BVH builder (device buffers, scratch).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "bvh_logic.h"
#include "common.h"

namespace solhip {

// one device allocation that only grows
struct DevBuf {
    void *ptr = nullptr;
    uint64_t bytes = 0;
    int reserve(uint64_t bytes_needed);
    void release();
};

// per-instance reference to a BLAS, read by the TLAS build kernels and by
// the pass that splits BLAS-wide triangle indices of hits into (geometry,
// triangle)
struct BlasRef {
    const WideNode *nodes;   // nullptr: empty BLAS
    const TriRec *tris;
    const float *root_box;   // 6 floats (lo xyz, hi xyz) in device memory
    const uint32_t *geo;     // [n_geo, first triangle of geometry 0 .. n_geo-1], or nullptr (one geometry)
};


static_assert(sizeof(BlasRef) == 32, "BlasRef layout");

// build scratch, kept by each BLAS / TLAS so that rebuilds allocate nothing
struct DevScratch {
    DevBuf boxes, bounds, keys[2], vals[2], hist, left, right, parent, rfirst, flags, nbox, ncount, counters,
        wsrc, leaf_ids, frontier;
    std::vector<DevBuf> scan_bufs;
    uint32_t *pinned = nullptr;
    int ensure(uint32_t n);
    uint32_t *scan_level(int level, uint32_t count);
    void release();
};

struct DeviceBvh {
    int kind = 0;          // 0 = BLAS (triangles), 1 = TLAS (instances)
    DevBuf nodes;          // WideNode[node_count]
    DevBuf leaves;         // TriRec[] or InstRec[]
    DevBuf meta;           // TLAS: uint32 inst_meta[]
    DevBuf root_box;       // 6 floats
    uint32_t node_count = 0, prim_count = 0, depth = 0, leaf_max = 0;
    void release()
    {
        nodes.release();
        leaves.release();
        meta.release();
        root_box.release();
    }
};

// flags: SOLHIP_BUILD_* (SOLHIP_BUILD_PREFER_FAST_TRACE: binned SAH, else Morton order).
// geo_dev: the n_geo descriptors (and, after them, the u32 table
// [n_geo, tri_begin...] that BlasRef::geo points to) in device memory; n
// triangles in all.  An index >= vertex_count fails the build with
// SOLHIP_ERROR_INVALID_ARGUMENT.
int build_blas(hipStream_t s, const GeomDesc *geo_dev, uint32_t n_geo, uint32_t n, uint32_t flags, DevScratch &sc,
               DeviceBvh &out);
int build_tlas(hipStream_t s, const solhip_instance *inst_dev, const BlasRef *refs_dev, uint32_t n, DevScratch &sc,
               DeviceBvh &out);

}  // namespace solhip
