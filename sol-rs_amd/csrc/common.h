// common.h -- internal layouts and helpers of libsolhip (gfx950 only).
//
// HBM layout of the acceleration structures
// -----------------------------------------
// A BVH (BLAS over triangles, TLAS over instances) is an array of 80-byte
// 8-wide nodes plus an array of leaf records stored in node order:
//
//   WideNode (80 B, 16-B aligned, fetched as 5 x dwordx4)
//     float    px, py, pz        quantisation origin (min corner of the node)
//     uint8    ex, ey, ez        per-axis scale 2^(e-127) (fp32 exponent bits)
//     uint8    imask             bit s: slot s holds an internal child
//     uint32   child_base        node index of the first internal child; the
//                                child in slot s is child_base + popc(imask & (2^s-1))
//     uint32   prim_base         leaf-record index of the node's first leaf primitive
//     uint8    meta[8]           leaf slot s: (count << 5) | offset from prim_base
//     uint8    qlo_x[8], qlo_y[8], qlo_z[8], qhi_x[8], qhi_y[8], qhi_z[8]
//                                child boxes, rounded outward and padded by
//                                one quantisation step
//   Slot order: slot s holds the child nearest along direction w_s with
//   w_s[a] = (bit a of s ? -1 : +1); a ray with octant bits oct (bit a set
//   when d[a] < 0) visits slots in the order s = i ^ oct, i = 0..7, which is
//   near-to-far without sorting.
//
//   BLAS leaf record: TriRec (48 B) = v0.xyz v1.xyz v2.xyz, BLAS-wide triangle
//                     index (geometries numbered one after another; the
//                     tie rule and the hit shaders use it), geometry index,
//                     triangle index within the geometry (40 B read per test).
//   TLAS leaf record: InstRec (64 B) = world-to-object 3x4 (48 B),
//                     BLAS node pointer, BLAS leaf pointer; plus a parallel
//                     uint32 inst_meta = instance index | mask << 24.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/solhip.h"

namespace solhip {

constexpr int kWideNodeBytes = 80;
constexpr uint32_t kLeafMaxTri = 3;
constexpr uint32_t kLeafMaxInst = 1;

struct alignas(16) WideNode {
    float px, py, pz;
    uint8_t ex, ey, ez, imask;
    uint32_t child_base;
    uint32_t prim_base;
    uint8_t meta[8];
    uint8_t qlo_x[8], qlo_y[8], qlo_z[8];
    uint8_t qhi_x[8], qhi_y[8], qhi_z[8];
};
static_assert(sizeof(WideNode) == kWideNodeBytes, "WideNode must be 80 bytes");

struct alignas(16) TriRec {
    float v[9];
    uint32_t prim;    // BLAS-wide triangle index
    uint32_t geom;    // geometry index (gl_GeometryIndexEXT)
    uint32_t lprim;   // triangle index within the geometry (gl_PrimitiveID)
};
static_assert(sizeof(TriRec) == 48, "TriRec must be 48 bytes");

struct alignas(16) InstRec {
    float inv[12];
    const WideNode *nodes;
    const TriRec *tris;
};
static_assert(sizeof(InstRec) == 64, "InstRec must be 64 bytes");

// device view of a two-level scene
struct SceneView {
    const WideNode *tnodes;     // TLAS nodes (nullptr: empty TLAS)
    const InstRec *insts;       // TLAS leaf records
    const uint32_t *inst_meta;  // index | mask << 24
    uint32_t stack_depth;       // LDS stack entries per lane (>= TLAS + BLAS depth)
    uint32_t single;            // 1: the TLAS holds one instance; rays enter it without the root test
};

struct Counters {
    unsigned long long rays, nodes, tris, insts, overflow;
};

}  // namespace solhip

#define SOLHIP_HIP_CHECK(expr)                                                                   \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess) {                                                                  \
            ::solhip::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));              \
            return _e == hipErrorOutOfMemory ? SOLHIP_ERROR_OUT_OF_MEMORY : SOLHIP_ERROR_DEVICE; \
        }                                                                                        \
    } while (0)

namespace solhip {
void set_error(const std::string &msg);
}
