// traverse.h -- two-level closest-hit / first-hit traversal of the 8-wide
// quantised BVH with a watertight fp32 ray/triangle test.
//
// Replaces the driver traversal behind traceRayEXT(topLevelAS,
// gl_RayFlagsOpaqueEXT, 0xFF, ...) (assets/glsl/ao.rgen:59-70,
// assets/glsl/pathtrace.rgen:65-76): closest opaque hit over TLAS -> BLAS,
// no face culling (TRIANGLE_FACING_CULL_DISABLE, src/ray/acceleration.rs:337),
// reporting (gl_InstanceID, gl_PrimitiveID, barycentrics, t).
//
// Closest hit is the lexicographic minimum of (t, instance, primitive), so
// the answer does not depend on traversal order; boxes are culled only when
// provably farther than the current best (outward-rounded, padded child
// boxes and a relative allowance on the t interval), so the answer equals a
// brute-force loop over all triangles with the same triangle test.
//
// Stack: node groups (child_base, remaining hit mask in visit order, imask)
// of 8 bytes in LDS, one column per lane.  Each tree level pushes at most one
// group, so the depth a launch needs is at most the TLAS depth plus the
// deepest BLAS depth (SceneView::stack_depth, known at build time); a push
// beyond it is counted as overflow and never happens on a consistent tree.
#pragma once

#include "common.h"
#include "devmath.h"

namespace solhip {

// largest stack (entries per lane) a launch may ask for
constexpr int kMaxStack = 64;

struct Hit {
    float t, u, v;
    uint32_t inst, prim;
};

struct TravCounts {
    uint32_t nodes, tris, insts;
    bool overflow;
};

// 1/d with zero components replaced by +-1e-18 (finite: no 0*inf in the
// box test; a near-zero component gives huge but finite slab distances)
SOLHIP_HD V3 safe_inv(V3 d)
{
    float x = fabsf(d.x) < 1e-18f ? (d.x < 0.0f ? -1e-18f : 1e-18f) : d.x;
    float y = fabsf(d.y) < 1e-18f ? (d.y < 0.0f ? -1e-18f : 1e-18f) : d.y;
    float z = fabsf(d.z) < 1e-18f ? (d.z < 0.0f ? -1e-18f : 1e-18f) : d.z;
    return v3(1.0f / x, 1.0f / y, 1.0f / z);
}
SOLHIP_HD uint32_t octant(V3 d)
{
    return (d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u);
}

SOLHIP_HD float sel3(float a, float b, float c, int k) { return k == 0 ? a : (k == 1 ? b : c); }

// per-ray constants of the watertight test (Woop, Benthin, Wald, JCGT 2013)
struct Wt {
    int kx, ky, kz;
    float sx, sy, sz;
    float ox, oy, oz;  // origin permuted to (kx, ky, kz)
};

SOLHIP_HD Wt wt_setup(V3 o, V3 d)
{
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = 0;
    float m = ax;
    if (ay > m) { kz = 1; m = ay; }
    if (az > m) { kz = 2; }
    int kx = kz + 1;
    if (kx == 3) kx = 0;
    int ky = kx + 1;
    if (ky == 3) ky = 0;
    float dkz = sel3(d.x, d.y, d.z, kz);
    if (dkz < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    Wt w;
    w.kx = kx;
    w.ky = ky;
    w.kz = kz;
    float dkx = sel3(d.x, d.y, d.z, kx), dky = sel3(d.x, d.y, d.z, ky);
    w.sx = dkx / dkz;
    w.sy = dky / dkz;
    w.sz = 1.0f / dkz;
    w.ox = sel3(o.x, o.y, o.z, kx);
    w.oy = sel3(o.x, o.y, o.z, ky);
    w.oz = sel3(o.x, o.y, o.z, kz);
    return w;
}

// identical operation sequence to the host restatement of the test
SOLHIP_HD bool wt_hit(const Wt &w, float tmin, float tmax, const float *v, float &t_out,
                                       float &u_out, float &v_out)
{
    float A_x = sel3(v[0], v[1], v[2], w.kx) - w.ox, A_y = sel3(v[0], v[1], v[2], w.ky) - w.oy,
          A_z = sel3(v[0], v[1], v[2], w.kz) - w.oz;
    float B_x = sel3(v[3], v[4], v[5], w.kx) - w.ox, B_y = sel3(v[3], v[4], v[5], w.ky) - w.oy,
          B_z = sel3(v[3], v[4], v[5], w.kz) - w.oz;
    float C_x = sel3(v[6], v[7], v[8], w.kx) - w.ox, C_y = sel3(v[6], v[7], v[8], w.ky) - w.oy,
          C_z = sel3(v[6], v[7], v[8], w.kz) - w.oz;
    float ax = A_x - w.sx * A_z, ay = A_y - w.sy * A_z;
    float bx = B_x - w.sx * B_z, by = B_y - w.sy * B_z;
    float cx = C_x - w.sx * C_z, cy = C_y - w.sy * C_z;
    float U = cx * by - cy * bx;
    float V = ax * cy - ay * cx;
    float W = bx * ay - by * ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        double cxby = (double)cx * (double)by, cybx = (double)cy * (double)bx;
        U = (float)(cxby - cybx);
        double axcy = (double)ax * (double)cy, aycx = (double)ay * (double)cx;
        V = (float)(axcy - aycx);
        double bxay = (double)bx * (double)ay, byax = (double)by * (double)ax;
        W = (float)(bxay - byax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float az = w.sz * A_z, bz = w.sz * B_z, cz = w.sz * C_z;
    float T = U * az + V * bz + W * cz;
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t >= tmin && t <= tmax)) return false;
    t_out = t;
    u_out = V * rcp;
    v_out = W * rcp;
    return true;
}

SOLHIP_HD bool better(float t, uint32_t inst, uint32_t prim, const Hit &h)
{
    if (t < h.t) return true;
    if (t > h.t) return false;
    if (inst != h.inst) return inst < h.inst;
    return prim < h.prim;
}

// world-to-object transform (row-major 3x4), same order as the host restatement
SOLHIP_HD V3 xform_point(const float *m, V3 p)
{
    return v3(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
              m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
}
SOLHIP_HD V3 xform_dir(const float *m, V3 p)
{
    return v3(m[0] * p.x + m[1] * p.y + m[2] * p.z, m[4] * p.x + m[5] * p.y + m[6] * p.z,
              m[8] * p.x + m[9] * p.y + m[10] * p.z);
}

SOLHIP_HD float byte_f(uint32_t w, int b) { return (float)((w >> (8 * b)) & 0xFFu); }

// loads through the global address space (the pointers come from memory, so
// the compiler cannot infer it and would otherwise emit flat loads).  The
// pointer goes through an integer: a direct cast from the generic pointer was
// folded back into a flat load for the first element of a record, whose
// s_waitcnt then also waited for the LDS stack operations.
#if defined(__HIP_DEVICE_COMPILE__)
typedef const __attribute__((address_space(1))) uint4 *gptr4;
typedef const __attribute__((address_space(1))) uint2 *gptr2;
typedef const __attribute__((address_space(1))) uint32_t *gptr1;
#define SOLHIP_G4(p) ((gptr4)(uintptr_t)(p))
#define SOLHIP_G2(p) ((gptr2)(uintptr_t)(p))
#define SOLHIP_G1(p) ((gptr1)(uintptr_t)(p))
#else
#define SOLHIP_G4(p) ((const uint4 *)(p))
#define SOLHIP_G2(p) ((const uint2 *)(p))
#define SOLHIP_G1(p) ((const uint32_t *)(p))
#endif

// reverse the 8-bit slot-order mask into visit order: bit i <-> slot i ^ oct
SOLHIP_HD uint32_t to_visit_order(uint32_t m, uint32_t oct)
{
    if (oct & 1u) m = ((m & 0x55u) << 1) | ((m & 0xAAu) >> 1);
    if (oct & 2u) m = ((m & 0x33u) << 2) | ((m & 0xCCu) >> 2);
    if (oct & 4u) m = ((m & 0x0Fu) << 4) | ((m & 0xF0u) >> 4);
    return m;
}

// count bits set from bit offset up: ((1 << count) - 1) << offset (count <=
// 31).  BFM: on the device one v_bfm_b32 (the compiler turns the shifts into
// a longer sequence).  The step traversal gains from it; trace_ray, whose
// loop the compiler schedules around it less well, does not.
template <bool BFM>
SOLHIP_HD uint32_t bit_range(uint32_t count, uint32_t offset)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if (BFM) {
        uint32_t m;
        asm("v_bfm_b32 %0, %1, %2" : "=v"(m) : "v"(count), "v"(offset));
        return m;
    }
#endif
    return ((1u << count) - 1u) << offset;
}

// Intersect the 8 child boxes of a node (branch-free).  Returns
// internal-child hits in visit order (bit i <-> slot i ^ oct) and the mask of
// leaf primitives (relative to prim_base) of the leaf children hit.
template <bool BFM = false>
SOLHIP_HD void intersect_node(const uint4 w0, const uint4 w1, const uint4 w2, const uint4 w3, const uint4 w4, V3 o,
                              V3 inv, uint32_t oct, float tlo, float thi, uint32_t &hits_int, uint32_t &prim_mask)
{
    const float px = as_f(w0.x), py = as_f(w0.y), pz = as_f(w0.z);
    const uint32_t e = w0.w;
    const float scx = as_f((e & 0xFFu) << 23);
    const float scy = as_f(((e >> 8) & 0xFFu) << 23);
    const float scz = as_f(((e >> 16) & 0xFFu) << 23);
    const uint32_t imask = e >> 24;
    const float sx = scx * inv.x, sy = scy * inv.y, sz = scz * inv.z;
    const float bx = (px - o.x) * inv.x, by = (py - o.y) * inv.y, bz = (pz - o.z) * inv.z;
    // near / far plane bytes by ray octant
    const bool fx = oct & 1u, fy = oct & 2u, fz = oct & 4u;
    const uint32_t nx0 = fx ? w3.z : w2.x, nx1 = fx ? w3.w : w2.y;
    const uint32_t ffx0 = fx ? w2.x : w3.z, ffx1 = fx ? w2.y : w3.w;
    const uint32_t ny0 = fy ? w4.x : w2.z, ny1 = fy ? w4.y : w2.w;
    const uint32_t ffy0 = fy ? w2.z : w4.x, ffy1 = fy ? w2.w : w4.y;
    const uint32_t nz0 = fz ? w4.z : w3.x, nz1 = fz ? w4.w : w3.y;
    const uint32_t ffz0 = fz ? w3.x : w4.z, ffz1 = fz ? w3.y : w4.w;
    uint32_t hm = 0, pm = 0;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int b = s & 3;
        const uint32_t anx = s < 4 ? nx0 : nx1, any_ = s < 4 ? ny0 : ny1, anz = s < 4 ? nz0 : nz1;
        const uint32_t afx = s < 4 ? ffx0 : ffx1, afy = s < 4 ? ffy0 : ffy1, afz = s < 4 ? ffz0 : ffz1;
        const float tnx = fmaf(byte_f(anx, b), sx, bx);
        const float tny = fmaf(byte_f(any_, b), sy, by);
        const float tnz = fmaf(byte_f(anz, b), sz, bz);
        const float tfx = fmaf(byte_f(afx, b), sx, bx);
        const float tfy = fmaf(byte_f(afy, b), sy, by);
        const float tfz = fmaf(byte_f(afz, b), sz, bz);
        const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tlo));
        const float tf = fminf(fminf(tfx, tfy), fminf(tfz, thi));
        const bool hit = tn <= tf;
        // leaf slot: (count << 5) | offset; internal and empty slots: 0
        const uint32_t meta = ((s < 4 ? w1.z : w1.w) >> (8 * b)) & 0xFFu;
        const uint32_t lm = bit_range<BFM>(meta >> 5, meta & 31u);
        hm |= hit ? (1u << s) : 0u;
        pm |= hit ? lm : 0u;
    }
    hits_int = to_visit_order(hm & imask, oct);
    prim_mask = pm;
}

// Two-level traversal.  lds: this lane's LDS stack column base, stride =
// number of lanes sharing the array, sv.stack_depth entries per lane.  Stack
// entries are 8-byte node groups (child_base, remaining hits in visit order |
// imask << 8).  A BLAS is traversed by a nested loop that uses the stack
// above the TLAS entries.
template <bool ANY, bool COUNT>
SOLHIP_HD Hit trace_ray(const SceneView sv, V3 wo, V3 wd, float tmin, float tmax, uint32_t cull, uint2 *lds,
                        int stride, TravCounts &cnt)
{
    Hit best;
    best.t = tmax;
    best.u = 0.0f;
    best.v = 0.0f;
    best.inst = SOLHIP_MISS;
    best.prim = SOLHIP_MISS;
    if (sv.tnodes != nullptr) {
        const int depth = (int)sv.stack_depth;
        int sp = 0;
        auto push = [&](uint32_t x, uint32_t y) {
            if (sp < depth) {
                lds[sp * stride] = make_uint2(x, y);
                ++sp;
            } else {
                cnt.overflow = true;
            }
        };
        auto pop = [&]() -> uint2 {
            --sp;
            return lds[sp * stride];
        };
        const float tlo = tmin - fabsf(tmin) * 0x1p-20f;

        // one BLAS in object space; true when ANY found a hit
        auto trace_blas = [&](const WideNode *nodes, const TriRec *tris, V3 o, V3 d, uint32_t inst) -> bool {
            const V3 inv = safe_inv(d);
            const uint32_t oct = octant(d);
            const Wt wt = wt_setup(o, d);
            const int base = sp;
            uint32_t ng_base = 0, ng_hits = 1u << oct, ng_imask = 1u;
            uint32_t tg_base = 0, tg_mask = 0;
            while (true) {
                if (tg_mask) {
                    // one triangle per iteration
                    const uint32_t k = __builtin_ctz(tg_mask);
                    tg_mask &= tg_mask - 1u;
                    const uint4 *tp = reinterpret_cast<const uint4 *>(tris + tg_base + k);
                    const uint4 a = SOLHIP_G4(tp)[0], b = SOLHIP_G4(tp)[1];
                    const uint2 c = SOLHIP_G2(tp + 2)[0];
                    const float vv[9] = {as_f(a.x), as_f(a.y), as_f(a.z), as_f(a.w), as_f(b.x),
                                         as_f(b.y), as_f(b.z), as_f(b.w), as_f(c.x)};
                    if (COUNT) cnt.tris++;
                    float t, u, v;
                    if (wt_hit(wt, tmin, best.t, vv, t, u, v) && better(t, inst, c.y, best)) {
                        best.t = t;
                        best.u = u;
                        best.v = v;
                        best.inst = inst;
                        best.prim = c.y;
                        if (ANY) {
                            sp = base;
                            return true;
                        }
                    }
                } else if (ng_hits) {
                    const uint32_t i = __builtin_ctz(ng_hits);
                    ng_hits &= ng_hits - 1u;
                    const uint32_t slot = i ^ oct;
                    const uint32_t child = ng_base + __builtin_popcount(ng_imask & ((1u << slot) - 1u));
                    if (ng_hits) push(ng_base, ng_hits | (ng_imask << 8));
                    const uint4 *np = reinterpret_cast<const uint4 *>(nodes + child);
                    const uint4 w0 = SOLHIP_G4(np)[0], w1 = SOLHIP_G4(np)[1], w2 = SOLHIP_G4(np)[2],
                                w3 = SOLHIP_G4(np)[3], w4 = SOLHIP_G4(np)[4];
                    if (COUNT) cnt.nodes++;
                    const float thi = best.t + fabsf(best.t) * 0x1p-20f;
                    intersect_node(w0, w1, w2, w3, w4, o, inv, oct, tlo, thi, ng_hits, tg_mask);
                    ng_base = w1.x;
                    ng_imask = w0.w >> 24;
                    tg_base = w1.y;
                } else if (sp > base) {
                    const uint2 e = pop();
                    ng_base = e.x;
                    ng_hits = e.y & 0xFFu;
                    ng_imask = (e.y >> 8) & 0xFFu;
                } else {
                    return false;
                }
            }
        };

        const V3 winv = safe_inv(wd);
        const uint32_t woct = octant(wd);
        // the TLAS root, or, with one instance, its leaf (record 0) at once
        uint32_t ng_base = 0, ng_hits = sv.single ? 0u : 1u << woct, ng_imask = 1u;
        uint32_t tg_base = 0, tg_mask = sv.single ? 1u : 0u;
        while (true) {
            if (tg_mask) {
                // instance leaf: transform the ray and trace its BLAS
                const uint32_t k = __builtin_ctz(tg_mask);
                tg_mask &= tg_mask - 1u;
                const uint32_t li = tg_base + k;
                const uint32_t meta = SOLHIP_G1(sv.inst_meta + li)[0];
                if (((meta >> 24) & cull) == 0u) continue;
                const uint4 *ip = reinterpret_cast<const uint4 *>(sv.insts + li);
                const uint4 i0 = SOLHIP_G4(ip)[0], i1 = SOLHIP_G4(ip)[1], i2 = SOLHIP_G4(ip)[2],
                            i3 = SOLHIP_G4(ip)[3];
                if (COUNT) cnt.insts++;
                const uint64_t bn = ((uint64_t)i3.y << 32) | i3.x;
                if (bn == 0) continue;
                const float m[12] = {as_f(i0.x), as_f(i0.y), as_f(i0.z), as_f(i0.w), as_f(i1.x), as_f(i1.y),
                                     as_f(i1.z), as_f(i1.w), as_f(i2.x), as_f(i2.y), as_f(i2.z), as_f(i2.w)};
                const V3 o = xform_point(m, wo);
                const V3 d = xform_dir(m, wd);
                const uint64_t tr = ((uint64_t)i3.w << 32) | i3.z;
                if (trace_blas(reinterpret_cast<const WideNode *>(bn), reinterpret_cast<const TriRec *>(tr), o, d,
                               meta & 0xFFFFFFu))
                    break;
            } else if (ng_hits) {
                const uint32_t i = __builtin_ctz(ng_hits);
                ng_hits &= ng_hits - 1u;
                const uint32_t slot = i ^ woct;
                const uint32_t child = ng_base + __builtin_popcount(ng_imask & ((1u << slot) - 1u));
                if (ng_hits) push(ng_base, ng_hits | (ng_imask << 8));
                const uint4 *np = reinterpret_cast<const uint4 *>(sv.tnodes + child);
                const uint4 w0 = SOLHIP_G4(np)[0], w1 = SOLHIP_G4(np)[1], w2 = SOLHIP_G4(np)[2],
                            w3 = SOLHIP_G4(np)[3], w4 = SOLHIP_G4(np)[4];
                if (COUNT) cnt.nodes++;
                const float thi = best.t + fabsf(best.t) * 0x1p-20f;
                intersect_node(w0, w1, w2, w3, w4, wo, winv, woct, tlo, thi, ng_hits, tg_mask);
                ng_base = w1.x;
                ng_imask = w0.w >> 24;
                tg_base = w1.y;
            } else if (sp > 0) {
                const uint2 e = pop();
                ng_base = e.x;
                ng_hits = e.y & 0xFFu;
                ng_imask = (e.y >> 8) & 0xFFu;
            } else {
                break;
            }
        }
    }
    if (best.inst == SOLHIP_MISS) {
        best.t = -1.0f;
        best.u = 0.0f;
        best.v = 0.0f;
        best.prim = SOLHIP_MISS;
    }
    return best;
}

}  // namespace solhip
