// traverse_step.h -- the two-level traversal of traverse.h as a resumable
// state machine: trav_step does one step of work (at most one triangle, one
// instance entry and one node, with the stack pops between them) and reports
// when the ray is done.
// A lane whose ray has finished can then take the next ray while the rest of
// its wave keeps traversing (persistent threads with per-lane refill), which
// is what incoherent rays (AO, bounces) need to keep the SIMD busy.
//
// The visit order is exactly trace_ray's: instance leaves before descending,
// a BLAS traversed to the end before the TLAS continues (the TLAS state is
// pushed as two stack entries, the pending node group and then the pending
// instance-leaf group, tagged in bit 31), so closest and first hits equal
// trace_ray's bit for bit.
#pragma once

#include "traverse.h"

namespace solhip {

struct TravState {
    V3 wo, wd, winv;           // world ray and its safe_inv direction
    V3 o, inv;                 // ray of the current level (object space inside a BLAS)
    Wt wt;                     // triangle test of the current BLAS ray
    float tmin, tlo;
    uint32_t oct;
    uint32_t level;            // 0: TLAS, 1: BLAS
    uint32_t inst;             // instance index of the BLAS being traversed
    const WideNode *nodes;     // BLAS nodes
    const TriRec *tris;        // BLAS leaf records
    uint32_t ng_base, ng_hits, ng_imask, tg_base, tg_mask;
    int sp, base;              // stack height; height at BLAS entry
    Hit best;
};

constexpr uint32_t kLeafGroupTag = 0x80000000u;

SOLHIP_HD void trav_init(TravState &s, const SceneView &sv, V3 wo, V3 wd, float tmin, float tmax)
{
    s.wo = wo;
    s.wd = wd;
    s.o = wo;
    s.winv = safe_inv(wd);
    s.inv = s.winv;
    s.oct = octant(wd);
    s.tmin = tmin;
    s.tlo = tmin - fabsf(tmin) * 0x1p-20f;
    s.level = 0;
    s.inst = 0;
    s.nodes = nullptr;
    s.tris = nullptr;
    // the TLAS root, or, with one instance, its leaf (record 0) at once
    s.ng_base = 0;
    s.ng_hits = sv.tnodes && !sv.single ? (1u << s.oct) : 0u;
    s.ng_imask = 1u;
    s.tg_base = 0;
    s.tg_mask = sv.tnodes && sv.single ? 1u : 0u;
    s.sp = 0;
    s.base = 0;
    s.best.t = tmax;
    s.best.u = 0.0f;
    s.best.v = 0.0f;
    s.best.inst = SOLHIP_MISS;
    s.best.prim = SOLHIP_MISS;
}

// the final hit record of a finished traversal (trace_ray's miss convention)
SOLHIP_HD Hit trav_result(const TravState &s)
{
    Hit h = s.best;
    if (h.inst == SOLHIP_MISS) {
        h.t = -1.0f;
        h.u = 0.0f;
        h.v = 0.0f;
        h.prim = SOLHIP_MISS;
    }
    return h;
}

// One traversal step; true when the ray is finished.  A step does, in this
// order and each at most once: one triangle of the current leaf group, the
// stack pops and BLAS exits that lead to the next piece of work, one instance
// entry, one node test.  Running them as phases of one step (rather than one
// of them per step) lets a wave's lanes share the expensive node test in the
// same pass whatever bookkeeping each needed first.  lds: this lane's stack
// column, stride = lanes sharing the array, sv.stack_depth entries.
//
// WINV: a BLAS exit restores the world ray's inverse direction from the
// state (three more registers) instead of recomputing it (three correctly
// rounded divisions).  Measured per kernel: the AO kernels run 6 % faster
// with it; the path tracer's extension kernel, whose rays leave a BLAS 3.9
// times on average, 6 % faster once its register budget holds the three
// registers (at 5 waves per SIMD they spilled and it ran 1 % slower).
template <bool ANY, bool COUNT, bool WINV = true>
SOLHIP_HD bool trav_step(TravState &s, const SceneView &sv, uint32_t cull, uint2 *lds, int stride, TravCounts &cnt)
{
    const int depth = (int)sv.stack_depth;
    // 1. one triangle
    if (s.level && s.tg_mask) {
        const uint32_t k = __builtin_ctz(s.tg_mask);
        s.tg_mask &= s.tg_mask - 1u;
        const uint4 *tp = reinterpret_cast<const uint4 *>(s.tris + s.tg_base + k);
        const uint4 a = SOLHIP_G4(tp)[0], b = SOLHIP_G4(tp)[1];
        const uint2 c = SOLHIP_G2(tp + 2)[0];
        const float vv[9] = {as_f(a.x), as_f(a.y), as_f(a.z), as_f(a.w), as_f(b.x),
                             as_f(b.y), as_f(b.z), as_f(b.w), as_f(c.x)};
        if (COUNT) cnt.tris++;
        float t, u, v;
        if (wt_hit(s.wt, s.tmin, s.best.t, vv, t, u, v) && better(t, s.inst, c.y, s.best)) {
            s.best.t = t;
            s.best.u = u;
            s.best.v = v;
            s.best.inst = s.inst;
            s.best.prim = c.y;
            if (ANY) return true;
        }
    }
    // 2. pops and BLAS exits up to the next node group or leaf group.  An
    // empty stack ends the ray at once, also inside a BLAS: nothing of the
    // TLAS is pending (its work is pushed below the BLAS entry), so the world
    // ray need not be restored.  Some lane of a wave finishes in almost every
    // step, so the restore would otherwise be issued in almost every step.
    while (!s.tg_mask && !s.ng_hits) {
        if (s.sp == 0) return true;
        if (s.level && s.sp == s.base) {
            // the BLAS is done: back to the world ray and the TLAS stack
            s.level = 0;
            s.o = s.wo;
            s.inv = WINV ? s.winv : safe_inv(s.wd);
            s.oct = octant(s.wd);
        } else {
            --s.sp;
            const uint2 e = lds[s.sp * stride];
            if (e.y & kLeafGroupTag) {
                s.tg_base = e.x;
                s.tg_mask = e.y & ~kLeafGroupTag;
            } else {
                s.ng_base = e.x;
                s.ng_hits = e.y & 0xFFu;
                s.ng_imask = (e.y >> 8) & 0xFFu;
            }
        }
    }
    // 3. an instance leaf: enter its BLAS
    if (!s.level && s.tg_mask) {
        const uint32_t k = __builtin_ctz(s.tg_mask);
        s.tg_mask &= s.tg_mask - 1u;
        const uint32_t li = s.tg_base + k;
        const uint32_t meta = SOLHIP_G1(sv.inst_meta + li)[0];
        if (((meta >> 24) & cull) != 0u) {
            const uint4 *ip = reinterpret_cast<const uint4 *>(sv.insts + li);
            const uint4 i0 = SOLHIP_G4(ip)[0], i1 = SOLHIP_G4(ip)[1], i2 = SOLHIP_G4(ip)[2], i3 = SOLHIP_G4(ip)[3];
            if (COUNT) cnt.insts++;
            const uint64_t bn = ((uint64_t)i3.y << 32) | i3.x;
            if (bn != 0) {
                // the TLAS continues later: pending node group, then pending leaves
                if (s.ng_hits) {
                    if (s.sp < depth) {
                        lds[s.sp * stride] = make_uint2(s.ng_base, s.ng_hits | (s.ng_imask << 8));
                        ++s.sp;
                    } else {
                        cnt.overflow = true;
                    }
                }
                if (s.tg_mask) {
                    if (s.sp < depth) {
                        lds[s.sp * stride] = make_uint2(s.tg_base, s.tg_mask | kLeafGroupTag);
                        ++s.sp;
                    } else {
                        cnt.overflow = true;
                    }
                }
                const float m[12] = {as_f(i0.x), as_f(i0.y), as_f(i0.z), as_f(i0.w), as_f(i1.x), as_f(i1.y),
                                     as_f(i1.z), as_f(i1.w), as_f(i2.x), as_f(i2.y), as_f(i2.z), as_f(i2.w)};
                const V3 o = xform_point(m, s.wo);
                const V3 d = xform_dir(m, s.wd);
                s.o = o;
                s.inv = safe_inv(d);
                s.oct = octant(d);
                s.wt = wt_setup(o, d);
                s.nodes = reinterpret_cast<const WideNode *>(bn);
                s.tris = reinterpret_cast<const TriRec *>(((uint64_t)i3.w << 32) | i3.z);
                s.inst = meta & 0xFFFFFFu;
                s.level = 1;
                s.base = s.sp;
                s.ng_base = 0;
                s.ng_hits = 1u << s.oct;
                s.ng_imask = 1u;
                s.tg_base = 0;
                s.tg_mask = 0;
            }
        }
    }
    // 4. one node
    if (!s.tg_mask && s.ng_hits) {
        const uint32_t i = __builtin_ctz(s.ng_hits);
        s.ng_hits &= s.ng_hits - 1u;
        const uint32_t slot = i ^ s.oct;
        const uint32_t child = s.ng_base + __builtin_popcount(s.ng_imask & ((1u << slot) - 1u));
        if (s.ng_hits) {
            if (s.sp < depth) {
                lds[s.sp * stride] = make_uint2(s.ng_base, s.ng_hits | (s.ng_imask << 8));
                ++s.sp;
            } else {
                cnt.overflow = true;
            }
        }
        const WideNode *nodes = s.level ? s.nodes : sv.tnodes;
        const uint4 *np = reinterpret_cast<const uint4 *>(nodes + child);
        const uint4 w0 = SOLHIP_G4(np)[0], w1 = SOLHIP_G4(np)[1], w2 = SOLHIP_G4(np)[2], w3 = SOLHIP_G4(np)[3],
                    w4 = SOLHIP_G4(np)[4];
        if (COUNT) cnt.nodes++;
        const float thi = s.best.t + fabsf(s.best.t) * 0x1p-20f;
        intersect_node<true>(w0, w1, w2, w3, w4, s.o, s.inv, s.oct, s.tlo, thi, s.ng_hits, s.tg_mask);
        s.ng_base = w1.x;
        s.ng_imask = w0.w >> 24;
        s.tg_base = w1.y;
    }
    return false;
}

}  // namespace solhip
