// bvh_cpu.hip -- host replay of the library's BVH build and traversal steps
// (bvh_logic.h, traverse.h) for CPU tests.  Compiled for the host only by
// tests/cpu_harness/build.py; it exercises the same per-item code the GPU
// kernels run, with a sequential sort / refit / allocator in place of the
// device-wide ones.  Test infrastructure: not part of libsolhip.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../sol-rs_amd/csrc/bvh_logic.h"
#include "../../sol-rs_amd/csrc/traverse.h"
#include "../../sol-rs_amd/csrc/traverse_step.h"

using namespace solhip;

namespace {

// a traversal result as the 24-B public hit record (one geometry per BLAS)
static void put_hit(void *hits, size_t i, const Hit &h)
{
    const uint32_t geom = h.inst == SOLHIP_MISS ? SOLHIP_MISS : 0u;
    std::memcpy((char *)hits + 24 * i, &h, 20);
    std::memcpy((char *)hits + 24 * i + 20, &geom, 4);
}

struct HostBvh {
    std::vector<WideNode> nodes;
    std::vector<uint32_t> leaf_ids;
    uint32_t depth = 0;
    float root_box[6];
};

uint32_t g_build_flags = 0;   // SOLHIP_BUILD_* for BLAS builds

struct HostBins {
    uint32_t c[3][kSahBins];
    float l[3][kSahBins][3], h[3][kSahBins][3];
    uint32_t count(int a, int b) const { return c[a][b]; }
    float lo(int a, int b, int k) const { return l[a][b][k]; }
    float hi(int a, int b, int k) const { return h[a][b][k]; }
};

// binned SAH binary tree, level by level as the GPU build (k_sah_level)
void sah_tree(const float *pbox, uint32_t n, std::vector<uint32_t> &vals, std::vector<uint32_t> &left,
              std::vector<uint32_t> &right, std::vector<uint32_t> &rfirst, std::vector<uint32_t> &ncount)
{
    std::vector<uint32_t> ids(n), out(n);
    std::iota(ids.begin(), ids.end(), 0u);
    vals.assign(n, 0u);
    struct Task {
        uint32_t lo, node;
    };
    std::vector<Task> tasks{{0u, 0u}}, next;
    ncount[0] = n;
    rfirst[0] = 0;
    uint32_t alloc = 1;
    auto cen = [&](uint32_t p, int a) { return 0.5f * pbox[6 * p + a] + 0.5f * pbox[6 * p + 3 + a]; };
    for (uint32_t level = 0; !tasks.empty(); ++level) {
        next.clear();
        for (const Task &t : tasks) {
            const uint32_t lo = t.lo, cnt = ncount[t.node], hi = lo + cnt;
            float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (uint32_t i = lo; i < hi; ++i)
                for (int a = 0; a < 3; ++a) {
                    cmin[a] = fminf(cmin[a], cen(ids[i], a));
                    cmax[a] = fmaxf(cmax[a], cen(ids[i], a));
                }
            float scale[3];
            for (int a = 0; a < 3; ++a) scale[a] = sah_scale(cmax[a] - cmin[a]);
            HostBins B;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < kSahBins; ++b) {
                    B.c[a][b] = 0;
                    for (int k = 0; k < 3; ++k) {
                        B.l[a][b][k] = INFINITY;
                        B.h[a][b][k] = -INFINITY;
                    }
                }
            for (uint32_t i = lo; i < hi; ++i) {
                const uint32_t p = ids[i];
                for (int a = 0; a < 3; ++a) {
                    if (!(scale[a] > 0.0f)) continue;
                    const int b = sah_bin(cen(p, a), cmin[a], scale[a]);
                    B.c[a][b] += 1;
                    for (int k = 0; k < 3; ++k) {
                        B.l[a][b][k] = fminf(B.l[a][b][k], pbox[6 * p + k]);
                        B.h[a][b][k] = fmaxf(B.h[a][b][k], pbox[6 * p + 3 + k]);
                    }
                }
            }
            const SahSplit sp = sah_choose(B, scale, cnt, level);
            uint32_t nl = 0, nr = 0;
            for (uint32_t i = lo; i < hi; ++i) {
                const uint32_t p = ids[i];
                const bool l = sp.axis >= 0 ? sah_bin(cen(p, sp.axis), cmin[sp.axis], scale[sp.axis]) <= sp.bin
                                            : (i - lo) < sp.nleft;
                if (l)
                    out[lo + nl++] = p;
                else
                    out[lo + sp.nleft + nr++] = p;
            }
            for (int side = 0; side < 2; ++side) {
                const uint32_t start = side ? lo + sp.nleft : lo, m = side ? cnt - sp.nleft : sp.nleft;
                uint32_t child;
                if (m == 1) {
                    child = n - 1 + start;
                    vals[start] = out[start];
                } else {
                    child = alloc++;
                    rfirst[child] = start;
                    ncount[child] = m;
                    next.push_back({start, child});
                }
                (side ? right : left)[t.node] = child;
            }
        }
        for (const Task &t : next)
            for (uint32_t i = t.lo; i < t.lo + ncount[t.node]; ++i) ids[i] = out[i];
        tasks.swap(next);
    }
}

// prim boxes: n x (lo3, hi3)
int build(const float *pbox, uint32_t n, uint32_t leaf_max, HostBvh &out, bool sah = false)
{
    out.nodes.clear();
    out.leaf_ids.assign(n, 0);
    if (n == 0) return 0;
    float blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            const float c = 0.5f * pbox[6 * i + a] + 0.5f * pbox[6 * i + 3 + a];
            blo[a] = fminf(blo[a], c);
            bhi[a] = fmaxf(bhi[a], c);
        }
    std::vector<uint32_t> keys(n), vals(n);
    for (uint32_t i = 0; i < n; ++i) keys[i] = morton30(pbox + 6 * i, pbox + 6 * i + 3, blo, bhi);
    std::iota(vals.begin(), vals.end(), 0u);
    std::stable_sort(vals.begin(), vals.end(), [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
    std::vector<uint32_t> skeys(n);
    for (uint32_t i = 0; i < n; ++i) skeys[i] = keys[vals[i]];
    const uint32_t nb = 2 * n - 1;
    std::vector<uint32_t> left(n), right(n), parent(nb, 0xFFFFFFFFu), rfirst(n), ncount(nb, 0);
    std::vector<float4> nbox(2 * (size_t)nb);
    if (sah && n > 1) {
        sah_tree(pbox, n, vals, left, right, rfirst, ncount);
        for (uint32_t i = 0; i + 1 < n; ++i) {
            parent[left[i]] = i;
            parent[right[i]] = i;
        }
    } else {
        for (int i = 0; i + 1 < (int)n; ++i)
            karras_node(skeys.data(), (int)n, i, left.data(), right.data(), parent.data(), rfirst.data());
    }
    if (n > 1) parent[0] = 0xFFFFFFFFu;
    // every node except the root must have exactly one parent
    for (uint32_t k = 1; k < nb; ++k)
        if (n > 1 && parent[k] >= n - 1) return -10;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t p = vals[k];
        const uint32_t node = n - 1 + k;
        nbox[2 * node] = make_float4(pbox[6 * p], pbox[6 * p + 1], pbox[6 * p + 2], 0.0f);
        nbox[2 * node + 1] = make_float4(pbox[6 * p + 3], pbox[6 * p + 4], pbox[6 * p + 5], 0.0f);
        ncount[node] = 1;
    }
    // internal nodes: children have larger indices than... not guaranteed, so
    // resolve by repeated passes over a post-order from the root
    if (n > 1) {
        std::vector<uint32_t> order, stack{0u};
        while (!stack.empty()) {
            const uint32_t x = stack.back();
            stack.pop_back();
            order.push_back(x);
            if (x < n - 1) {
                stack.push_back(left[x]);
                stack.push_back(right[x]);
            }
        }
        if (order.size() != nb) return -11;
        for (auto it = order.rbegin(); it != order.rend(); ++it) {
            const uint32_t x = *it;
            if (x >= n - 1) continue;
            const float4 al = nbox[2 * left[x]], ah = nbox[2 * left[x] + 1];
            const float4 bl = nbox[2 * right[x]], bh = nbox[2 * right[x] + 1];
            nbox[2 * x] = make_float4(fminf(al.x, bl.x), fminf(al.y, bl.y), fminf(al.z, bl.z), 0.0f);
            nbox[2 * x + 1] = make_float4(fmaxf(ah.x, bh.x), fmaxf(ah.y, bh.y), fmaxf(ah.z, bh.z), 0.0f);
            ncount[x] = ncount[left[x]] + ncount[right[x]];
        }
    }
    const float4 rl = nbox[0], rh = nbox[1];
    const float rb[6] = {rl.x, rl.y, rl.z, rh.x, rh.y, rh.z};
    std::memcpy(out.root_box, rb, sizeof(rb));
    out.nodes.assign(n, WideNode{});
    std::vector<uint32_t> wsrc(n, 0);
    uint32_t counters[3] = {1u, 0u, 0u};
    CollapseArgs A;
    A.n = (int)n;
    A.leaf_max = leaf_max;
    A.node_cap = n;
    A.left = left.data();
    A.right = right.data();
    A.rfirst = rfirst.data();
    A.ncount = ncount.data();
    A.nbox = nbox.data();
    A.vals = vals.data();
    A.wnodes = out.nodes.data();
    A.wsrc = wsrc.data();
    A.counters = counters;
    A.leaf_ids = out.leaf_ids.data();
    uint32_t begin = 0, end = 1;
    out.depth = 0;
    auto alloc = [&](int which, uint32_t amount) {
        const uint32_t old = counters[which];
        counters[which] += amount;
        return old;
    };
    while (begin < end) {
        A.begin = begin;
        A.end = end;
        for (uint32_t w = begin; w < end; ++w)
            if (collapse_node(A, w, alloc)) return -12;
        begin = end;
        end = counters[0];
        ++out.depth;
        if (end > n) return -13;
    }
    if (counters[1] != n) return -14;
    out.nodes.resize(end);
    for (uint32_t w = 0; w < end; ++w)
        if (!node_valid(out.nodes[w], w, end, n)) return -15;
    return 0;
}

struct HostBlas {
    HostBvh bvh;
    std::vector<TriRec> tris;
};
struct HostScene {
    std::vector<HostBlas> blas;
    HostBvh tlas;
    std::vector<InstRec> insts;
    std::vector<uint32_t> meta;
};

}  // namespace

extern "C" {

// tris: per BLAS, 9 floats per triangle; inst: 12-float row-major transforms,
// blas index, mask.  Returns a handle (heap pointer) or null on error.
void *hc_scene_create(const float *const *tris, const uint32_t *ntris, uint32_t nblas, const float *xf,
                      const uint32_t *inst_blas, const uint32_t *inst_mask, uint32_t ninst, int *err)
{
    HostScene *s = new HostScene();
    s->blas.resize(nblas);
    for (uint32_t b = 0; b < nblas; ++b) {
        const uint32_t n = ntris[b];
        std::vector<float> pbox(6 * (size_t)n);
        for (uint32_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) {
                const float *t = tris[b] + 9 * (size_t)i;
                pbox[6 * i + a] = fminf(fminf(t[a], t[3 + a]), t[6 + a]);
                pbox[6 * i + 3 + a] = fmaxf(fmaxf(t[a], t[3 + a]), t[6 + a]);
            }
        int rc = build(pbox.data(), n, kLeafMaxTri, s->blas[b].bvh, (g_build_flags & 1u) != 0);
        if (rc) {
            *err = rc;
            delete s;
            return nullptr;
        }
        s->blas[b].tris.resize(n);
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t p = s->blas[b].bvh.leaf_ids[i];
            TriRec r;
            std::memcpy(r.v, tris[b] + 9 * (size_t)p, 36);
            r.prim = p;
            r.geom = 0;
            r.lprim = p;
            s->blas[b].tris[i] = r;
        }
    }
    std::vector<float> pbox(6 * (size_t)ninst);
    for (uint32_t i = 0; i < ninst; ++i) {
        const HostBlas &hb = s->blas[inst_blas[i]];
        instance_box(xf + 12 * (size_t)i, !hb.tris.empty(), hb.bvh.root_box, &pbox[6 * i], &pbox[6 * i + 3]);
    }
    int rc = build(pbox.data(), ninst, kLeafMaxInst, s->tlas);
    if (rc) {
        *err = rc - 100;
        delete s;
        return nullptr;
    }
    s->insts.resize(ninst);
    s->meta.resize(ninst);
    for (uint32_t i = 0; i < ninst; ++i) {
        const uint32_t p = s->tlas.leaf_ids[i];
        InstRec r;
        inverse3x4(xf + 12 * (size_t)p, r.inv);
        const HostBlas &hb = s->blas[inst_blas[p]];
        r.nodes = hb.tris.empty() ? nullptr : hb.bvh.nodes.data();
        r.tris = hb.tris.data();
        s->insts[i] = r;
        s->meta[i] = (p & 0xFFFFFFu) | (inst_mask[p] << 24);
    }
    *err = 0;
    return s;
}

void hc_scene_destroy(void *h) { delete static_cast<HostScene *>(h); }

// BLAS build flags of the next hc_scene_create (1: binned SAH, as
// SOLHIP_BUILD_PREFER_FAST_TRACE)
void hc_set_build_flags(uint32_t flags) { g_build_flags = flags; }

void hc_scene_info(void *h, uint32_t blas, uint32_t *nodes, uint32_t *depth)
{
    HostScene *s = static_cast<HostScene *>(h);
    const HostBvh &b = blas == 0xFFFFFFFFu ? s->tlas : s->blas[blas].bvh;
    *nodes = (uint32_t)b.nodes.size();
    *depth = b.depth;
}

// rays: 8 floats each; hits: 5 words each; counts: nodes, tris, insts, overflow
void hc_trace(void *h, const float *rays, uint32_t n, uint32_t any, void *hits, uint64_t *counts)
{
    HostScene *s = static_cast<HostScene *>(h);
    SceneView sv;
    sv.tnodes = s->insts.empty() ? nullptr : s->tlas.nodes.data();
    sv.insts = s->insts.data();
    sv.single = s->insts.size() == 1 ? 1u : 0u;
    sv.inst_meta = s->meta.data();
    uint32_t max_blas_depth = 0;
    for (const auto &b : s->blas) max_blas_depth = std::max(max_blas_depth, b.bvh.depth);
    sv.stack_depth = std::min<uint32_t>(kMaxStack, s->tlas.depth + max_blas_depth + 1);
    uint2 lds[kMaxStack];
    for (uint32_t i = 0; i < n; ++i) {
        const float *r = rays + 8 * (size_t)i;
        TravCounts tc{0, 0, 0, false};
        Hit hit = any ? trace_ray<true, true>(sv, v3(r[0], r[1], r[2]), v3(r[4], r[5], r[6]), r[3], r[7], 0xFFu, lds, 1, tc)
                      : trace_ray<false, true>(sv, v3(r[0], r[1], r[2]), v3(r[4], r[5], r[6]), r[3], r[7], 0xFFu, lds, 1, tc);
        put_hit(hits, i, hit);
        counts[0] += tc.nodes;
        counts[1] += tc.tris;
        counts[2] += tc.insts;
        counts[3] += tc.overflow ? 1 : 0;
    }
}

// the resumable traversal as the persistent AO kernel drives it: 64 lanes,
// each with its own stack column, take rays from a shared cursor and advance
// one step per round; a lane that finishes takes the next ray at once
void hc_trace_steps(void *h, const float *rays, uint32_t n, uint32_t any, void *hits, uint64_t *counts)
{
    HostScene *s = static_cast<HostScene *>(h);
    SceneView sv;
    sv.tnodes = s->insts.empty() ? nullptr : s->tlas.nodes.data();
    sv.insts = s->insts.data();
    sv.single = s->insts.size() == 1 ? 1u : 0u;
    sv.inst_meta = s->meta.data();
    uint32_t max_blas_depth = 0;
    for (const auto &b : s->blas) max_blas_depth = std::max(max_blas_depth, b.bvh.depth);
    sv.stack_depth = std::min<uint32_t>(kMaxStack, s->tlas.depth + max_blas_depth + 1) + 2;
    constexpr int kLanes = 64;
    std::vector<uint2> lds((size_t)(kMaxStack + 2) * kLanes);
    TravState st[kLanes];
    int64_t rid[kLanes];
    TravCounts tc[kLanes];
    uint32_t cursor = 0;
    for (int l = 0; l < kLanes; ++l) rid[l] = -1;
    while (true) {
        bool any_live = false;
        for (int l = 0; l < kLanes; ++l) {
            if (rid[l] < 0 && cursor < n) {
                rid[l] = cursor++;
                const float *r = rays + 8 * (size_t)rid[l];
                trav_init(st[l], sv, v3(r[0], r[1], r[2]), v3(r[4], r[5], r[6]), r[3], r[7]);
                tc[l] = TravCounts{0, 0, 0, false};
            }
            if (rid[l] < 0) continue;
            any_live = true;
            const bool done = any ? trav_step<true, true>(st[l], sv, 0xFFu, lds.data() + l, kLanes, tc[l])
                                  : trav_step<false, true>(st[l], sv, 0xFFu, lds.data() + l, kLanes, tc[l]);
            if (done) {
                const Hit hit = trav_result(st[l]);
                put_hit(hits, rid[l], hit);
                counts[0] += tc[l].nodes;
                counts[1] += tc[l].tris;
                counts[2] += tc[l].insts;
                counts[3] += tc[l].overflow ? 1 : 0;
                rid[l] = -1;
            }
        }
        if (!any_live && cursor >= n) break;
    }
}

// The BLAS build's vertex fetch replayed on the host over n_geo geometry
// descriptors (host pointers): positions of every BLAS-wide triangle (9
// floats each, as k_tri_boxes / k_write_tris read them) and the geometry and
// in-geometry index of each; returns the number of out-of-range references
// the build would count.
uint32_t hc_geometry_fetch(const GeomDesc *geo, uint32_t n_geo, uint32_t n_tris, float *pos, uint32_t *geom,
                           uint32_t *lprim)
{
    uint32_t bad = 0;
    for (uint32_t t = 0; t < n_tris; ++t) {
        const uint32_t g = geom_index(geo, n_geo, t);
        const GeomDesc &G = geo[g];
        for (int k = 0; k < 3; ++k) {
            uint32_t iv;
            if (!tri_vertex_index(G, t, k, iv)) ++bad;
            const float *q = reinterpret_cast<const float *>(G.vbase + (uint64_t)(G.first_vertex + iv) * G.stride);
            std::memcpy(pos + 9 * (size_t)t + 3 * k, q, 12);
        }
        geom[t] = g;
        lprim[t] = t - G.tri_begin;
    }
    return bad;
}

}  // extern "C"
