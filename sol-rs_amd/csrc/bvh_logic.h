// bvh_logic.h -- per-item steps of the BVH build, shared by the GPU kernels
// (bvh_build.hip) and any host replay of the same steps.  Every index that
// is computed from data is range-checked: a violation is reported through
// the returned error count instead of touching memory out of bounds.
#pragma once

#include "common.h"
#include "devmath.h"

namespace solhip {

SOLHIP_HD uint32_t clz32(uint32_t x) { return x ? (uint32_t)__builtin_clz(x) : 32u; }

// one geometry of a BLAS build, as the build kernels read it: vertex k of
// BLAS-wide triangle t (tri_begin <= t < tri_begin + tri_count) is at
// vbase + (first_vertex + i_k) * stride with i_k = idx[3 (t - tri_begin) + k],
// or i_k = 3 (t - tri_begin) + k when idx is null (vbase then includes the
// primitive offset); indices >= vertex_count are counted as errors
struct GeomDesc {
    const uint8_t *vbase;
    const uint32_t *idx;
    uint64_t stride;
    uint32_t first_vertex, vertex_count;
    uint32_t tri_begin, tri_count;
};
static_assert(sizeof(GeomDesc) == 40, "GeomDesc layout");

// geometry of BLAS-wide triangle t: the last descriptor whose tri_begin <= t
// (descriptors are in tri_begin order; one geometry needs no search)
SOLHIP_HD uint32_t geom_index(const GeomDesc *geo, uint32_t n_geo, uint32_t t)
{
    uint32_t lo = 0, hi = n_geo;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (geo[mid].tri_begin <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

// vertex of corner k of BLAS-wide triangle t in geometry G, relative to
// G.first_vertex.  False when the index is >= vertex_count: iv is then the
// last vertex (0 without vertices), so that malformed input cannot make the
// build read outside the geometry's vertices, and the build fails.
SOLHIP_HD bool tri_vertex_index(const GeomDesc &G, uint32_t t, int k, uint32_t &iv)
{
    const uint32_t local = t - G.tri_begin;
    iv = G.idx ? G.idx[3ull * local + k] : 3u * local + (uint32_t)k;
    if (iv < G.vertex_count) return true;
    iv = G.vertex_count ? G.vertex_count - 1 : 0;
    return false;
}

SOLHIP_HD uint32_t expand10(uint32_t v)
{
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// 30-bit Morton code of the box centroid within the centroid bounds.  All
// three axes share one scale, 1024 cells over the largest extent: scaling each
// axis to its own extent would give a thin axis (the height of a terrain) as
// many top-level splits as the wide ones, and those slabs cut across the whole
// surface (a 1024^2 terrain: 16.6 -> 12.8 node tests per primary ray).
SOLHIP_HD uint32_t morton30(const float lo[3], const float hi[3], const float blo[3], const float bhi[3])
{
    const float ext = fmaxf(fmaxf(bhi[0] - blo[0], bhi[1] - blo[1]), bhi[2] - blo[2]);
    const float scale = ext > 0.0f ? 1024.0f / ext : 0.0f;
    uint32_t code = 0;
    for (int a = 0; a < 3; ++a) {
        const float c = 0.5f * lo[a] + 0.5f * hi[a];
        const float q = (c - blo[a]) * scale;
        const uint32_t qi = q <= 0.0f ? 0u : (q >= 1023.0f ? 1023u : (uint32_t)q);
        code |= expand10(qi) << (2 - a);
    }
    return code;
}

// Karras 2012 common-prefix length with the index tie-break for equal codes
SOLHIP_HD int karras_delta(const uint32_t *k, int n, int i, int j)
{
    if (j < 0 || j >= n) return -1;
    const uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + (int)clz32((uint32_t)i ^ (uint32_t)j);
    return (int)clz32(a ^ b);
}

// internal node i in [0, n-2] of the radix tree over n sorted codes; leaf k
// is node n-1+k.  Writes children, parents and the covered leaf range.
SOLHIP_HD void karras_node(const uint32_t *keys, int n, int i, uint32_t *left, uint32_t *right, uint32_t *parent,
                           uint32_t *rfirst, uint32_t *rcount = nullptr)
{
    const int d = (karras_delta(keys, n, i, i + 1) - karras_delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = karras_delta(keys, n, i, i - d);
    int lmax = 2;
    while (karras_delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (karras_delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = karras_delta(keys, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (karras_delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + (d < 0 ? d : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const uint32_t lc = (lo == gamma) ? (uint32_t)(n - 1 + gamma) : (uint32_t)gamma;
    const uint32_t rc = (hi == gamma + 1) ? (uint32_t)(n - 1 + gamma + 1) : (uint32_t)(gamma + 1);
    left[i] = lc;
    right[i] = rc;
    parent[lc] = (uint32_t)i;
    parent[rc] = (uint32_t)i;
    rfirst[i] = (uint32_t)lo;
    if (rcount) rcount[i] = (uint32_t)(hi - lo + 1);
}

struct CollapseArgs {
    int n;                     // primitives
    uint32_t leaf_max;
    uint32_t node_cap;         // capacity of wnodes / wsrc
    const uint32_t *left, *right, *rfirst, *ncount;
    const float4 *nbox;        // 2 float4 (lo, hi) per binary node
    const uint32_t *vals;      // sorted primitive ids
    WideNode *wnodes;
    uint32_t *wsrc;            // binary node of each wide node
    uint32_t *counters;        // [0] wide nodes allocated, [1] leaf prims allocated, [2] errors
    uint32_t *leaf_ids;        // original primitive id of each leaf record
    uint32_t begin, end;       // wide nodes of this level
};

SOLHIP_HD float half_area(float4 lo, float4 hi)
{
    const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    return dx * dy + dy * dz + dz * dx;
}

// ---- binned SAH split (the fast-trace BLAS build) ------------------------
// A node's primitives are binned by centroid into kSahBins equal bins per
// axis over the node's centroid range; the split between two bins with the
// least cost A(left) * N(left) + A(right) * N(right) (A = half area of the
// union of the side's primitive boxes) is taken; ties keep the first found
// (axis x, y, z, bin ascending).  Without a split that leaves both sides
// non-empty (all centroids in one bin, or a zero range), and at every level
// from kSahMedianLevel on (a depth bound), the range is halved in its current
// order instead.
constexpr int kSahBins = 32;
constexpr uint32_t kSahMedianLevel = 48;

SOLHIP_HD float half_area3(const float l[3], const float h[3])
{
    const float dx = h[0] - l[0], dy = h[1] - l[1], dz = h[2] - l[2];
    return dx * dy + dy * dz + dz * dx;
}

// scale of an axis with centroid extent ext (0: the axis cannot be split)
SOLHIP_HD float sah_scale(float ext) { return ext > 0.0f && ext < INFINITY ? (float)kSahBins / ext : 0.0f; }

SOLHIP_HD int sah_bin(float c, float cmin, float scale)
{
    const float q = (c - cmin) * scale;
    if (!(q > 0.0f)) return 0;
    if (q >= (float)(kSahBins - 1)) return kSahBins - 1;
    return (int)q;
}

struct SahSplit {
    int axis;        // -1: halve the range in its current order
    int bin;         // left side = bins 0..bin of `axis`
    uint32_t nleft;
};

// Bins: count(a, b), lo(a, b, k), hi(a, b, k) (empty bins: +inf / -inf).
// Best split along axis a; updates (best_cost, best) only on a strictly lower
// cost, so running the axes in order x, y, z (or each alone and keeping the
// first strict minimum in that order) keeps the first split found.
template <class Bins>
SOLHIP_HD void sah_axis(const Bins &B, int a, float &best_cost, SahSplit &best)
{
    float ra[kSahBins];
    uint32_t rc[kSahBins];
    float l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
    uint32_t c = 0;
    for (int b = kSahBins - 1; b > 0; --b) {
        c += B.count(a, b);
        for (int k = 0; k < 3; ++k) {
            l[k] = fminf(l[k], B.lo(a, b, k));
            h[k] = fmaxf(h[k], B.hi(a, b, k));
        }
        ra[b] = c ? half_area3(l, h) : 0.0f;
        rc[b] = c;
    }
    for (int k = 0; k < 3; ++k) {
        l[k] = INFINITY;
        h[k] = -INFINITY;
    }
    c = 0;
    for (int b = 0; b < kSahBins - 1; ++b) {
        c += B.count(a, b);
        for (int k = 0; k < 3; ++k) {
            l[k] = fminf(l[k], B.lo(a, b, k));
            h[k] = fmaxf(h[k], B.hi(a, b, k));
        }
        if (c == 0 || rc[b + 1] == 0) continue;
        const float cost = half_area3(l, h) * (float)c + ra[b + 1] * (float)rc[b + 1];
        if (cost < best_cost) {
            best_cost = cost;
            best.axis = a;
            best.bin = b;
            best.nleft = c;
        }
    }
}

template <class Bins>
SOLHIP_HD SahSplit sah_choose(const Bins &B, const float scale[3], uint32_t count, uint32_t level)
{
    SahSplit best{-1, 0, count / 2};
    if (level >= kSahMedianLevel) return best;
    float best_cost = INFINITY;
    for (int a = 0; a < 3; ++a)
        if (scale[a] > 0.0f) sah_axis(B, a, best_cost, best);
    return best;
}

// smallest power of two 2^e (e >= -126) with 2^e >= x, as biased exponent
SOLHIP_HD uint32_t pow2_ceil_exp(float x)
{
    if (!(x > 0x1p-126f)) return 1u;
    const uint32_t b = as_u(x);
    uint32_t e = (b >> 23) & 0xFFu;
    if (b & 0x7FFFFFu) e += 1u;
    return e < 254u ? e : 254u;
}

// Build wide node w from its binary source node.  `alloc(counter, amount)`
// returns the previous value of counters[counter] after adding amount.
// Returns the number of range violations (0 on a consistent tree).
template <class Alloc>
SOLHIP_HD uint32_t collapse_node(const CollapseArgs &A, uint32_t w, Alloc alloc)
{
    const int n = A.n;
    const uint32_t nbin = 2u * (uint32_t)n - 1u;
    const uint32_t b = A.wsrc[w];
    if (b >= nbin) return 1;
    uint32_t ch[8];
    int nc;
    if (b >= (uint32_t)(n - 1)) {
        ch[0] = b;  // single-leaf root
        nc = 1;
    } else {
        ch[0] = A.left[b];
        ch[1] = A.right[b];
        nc = 2;
        if (ch[0] >= nbin || ch[1] >= nbin) return 1;
    }
    while (nc < 8) {
        int best = -1;
        float best_a = -1.0f;
        for (int c = 0; c < nc; ++c) {
            if (ch[c] >= (uint32_t)(n - 1)) continue;  // binary leaf
            const float a = half_area(A.nbox[2ull * ch[c]], A.nbox[2ull * ch[c] + 1]);
            if (best < 0 || a > best_a) {
                best_a = a;
                best = c;
            }
        }
        if (best < 0) break;
        const uint32_t x = ch[best];
        const uint32_t l = A.left[x], r = A.right[x];
        if (l >= nbin || r >= nbin) return 1;
        ch[best] = l;
        ch[nc++] = r;
    }
    // node box and child centroids
    float4 lo = make_float4(INFINITY, INFINITY, INFINITY, 0.0f), hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
    float cc[8][3];
    for (int c = 0; c < nc; ++c) {
        const float4 l = A.nbox[2ull * ch[c]], h = A.nbox[2ull * ch[c] + 1];
        lo = make_float4(fminf(lo.x, l.x), fminf(lo.y, l.y), fminf(lo.z, l.z), 0.0f);
        hi = make_float4(fmaxf(hi.x, h.x), fmaxf(hi.y, h.y), fmaxf(hi.z, h.z), 0.0f);
        cc[c][0] = 0.5f * l.x + 0.5f * h.x;
        cc[c][1] = 0.5f * l.y + 0.5f * h.y;
        cc[c][2] = 0.5f * l.z + 0.5f * h.z;
    }
    const float pc[3] = {0.5f * lo.x + 0.5f * hi.x, 0.5f * lo.y + 0.5f * hi.y, 0.5f * lo.z + 0.5f * hi.z};
    // octant slot assignment: greedily give slot s the child minimising
    // dot(c - p, w_s), w_s[a] = bit a of s ? -1 : +1
    int slot_of[8];
    int child_in[8];
    for (int s = 0; s < 8; ++s) child_in[s] = -1;
    for (int c = 0; c < 8; ++c) slot_of[c] = -1;
    for (int it = 0; it < nc; ++it) {
        float bc = 0.0f;
        int bcc = -1, bs = -1;
        for (int c = 0; c < nc; ++c) {
            if (slot_of[c] >= 0) continue;
            for (int s = 0; s < 8; ++s) {
                if (child_in[s] >= 0) continue;
                float cost = 0.0f;
                for (int a = 0; a < 3; ++a) {
                    const float d = cc[c][a] - pc[a];
                    cost += ((s >> a) & 1) ? -d : d;
                }
                if (bcc < 0 || cost < bc) {
                    bc = cost;
                    bcc = c;
                    bs = s;
                }
            }
        }
        if (bcc < 0) return 1;
        slot_of[bcc] = bs;
        child_in[bs] = bcc;
    }
    // classify and allocate
    uint32_t imask = 0, nint = 0, nprim = 0;
    for (int s = 0; s < 8; ++s) {
        const int c = child_in[s];
        if (c < 0) continue;
        const uint32_t cnt = A.ncount[ch[c]];
        if (cnt > A.leaf_max) {
            imask |= 1u << s;
            ++nint;
        } else {
            nprim += cnt;
        }
    }
    const uint32_t child_base = nint ? alloc(0, nint) : 0u;
    const uint32_t prim_base = nprim ? alloc(1, nprim) : 0u;
    if ((uint64_t)child_base + nint > A.node_cap || (uint64_t)prim_base + nprim > (uint64_t)n) return 1;
    // quantisation frame: scale 2^e >= extent / 253 (with a floor of 2^-8 of
    // the largest extent so that flat boxes keep a padding step that is
    // large against fp32 error in the slab distances), origin one step below
    WideNode nd;
    const float ext[3] = {hi.x - lo.x, hi.y - lo.y, hi.z - lo.z};
    const float los[3] = {lo.x, lo.y, lo.z};
    const float maxext = fmaxf(fmaxf(ext[0], ext[1]), ext[2]);
    uint32_t eb[3];
    float sc[3], org[3];
    for (int a = 0; a < 3; ++a) {
        eb[a] = pow2_ceil_exp(fmaxf(ext[a], maxext * 0x1p-8f) / 253.0f);
        sc[a] = as_f(eb[a] << 23);
        org[a] = los[a] - sc[a];
    }
    nd.px = org[0];
    nd.py = org[1];
    nd.pz = org[2];
    nd.ex = (uint8_t)eb[0];
    nd.ey = (uint8_t)eb[1];
    nd.ez = (uint8_t)eb[2];
    nd.imask = (uint8_t)imask;
    nd.child_base = child_base;
    nd.prim_base = prim_base;
    uint32_t rank = 0, poff = 0;
    for (int s = 0; s < 8; ++s) {
        const int c = child_in[s];
        uint8_t *qlo[3] = {nd.qlo_x, nd.qlo_y, nd.qlo_z};
        uint8_t *qhi[3] = {nd.qhi_x, nd.qhi_y, nd.qhi_z};
        if (c < 0) {
            nd.meta[s] = 0;
            for (int a = 0; a < 3; ++a) {
                qlo[a][s] = 255;
                qhi[a][s] = 0;
            }
            continue;
        }
        const uint32_t bn = ch[c];
        const float4 l = A.nbox[2ull * bn], h = A.nbox[2ull * bn + 1];
        const float cl[3] = {l.x, l.y, l.z}, chh[3] = {h.x, h.y, h.z};
        for (int a = 0; a < 3; ++a) {
            const double o = (double)org[a], s_ = (double)sc[a];
            float fl = floorf((cl[a] - org[a]) / sc[a]) - 1.0f;
            float fh = ceilf((chh[a] - org[a]) / sc[a]) + 1.0f;
            int ql = fl <= 0.0f ? 0 : (fl >= 255.0f ? 255 : (int)fl);
            int qh = fh <= 0.0f ? 0 : (fh >= 255.0f ? 255 : (int)fh);
            while (ql > 0 && o + ql * s_ > (double)cl[a]) --ql;
            while (qh < 255 && o + qh * s_ < (double)chh[a]) ++qh;
            qlo[a][s] = (uint8_t)ql;
            qhi[a][s] = (uint8_t)qh;
        }
        const uint32_t cnt = A.ncount[bn];
        if ((imask >> s) & 1u) {
            nd.meta[s] = 0;
            A.wsrc[child_base + rank] = bn;
            ++rank;
        } else {
            nd.meta[s] = (uint8_t)((cnt << 5) | poff);
            const uint32_t first = bn >= (uint32_t)(n - 1) ? bn - (uint32_t)(n - 1) : A.rfirst[bn];
            if ((uint64_t)first + cnt > (uint64_t)n) return 1;
            for (uint32_t j = 0; j < cnt; ++j) A.leaf_ids[prim_base + poff + j] = A.vals[first + j];
            poff += cnt;
        }
    }
    A.wnodes[w] = nd;
    return 0;
}

// structural check of a collapsed node: children are allocated after their
// parent (so the node graph is acyclic) and every index is in range
SOLHIP_HD bool node_valid(const WideNode &nd, uint32_t w, uint32_t nodes, uint32_t prims)
{
    uint32_t nint = 0, nprim = 0;
    for (int s = 0; s < 8; ++s) {
        if ((nd.imask >> s) & 1u) {
            ++nint;
            continue;
        }
        const uint32_t m = nd.meta[s];
        if (!m) continue;
        const uint32_t cnt = m >> 5, off = m & 31u;
        if (cnt == 0 || (uint64_t)nd.prim_base + off + cnt > prims) return false;
        nprim += cnt;
    }
    if (nint && (nd.child_base <= w || (uint64_t)nd.child_base + nint > nodes)) return false;
    if (nprim && (uint64_t)nd.prim_base + nprim > prims) return false;
    return true;
}

// world-to-object inverse of a row-major 3x4 affine transform by adjugate /
// determinant (the host restatement evaluates the identical expression)
SOLHIP_HD void inverse3x4(const float *m, float *r)
{
    const float a = m[0], b = m[1], c = m[2], tx = m[3];
    const float d = m[4], e = m[5], f = m[6], ty = m[7];
    const float g = m[8], h = m[9], i = m[10], tz = m[11];
    const float c00 = e * i - f * h, c01 = c * h - b * i, c02 = b * f - c * e;
    const float c10 = f * g - d * i, c11 = a * i - c * g, c12 = c * d - a * f;
    const float c20 = d * h - e * g, c21 = b * g - a * h, c22 = a * e - b * d;
    const float det = a * c00 + b * c10 + c * c20;
    const float rd = 1.0f / det;
    r[0] = c00 * rd;
    r[1] = c01 * rd;
    r[2] = c02 * rd;
    r[4] = c10 * rd;
    r[5] = c11 * rd;
    r[6] = c12 * rd;
    r[8] = c20 * rd;
    r[9] = c21 * rd;
    r[10] = c22 * rd;
    r[3] = -(r[0] * tx + r[1] * ty + r[2] * tz);
    r[7] = -(r[4] * tx + r[5] * ty + r[6] * tz);
    r[11] = -(r[8] * tx + r[9] * ty + r[10] * tz);
}

// world box of an instance: the 8 corners of the BLAS root box through the
// 3x4 object-to-world transform, padded relatively (an empty BLAS gives the
// degenerate box at the instance origin)
SOLHIP_HD void instance_box(const float *m, bool has_blas, const float *rb, float lo[3], float hi[3])
{
    for (int a = 0; a < 3; ++a) {
        lo[a] = INFINITY;
        hi[a] = -INFINITY;
    }
    if (!has_blas) {
        for (int a = 0; a < 3; ++a) lo[a] = hi[a] = m[4 * a + 3];
        return;
    }
    for (int k = 0; k < 8; ++k) {
        const float p[3] = {(k & 1) ? rb[3] : rb[0], (k & 2) ? rb[4] : rb[1], (k & 4) ? rb[5] : rb[2]};
        for (int a = 0; a < 3; ++a) {
            const float w = m[4 * a] * p[0] + m[4 * a + 1] * p[1] + m[4 * a + 2] * p[2] + m[4 * a + 3];
            lo[a] = fminf(lo[a], w);
            hi[a] = fmaxf(hi[a], w);
        }
    }
    for (int a = 0; a < 3; ++a) {
        const float pad = (hi[a] - lo[a]) * 1e-5f + fmaxf(fabsf(lo[a]), fabsf(hi[a])) * 1e-6f;
        lo[a] -= pad;
        hi[a] += pad;
    }
}

}  // namespace solhip
