// pathshade.h -- one path vertex of the wavefront path tracer: the
// pathtrace.rchit / pathtrace.rmiss material model on a closest hit, with
// next-event estimation of the diffuse lobe and Russian roulette.  Shared by
// the per-frame wavefront (wavefront.hip) and the multi-job server
// (server.hip), so both produce the same bits.
#pragma once

#include "kernels.h"
#include "shading.h"

namespace solhip {

constexpr float kPi = 0x1.921fb6p+1f;

__device__ __forceinline__ bool is_emitter(const float *mat)
{
    return mat[4] >= 1.0f || mat[5] >= 1.0f || mat[6] >= 1.0f;   // pathtrace.rchit:71
}

// first light whose inclusive area sum exceeds u (the last one if none)
__device__ __forceinline__ uint32_t pick_light(const LightTri *L, uint32_t n, float u)
{
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (L[mid].cdf > u) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

struct ShadeParams {
    DevTables tb;
    uint32_t enable_sky, max_depth, rr_start;
    bool nee;                 // next-event estimation active (lights present, total area > 0)
    const LightTri *lights;
    uint32_t nlights;
    float total;              // emitter area
    float rmax;
};

struct ShadeOut {
    bool cont = false, shadow = false;
    QRay nr, sr;              // next extension ray / shadow ray
    float4 sc;                // shadow-ray contribution
};

// Shades the hit h of ray r of path p.  s0 = (T, rng), s1 = (L, depth |
// count_emission << 16) are updated in place; o_ gets the rays to queue.
__device__ __forceinline__ void shade_vertex(const ShadeParams &P, uint32_t p, const QRay &r, const QHit &h,
                                             float4 &s0, float4 &s1, ShadeOut &o_)
{
    o_.sc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    V3 T = v3(s0.x, s0.y, s0.z);
    uint32_t rng = as_u(s0.w);
    V3 L = v3(s1.x, s1.y, s1.z);
    const uint32_t fl = as_u(s1.w);
    uint32_t depth = fl & 0xFFFFu;
    bool count_emit = (fl >> 16) & 1u;
    const V3 d = v3(r.dx, r.dy, r.dz);
    if (h.inst == SOLHIP_MISS) {
        const V3 sky = pt_miss(P.enable_sky, P.enable_sky != 2, d);
        L = v3(L.x + T.x * sky.x, L.y + T.y * sky.y, L.z + T.z * sky.z);
    } else {
        const float *mat = reinterpret_cast<const float *>(P.tb.material_table[h.inst]);
        if (is_emitter(mat)) {
            if (count_emit) L = v3(L.x + T.x * mat[4], L.y + T.y * mat[5], L.z + T.z * mat[6]);
        } else {
            const Surf sf = fetch_surface(P.tb, h.inst, h.prim, h.u, h.v);
            const V3 wI = normalize(d);
            const float sg = sgn(dot(sf.nrm, neg(wI)));
            const V3 nO = scale(sf.nrm, sg);
            const float metallic = mat[8], roughness = mat[9];
            const float alphaSquared = roughness * roughness;
            const float xi0 = next_rand(rng);
            const float xi1 = next_rand(rng);
            const float rnd = next_rand(rng);
            const V3 o = v3(sf.pos.x + 0.0001f * nO.x, sf.pos.y + 0.0001f * nO.y,
                            sf.pos.z + 0.0001f * nO.z);
            const V3 base_c = v3(mat[0] * sf.col.x, mat[1] * sf.col.y, mat[2] * sf.col.z);
            V3 nd, hv;
            bool diffuse = false;
            if (rnd < metallic) {
                nd = sample_ggx(reflect(d, nO), xi0, xi1, alphaSquared);
                hv = base_c;
            } else {
                const V3 m = sample_ggx(nO, xi0, xi1, alphaSquared);
                if (rnd < fresnel_dielectric(nO, m, 1.0f / 1.5f)) {
                    nd = reflect(d, m);
                    hv = v3(1.0f, 1.0f, 1.0f);
                } else {
                    nd = sample_cosine(nO, xi0, xi1);
                    hv = base_c;
                    diffuse = true;
                }
            }
            // throughput into this vertex times its diffuse albedo
            const V3 Tb = v3(T.x * base_c.x, T.y * base_c.y, T.z * base_c.z);
            T = v3(T.x * hv.x, T.y * hv.y, T.z * hv.z);
            depth++;
            if (depth <= P.max_depth) {
                count_emit = true;
                if (P.nee) {
                    // the diffuse lobe's share of direct light: pathtrace.rchit:93-110
                    // takes it when rand >= max(metallic, F(m)), m drawn from the same
                    // Xi as the cosine direction (same azimuth, cos^2 theta_m =
                    // (1 - c^2) / (c^2 (a^2 - 1) + 1) for a direction at cos theta = c)
                    const float ul = next_rand(rng) * P.total;
                    const float b1 = next_rand(rng), b2 = next_rand(rng);
                    const LightTri &lt = P.lights[pick_light(P.lights, P.nlights, ul)];
                    const float su = sqrtf(b1), ba = 1.0f - su, bb = su * (1.0f - b2), bc = su * b2;
                    const V3 dl = v3((lt.p0[0] * ba + lt.p1[0] * bb + lt.p2[0] * bc) - o.x,
                                     (lt.p0[1] * ba + lt.p1[1] * bb + lt.p2[1] * bc) - o.y,
                                     (lt.p0[2] * ba + lt.p1[2] * bb + lt.p2[2] * bc) - o.z);
                    const float dist2 = dot(dl, dl), dist = sqrtf(dist2);
                    const V3 wl = v3(dl.x / dist, dl.y / dist, dl.z / dist);
                    const float cos_x = dot(nO, wl);
                    const float cos_l = fabsf(dot(v3(lt.n[0], lt.n[1], lt.n[2]), wl));
                    if (dist2 > 0.0f && cos_x > 0.0f && cos_l > 0.0f) {
                        const float x0 = cos_x * cos_x;
                        const float cm =
                            sqrtf(saturate((1.0f - x0) / (x0 * (alphaSquared - 1.0f) + 1.0f)));
                        const float pd = 1.0f - fmaxf(metallic, fresnel_cos(cm, 1.0f / 1.5f));
                        if (pd > 0.0f) {
                            const float g = (pd * (cos_x * cos_l)) * P.total / (kPi * dist2);
                            o_.sc = make_float4((Tb.x * lt.e[0]) * g, (Tb.y * lt.e[1]) * g,
                                             (Tb.z * lt.e[2]) * g, 0.0f);
                            o_.sr.ox = o.x;
                            o_.sr.oy = o.y;
                            o_.sr.oz = o.z;
                            o_.sr.path = p;
                            o_.sr.dx = wl.x;
                            o_.sr.dy = wl.y;
                            o_.sr.dz = wl.z;
                            o_.sr.tmax = dist * 0.9999f;
                            o_.shadow = true;
                        }
                    }
                    // emitters met by the diffuse lobe were sampled above
                    if (diffuse) count_emit = false;
                }
                o_.cont = true;
                if (P.rr_start && depth >= P.rr_start) {
                    const float q = fminf(fmaxf(T.x, fmaxf(T.y, T.z)), 0.95f);
                    if (next_rand(rng) >= q) o_.cont = false;
                    else T = v3(T.x / q, T.y / q, T.z / q);
                }
                if (o_.cont) {
                    o_.nr.ox = o.x;
                    o_.nr.oy = o.y;
                    o_.nr.oz = o.z;
                    o_.nr.path = p;
                    o_.nr.dx = nd.x;
                    o_.nr.dy = nd.y;
                    o_.nr.dz = nd.z;
                    o_.nr.tmax = P.rmax;
                }
            }
        }
    }
    s0 = make_float4(T.x, T.y, T.z, as_f(rng));
    s1 = make_float4(L.x, L.y, L.z, as_f((depth & 0xFFFFu) | (count_emit ? 1u << 16 : 0u)));
}

}  // namespace solhip
