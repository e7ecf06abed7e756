// devmath.h -- the GLSL sampling / RNG / camera / hit arithmetic as device
// functions.
//
// Floating-point contract: this library is compiled with -ffp-contract=off,
// so every expression below is evaluated in fp32 exactly as written (left to
// right, one rounding per operation); division and sqrt are correctly rounded
// (HIP default).  sin/cos come from own_sincos (no ocml), so results do not
// depend on the device math library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef SOLHIP_HD
#define SOLHIP_HD __host__ __device__ __forceinline__
#endif

namespace solhip {

SOLHIP_HD float as_f(uint32_t u) { return __builtin_bit_cast(float, u); }
SOLHIP_HD uint32_t as_u(float f) { return __builtin_bit_cast(uint32_t, f); }

struct V3 {
    float x, y, z;
};
SOLHIP_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
SOLHIP_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SOLHIP_HD V3 scale(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
SOLHIP_HD V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
// GLSL normalize: v * (1 / sqrt(dot(v, v)))
SOLHIP_HD V3 normalize(V3 a)
{
    float inv = 1.0f / sqrtf(dot(a, a));
    return v3(a.x * inv, a.y * inv, a.z * inv);
}
SOLHIP_HD float saturate(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
SOLHIP_HD float sgn(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
// GLSL mix(x, y, a) = x * (1 - a) + y * a
SOLHIP_HD float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
// GLSL reflect(I, N) = I - 2 * dot(N, I) * N
SOLHIP_HD V3 reflect(V3 I, V3 N)
{
    float k = 2.0f * dot(N, I);
    return v3(I.x - k * N.x, I.y - k * N.y, I.z - k * N.z);
}
// column-major mat4 * vec4 row i: ((m[i]*x + m[4+i]*y) + m[8+i]*z) + m[12+i]*w
SOLHIP_HD void mat4_mul(const float *m, float x, float y, float z, float w, float r[4])
{
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = m[i] * x + m[4 + i] * y + m[8 + i] * z + m[12 + i] * w;
}

// ---- sampling.glsl -------------------------------------------------------
// assets/glsl/sampling.glsl:18-32
SOLHIP_HD uint32_t tea(uint32_t val0, uint32_t val1)
{
    uint32_t v0 = val0, v1 = val1, s0 = 0;
#pragma unroll
    for (uint32_t n = 0; n < 16; n++) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}

// assets/glsl/sampling.glsl:35-42; float(word) / 4294967295.0f == word * 2^-32
SOLHIP_HD float next_rand(uint32_t &rng)
{
    rng = rng * 747796405u + 1u;
    uint32_t word = ((rng >> ((rng >> 28) + 4u)) ^ rng) * 277803737u;
    word = (word >> 22) ^ word;
    return (float)word * 0x1p-32f;
}

// n LCG steps at once (arbitrary-stride jump of the same generator)
SOLHIP_HD uint32_t pcg_advance(uint32_t rng, uint32_t steps)
{
    uint32_t am = 1u, ap = 0u, cm = 747796405u, cp = 1u;
    while (steps) {
        if (steps & 1u) {
            am *= cm;
            ap = ap * cm + cp;
        }
        cp = (cm + 1u) * cp;
        cm *= cm;
        steps >>= 1;
    }
    return am * rng + ap;
}

// sin/cos: Cody-Waite reduction by pi/2 (3-part constant), Taylor polynomials
// on [-pi/4, pi/4]; the same operation sequence as the host restatement.
SOLHIP_HD void own_sincos(float x, float &s, float &c)
{
    float k = floorf(x * 0x1.45f306p-1f + 0.5f);
    float r = x - k * 0x1.92p0f;
    r = r - k * 0x1.fb4p-12f;
    r = r - k * 0x1.4442d2p-24f;
    float r2 = r * r;
    float sp = -0x1.a01a02p-13f + r2 * 0x1.71de3ap-19f;
    sp = 0x1.111112p-7f + r2 * sp;
    sp = -0x1.555556p-3f + r2 * sp;
    float sn = r + (r * r2) * sp;
    float cpp = 0x1.a01a02p-16f + r2 * -0x1.27e4fcp-22f;
    cpp = -0x1.6c16c2p-10f + r2 * cpp;
    cpp = 0x1.555556p-5f + r2 * cpp;
    cpp = -0.5f + r2 * cpp;
    float cs = 1.0f + r2 * cpp;
    int q = ((int)k) & 3;
    float s0 = (q & 1) ? cs : sn;
    float c0 = (q & 1) ? sn : cs;
    s = (q & 2) ? -s0 : s0;
    c = ((q + 1) & 2) ? -c0 : c0;
}

// assets/glsl/sampling.glsl:52-64, as a function of cos(theta_i)
SOLHIP_HD float fresnel_cos(float cosThetaI, float eta)
{
    float result = 1.0f;
    float sinThetaOSquared = (eta * eta) * (1.0f - cosThetaI * cosThetaI);
    if (sinThetaOSquared <= 1.0f) {
        float cosThetaO = sqrtf(saturate(1.0f - sinThetaOSquared));
        float Rs = (cosThetaI - eta * cosThetaO) / (cosThetaI + eta * cosThetaO);
        float Rp = (eta * cosThetaI - cosThetaO) / (eta * cosThetaI + cosThetaO);
        result = 0.5f * (Rs * Rs + Rp * Rp);
    }
    return result;
}

// assets/glsl/sampling.glsl:52-64
SOLHIP_HD float fresnel_dielectric(V3 i, V3 m, float eta) { return fresnel_cos(fabsf(dot(i, m)), eta); }

// assets/glsl/sampling.glsl:66-84 (basis of Duff et al. 2017)
SOLHIP_HD V3 align_to_direction(V3 n, float cosTheta, float phi)
{
    float sinTheta = sqrtf(saturate(1.0f - cosTheta * cosTheta));
    float s = (n.z < 0.0f ? -1.0f : 1.0f);
    float a = -1.0f / (s + n.z);
    float b = n.x * n.y * a;
    V3 u = v3(1.0f + s * n.x * n.x * a, s * b, -s * n.x);
    V3 v = v3(b, s + n.y * n.y * a, -n.y);
    float sp, cp;
    own_sincos(phi, sp, cp);
    return v3((u.x * cp + v.x * sp) * sinTheta + n.x * cosTheta,
              (u.y * cp + v.y * sp) * sinTheta + n.y * cosTheta,
              (u.z * cp + v.z * sp) * sinTheta + n.z * cosTheta);
}

// assets/glsl/sampling.glsl:86-90 (TWO_PI literal rounds to 0x1.921fb6p+2f)
SOLHIP_HD V3 sample_ggx(V3 n, float xi0, float xi1, float alphaSquared)
{
    float cosTheta = sqrtf(saturate((1.0f - xi0) / (xi0 * (alphaSquared - 1.0f) + 1.0f)));
    return align_to_direction(n, cosTheta, xi1 * 0x1.921fb6p+2f);
}

// assets/glsl/sampling.glsl:92-96
SOLHIP_HD V3 sample_cosine(V3 n, float xi0, float xi1)
{
    float cosTheta = sqrtf(xi0);
    return align_to_direction(n, cosTheta, xi1 * 0x1.921fb6p+2f);
}

// ---- camera: ao.rgen:49-54, pathtrace.rgen:52-58 ---------------------------
SOLHIP_HD void camera_ray(const float *view_inverse, const float *projection_inverse,
                                           uint32_t x, uint32_t y, uint32_t W, uint32_t H, float jx, float jy,
                                           V3 &origin, V3 &dir)
{
    float pcx = (float)x + jx, pcy = (float)y + jy;
    float uvx = pcx / (float)W, uvy = pcy / (float)H;
    float dx = uvx * 2.0f - 1.0f, dy = uvy * 2.0f - 1.0f;
    float o4[4], t4[4], d4[4];
    mat4_mul(view_inverse, 0.0f, 0.0f, 0.0f, 1.0f, o4);
    mat4_mul(projection_inverse, dx, dy, 1.0f, 1.0f, t4);
    V3 tn = normalize(v3(t4[0], t4[1], t4[2]));
    mat4_mul(view_inverse, tn.x, tn.y, tn.z, 0.0f, d4);
    origin = v3(o4[0], o4[1], o4[2]);
    dir = v3(d4[0], d4[1], d4[2]);
}

// assets/glsl/postprocess.glsl:38-41 gammaCorrect, then an RGBA8 UNORM store
SOLHIP_HD uint32_t to_unorm8(float c)
{
    float v = c;
    if (!(v > 0.0f)) v = 0.0f;
    if (v > 1.0f) v = 1.0f;
    return (uint32_t)floorf(v * 255.0f + 0.5f);
}

}  // namespace solhip
