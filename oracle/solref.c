/*
 * solref.c -- CPU oracle for the sol-rs hardware ray-tracing path.
 *
 * TEST INFRASTRUCTURE ONLY (see solref.h).  Built with gcc -ffp-contract=off;
 * the shipped HIP library never links or loads it.
 *
 * What pins each part of this restatement (checked in tests/):
 *   sr_tea ..................... TEA (Wheeler and Needham 1994) with 16 cycles
 *                                and the shader's key: a cipher pinned by the
 *                                published 32-cycle zero-key vector
 *                                (41ea3a0a 94baa940) gives the known answers.
 *   sr_next_rand ............... golden streams written by an independent
 *                                numpy-uint32 restatement
 *                                (tests/golden/make_golden.py) and the PCG
 *                                RXS-M-XS 32/32 output function (O'Neill 2014);
 *                                float(word)/4294967295.0f == word * 2^-32.
 *   sr_pcg_advance ............. equality with repeated single steps.
 *   sr_sinf, sr_cosf ........... libm double sin/cos within 2 ulp on [0, 2pi].
 *   sampling ................... closed forms: unit length, E[cos]=2/3 for the
 *                                cosine lobe, GGX alpha->0 gives n, ONB is
 *                                orthonormal, Fresnel(n,n,1/1.5)=0.04.
 *   sr_camera_ray .............. centre pixel ray == view axis (closed form
 *                                look_at/perspective).
 *   sr_intersect_triangle ...... exact hits on axis-aligned quads, edge
 *                                watertightness, t = distance on known cases.
 *   sr_trace(use_bvh=1) ........ equality with brute force (use_bvh=0).
 *   sr_render_ao ............... an open plane gives exactly 1; a floor
 *                                under a parallel ceiling gives occlusion
 *                                1 - (h/R)^2.
 *   sr_render_ao_glsl .......... an open plane gives exactly 1 (the chain
 *                                adds from its second hit on); between two
 *                                parallel planes the image mean is
 *                                1 - E[K]/4 with P(K >= j) = p^j, p the
 *                                share of blue-noise draws whose cosine
 *                                direction reaches the other plane.
 *   sr_render_pathtrace_glsl ... sky gradient at zenith / nadir; emitters
 *                                seen directly; a roughness-0 metal plane
 *                                gives base * sky(reflect(d, n)) per pixel;
 *                                a roughness-0 black dielectric reflects
 *                                with probability F = 0.04.
 *   sr_render_pathtrace ........ bit-identical to sr_render_pathtrace_glsl
 *                                without extensions; NEE / roulette /
 *                                per-sample streams keep the expectation.
 *   sr_render_debug ............ barycentrics of the ray/plane point solved
 *                                in float64; misses return the camera ray;
 *                                nothing beyond the range 1000 is seen.
 */
#include "solref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* small vector helpers (left-to-right fp32, no contraction)                */
/* ------------------------------------------------------------------------ */
static inline float dot3(const float a[3], const float b[3])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void normalize3(const float a[3], float r[3])
{
    float inv = 1.0f / sqrtf(dot3(a, a));
    r[0] = a[0] * inv; r[1] = a[1] * inv; r[2] = a[2] * inv;
}
static inline float saturate(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
static inline float signf_(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
/* GLSL mix(x, y, a) = x * (1 - a) + y * a */
static inline float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
/* GLSL reflect(I, N) = I - 2 * dot(N, I) * N */
static inline void reflect3(const float I[3], const float N[3], float r[3])
{
    float k = 2.0f * dot3(N, I);
    r[0] = I[0] - k * N[0]; r[1] = I[1] - k * N[1]; r[2] = I[2] - k * N[2];
}
/* column-major mat4 * vec4, row i = ((m0i*x + m1i*y) + m2i*z) + m3i*w */
static inline void mat4_mul_vec4(const float m[16], const float v[4], float r[4])
{
    for (int i = 0; i < 4; ++i)
        r[i] = m[0 + i] * v[0] + m[4 + i] * v[1] + m[8 + i] * v[2] + m[12 + i] * v[3];
}

/* ------------------------------------------------------------------------ */
/* sampling.glsl                                                            */
/* ------------------------------------------------------------------------ */

/* assets/glsl/sampling.glsl:18-32 (TEA, 16 rounds) */
uint32_t sr_tea(uint32_t val0, uint32_t val1)
{
    uint32_t v0 = val0, v1 = val1, s0 = 0;
    for (uint32_t n = 0; n < 16; n++) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}

/* assets/glsl/sampling.glsl:35-42.  4294967295.0f rounds to 2^32 in fp32, so
 * the division is the exact scaling word * 2^-32 of the RNE-converted word. */
float sr_next_rand(uint32_t *rng)
{
    *rng = *rng * 747796405u + 1u;
    uint32_t word = ((*rng >> ((*rng >> 28) + 4u)) ^ *rng) * 277803737u;
    word = (word >> 22) ^ word;
    return (float)word * 0x1p-32f;
}

/* n steps of the LCG in sr_next_rand in O(log n) (Brown 1994, "Random
 * number generation with arbitrary strides"). */
uint32_t sr_pcg_advance(uint32_t rng, uint32_t steps)
{
    uint32_t acc_mult = 1u, acc_plus = 0u, cur_mult = 747796405u, cur_plus = 1u;
    while (steps) {
        if (steps & 1u) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1u) * cur_plus;
        cur_mult *= cur_mult;
        steps >>= 1;
    }
    return acc_mult * rng + acc_plus;
}

/* sin/cos used for phi = 2*pi*xi in [0, 2*pi].  Own implementation so that
 * CPU and GPU agree bit for bit: Cody-Waite reduction by pi/2 with a 3-part
 * constant (hi, mid have few enough bits that k*hi, k*mid are exact for
 * |k| < 2^11), then Taylor polynomials on [-pi/4, pi/4] (truncation error
 * below 2^-27).  Accurate to ~1 ulp over the sampling range. */
static void sr_sincos(float x, float *s, float *c)
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
    float cp = 0x1.a01a02p-16f + r2 * -0x1.27e4fcp-22f;
    cp = -0x1.6c16c2p-10f + r2 * cp;
    cp = 0x1.555556p-5f + r2 * cp;
    cp = -0.5f + r2 * cp;
    float cs = 1.0f + r2 * cp;
    int q = ((int)k) & 3;
    switch (q) {
    case 0: *s = sn; *c = cs; break;
    case 1: *s = cs; *c = -sn; break;
    case 2: *s = -sn; *c = -cs; break;
    default: *s = -cs; *c = sn; break;
    }
}
float sr_sinf(float x) { float s, c; sr_sincos(x, &s, &c); return s; }
float sr_cosf(float x) { float s, c; sr_sincos(x, &s, &c); return c; }

/* assets/glsl/sampling.glsl:52-64, as a function of cos(theta_i) */
static float fresnel_cos(float cosThetaI, float eta)
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

/* assets/glsl/sampling.glsl:52-64 */
float sr_fresnel_dielectric(const float i[3], const float m[3], float eta)
{
    return fresnel_cos(fabsf(dot3(i, m)), eta);
}

/* assets/glsl/sampling.glsl:66-73 (Duff et al. 2017 basis) */
void sr_onb(const float n[3], float u[3], float v[3])
{
    float s = (n[2] < 0.0f ? -1.0f : 1.0f);
    float a = -1.0f / (s + n[2]);
    float b = n[0] * n[1] * a;
    u[0] = 1.0f + s * n[0] * n[0] * a; u[1] = s * b; u[2] = -s * n[0];
    v[0] = b; v[1] = s + n[1] * n[1] * a; v[2] = -n[1];
}

/* assets/glsl/sampling.glsl:75-84 */
void sr_align_to_direction(const float n[3], float cosTheta, float phi, float out[3])
{
    float sinTheta = sqrtf(saturate(1.0f - cosTheta * cosTheta));
    float u[3], v[3];
    sr_onb(n, u, v);
    float sp, cp;
    sr_sincos(phi, &sp, &cp);
    for (int k = 0; k < 3; ++k)
        out[k] = (u[k] * cp + v[k] * sp) * sinTheta + n[k] * cosTheta;
}

/* assets/glsl/sampling.glsl:86-90; TWO_PI = 6.28318530718 -> 0x1.921fb6p+2f */
void sr_sample_ggx(const float n[3], const float xi[2], float alphaSquared, float out[3])
{
    float cosTheta = sqrtf(saturate((1.0f - xi[0]) / (xi[0] * (alphaSquared - 1.0f) + 1.0f)));
    sr_align_to_direction(n, cosTheta, xi[1] * 0x1.921fb6p+2f, out);
}

/* assets/glsl/sampling.glsl:92-96 */
void sr_sample_cosine(const float n[3], const float xi[2], float out[3])
{
    float cosTheta = sqrtf(xi[0]);
    sr_align_to_direction(n, cosTheta, xi[1] * 0x1.921fb6p+2f, out);
}

/* assets/glsl/postprocess.glsl:38-41 gammaCorrect + RGBA8 UNORM store */
static inline uint8_t to_unorm8(float c)
{
    float v = c;
    if (!(v > 0.0f)) v = 0.0f;
    if (v > 1.0f) v = 1.0f;
    return (uint8_t)floorf(v * 255.0f + 0.5f);
}

/* ------------------------------------------------------------------------ */
/* camera (ao.rgen:49-54, pathtrace.rgen:52-58)                             */
/* ------------------------------------------------------------------------ */
void sr_camera_ray(const sr_uniforms *un, uint32_t x, uint32_t y, uint32_t W, uint32_t H,
                   float jx, float jy, float origin[3], float dir[3])
{
    float pcx = (float)x + jx, pcy = (float)y + jy;
    float uvx = pcx / (float)W, uvy = pcy / (float)H;
    float dx = uvx * 2.0f - 1.0f, dy = uvy * 2.0f - 1.0f;
    float o4[4], t4[4], d4[4];
    const float zero_w1[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    mat4_mul_vec4(un->view_inverse, zero_w1, o4);
    const float tv[4] = {dx, dy, 1.0f, 1.0f};
    mat4_mul_vec4(un->projection_inverse, tv, t4);
    float tn[3];
    normalize3(t4, tn);
    const float dv[4] = {tn[0], tn[1], tn[2], 0.0f};
    mat4_mul_vec4(un->view_inverse, dv, d4);
    origin[0] = o4[0]; origin[1] = o4[1]; origin[2] = o4[2];
    dir[0] = d4[0]; dir[1] = d4[1]; dir[2] = d4[2];
}

/* ------------------------------------------------------------------------ */
/* watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013)            */
/* ------------------------------------------------------------------------ */
typedef struct { int kx, ky, kz; float sx, sy, sz; } wt_ray;

static inline void wt_setup(const float d[3], wt_ray *w)
{
    float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
    int kz = 0;
    float m = ax;
    if (ay > m) { kz = 1; m = ay; }
    if (az > m) { kz = 2; }
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (d[kz] < 0.0f) { int t = kx; kx = ky; ky = t; }
    w->kx = kx; w->ky = ky; w->kz = kz;
    w->sx = d[kx] / d[kz];
    w->sy = d[ky] / d[kz];
    w->sz = 1.0f / d[kz];
}

static inline int wt_hit(const wt_ray *w, const float o[3], float tmin, float tmax,
                         const float v0[3], const float v1[3], const float v2[3],
                         float *t_out, float *u_out, float *v_out)
{
    const int kx = w->kx, ky = w->ky, kz = w->kz;
    float A_x = v0[kx] - o[kx], A_y = v0[ky] - o[ky], A_z = v0[kz] - o[kz];
    float B_x = v1[kx] - o[kx], B_y = v1[ky] - o[ky], B_z = v1[kz] - o[kz];
    float C_x = v2[kx] - o[kx], C_y = v2[ky] - o[ky], C_z = v2[kz] - o[kz];
    float ax = A_x - w->sx * A_z, ay = A_y - w->sy * A_z;
    float bx = B_x - w->sx * B_z, by = B_y - w->sy * B_z;
    float cx = C_x - w->sx * C_z, cy = C_y - w->sy * C_z;
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
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f))
        return 0;
    float det = U + V + W;
    if (det == 0.0f)
        return 0;
    float az = w->sz * A_z, bz = w->sz * B_z, cz = w->sz * C_z;
    float T = U * az + V * bz + W * cz;
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t >= tmin && t <= tmax))
        return 0;
    *t_out = t;
    *u_out = V * rcp;
    *v_out = W * rcp;
    return 1;
}

int sr_intersect_triangle(const float o[3], const float d[3], float tmin, float tmax,
                          const float v0[3], const float v1[3], const float v2[3],
                          float *t, float *u, float *v)
{
    wt_ray w;
    wt_setup(d, &w);
    return wt_hit(&w, o, tmin, tmax, v0, v1, v2, t, u, v);
}

/* world-to-object inverse of a row-major 3x4 affine transform (adjugate /
 * determinant; the HIP TLAS build evaluates the same expression) */
void sr_instance_inverse(const float m[12], float r[12])
{
    float a = m[0], b = m[1], c = m[2], tx = m[3];
    float d = m[4], e = m[5], f = m[6], ty = m[7];
    float g = m[8], h = m[9], i = m[10], tz = m[11];
    float c00 = e * i - f * h, c01 = c * h - b * i, c02 = b * f - c * e;
    float c10 = f * g - d * i, c11 = a * i - c * g, c12 = c * d - a * f;
    float c20 = d * h - e * g, c21 = b * g - a * h, c22 = a * e - b * d;
    float det = a * c00 + b * c10 + c * c20;
    float rd = 1.0f / det;
    r[0] = c00 * rd; r[1] = c01 * rd; r[2] = c02 * rd;
    r[4] = c10 * rd; r[5] = c11 * rd; r[6] = c12 * rd;
    r[8] = c20 * rd; r[9] = c21 * rd; r[10] = c22 * rd;
    r[3] = -(r[0] * tx + r[1] * ty + r[2] * tz);
    r[7] = -(r[4] * tx + r[5] * ty + r[6] * tz);
    r[11] = -(r[8] * tx + r[9] * ty + r[10] * tz);
}

static inline void xform_point(const float m[12], const float p[3], float r[3])
{
    r[0] = m[0] * p[0] + m[1] * p[1] + m[2] * p[2] + m[3];
    r[1] = m[4] * p[0] + m[5] * p[1] + m[6] * p[2] + m[7];
    r[2] = m[8] * p[0] + m[9] * p[1] + m[10] * p[2] + m[11];
}
static inline void xform_dir(const float m[12], const float p[3], float r[3])
{
    r[0] = m[0] * p[0] + m[1] * p[1] + m[2] * p[2];
    r[1] = m[4] * p[0] + m[5] * p[1] + m[6] * p[2];
    r[2] = m[8] * p[0] + m[9] * p[1] + m[10] * p[2];
}

/* ------------------------------------------------------------------------ */
/* scene + acceleration (oracle's own binary BVH; exactness is by           */
/* conservative culling: it must return exactly the brute-force answer)     */
/* ------------------------------------------------------------------------ */
typedef struct { float lo[3], hi[3]; uint32_t first, count; uint32_t left, right; } bnode;

typedef struct {
    uint32_t ntri;
    float *tri;          /* 9 floats per triangle, BLAS-wide primitive order */
    bnode *nodes;
    uint32_t nnodes;
    uint32_t *order;     /* leaf ranges index into this permutation */
    uint32_t ngeo;
    uint32_t *geo_begin; /* first BLAS-wide triangle of each geometry */
} blas_t;

struct sr_scene {
    uint32_t nblas, ninst;
    blas_t *blas;
    sr_instance_desc *inst;
    float (*inv)[12];
    float (*wlo)[3], (*whi)[3];    /* world boxes of instances */
    bnode *tnodes;                 /* TLAS over instances */
    uint32_t ntnodes;
    uint32_t *torder;
};

static uint32_t expand_bits10(uint32_t v)
{
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

typedef struct { uint64_t key; } mkey;
static int cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* build a binary BVH over n boxes (lo/hi per item); leaf size <= leafmax */
typedef struct {
    const float (*lo)[3];
    const float (*hi)[3];
    uint64_t *keys;
    bnode *nodes;
    uint32_t nnodes;
    uint32_t *order;
    uint32_t leafmax;
} bbuild;

static uint32_t bb_rec(bbuild *b, uint32_t first, uint32_t count)
{
    uint32_t idx = b->nnodes++;
    bnode *nd = &b->nodes[idx];
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t k = first; k < first + count; ++k) {
        uint32_t p = (uint32_t)(b->keys[k] & 0xFFFFFFFFu);
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], b->lo[p][a]);
            hi[a] = fmaxf(hi[a], b->hi[p][a]);
        }
    }
    /* pad the box outward (relative + absolute) so that float rounding in the
     * slab test can never cull a box whose content the triangle test hits */
    for (int a = 0; a < 3; ++a) {
        float ext = hi[a] - lo[a];
        float mag = fmaxf(fabsf(lo[a]), fabsf(hi[a]));
        float pad = ext * 1e-5f + mag * 1e-6f + 1e-30f;
        lo[a] -= pad; hi[a] += pad;
        nd->lo[a] = lo[a]; nd->hi[a] = hi[a];
    }
    if (count <= b->leafmax) {
        nd->first = first; nd->count = count; nd->left = nd->right = 0;
        return idx;
    }
    /* split at the highest differing Morton bit, else in the middle */
    uint32_t ka = (uint32_t)(b->keys[first] >> 32), kb = (uint32_t)(b->keys[first + count - 1] >> 32);
    uint32_t split = first + count / 2;
    if (ka != kb) {
        int bit = 31 - __builtin_clz(ka ^ kb);
        uint32_t lo_i = first, hi_i = first + count - 1;
        while (lo_i + 1 < hi_i) {
            uint32_t mid = (lo_i + hi_i) / 2;
            if (((uint32_t)(b->keys[mid] >> 32) >> bit) & 1u) hi_i = mid; else lo_i = mid;
        }
        split = hi_i;
    }
    nd->first = 0; nd->count = 0;
    uint32_t l = bb_rec(b, first, split - first);
    uint32_t r = bb_rec(b, split, first + count - split);
    b->nodes[idx].left = l;
    b->nodes[idx].right = r;
    return idx;
}

static void build_bvh(const float (*lo)[3], const float (*hi)[3], uint32_t n, uint32_t leafmax,
                      bnode **nodes_out, uint32_t *nnodes_out, uint32_t **order_out)
{
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            float c = 0.5f * (lo[i][a] + hi[i][a]);
            clo[a] = fminf(clo[a], c); chi[a] = fmaxf(chi[a], c);
        }
    uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t code = 0;
        for (int a = 0; a < 3; ++a) {
            float ext = chi[a] - clo[a];
            float c = 0.5f * (lo[i][a] + hi[i][a]);
            float q = ext > 0.0f ? (c - clo[a]) / ext * 1023.0f : 0.0f;
            uint32_t qi = q <= 0.0f ? 0u : (q >= 1023.0f ? 1023u : (uint32_t)q);
            code |= expand_bits10(qi) << (2 - a);
        }
        keys[i] = ((uint64_t)code << 32) | i;
    }
    qsort(keys, n, sizeof(uint64_t), cmp_u64);
    bbuild b;
    b.lo = lo; b.hi = hi; b.keys = keys; b.leafmax = leafmax;
    b.nodes = (bnode *)calloc(2 * (size_t)(n ? n : 1), sizeof(bnode));
    b.nnodes = 0;
    if (n) bb_rec(&b, 0, n);
    uint32_t *order = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; ++i) order[i] = (uint32_t)(keys[i] & 0xFFFFFFFFu);
    free(keys);
    *nodes_out = b.nodes; *nnodes_out = b.nnodes; *order_out = order;
}

static uint32_t geo_tris(const sr_geometry *g)
{
    return g->indices ? g->index_count / 3 : g->vertex_count / 3;
}

sr_scene *sr_scene_create(const sr_geometry *g, uint32_t nblas,
                          const sr_instance_desc *inst, uint32_t ninst)
{
    uint32_t *ones = (uint32_t *)malloc(sizeof(uint32_t) * (nblas ? nblas : 1));
    for (uint32_t b = 0; b < nblas; ++b) ones[b] = 1;
    sr_scene *s = sr_scene_create_multi(g, ones, nblas, inst, ninst);
    free(ones);
    return s;
}

sr_scene *sr_scene_create_multi(const sr_geometry *geos, const uint32_t *geo_count, uint32_t nblas,
                                const sr_instance_desc *inst, uint32_t ninst)
{
    sr_scene *s = (sr_scene *)calloc(1, sizeof(sr_scene));
    s->nblas = nblas; s->ninst = ninst;
    s->blas = (blas_t *)calloc(nblas ? nblas : 1, sizeof(blas_t));
    const sr_geometry *g = geos;
    for (uint32_t b = 0; b < nblas; ++b) {
        blas_t *bl = &s->blas[b];
        const uint32_t ng = geo_count[b];
        uint32_t nt = 0;
        bl->ngeo = ng;
        bl->geo_begin = (uint32_t *)malloc(sizeof(uint32_t) * (ng ? ng : 1));
        for (uint32_t k = 0; k < ng; ++k) {
            bl->geo_begin[k] = nt;
            nt += geo_tris(&g[k]);
        }
        bl->ntri = nt;
        bl->tri = (float *)malloc(sizeof(float) * 9 * (nt ? nt : 1));
        float (*lo)[3] = malloc(sizeof(float) * 3 * (nt ? nt : 1));
        float (*hi)[3] = malloc(sizeof(float) * 3 * (nt ? nt : 1));
        uint32_t gk = 0;
        for (uint32_t t = 0; t < nt; ++t) {
            while (gk + 1 < ng && bl->geo_begin[gk + 1] <= t) ++gk;
            const sr_geometry *G = &g[gk];
            const uint32_t lt = t - bl->geo_begin[gk];
            for (int k = 0; k < 3; ++k) {
                uint32_t vi = G->indices ? G->indices[3 * lt + k] : 3 * lt + (uint32_t)k;
                const float *p = (const float *)((const char *)G->vertices + (uint64_t)vi * G->vertex_stride);
                bl->tri[9 * t + 3 * k + 0] = p[0];
                bl->tri[9 * t + 3 * k + 1] = p[1];
                bl->tri[9 * t + 3 * k + 2] = p[2];
            }
            for (int a = 0; a < 3; ++a) {
                float x0 = bl->tri[9 * t + a], x1 = bl->tri[9 * t + 3 + a], x2 = bl->tri[9 * t + 6 + a];
                lo[t][a] = fminf(fminf(x0, x1), x2);
                hi[t][a] = fmaxf(fmaxf(x0, x1), x2);
            }
        }
        build_bvh((const float (*)[3])lo, (const float (*)[3])hi, nt, 4, &bl->nodes, &bl->nnodes, &bl->order);
        free(lo); free(hi);
        g += ng;
    }
    s->inst = (sr_instance_desc *)malloc(sizeof(sr_instance_desc) * (ninst ? ninst : 1));
    memcpy(s->inst, inst, sizeof(sr_instance_desc) * ninst);
    s->inv = malloc(sizeof(float) * 12 * (ninst ? ninst : 1));
    s->wlo = malloc(sizeof(float) * 3 * (ninst ? ninst : 1));
    s->whi = malloc(sizeof(float) * 3 * (ninst ? ninst : 1));
    for (uint32_t i = 0; i < ninst; ++i) {
        sr_instance_inverse(inst[i].transform, s->inv[i]);
        const blas_t *bl = &s->blas[inst[i].blas];
        for (int a = 0; a < 3; ++a) { s->wlo[i][a] = INFINITY; s->whi[i][a] = -INFINITY; }
        if (bl->nnodes == 0) {
            for (int a = 0; a < 3; ++a) { s->wlo[i][a] = s->whi[i][a] = inst[i].transform[4 * a + 3]; }
            continue;
        }
        const bnode *r = &bl->nodes[0];
        for (int c = 0; c < 8; ++c) {
            float p[3] = {(c & 1) ? r->hi[0] : r->lo[0], (c & 2) ? r->hi[1] : r->lo[1], (c & 4) ? r->hi[2] : r->lo[2]};
            float w[3];
            xform_point(inst[i].transform, p, w);
            for (int a = 0; a < 3; ++a) {
                s->wlo[i][a] = fminf(s->wlo[i][a], w[a]);
                s->whi[i][a] = fmaxf(s->whi[i][a], w[a]);
            }
        }
    }
    build_bvh((const float (*)[3])s->wlo, (const float (*)[3])s->whi, ninst, 1, &s->tnodes, &s->ntnodes, &s->torder);
    return s;
}

void sr_scene_destroy(sr_scene *s)
{
    if (!s) return;
    for (uint32_t b = 0; b < s->nblas; ++b) {
        free(s->blas[b].tri); free(s->blas[b].nodes); free(s->blas[b].order); free(s->blas[b].geo_begin);
    }
    free(s->blas); free(s->inst); free(s->inv); free(s->wlo); free(s->whi);
    free(s->tnodes); free(s->torder);
    free(s);
}

/* (t, inst, prim) lexicographic order with prim the BLAS-wide triangle
 * index, i.e. (t, inst, geometry, triangle in geometry): the documented tie
 * rule */
static inline int better_hit(float t, uint32_t inst, uint32_t prim, const sr_hit *h)
{
    if (t < h->t) return 1;
    if (t > h->t) return 0;
    if (inst != h->inst) return inst < h->inst;
    return prim < h->prim;
}

typedef struct {
    float o[3], d[3];
    double id[3];
    wt_ray w;
} oray;

static inline void oray_setup(oray *r, const float o[3], const float d[3])
{
    for (int a = 0; a < 3; ++a) {
        r->o[a] = o[a]; r->d[a] = d[a];
        double dd = d[a];
        if (fabs(dd) < 1e-30) dd = dd < 0 ? -1e-30 : 1e-30;
        r->id[a] = 1.0 / dd;
    }
    wt_setup(d, &r->w);
}

/* slab test in double on padded boxes, with a relative allowance on t */
static inline int box_hit(const oray *r, const float lo[3], const float hi[3], float tmin, float tmax, double *tnear)
{
    double tn = -INFINITY, tf = INFINITY;
    for (int a = 0; a < 3; ++a) {
        double inv = r->id[a];
        double t0 = ((double)lo[a] - r->o[a]) * inv;
        double t1 = ((double)hi[a] - r->o[a]) * inv;
        if (t0 > t1) { double x = t0; t0 = t1; t1 = x; }
        if (t0 > tn) tn = t0;
        if (t1 < tf) tf = t1;
    }
    double lo_t = (double)tmin - fabs((double)tmin) * 1e-5;
    double hi_t = (double)tmax + fabs((double)tmax) * 1e-5;
    if (tn < lo_t) tn = lo_t;
    if (tf > hi_t) tf = hi_t;
    *tnear = tn;
    return tn <= tf;
}

static void blas_closest(const blas_t *bl, const oray *r, float tmin, uint32_t inst, sr_hit *h, int any)
{
    if (bl->nnodes == 0) return;
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const bnode *n = &bl->nodes[stack[--sp]];
        double tn;
        if (!box_hit(r, n->lo, n->hi, tmin, h->t, &tn)) continue;
        if (n->count) {
            for (uint32_t k = n->first; k < n->first + n->count; ++k) {
                uint32_t p = bl->order[k];
                const float *T = &bl->tri[9 * p];
                float t, u, v;
                if (wt_hit(&r->w, r->o, tmin, h->t, T, T + 3, T + 6, &t, &u, &v) &&
                    better_hit(t, inst, p, h)) {
                    h->t = t; h->u = u; h->v = v; h->inst = inst; h->prim = p;
                    if (any) return;
                }
            }
            continue;
        }
        double tl, tr;
        int hl = box_hit(r, bl->nodes[n->left].lo, bl->nodes[n->left].hi, tmin, h->t, &tl);
        int hr = box_hit(r, bl->nodes[n->right].lo, bl->nodes[n->right].hi, tmin, h->t, &tr);
        if (hl && hr) {
            if (tl <= tr) { stack[sp++] = n->right; stack[sp++] = n->left; }
            else { stack[sp++] = n->left; stack[sp++] = n->right; }
        } else if (hl) stack[sp++] = n->left;
        else if (hr) stack[sp++] = n->right;
    }
}

static void trace_one(const sr_scene *s, const sr_ray *ray, sr_hit *h, int use_bvh, uint32_t cull, int any)
{
    h->t = ray->tmax; h->u = 0.0f; h->v = 0.0f; h->inst = SR_MISS; h->prim = SR_MISS; h->geom = 0;
    const float o[3] = {ray->ox, ray->oy, ray->oz}, d[3] = {ray->dx, ray->dy, ray->dz};
    if (!use_bvh) {
        for (uint32_t i = 0; i < s->ninst; ++i) {
            if (((s->inst[i].id_mask >> 24) & cull) == 0) continue;
            const blas_t *bl = &s->blas[s->inst[i].blas];
            oray r;
            float oo[3], od[3];
            xform_point(s->inv[i], o, oo);
            xform_dir(s->inv[i], d, od);
            oray_setup(&r, oo, od);
            for (uint32_t p = 0; p < bl->ntri; ++p) {
                const float *T = &bl->tri[9 * p];
                float t, u, v;
                if (wt_hit(&r.w, r.o, ray->tmin, h->t, T, T + 3, T + 6, &t, &u, &v) &&
                    better_hit(t, i, p, h)) {
                    h->t = t; h->u = u; h->v = v; h->inst = i; h->prim = p;
                }
            }
        }
    } else if (s->ntnodes) {
        oray wr;
        oray_setup(&wr, o, d);
        uint32_t stack[128];
        int sp = 0;
        stack[sp++] = 0;
        while (sp) {
            const bnode *n = &s->tnodes[stack[--sp]];
            double tn;
            if (!box_hit(&wr, n->lo, n->hi, ray->tmin, h->t, &tn)) continue;
            if (n->count) {
                for (uint32_t k = n->first; k < n->first + n->count; ++k) {
                    uint32_t i = s->torder[k];
                    if (((s->inst[i].id_mask >> 24) & cull) == 0) continue;
                    oray r;
                    float oo[3], od[3];
                    xform_point(s->inv[i], o, oo);
                    xform_dir(s->inv[i], d, od);
                    oray_setup(&r, oo, od);
                    blas_closest(&s->blas[s->inst[i].blas], &r, ray->tmin, i, h, any);
                    if (any && h->inst != SR_MISS) goto done;
                }
                continue;
            }
            stack[sp++] = n->right;
            stack[sp++] = n->left;
        }
    }
done:
    if (h->inst == SR_MISS) { h->t = -1.0f; h->u = 0.0f; h->v = 0.0f; }
}

/* a renderer-internal hit (BLAS-wide prim, geom 0) as a public record:
 * (geometry, triangle in geometry), geom = SR_MISS on a miss */
static sr_hit public_hit(const sr_scene *s, sr_hit h)
{
    if (h.inst == SR_MISS) { h.geom = SR_MISS; return h; }
    const blas_t *bl = &s->blas[s->inst[h.inst].blas];
    uint32_t gk = 0;
    while (gk + 1 < bl->ngeo && bl->geo_begin[gk + 1] <= h.prim) ++gk;
    h.geom = gk;
    h.prim -= bl->geo_begin[gk];
    return h;
}

void sr_trace(const sr_scene *s, const sr_ray *rays, sr_hit *hits, uint64_t n,
              int use_bvh, uint32_t cull_mask, int nthreads)
{
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
#endif
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        trace_one(s, &rays[i], &hits[i], use_bvh, cull_mask, 0);
        hits[i] = public_hit(s, hits[i]);
    }
}

/* ------------------------------------------------------------------------ */
/* hit-shader helpers                                                       */
/* ------------------------------------------------------------------------ */
static inline uint32_t fetch_index(const sr_tables *tb, uint32_t obj, uint32_t k)
{
    if (tb->index_size == 8) return (uint32_t)((const uint64_t *)tb->indices[obj])[k];
    return ((const uint32_t *)tb->indices[obj])[k];
}

/* world-space position / normal / vertex colour at a hit; shared by
 * ao.rchit:56-77 and pathtrace.rchit:59-88 */
typedef struct { float pos[3], nrm[3], col[3]; } surf_t;
static void fetch_surface(const sr_tables *tb, const sr_hit *h, surf_t *s)
{
    const sr_scene_instance *si = &tb->instances[h->inst];
    uint32_t obj = si->id;
    uint32_t i0 = fetch_index(tb, obj, 3 * h->prim + 0);
    uint32_t i1 = fetch_index(tb, obj, 3 * h->prim + 1);
    uint32_t i2 = fetch_index(tb, obj, 3 * h->prim + 2);
    const sr_model_vertex *v0 = &tb->vertices[obj][i0];
    const sr_model_vertex *v1 = &tb->vertices[obj][i1];
    const sr_model_vertex *v2 = &tb->vertices[obj][i2];
    float bx = 1.0f - h->u - h->v, by = h->u, bz = h->v;
    float n[4], p[4], nw[4], pw[4];
    for (int a = 0; a < 3; ++a) {
        n[a] = v0->normal[a] * bx + v1->normal[a] * by + v2->normal[a] * bz;
        p[a] = v0->pos[a] * bx + v1->pos[a] * by + v2->pos[a] * bz;
        s->col[a] = v0->color[a] * bx + v1->color[a] * by + v2->color[a] * bz;
    }
    n[3] = 0.0f; p[3] = 1.0f;
    mat4_mul_vec4(si->transform_it, n, nw);
    normalize3(nw, s->nrm);
    mat4_mul_vec4(si->transform, p, pw);
    s->pos[0] = pw[0]; s->pos[1] = pw[1]; s->pos[2] = pw[2];
}

static inline void mk_ray(sr_ray *r, const float o[3], float tmin, const float d[3], float tmax)
{
    r->ox = o[0]; r->oy = o[1]; r->oz = o[2]; r->tmin = tmin;
    r->dx = d[0]; r->dy = d[1]; r->dz = d[2]; r->tmax = tmax;
}

static inline float length3(const float a[3]) { return sqrtf(dot3(a, a)); }

/* ------------------------------------------------------------------------ */
/* ao.rgen / ao.rchit / ao.rmiss                                            */
/* ------------------------------------------------------------------------ */
static void ao_glsl_pixel(const sr_scene *s, const sr_tables *tb, const sr_uniforms *un,
                          int32_t accum_start, float *accum, const uint8_t *blue,
                          uint32_t bw, uint32_t bh, uint32_t x, uint32_t y, uint32_t W, uint32_t H)
{
    const int max_samples = 4, sample_count = 4;          /* ao.rgen:42-43 */
    float ao = 0.0f;
    uint32_t rng = sr_tea(x + y * W, un->frame[2]);        /* ao.rgen:45 */
    for (int i = 0; i < sample_count; i++) {
        float jx = sr_next_rand(&rng);
        float jy = sr_next_rand(&rng);                     /* ao.rgen:49 */
        float o[3], d[3];
        sr_camera_ray(un, x, y, W, H, jx, jy, o, d);
        /* preparePayload, ao.rgen:26-35 */
        float hitValue = 0.0f;
        uint32_t depth = 0, done = 0, sampleId = (uint32_t)i;
        float rmin = fmaxf(1.0f, length3(o)) * 1e-3f, rmax = 10000.0f;
        for (;;) {
            sr_ray ray;
            sr_hit h;
            mk_ray(&ray, o, rmin, d, rmax);
            trace_one(s, &ray, &h, 1, 0xFFu, 0);
            if (h.inst == SR_MISS) {
                done = 1;                                   /* ao.rmiss:9 */
            } else {
                surf_t sf;
                fetch_surface(tb, &h, &sf);                 /* ao.rchit:56-77 */
                float no[3];
                for (int a = 0; a < 3; ++a) no[a] = sf.pos[a] + 0.00001f * d[a];   /* :79 */
                /* getBlueRand2(depth + depth * sampleId), ao.rchit:45-51, 80 */
                uint32_t bi = depth + depth * sampleId;
                float r0 = sr_next_rand(&rng), r1 = sr_next_rand(&rng);
                float fx = (float)x + r0 * (float)bw, fy = (float)y + r1 * (float)bh;
                float mx = fx - (float)bw * floorf(fx / (float)bw);
                float my = fy - (float)bh * floorf(fy / (float)bh);
                int cx = (int)mx, cy = (int)my;
                if (cx >= (int)bw) cx -= (int)bw;
                if (cy >= (int)bh) cy -= (int)bh;
                if (cx < 0) cx += (int)bw;
                if (cy < 0) cy += (int)bh;
                const uint8_t *texel = &blue[4 * ((size_t)cy * bw + (size_t)cx)];
                float xi[2] = {(float)texel[bi % 4u] / 255.0f, (float)texel[(bi + 1u) % 4u] / 255.0f};
                float nd[3] = {-d[0], -d[1], -d[2]};
                float sg = signf_(dot3(nd, sf.nrm));
                float hn[3] = {sf.nrm[0] * sg, sf.nrm[1] * sg, sf.nrm[2] * sg};   /* :81 */
                sr_sample_cosine(hn, xi, d);                /* :82 */
                o[0] = no[0]; o[1] = no[1]; o[2] = no[2];
                rmin = 0.001f; rmax = 10.0f;                /* :83 */
                if (depth > 0) hitValue += 1.0f;            /* :84-86 */
                depth++;                                    /* :87 */
            }
            if (done == 1 || depth > (uint32_t)max_samples) break;   /* ao.rgen:71 */
        }
        ao += 1.0f / (float)max_samples * hitValue;         /* ao.rgen:74 */
    }
    float color = 1.0f - ao / (float)sample_count;          /* ao.rgen:76 */
    float a = 1.0f / (float)(un->frame[2] - (uint32_t)accum_start + 1u);   /* :78 */
    float *px = &accum[4 * ((size_t)y * W + x)];
    for (int c = 0; c < 3; ++c) px[c] = mixf(px[c], color, a);
    px[3] = 1.0f;
}

void sr_render_ao_glsl(const sr_scene *s, const sr_tables *tb, const sr_uniforms *u,
                       int32_t accum_start_frame, float *accum,
                       const uint8_t *blue, uint32_t bw, uint32_t bh,
                       uint32_t W, uint32_t H, uint32_t y0, uint32_t y1, int nthreads)
{
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int64_t y = y0; y < (int64_t)y1; ++y)
        for (uint32_t x = 0; x < W; ++x)
            ao_glsl_pixel(s, tb, u, accum_start_frame, accum, blue, bw, bh, x, (uint32_t)y, W, H);
}

/* ------------------------------------------------------------------------ */
/* plain ray-traced AO (the 2 M-triangle AO workload, built on ao.rgen's   */
/* camera/seed and ao.rchit's surface reconstruction):                      */
/*   rng = tea(pixel, frame); jitter = 2 rands; primary ray with            */
/*   preparePayload's range; on a hit, n_ao cosine directions (2 rands      */
/*   each) about the normal flipped toward the viewer, origin = hit point,  */
/*   range [ao_tmin, ao_tmax]; value = 1 - occluded / n_ao (1 on a miss);   */
/*   running mean as ao.rgen:77-81.                                          */
/* ------------------------------------------------------------------------ */
void sr_render_ao(const sr_scene *s, const sr_tables *tb, const sr_uniforms *un,
                  const sr_ao_params *p, int32_t accum_start, float *accum,
                  sr_hit *prim_hits, sr_hit *ao_hits, uint64_t *ray_count,
                  uint32_t W, uint32_t H, uint32_t y0, uint32_t y1, int nthreads)
{
    uint64_t rays = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(+ : rays)
#endif
    for (int64_t yy = y0; yy < (int64_t)y1; ++yy) {
        uint32_t y = (uint32_t)yy;
        for (uint32_t x = 0; x < W; ++x) {
            size_t pix = (size_t)y * W + x;
            uint32_t rng = sr_tea(x + y * W, un->frame[2]);
            float jx = sr_next_rand(&rng);
            float jy = sr_next_rand(&rng);
            float o[3], d[3];
            sr_camera_ray(un, x, y, W, H, jx, jy, o, d);
            sr_ray ray;
            sr_hit h;
            mk_ray(&ray, o, fmaxf(1.0f, length3(o)) * 1e-3f, d, 10000.0f);
            trace_one(s, &ray, &h, 1, 0xFFu, 0);
            rays += 1;
            if (prim_hits) prim_hits[pix] = public_hit(s, h);
            float color = 1.0f;
            if (h.inst != SR_MISS) {
                surf_t sf;
                fetch_surface(tb, &h, &sf);
                float nd[3] = {-d[0], -d[1], -d[2]};
                float sg = signf_(dot3(nd, sf.nrm));
                float hn[3] = {sf.nrm[0] * sg, sf.nrm[1] * sg, sf.nrm[2] * sg};
                uint32_t occl = 0;
                for (uint32_t k = 0; k < p->n_ao; ++k) {
                    float xi[2];
                    xi[0] = sr_next_rand(&rng);
                    xi[1] = sr_next_rand(&rng);
                    float ad[3];
                    sr_sample_cosine(hn, xi, ad);
                    sr_ray ar;
                    sr_hit ah;
                    mk_ray(&ar, sf.pos, p->ao_tmin, ad, p->ao_tmax);
                    trace_one(s, &ar, &ah, 1, 0xFFu, p->any_hit ? 1 : 0);
                    if (ah.inst != SR_MISS) occl++;
                    if (ao_hits) ao_hits[pix * p->n_ao + k] = public_hit(s, ah);
                }
                rays += p->n_ao;
                color = 1.0f - (float)occl / (float)p->n_ao;
            } else if (ao_hits) {
                for (uint32_t k = 0; k < p->n_ao; ++k) {
                    sr_hit *m = &ao_hits[pix * p->n_ao + k];
                    m->t = -1.0f; m->u = 0.0f; m->v = 0.0f; m->inst = SR_MISS; m->prim = SR_MISS; m->geom = SR_MISS;
                }
            }
            float a = 1.0f / (float)(un->frame[2] - (uint32_t)accum_start + 1u);
            float *px = &accum[4 * pix];
            for (int c = 0; c < 3; ++c) px[c] = mixf(px[c], color, a);
            px[3] = 1.0f;
        }
    }
    if (ray_count) *ray_count = rays;
}

/* ------------------------------------------------------------------------ */
/* pathtrace.rgen / pathtrace.rchit / pathtrace.rmiss                       */
/* ------------------------------------------------------------------------ */
static void pt_miss_sun(uint32_t enable_sky, int sun, const float d[3], float hv[3])
{
    if (enable_sky) {                                      /* pathtrace.rmiss:10-16 */
        float wI[3];
        normalize3(d, wI);
        float e = 0.5f * (wI[1] + 1.0f);
        float tt = saturate((e - 0.35f) / (0.65f - 0.35f));
        float t = tt * tt * (3.0f - 2.0f * tt);
        const float c0[3] = {0.58f, 0.45f, 0.25f}, c1[3] = {0.3f, 0.4f, 0.5f};
        const float sd0[3] = {0.0f, 1.0f, -0.25f};
        float sd[3];
        normalize3(sd0, sd);
        float isSun = (sun && dot3(wI, sd) > 0.99f) ? 1.0f : 0.0f;
        const float sun[3] = {120.0f, 100.0f, 50.0f};
        for (int a = 0; a < 3; ++a) {
            float sky = mixf(c0[a], c1[a], t);
            hv[a] = mixf(sky, sun[a], isSun);
        }
    } else {
        hv[0] = hv[1] = hv[2] = 0.0f;
    }
}

static void pt_miss(uint32_t enable_sky, const float d[3], float hv[3]) { pt_miss_sun(enable_sky, 1, d, hv); }

static void pt_glsl_pixel(const sr_scene *s, const sr_tables *tb, const sr_uniforms *un,
                          const sr_pt_params *p, int32_t accum_start, float *accum,
                          uint8_t *out, uint64_t *rays, uint32_t x, uint32_t y, uint32_t W, uint32_t H)
{
    const int maxBounces = (int)p->max_bounces, sampleCount = (int)p->samples;
    float pixelColor[3] = {0.0f, 0.0f, 0.0f};
    uint32_t rng = sr_tea(x + y * W, un->frame[2]);         /* pathtrace.rgen:47 */
    for (int i = 0; i < sampleCount; i++) {
        float jx = sr_next_rand(&rng);
        float jy = sr_next_rand(&rng);                      /* :52 */
        float o[3], d[3];
        sr_camera_ray(un, x, y, W, H, jx, jy, o, d);
        float rmin = fmaxf(1.0f, length3(o)) * 1e-3f, rmax = 10000.0f;   /* :35 */
        uint32_t depth = 0, done = 0;
        float acc[3] = {1.0f, 1.0f, 1.0f};
        for (;;) {
            sr_ray ray;
            sr_hit h;
            float hv[3];
            mk_ray(&ray, o, rmin, d, rmax);
            trace_one(s, &ray, &h, 1, 0xFFu, 0);
            *rays += 1;
            if (h.inst == SR_MISS) {
                pt_miss(p->enable_sky, d, hv);
                done = 1;
            } else {
                const sr_material *mat = tb->materials[h.inst];   /* pathtrace.rchit:69 */
                if (mat->emissive[0] >= 1.0f || mat->emissive[1] >= 1.0f || mat->emissive[2] >= 1.0f) {
                    hv[0] = mat->emissive[0]; hv[1] = mat->emissive[1]; hv[2] = mat->emissive[2];
                    done = 1;
                    depth++;                                 /* :71-76 */
                } else {
                    surf_t sf;
                    fetch_surface(tb, &h, &sf);             /* :78-88 */
                    float wI[3];
                    normalize3(d, wI);                       /* :90 */
                    float nwI[3] = {-wI[0], -wI[1], -wI[2]};
                    float sg = signf_(dot3(sf.nrm, nwI));
                    float nO[3] = {sf.nrm[0] * sg, sf.nrm[1] * sg, sf.nrm[2] * sg};
                    float alphaSquared = mat->roughness * mat->roughness;
                    float xi[2];
                    xi[0] = sr_next_rand(&rng);
                    xi[1] = sr_next_rand(&rng);              /* :93 */
                    float rnd = sr_next_rand(&rng);          /* :94 */
                    for (int a = 0; a < 3; ++a) o[a] = sf.pos[a] + 0.0001f * nO[a];   /* :96 */
                    float nd[3];
                    if (rnd < mat->metallic) {               /* :97-100 */
                        float rf[3];
                        reflect3(d, nO, rf);
                        sr_sample_ggx(rf, xi, alphaSquared, nd);
                        for (int a = 0; a < 3; ++a) hv[a] = mat->base_color[a] * sf.col[a];
                    } else {
                        float m[3];
                        sr_sample_ggx(nO, xi, alphaSquared, m);   /* :102 */
                        if (rnd < sr_fresnel_dielectric(nO, m, 1.0f / 1.5f)) {
                            reflect3(d, m, nd);
                            hv[0] = hv[1] = hv[2] = 1.0f;
                        } else {
                            sr_sample_cosine(nO, xi, nd);
                            for (int a = 0; a < 3; ++a) hv[a] = mat->base_color[a] * sf.col[a];
                        }
                    }
                    d[0] = nd[0]; d[1] = nd[1]; d[2] = nd[2];
                    depth++;                                 /* :112 */
                }
            }
            for (int a = 0; a < 3; ++a) acc[a] = acc[a] * hv[a];   /* pathtrace.rgen:77 */
            if (done == 1) break;
            if ((int)depth > maxBounces) {
                acc[0] = acc[1] = acc[2] = 0.0f;
                break;
            }
        }
        for (int a = 0; a < 3; ++a) pixelColor[a] = pixelColor[a] + acc[a];
    }
    float inv = 1.0f / (float)sampleCount;                  /* :88 */
    for (int a = 0; a < 3; ++a) pixelColor[a] = inv * pixelColor[a];
    float alpha = 1.0f / (float)(un->frame[2] + 1u - (uint32_t)accum_start);   /* :90 */
    float *px = &accum[4 * ((size_t)y * W + x)];
    float mixed[3];
    int bad = 0;
    for (int a = 0; a < 3; ++a) {
        mixed[a] = mixf(px[a], pixelColor[a], alpha);
        if (isnan(mixed[a])) bad = 1;
    }
    for (int a = 0; a < 3; ++a) if (isinf(mixed[a])) bad = 1;
    if (bad) { mixed[0] = px[0]; mixed[1] = px[1]; mixed[2] = px[2]; }   /* :94-99 */
    px[0] = mixed[0]; px[1] = mixed[1]; px[2] = mixed[2]; px[3] = 1.0f;
    if (out) {
        uint8_t *q = &out[4 * ((size_t)y * W + x)];
        for (int a = 0; a < 3; ++a) q[a] = to_unorm8(powf(mixed[a], 1.0f / 2.2f));   /* :102 */
        q[3] = 255;
    }
}

void sr_render_pathtrace_glsl(const sr_scene *s, const sr_tables *tb, const sr_uniforms *u,
                              const sr_pt_params *p, int32_t accum_start_frame,
                              float *accum, uint8_t *out_rgba8, uint64_t *ray_count,
                              uint32_t W, uint32_t H, uint32_t y0, uint32_t y1, int nthreads)
{
    uint64_t rays = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(+ : rays)
#endif
    for (int64_t y = y0; y < (int64_t)y1; ++y) {
        uint64_t local = 0;
        for (uint32_t x = 0; x < W; ++x)
            pt_glsl_pixel(s, tb, u, p, accum_start_frame, accum, out_rgba8, &local, x, (uint32_t)y, W, H);
        rays += local;
    }
    if (ray_count) *ray_count = rays;
}

/* ------------------------------------------------------------------------ */
/* wavefront path tracer's estimator (SOLHIP_PROG_PATHTRACE)                */
/* ------------------------------------------------------------------------ */
static inline int is_emitter(const sr_material *m)
{
    return m->emissive[0] >= 1.0f || m->emissive[1] >= 1.0f || m->emissive[2] >= 1.0f;   /* pathtrace.rchit:71 */
}

uint32_t sr_lights(const sr_scene *s, const sr_tables *tb, sr_light *out, uint32_t cap, float *total_area)
{
    uint32_t n = 0;
    float run = 0.0f;
    uint32_t ninst = s->ninst < tb->instance_count ? s->ninst : tb->instance_count;
    for (uint32_t i = 0; tb->materials && i < ninst; ++i) {
        const sr_material *m = tb->materials[i];
        if (!is_emitter(m)) continue;
        const sr_scene_instance *si = &tb->instances[i];
        const uint32_t obj = si->id, nt = s->blas[s->inst[i].blas].ntri;
        for (uint32_t t = 0; t < nt; ++t) {
            float w[3][4];
            for (int k = 0; k < 3; ++k) {
                const sr_model_vertex *v = &tb->vertices[obj][fetch_index(tb, obj, 3 * t + k)];
                const float p[4] = {v->pos[0], v->pos[1], v->pos[2], 1.0f};
                mat4_mul_vec4(si->transform, p, w[k]);
            }
            const float e1[3] = {w[1][0] - w[0][0], w[1][1] - w[0][1], w[1][2] - w[0][2]};
            const float e2[3] = {w[2][0] - w[0][0], w[2][1] - w[0][1], w[2][2] - w[0][2]};
            const float c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                e1[0] * e2[1] - e1[1] * e2[0]};
            const float len = sqrtf(dot3(c, c));
            run = run + 0.5f * len;
            if (n < cap) {
                sr_light *L = &out[n];
                for (int a = 0; a < 3; ++a) {
                    L->p0[a] = w[0][a];
                    L->p1[a] = w[1][a];
                    L->p2[a] = w[2][a];
                    L->n[a] = len > 0.0f ? c[a] / len : 0.0f;
                    L->e[a] = m->emissive[a];
                }
                L->cdf = run;
            }
            ++n;
        }
    }
    if (total_area) *total_area = run;
    return n;
}

/* first light whose inclusive area sum exceeds u (the last one if none) */
static inline uint32_t pick_light(const sr_light *L, uint32_t n, float u)
{
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (L[mid].cdf > u) hi = mid; else lo = mid + 1;
    }
    return lo;
}

static void pt_cfg_pixel(const sr_scene *s, const sr_tables *tb, const sr_uniforms *un,
                         const sr_pt_config *p, const sr_light *lights, uint32_t nl, float total,
                         int32_t accum_start, float *accum, uint8_t *out, uint64_t *rays,
                         uint32_t x, uint32_t y, uint32_t W, uint32_t H)
{
    const int maxBounces = (int)p->max_bounces, sampleCount = (int)p->samples;
    const int sun = p->enable_sky != 2;
    const int nee = p->nee && nl > 0 && total > 0.0f;
    float pixelColor[3] = {0.0f, 0.0f, 0.0f};
    int64_t fix[3] = {0, 0, 0};
    int fix_bad = 0;
    uint32_t rng = sr_tea(x + y * W, un->frame[2]);
    for (int i = 0; i < sampleCount; i++) {
        /* pathtrace.rgen:47 seeds once per pixel; with sample_streams every
         * sample gets its own seed of the same hash (sample 0 unchanged) */
        if (p->sample_streams) rng = sr_tea(x + y * W + (uint32_t)i * W * H, un->frame[2]);
        float jx = sr_next_rand(&rng);
        float jy = sr_next_rand(&rng);
        float o[3], d[3];
        sr_camera_ray(un, x, y, W, H, jx, jy, o, d);
        const float rmin = fmaxf(1.0f, length3(o)) * 1e-3f, rmax = 10000.0f;
        uint32_t depth = 0;
        int count_emit = 1;
        float T[3] = {1.0f, 1.0f, 1.0f}, L[3] = {0.0f, 0.0f, 0.0f};
        for (;;) {
            sr_ray ray;
            sr_hit h;
            mk_ray(&ray, o, rmin, d, rmax);
            trace_one(s, &ray, &h, 1, 0xFFu, 0);
            *rays += 1;
            if (h.inst == SR_MISS) {
                float sv[3];
                pt_miss_sun(p->enable_sky, sun, d, sv);
                for (int a = 0; a < 3; ++a) L[a] = L[a] + T[a] * sv[a];
                break;
            }
            const sr_material *mat = tb->materials[h.inst];
            if (is_emitter(mat)) {
                if (count_emit)
                    for (int a = 0; a < 3; ++a) L[a] = L[a] + T[a] * mat->emissive[a];
                break;
            }
            surf_t sf;
            fetch_surface(tb, &h, &sf);
            float wI[3];
            normalize3(d, wI);
            float nwI[3] = {-wI[0], -wI[1], -wI[2]};
            float sg = signf_(dot3(sf.nrm, nwI));
            float nO[3] = {sf.nrm[0] * sg, sf.nrm[1] * sg, sf.nrm[2] * sg};
            float alphaSquared = mat->roughness * mat->roughness;
            float xi[2];
            xi[0] = sr_next_rand(&rng);
            xi[1] = sr_next_rand(&rng);
            float rnd = sr_next_rand(&rng);
            for (int a = 0; a < 3; ++a) o[a] = sf.pos[a] + 0.0001f * nO[a];
            float nd[3], hv[3];
            int diffuse = 0;
            if (rnd < mat->metallic) {
                float rf[3];
                reflect3(d, nO, rf);
                sr_sample_ggx(rf, xi, alphaSquared, nd);
                for (int a = 0; a < 3; ++a) hv[a] = mat->base_color[a] * sf.col[a];
            } else {
                float m[3];
                sr_sample_ggx(nO, xi, alphaSquared, m);
                if (rnd < sr_fresnel_dielectric(nO, m, 1.0f / 1.5f)) {
                    reflect3(d, m, nd);
                    hv[0] = hv[1] = hv[2] = 1.0f;
                } else {
                    sr_sample_cosine(nO, xi, nd);
                    for (int a = 0; a < 3; ++a) hv[a] = mat->base_color[a] * sf.col[a];
                    diffuse = 1;
                }
            }
            float Tb[3];   /* throughput into this vertex times its diffuse albedo */
            for (int a = 0; a < 3; ++a) Tb[a] = T[a] * (mat->base_color[a] * sf.col[a]);
            for (int a = 0; a < 3; ++a) T[a] = T[a] * hv[a];
            d[0] = nd[0]; d[1] = nd[1]; d[2] = nd[2];
            depth++;
            if ((int)depth > maxBounces) break;
            count_emit = 1;
            if (nee) {
                /* pathtrace.rchit:93-110 picks the diffuse lobe when
                 * rand >= max(metallic, F(m)), with m drawn from the same Xi
                 * as the cosine direction: m shares its azimuth and has
                 * cos^2 theta_m = (1 - c^2) / (c^2 (a^2 - 1) + 1) for a
                 * direction at cos theta = c.  The diffuse lobe's share of
                 * direct light is therefore estimated with weight
                 * 1 - max(metallic, F(theta_m)) at every vertex. */
                const float ul = sr_next_rand(&rng) * total;
                const float b1 = sr_next_rand(&rng), b2 = sr_next_rand(&rng);
                const sr_light *Lt = &lights[pick_light(lights, nl, ul)];
                const float su = sqrtf(b1), ba = 1.0f - su, bb = su * (1.0f - b2), bc = su * b2;
                float dl[3];
                for (int a = 0; a < 3; ++a) dl[a] = (Lt->p0[a] * ba + Lt->p1[a] * bb + Lt->p2[a] * bc) - o[a];
                const float dist2 = dot3(dl, dl), dist = sqrtf(dist2);
                float wl[3];
                for (int a = 0; a < 3; ++a) wl[a] = dl[a] / dist;
                const float cos_x = dot3(nO, wl), cos_l = fabsf(dot3(Lt->n, wl));
                if (dist2 > 0.0f && cos_x > 0.0f && cos_l > 0.0f) {
                    const float x0 = cos_x * cos_x;
                    const float cm = sqrtf(saturate((1.0f - x0) / (x0 * (alphaSquared - 1.0f) + 1.0f)));
                    const float pd = 1.0f - fmaxf(mat->metallic, fresnel_cos(cm, 1.0f / 1.5f));
                    if (pd > 0.0f) {
                        const float g = (pd * (cos_x * cos_l)) * total / (0x1.921fb6p+1f * dist2);
                        sr_ray sray;
                        sr_hit sh;
                        mk_ray(&sray, o, rmin, wl, dist * 0.9999f);
                        trace_one(s, &sray, &sh, 1, 0xFFu, 1);
                        *rays += 1;
                        if (sh.inst == SR_MISS)
                            for (int a = 0; a < 3; ++a) L[a] = L[a] + (Tb[a] * Lt->e[a]) * g;
                    }
                }
                /* emitters met by the diffuse lobe were sampled above */
                if (diffuse) count_emit = 0;
            }
            if (p->rr_start && depth >= p->rr_start) {
                const float q = fminf(fmaxf(T[0], fmaxf(T[1], T[2])), 0.95f);
                if (sr_next_rand(&rng) >= q) break;
                for (int a = 0; a < 3; ++a) T[a] = T[a] / q;
            }
        }
        if (p->sample_streams) {
            /* order-free sum: each sample in fixed point, 2^20 units per 1.0,
             * rounded to nearest even; 2^19 or more (or NaN / Inf) is bad */
            for (int a = 0; a < 3; ++a) {
                if (!(fabsf(L[a]) < 0x1p19f)) fix_bad = 1;
                else fix[a] += (int64_t)rintf(L[a] * 0x1p20f);
            }
        } else {
            for (int a = 0; a < 3; ++a) pixelColor[a] = pixelColor[a] + L[a];
        }
    }
    if (p->sample_streams)
        for (int a = 0; a < 3; ++a) pixelColor[a] = (float)((double)fix[a] * 0x1p-20);
    float inv = 1.0f / (float)sampleCount;
    for (int a = 0; a < 3; ++a) pixelColor[a] = inv * pixelColor[a];
    float alpha = 1.0f / (float)(un->frame[2] + 1u - (uint32_t)accum_start);
    float *px = &accum[4 * ((size_t)y * W + x)];
    float mixed[3];
    int bad = fix_bad;
    for (int a = 0; a < 3; ++a) {
        mixed[a] = mixf(px[a], pixelColor[a], alpha);
        if (isnan(mixed[a]) || isinf(mixed[a])) bad = 1;
    }
    if (bad) { mixed[0] = px[0]; mixed[1] = px[1]; mixed[2] = px[2]; }
    px[0] = mixed[0]; px[1] = mixed[1]; px[2] = mixed[2]; px[3] = 1.0f;
    if (out) {
        uint8_t *q = &out[4 * ((size_t)y * W + x)];
        for (int a = 0; a < 3; ++a) q[a] = to_unorm8(powf(mixed[a], 1.0f / 2.2f));
        q[3] = 255;
    }
}

void sr_render_pathtrace(const sr_scene *s, const sr_tables *tb, const sr_uniforms *u,
                         const sr_pt_config *p, int32_t accum_start_frame,
                         float *accum, uint8_t *out_rgba8, uint64_t *ray_count,
                         uint32_t W, uint32_t H, uint32_t y0, uint32_t y1, int nthreads)
{
    float total = 0.0f;
    uint32_t nl = p->nee ? sr_lights(s, tb, NULL, 0, &total) : 0;
    sr_light *lights = (sr_light *)malloc(sizeof(sr_light) * (nl ? nl : 1));
    if (nl) sr_lights(s, tb, lights, nl, &total);
    uint64_t rays = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(+ : rays)
#endif
    for (int64_t y = y0; y < (int64_t)y1; ++y) {
        uint64_t local = 0;
        for (uint32_t x = 0; x < W; ++x)
            pt_cfg_pixel(s, tb, u, p, lights, nl, total, accum_start_frame, accum, out_rgba8, &local, x,
                         (uint32_t)y, W, H);
        rays += local;
    }
    free(lights);
    if (ray_count) *ray_count = rays;
}

/* ------------------------------------------------------------------------ */
/* debug.rgen / debug.rchit / debug.rmiss                                   */
/* ------------------------------------------------------------------------ */
void sr_render_debug(const sr_scene *s, const sr_uniforms *un, float *out, uint32_t W, uint32_t H,
                     int nthreads)
{
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int64_t yy = 0; yy < (int64_t)H; ++yy) {
        uint32_t y = (uint32_t)yy;
        for (uint32_t x = 0; x < W; ++x) {
            float o[3], d[3];
            sr_camera_ray(un, x, y, W, H, 0.5f, 0.5f, o, d);    /* debug.rgen:20-26 */
            sr_ray ray;
            sr_hit h;
            mk_ray(&ray, o, 0.001f, d, 1000.0f);                /* :32-34 */
            trace_one(s, &ray, &h, 1, 0xFFu, 0);
            float *q = &out[4 * ((size_t)y * W + x)];
            if (h.inst == SR_MISS) { q[0] = d[0]; q[1] = d[1]; q[2] = d[2]; }
            else { q[0] = 1.0f - h.u - h.v; q[1] = h.u; q[2] = h.v; }   /* debug.rchit:11-12 */
            q[3] = 0.0f;
        }
    }
}

double sr_wall_seconds(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}
