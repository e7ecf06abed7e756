// mlp.hip -- the radiance cache's network: a 64-input MLP with four hidden
// ReLU layers of width 64 and 3 linear outputs, no biases, BF16 operands on
// MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation, trained by
// hand-written backward kernels and Adam with FP32 master weights.
//
// Parameters (FP32 master, P = 4*64*64 + 64*16 = 17408 floats):
//   W_l[k][n] at l*4096 + k*64 + n        (l = 0..3, k = input, n = output)
//   W_4[k][n] at 16384 + k*16 + n         (n < 3 used; columns 3..15 stay 0)
// BF16 copies: "forward" layout Wt_l[n][k] (operand B of X*W with k
// contiguous) and "backward" layout = the master layout (operand B of dY*W^T).
//
// MFMA 16x16x32 lane layout (lane l, r = l % 16, q = l / 16):
//   A (16 x 32): row r, k = 8q .. 8q+7      B (32 x 16): col r, k = 8q .. 8q+7
//   D (16 x 16): col r, rows 4q .. 4q+3 (registers 0..3)
//
// Kernels: forward (optionally saving every layer's input transposed,
// feature-major, for the weight gradients), backward-data (loss gradient,
// ReLU masks, dZ_l per layer, transposed), backward-weights (X_l^T dZ_l over
// sample chunks into per-chunk partials), reduce + Adam.  Loss: relative L2,
// sum_c (p_c - t_c)^2 / (sg(p_c)^2 + 0.01) / N (Mueller et al. 2021).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "objects.h"

namespace solhip {

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kOutPad = 16;          // output layer columns in the MFMA tile
constexpr int kOut = 3;
constexpr int kHidden = 4;           // layers 0..3 are 64x64 + ReLU, layer 4 is 64x16 linear
constexpr uint32_t kP = 4 * 4096 + 64 * 16;
constexpr uint32_t kChunk = 1024;    // samples per backward-weights partial
constexpr int kMlpBlock = 256;       // 4 waves

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ bf16x8 zero8()
{
    bf16x8 z;
    for (int j = 0; j < 8; ++j) z[j] = (bf16)0.0f;
    return z;
}

__device__ __forceinline__ uint32_t wt_offset(int l, int n) { return (uint32_t)l * 4096u + (uint32_t)n * 64u; }

struct FwdArgs {
    const bf16 *x;          // N x 64
    uint32_t n, npad;
    const uint32_t *n_dev;  // if set, N is read from device memory (at most n)
    const bf16 *wt;         // forward layout
    float4 *y;              // N outputs (x, y, z, 0), or null
    bf16 *actT;             // 5 x [64][npad] layer inputs, feature-major, or null
    // cache-query epilogue: st1[rec.path].xyz += rec.T * max(y, 0)
    const CacheQuery *qrec;
    float4 *st1;
};

// one wave = 16 rows per band; weights staged in LDS (34 KB)
__global__ void __launch_bounds__(kMlpBlock) k_mlp_fwd(FwdArgs A)
{
    __shared__ bf16 wl[4 * 4096 + 16 * 64];
    __shared__ bf16 hbuf[kMlpBlock / 64][16 * 64];
    for (uint32_t i = threadIdx.x; i < (4 * 4096 + 16 * 64) / 8; i += kMlpBlock)
        reinterpret_cast<bf16x8 *>(wl)[i] = reinterpret_cast<const bf16x8 *>(A.wt)[i];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, r = lane & 15u, q = lane >> 4, wv = threadIdx.x >> 6;
    bf16 *H = hbuf[wv];
    const uint32_t n = A.n_dev ? min(*A.n_dev, A.n) : A.n;
    const uint32_t nbands = A.n_dev ? (n + 15) / 16 : A.npad / 16, waves = gridDim.x * (kMlpBlock / 64);
    for (uint32_t band = blockIdx.x * (kMlpBlock / 64) + wv; band < nbands; band += waves) {
        const uint32_t row0 = band * 16;
        const bool rv = row0 + r < n;
        bf16x8 a0 = rv ? *reinterpret_cast<const bf16x8 *>(A.x + (uint64_t)(row0 + r) * 64 + 8 * q) : zero8();
        bf16x8 a1 = rv ? *reinterpret_cast<const bf16x8 *>(A.x + (uint64_t)(row0 + r) * 64 + 32 + 8 * q) : zero8();
        if (A.actT) {
            // the input, feature-major, through LDS
            *reinterpret_cast<bf16x8 *>(H + r * 64 + 8 * q) = a0;
            *reinterpret_cast<bf16x8 *>(H + r * 64 + 32 + 8 * q) = a1;
            wave_sync();
            for (int nt = 0; nt < 4; ++nt) {
                const uint32_t c = nt * 16 + r;
                bf16x4 v;
                for (int i = 0; i < 4; ++i) v[i] = H[(4 * q + i) * 64 + c];
                *reinterpret_cast<bf16x4 *>(A.actT + (uint64_t)c * A.npad + row0 + 4 * q) = v;
            }
            wave_sync();
        }
        for (int l = 0; l < kHidden; ++l) {
            for (int nt = 0; nt < 4; ++nt) {
                const bf16 *wb = wl + wt_offset(l, nt * 16 + (int)r);
                f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                acc = mfma(a0, *reinterpret_cast<const bf16x8 *>(wb + 8 * q), acc);
                acc = mfma(a1, *reinterpret_cast<const bf16x8 *>(wb + 32 + 8 * q), acc);
                const uint32_t c = nt * 16 + r;
                bf16x4 v;
                for (int i = 0; i < 4; ++i) {
                    const bf16 h = (bf16)fmaxf(acc[i], 0.0f);
                    v[i] = h;
                    H[(4 * q + i) * 64 + c] = h;
                }
                if (A.actT)
                    *reinterpret_cast<bf16x4 *>(A.actT + (uint64_t)(l + 1) * 64 * A.npad + (uint64_t)c * A.npad + row0 +
                                                4 * q) = v;
            }
            wave_sync();
            a0 = *reinterpret_cast<const bf16x8 *>(H + r * 64 + 8 * q);
            a1 = *reinterpret_cast<const bf16x8 *>(H + r * 64 + 32 + 8 * q);
            wave_sync();
        }
        const bf16 *wb = wl + 4 * 4096 + r * 64;
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mfma(a0, *reinterpret_cast<const bf16x8 *>(wb + 8 * q), acc);
        acc = mfma(a1, *reinterpret_cast<const bf16x8 *>(wb + 32 + 8 * q), acc);
        if (A.y && r < kOut) {
            for (int i = 0; i < 4; ++i) {
                const uint32_t row = row0 + 4 * q + i;
                if (row < n) reinterpret_cast<float *>(A.y + row)[r] = acc[i];
            }
        }
        if (A.qrec && r < kOut) {
            // one lane per (row, component): distinct addresses, no atomics
            for (int i = 0; i < 4; ++i) {
                const uint32_t row = row0 + 4 * q + i;
                if (row < n) {
                    const CacheQuery rec = A.qrec[row];
                    float *L = reinterpret_cast<float *>(A.st1 + rec.path) + r;
                    *L = *L + rec.T[r] * fmaxf(acc[i], 0.0f);
                }
            }
        }
    }
}

struct BwdArgs {
    const float4 *y, *target;   // prediction, target (x, y, z, -)
    uint32_t n, npad;
    const bf16 *wb;             // backward layout (= master layout)
    const bf16 *actT;           // layer inputs, feature-major
    bf16 *dzT;                  // 5 x [64][npad]: gradient at each layer's output, feature-major
    float *loss;                // sum of the loss (atomic)
};

__global__ void __launch_bounds__(kMlpBlock) k_mlp_bwd_data(BwdArgs A)
{
    __shared__ bf16 wl[4 * 4096 + 64 * 16];
    __shared__ bf16 hbuf[kMlpBlock / 64][16 * 64];
    for (uint32_t i = threadIdx.x; i < (4 * 4096 + 64 * 16) / 8; i += kMlpBlock)
        reinterpret_cast<bf16x8 *>(wl)[i] = reinterpret_cast<const bf16x8 *>(A.wb)[i];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, r = lane & 15u, q = lane >> 4, wv = threadIdx.x >> 6;
    bf16 *H = hbuf[wv];
    const uint32_t nbands = A.npad / 16, waves = gridDim.x * (kMlpBlock / 64);
    const float invn = 1.0f / (float)A.n;
    float lsum = 0.0f;
    for (uint32_t band = blockIdx.x * (kMlpBlock / 64) + wv; band < nbands; band += waves) {
        const uint32_t row0 = band * 16, row = row0 + r;
        // output gradient: lanes q == 0 hold row r's 3 components in k = 0..2
        bf16x8 a0 = zero8();
        if (q == 0 && row < A.n) {
            const float4 p = A.y[row], t = A.target[row];
            const float pc[3] = {p.x, p.y, p.z}, tc[3] = {t.x, t.y, t.z};
            for (int c = 0; c < kOut; ++c) {
                const float den = pc[c] * pc[c] + 0.01f, d = pc[c] - tc[c];
                lsum += d * d / den * invn;
                a0[c] = (bf16)(2.0f * d / den * invn);
            }
        }
        // dZ_4 feature-major (16 columns)
        if (q == 0) {
            for (int c = 0; c < kOutPad; ++c)
                A.dzT[(uint64_t)4 * 64 * A.npad + (uint64_t)c * A.npad + row] = c < 8 ? a0[c] : (bf16)0.0f;
        }
        // back through layer 4: dH4 = dY4 W4^T (k = output column, padded to 32)
        bf16x8 g0 = zero8(), g1 = zero8();
        {
            for (int nt = 0; nt < 4; ++nt) {
                const uint32_t i_in = nt * 16 + r;
                const bf16x8 b = q < 2 ? *reinterpret_cast<const bf16x8 *>(wl + 4 * 4096 + i_in * 16 + 8 * q) : zero8();
                f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                acc = mfma(a0, b, acc);
                // ReLU mask of layer 3's output (= layer 4's input)
                const bf16x4 h = *reinterpret_cast<const bf16x4 *>(A.actT + (uint64_t)4 * 64 * A.npad +
                                                                    (uint64_t)i_in * A.npad + row0 + 4 * q);
                bf16x4 dz;
                for (int i = 0; i < 4; ++i) {
                    const bf16 v = (float)h[i] > 0.0f ? (bf16)acc[i] : (bf16)0.0f;
                    dz[i] = v;
                    H[(4 * q + i) * 64 + i_in] = v;
                }
                *reinterpret_cast<bf16x4 *>(A.dzT + (uint64_t)3 * 64 * A.npad + (uint64_t)i_in * A.npad + row0 +
                                            4 * q) = dz;
            }
            wave_sync();
            g0 = *reinterpret_cast<const bf16x8 *>(H + r * 64 + 8 * q);
            g1 = *reinterpret_cast<const bf16x8 *>(H + r * 64 + 32 + 8 * q);
            wave_sync();
        }
        // layers 3..1: dH_l = dZ_l W_l^T, masked by layer l-1's output
        for (int l = 3; l >= 1; --l) {
            for (int nt = 0; nt < 4; ++nt) {
                const uint32_t i_in = nt * 16 + r;
                const bf16 *wrow = wl + (uint32_t)l * 4096 + i_in * 64;
                f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                acc = mfma(g0, *reinterpret_cast<const bf16x8 *>(wrow + 8 * q), acc);
                acc = mfma(g1, *reinterpret_cast<const bf16x8 *>(wrow + 32 + 8 * q), acc);
                const bf16x4 h = *reinterpret_cast<const bf16x4 *>(A.actT + (uint64_t)l * 64 * A.npad +
                                                                    (uint64_t)i_in * A.npad + row0 + 4 * q);
                bf16x4 dz;
                for (int i = 0; i < 4; ++i) {
                    const bf16 v = (float)h[i] > 0.0f ? (bf16)acc[i] : (bf16)0.0f;
                    dz[i] = v;
                    H[(4 * q + i) * 64 + i_in] = v;
                }
                *reinterpret_cast<bf16x4 *>(A.dzT + (uint64_t)(l - 1) * 64 * A.npad + (uint64_t)i_in * A.npad +
                                            row0 + 4 * q) = dz;
            }
            wave_sync();
            g0 = *reinterpret_cast<const bf16x8 *>(H + r * 64 + 8 * q);
            g1 = *reinterpret_cast<const bf16x8 *>(H + r * 64 + 32 + 8 * q);
            wave_sync();
        }
    }
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
    if (lane == 0 && A.loss) atomicAdd(A.loss, lsum);
}

struct BwdWArgs {
    const bf16 *actT, *dzT;
    uint32_t npad;
    float *partials;   // [nchunks][kP]
};

// grid (chunks, 5 layers); wave w owns input features 16w .. 16w+15
__global__ void __launch_bounds__(kMlpBlock) k_mlp_bwd_w(BwdWArgs A)
{
    const uint32_t lane = threadIdx.x & 63u, r = lane & 15u, q = lane >> 4, wv = threadIdx.x >> 6;
    const uint32_t layer = blockIdx.y, chunk = blockIdx.x;
    const uint32_t s0 = chunk * kChunk, s1 = min(A.npad, s0 + kChunk);
    const int ntiles = layer < 4 ? 4 : 1;
    const bf16 *X = A.actT + (uint64_t)layer * 64 * A.npad;
    const bf16 *D = A.dzT + (uint64_t)layer * 64 * A.npad;
    const uint32_t i_row = wv * 16 + r;
    f32x4 acc[4];
    for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t s = s0; s + 32 <= s1; s += 32) {
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(X + (uint64_t)i_row * A.npad + s + 8 * q);
        for (int t = 0; t < ntiles; ++t) {
            const bf16x8 b = *reinterpret_cast<const bf16x8 *>(D + (uint64_t)(t * 16 + r) * A.npad + s + 8 * q);
            acc[t] = mfma(a, b, acc[t]);
        }
    }
    float *P = A.partials + (uint64_t)chunk * kP;
    for (int t = 0; t < ntiles; ++t) {
        for (int i = 0; i < 4; ++i) {
            const uint32_t k_in = wv * 16 + 4 * q + i, n_out = t * 16 + r;
            const uint32_t idx = layer < 4 ? layer * 4096 + k_in * 64 + n_out : 16384 + k_in * 16 + n_out;
            P[idx] = acc[t][i];
        }
    }
}

struct AdamArgs {
    const float *partials;
    uint32_t nchunks;
    float *w, *m, *v;
    bf16 *wt, *wb;
    float *grad_out;   // optional: the summed gradient
    float lr, b1, b2, eps, c1, c2;   // c1 = 1 - b1^t, c2 = 1 - b2^t
    uint32_t update;
};

__device__ __forceinline__ bool padded(uint32_t p) { return p >= 16384 && ((p - 16384) & 15u) >= (uint32_t)kOut; }

__device__ __forceinline__ void write_bf16(bf16 *wt, bf16 *wb, uint32_t p, float w)
{
    const bf16 h = (bf16)w;
    wb[p] = h;
    if (p < 16384) {
        const uint32_t l = p >> 12, k = (p >> 6) & 63u, n = p & 63u;
        wt[l * 4096 + n * 64 + k] = h;
    } else {
        const uint32_t k = (p - 16384) >> 4, n = (p - 16384) & 15u;
        wt[16384 + n * 64 + k] = h;
    }
}

__global__ void k_mlp_adam(AdamArgs A)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kP) return;
    float g = 0.0f;
    for (uint32_t c = 0; c < A.nchunks; ++c) g += A.partials[(uint64_t)c * kP + p];
    if (padded(p)) g = 0.0f;
    if (A.grad_out) A.grad_out[p] = g;
    if (!A.update) return;
    const float m = A.b1 * A.m[p] + (1.0f - A.b1) * g;
    const float v = A.b2 * A.v[p] + (1.0f - A.b2) * g * g;
    A.m[p] = m;
    A.v[p] = v;
    float w = A.w[p];
    if (!padded(p)) w -= A.lr * (m / A.c1) / (sqrtf(v / A.c2) + A.eps);
    A.w[p] = w;
    write_bf16(A.wt, A.wb, p, w);
}

__global__ void k_mlp_sync_bf16(const float *w, bf16 *wt, bf16 *wb)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kP) write_bf16(wt, wb, p, w[p]);
}

int mlp_status(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string(what) + " launch: " + hipGetErrorString(e));
        return SOLHIP_ERROR_DEVICE;
    }
    return 0;
}

}  // namespace

struct Mlp : Obj {
    DevBuf w, m, v, wt, wb, actT, dzT, partials, loss, grad, ybuf;
    float lr = 1e-2f, b1 = 0.9f, b2 = 0.99f, eps = 1e-8f;
    uint32_t step = 0, ncu = 256;
    uint32_t cap = 0;   // npad the activation buffers hold
    void release_all()
    {
        for (DevBuf *b : {&w, &m, &v, &wt, &wb, &actT, &dzT, &partials, &loss, &grad, &ybuf}) b->release();
    }
    int reserve_rows(uint32_t npad)
    {
        if (npad <= cap) return 0;
        const uint64_t nchunks = (npad + kChunk - 1) / kChunk;
        if (actT.reserve((uint64_t)5 * 64 * npad * 2) || dzT.reserve((uint64_t)5 * 64 * npad * 2) ||
            partials.reserve(nchunks * kP * 4))
            return SOLHIP_ERROR_OUT_OF_MEMORY;
        cap = npad;
        return 0;
    }
    int forward(hipStream_t s, const bf16 *x, uint32_t n, float4 *y, bool save)
    {
        const uint32_t npad = (n + 31u) & ~31u;
        if (save && reserve_rows(npad)) return SOLHIP_ERROR_OUT_OF_MEMORY;
        FwdArgs a;
        std::memset(&a, 0, sizeof(a));
        a.x = x;
        a.n = n;
        a.npad = npad;
        a.wt = (const bf16 *)wt.ptr;
        a.y = y;
        a.actT = save ? (bf16 *)actT.ptr : nullptr;
        const uint32_t blocks = std::min<uint32_t>((npad / 16 + 3) / 4, ncu * 4);
        hipLaunchKernelGGL(k_mlp_fwd, dim3(std::max(1u, blocks)), dim3(kMlpBlock), 0, s, a);
        return mlp_status("mlp forward");
    }
    // loss gradient of (x, target) summed into grad (and applied by Adam when update)
    int backward(hipStream_t s, uint32_t n, const float4 *y, const float4 *target, bool update, float *grad_host)
    {
        // rows n .. npad (a multiple of 32) carry zero inputs and zero
        // gradients, so they add nothing to the weight gradients
        const uint32_t npad = (n + 31u) & ~31u;
        const uint32_t nchunks = (npad + kChunk - 1) / kChunk;
        BwdArgs b;
        b.y = y;
        b.target = target;
        b.n = n;
        b.npad = npad;
        b.wb = (const bf16 *)wb.ptr;
        b.actT = (const bf16 *)actT.ptr;
        b.dzT = (bf16 *)dzT.ptr;
        b.loss = (float *)loss.ptr;
        const uint32_t blocks = std::min<uint32_t>((npad / 16 + 3) / 4, ncu * 4);
        hipLaunchKernelGGL(k_mlp_bwd_data, dim3(std::max(1u, blocks)), dim3(kMlpBlock), 0, s, b);
        BwdWArgs c;
        c.actT = (const bf16 *)actT.ptr;
        c.dzT = (const bf16 *)dzT.ptr;
        c.npad = npad;
        c.partials = (float *)partials.ptr;
        hipLaunchKernelGGL(k_mlp_bwd_w, dim3(nchunks, 5), dim3(kMlpBlock), 0, s, c);
        AdamArgs d;
        d.partials = (const float *)partials.ptr;
        d.nchunks = nchunks;
        d.w = (float *)w.ptr;
        d.m = (float *)m.ptr;
        d.v = (float *)v.ptr;
        d.wt = (bf16 *)wt.ptr;
        d.wb = (bf16 *)wb.ptr;
        d.grad_out = grad_host ? (float *)grad.ptr : nullptr;
        d.update = update ? 1u : 0u;
        if (update) ++step;
        d.lr = lr;
        d.b1 = b1;
        d.b2 = b2;
        d.eps = eps;
        d.c1 = 1.0f - std::pow(b1, (float)std::max(1u, step));
        d.c2 = 1.0f - std::pow(b2, (float)std::max(1u, step));
        hipLaunchKernelGGL(k_mlp_adam, dim3((kP + 255) / 256), dim3(256), 0, s, d);
        int rc = mlp_status("mlp backward");
        if (rc) return rc;
        if (grad_host) {
            SOLHIP_HIP_CHECK(hipMemcpyAsync(grad_host, grad.ptr, kP * 4, hipMemcpyDeviceToHost, s));
            SOLHIP_HIP_CHECK(hipStreamSynchronize(s));
        }
        return 0;
    }
};


Mlp *mlp_new(Context *c, uint32_t seed, float lr);

int mlp_query(hipStream_t s, Mlp *m, const uint16_t *feat, const uint32_t *n_dev, uint32_t n_max,
              const CacheQuery *qrec, float4 *st1)
{
    FwdArgs a;
    std::memset(&a, 0, sizeof(a));
    a.x = (const bf16 *)feat;
    a.n = n_max;
    a.npad = (n_max + 31u) & ~31u;
    a.n_dev = n_dev;
    a.wt = (const bf16 *)m->wt.ptr;
    a.qrec = qrec;
    a.st1 = st1;
    const uint32_t blocks = std::min<uint32_t>((a.npad / 16 + 3) / 4, m->ncu * 4);
    hipLaunchKernelGGL(k_mlp_fwd, dim3(std::max(1u, blocks)), dim3(kMlpBlock), 0, s, a);
    return mlp_status("cache query");
}

int mlp_train_dev(hipStream_t s, Mlp *m, const uint16_t *feat, const float4 *target, uint32_t n, uint32_t steps)
{
    if (n == 0) return 0;
    if (m->ybuf.reserve((uint64_t)n * 16)) return SOLHIP_ERROR_OUT_OF_MEMORY;
    for (uint32_t k = 0; k < steps; ++k) {
        int rc = m->forward(s, (const bf16 *)feat, n, (float4 *)m->ybuf.ptr, true);
        if (rc) return rc;
        rc = m->backward(s, n, (const float4 *)m->ybuf.ptr, target, true, nullptr);
        if (rc) return rc;
    }
    return 0;
}

void mlp_delete(Mlp *m)
{
    if (!m) return;
    m->release_all();
    delete m;
}

// He-uniform hidden layers, Glorot-uniform output layer, from a PCG stream
static void init_params(uint32_t seed, std::vector<float> &w)
{
    w.assign(kP, 0.0f);
    uint64_t st = 0x853c49e6748fea9bull ^ ((uint64_t)seed << 1);
    auto next = [&]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        const uint32_t x = (uint32_t)(((st >> 18) ^ st) >> 27), rot = (uint32_t)(st >> 59);
        return ((x >> rot) | (x << ((32 - rot) & 31))) * 0x1p-32f;   // [0, 1)
    };
    const float bh = std::sqrt(6.0f / 64.0f), bo = std::sqrt(6.0f / (64.0f + 3.0f));
    for (uint32_t p = 0; p < 16384; ++p) w[p] = (2.0f * next() - 1.0f) * bh;
    for (uint32_t k = 0; k < 64; ++k)
        for (uint32_t n = 0; n < (uint32_t)kOut; ++n) w[16384 + k * 16 + n] = (2.0f * next() - 1.0f) * bo;
}

static int mlp_sync_from_master(hipStream_t s, Mlp *m)
{
    hipLaunchKernelGGL(k_mlp_sync_bf16, dim3((kP + 255) / 256), dim3(256), 0, s, (const float *)m->w.ptr,
                       (bf16 *)m->wt.ptr, (bf16 *)m->wb.ptr);
    return mlp_status("mlp weights");
}

Mlp *mlp_new(Context *c, uint32_t seed, float lr)
{
    Mlp *m = new Mlp();
    m->kind = Kind::Mlp;
    m->lr = lr;
    m->ncu = c->ncu;
    if (m->w.reserve(kP * 4) || m->m.reserve(kP * 4) || m->v.reserve(kP * 4) || m->wt.reserve(kP * 2) ||
        m->wb.reserve(kP * 2) || m->loss.reserve(256) || m->grad.reserve(kP * 4)) {
        mlp_delete(m);
        set_error("mlp allocation failed");
        return nullptr;
    }
    std::vector<float> w;
    init_params(seed, w);
    std::vector<float> z(kP, 0.0f);
    if (hipMemcpy(m->w.ptr, w.data(), kP * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->m.ptr, z.data(), kP * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->v.ptr, z.data(), kP * 4, hipMemcpyHostToDevice) != hipSuccess ||
        mlp_sync_from_master(c->stream, m) || hipStreamSynchronize(c->stream) != hipSuccess) {
        mlp_delete(m);
        set_error("mlp initialisation failed");
        return nullptr;
    }
    return m;
}

}  // namespace solhip

using namespace solhip;

namespace {

Mlp *mlp_of(solhip_mlp h) { return get<Mlp>(h, Kind::Mlp); }

}  // namespace

extern "C" {

int32_t solhip_mlp_create(solhip_ctx ctx, uint32_t seed, float learning_rate, solhip_mlp *out)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    if (!out || !(learning_rate > 0.0f)) return bad_arg("out / learning_rate");
    Mlp *m = mlp_new(c, seed, learning_rate);
    if (!m) return SOLHIP_ERROR_DEVICE;
    m->ctx = ctx;
    *out = reg(m);
    return SOLHIP_OK;
}

int32_t solhip_mlp_destroy(solhip_ctx ctx, solhip_mlp h)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    Mlp *m = static_cast<Mlp *>(unreg(h, Kind::Mlp));
    if (!m) return bad_handle("mlp");
    (void)hipDeviceSynchronize();
    mlp_delete(m);
    return SOLHIP_OK;
}

int32_t solhip_mlp_params(solhip_ctx ctx, solhip_mlp h, float *get_out, const float *set_in)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    Mlp *m = mlp_of(h);
    if (!m) return bad_handle("mlp");
    SOLHIP_HIP_CHECK(hipDeviceSynchronize());
    if (set_in) {
        std::vector<float> w(set_in, set_in + kP);
        for (uint32_t p = 16384; p < kP; ++p)
            if (((p - 16384) & 15u) >= (uint32_t)kOut) w[p] = 0.0f;
        SOLHIP_HIP_CHECK(hipMemcpy(m->w.ptr, w.data(), kP * 4, hipMemcpyHostToDevice));
        int rc = mlp_sync_from_master(c->stream, m);
        if (rc) return rc;
        SOLHIP_HIP_CHECK(hipStreamSynchronize(c->stream));
    }
    if (get_out) SOLHIP_HIP_CHECK(hipMemcpy(get_out, m->w.ptr, kP * 4, hipMemcpyDeviceToHost));
    return SOLHIP_OK;
}

static int mlp_io(Context *c, solhip_buffer in, uint64_t in_bytes, solhip_buffer o, uint64_t o_bytes, void **pin,
                  void **po)
{
    Buffer *bi = get<Buffer>(in, Kind::Buffer);
    Buffer *bo = get<Buffer>(o, Kind::Buffer);
    (void)c;
    if (!bi || !bo) return bad_handle("buffer");
    if (bi->bytes < in_bytes || bo->bytes < o_bytes) return bad_arg("buffer too small");
    *pin = bi->ptr;
    *po = bo->ptr;
    return 0;
}

int32_t solhip_mlp_forward(solhip_ctx ctx, solhip_stream st, solhip_mlp h, solhip_buffer in, uint32_t n,
                           solhip_buffer out)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    bool ok;
    hipStream_t s = stream_of(c, st, ok);
    if (!ok) return bad_handle("stream");
    Mlp *m = mlp_of(h);
    if (!m) return bad_handle("mlp");
    if (n == 0) return SOLHIP_OK;
    void *pi, *po;
    int rc = mlp_io(c, in, (uint64_t)n * 128, out, (uint64_t)n * 16, &pi, &po);
    if (rc) return rc;
    return m->forward(s, (const bf16 *)pi, n, (float4 *)po, false);
}

static int32_t mlp_fit(solhip_ctx ctx, solhip_stream st, solhip_mlp h, solhip_buffer in, solhip_buffer target,
                       uint32_t n, uint32_t steps, bool update, float *grad_host, float *loss_out)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    bool ok;
    hipStream_t s = stream_of(c, st, ok);
    if (!ok) return bad_handle("stream");
    Mlp *m = mlp_of(h);
    if (!m) return bad_handle("mlp");
    if (n == 0 || steps == 0) return bad_arg("n / steps");
    void *pi, *pt;
    int rc = mlp_io(c, in, (uint64_t)n * 128, target, (uint64_t)n * 16, &pi, &pt);
    if (rc) return rc;
    if (m->ybuf.reserve((uint64_t)n * 16)) return SOLHIP_ERROR_OUT_OF_MEMORY;
    for (uint32_t k = 0; k < steps; ++k) {
        rc = launch_fill(s, (uint32_t *)m->loss.ptr, 1, 0u);
        if (rc) return rc;
        rc = m->forward(s, (const bf16 *)pi, n, (float4 *)m->ybuf.ptr, true);
        if (rc) return rc;
        rc = m->backward(s, n, (const float4 *)m->ybuf.ptr, (const float4 *)pt, update,
                         k + 1 == steps ? grad_host : nullptr);
        if (rc) return rc;
    }
    if (loss_out) {
        SOLHIP_HIP_CHECK(hipMemcpyAsync(loss_out, m->loss.ptr, 4, hipMemcpyDeviceToHost, s));
        SOLHIP_HIP_CHECK(hipStreamSynchronize(s));
    }
    return SOLHIP_OK;
}

int32_t solhip_mlp_train(solhip_ctx ctx, solhip_stream st, solhip_mlp h, solhip_buffer in, solhip_buffer target,
                         uint32_t n, uint32_t steps, float *loss)
{
    return mlp_fit(ctx, st, h, in, target, n, steps, true, nullptr, loss);
}

int32_t solhip_mlp_gradient(solhip_ctx ctx, solhip_stream st, solhip_mlp h, solhip_buffer in, solhip_buffer target,
                            uint32_t n, float *grad, float *loss)
{
    if (!grad) return bad_arg("grad");
    return mlp_fit(ctx, st, h, in, target, n, 1, false, grad, loss);
}

}  // extern "C"
