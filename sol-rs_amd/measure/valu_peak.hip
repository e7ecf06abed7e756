// valu_peak.hip -- the VALU issue rate a gfx950 SIMD sustains on fp32 FMAs:
// 8 independent FMA chains per lane, 8 waves per SIMD, 65536 iterations, in
// three instruction forms.  Prints wave-instructions per second and per SIMD
// per cycle at the nominal 2.4 GHz; the fastest form is the peak that the
// traversal kernels' SQ_INSTS_VALU is compared with (their issue fraction).
//
//     sol-rs_amd/bin/valu_peak      (built by sol-rs_amd/build.py)
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#define CHECK(x)                                                                       \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            return 1;                                                                  \
        }                                                                              \
    } while (0)

constexpr int kIters = 65536;
constexpr int kChains = 8;
constexpr int kBlock = 256;

// FORM 0: v_fma_f32 with the multiplier and addend from kernel arguments
// (VOP3); 1: v_fmac_f32 with them in VGPRs (VOP2); 2: v_fmaak_f32 with a
// literal addend (VOP2 + literal).  The peak is the fastest form.
template <int FORM>
__global__ void __launch_bounds__(kBlock) k_fma_chains(float *out, float m, float c)
{
    const uint32_t gid = blockIdx.x * kBlock + threadIdx.x;
    float a[kChains], mv[kChains], cv[kChains];
    for (int k = 0; k < kChains; ++k) {
        a[k] = (float)(gid + k) * 1e-3f;
        mv[k] = m + (float)k * 1e-9f;
        cv[k] = c + (float)(gid & 3u) * 1e-9f;
    }
    for (int i = 0; i < kIters; ++i) {
#pragma unroll
        for (int k = 0; k < kChains; ++k) {
            if (FORM == 0) a[k] = fmaf(a[k], m, c);
            if (FORM == 1) a[k] = fmaf(mv[k], cv[k], a[k]) * 0.5f;
            if (FORM == 2) a[k] = fmaf(a[k], m, 1e-4f);
        }
    }
    float s = 0.0f;
    for (int k = 0; k < kChains; ++k) s += a[k];
    out[gid] = s;
}

template <int FORM>
static int time_form(int blocks, float *out, hipEvent_t e0, hipEvent_t e1, float &best)
{
    best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
        CHECK(hipEventRecord(e0));
        hipLaunchKernelGGL(k_fma_chains<FORM>, dim3(blocks), dim3(kBlock), 0, 0, out, 0.999f, 1e-4f);
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        float ms = 0.0f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (rep && ms < best) best = ms;
    }
    return 0;
}

int main()
{
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, 0));
    const int simds = prop.multiProcessorCount * 4;
    const int waves_per_simd = 8;
    const int blocks = simds * waves_per_simd / (kBlock / 64);
    float *out = nullptr;
    CHECK(hipMalloc(&out, (size_t)blocks * kBlock * sizeof(float)));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    float ms[3];
    if (time_form<0>(blocks, out, e0, e1, ms[0]) || time_form<1>(blocks, out, e0, e1, ms[1]) ||
        time_form<2>(blocks, out, e0, e1, ms[2]))
        return 1;
    // FMA wave-instructions per launch: 8 chains per lane, one FMA each per
    // iteration (FORM 1 adds one multiply per FMA, counted too)
    const double waves = (double)blocks * (kBlock / 64);
    const double inst[3] = {waves * kIters * kChains, waves * kIters * kChains * 2, waves * kIters * kChains};
    double best = 0.0;
    std::printf("{\"simds\": %d, \"waves\": %.0f, \"forms\": [", simds, waves);
    for (int f = 0; f < 3; ++f) {
        const double r = inst[f] / (ms[f] * 1e-3);
        if (r > best) best = r;
        std::printf("%s{\"form\": %d, \"wave_instructions\": %.0f, \"ms\": %.4f, \"per_simd_cycle_at_2400MHz\": %.4f}",
                    f ? ", " : "", f, inst[f], ms[f], r / simds / 2.4e9);
    }
    std::printf("], \"peak_wave_instr_per_s\": %.4e, \"peak_per_simd_cycle_at_2400MHz\": %.4f}\n", best,
                best / simds / 2.4e9);
    CHECK(hipFree(out));
    return 0;
}
