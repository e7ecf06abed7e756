// noalloc.hip -- solhip_trace_noalloc: solhip_trace with the AO wavefront's
// scratch (queue counter, fetch heads, queue entries, per-ray occlusion bytes)
// in one fixed device array instead of the program's own buffers, for callers
// that cannot allocate device memory (code under HIP graph capture, serving
// loops that forbid allocation).
//
// The array is a file-scope __device__ variable: every GPU the library runs on
// holds its own copy, at its own address, until the process ends.  In host
// code `scratch` names a host-side stand-in, not device memory, so its device
// address is taken once per device with hipGetSymbolAddress (at context
// creation, with that device current) and only that address is handed to
// kernels.  All calls on one GPU share the array: calls that overlap in time on
// one GPU (two streams, two host threads) race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "kernels.h"
#include "objects.h"

namespace solhip {

// a 1920x1080 AO frame with up to 32 AO rays per pixel fits: it needs
// 256 + 8 KiB + W*H*(32 + n_ao) bytes
constexpr uint64_t kNoAllocScratchBytes = SOLHIP_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

// device address of this device's copy of `scratch`; `device` must be the
// current device.  Called by solhip_context_create.
int noalloc_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) {
        set_error("device index beyond the fixed-scratch table");
        return SOLHIP_ERROR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> l(g_scratch_mu);
    if (g_scratch_ptr[device]) return SOLHIP_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        return SOLHIP_ERROR_DEVICE;
    }
    g_scratch_ptr[device] = ptr;
    return SOLHIP_OK;
}

static void *scratch_of(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> l(g_scratch_mu);
    return g_scratch_ptr[device];
}

}  // namespace solhip

using namespace solhip;

extern "C" {

int32_t solhip_trace_noalloc(solhip_ctx ctx, solhip_stream st, solhip_program p, solhip_tlas t,
                             const solhip_scene_tables *tables, const solhip_scene_uniforms *uniforms,
                             int32_t accum_start_frame, const solhip_images *images, uint32_t W, uint32_t H)
{
    Context *c = ctx_of(ctx);
    if (!c) return bad_handle("context");
    FixedScratch fixed;
    fixed.ptr = scratch_of(c->device);
    fixed.bytes = kNoAllocScratchBytes;
    if (!fixed.ptr) {
        set_error("the fixed scratch of this context's device was not set up");
        return SOLHIP_ERROR_DEVICE;
    }
    return trace_program(ctx, st, p, t, tables, uniforms, accum_start_frame, images, W, H, &fixed);
}

}  // extern "C"
