// Progressive accumulation and 8-bit display of rendered frames: the tail of the
// reference's pathTracingKernel (src/core/RenderKernel.cu:202-217) and its buffer reset
// (src/core/Renderer.cpp:86-93). One thread per pixel over a grid-stride loop; the
// kernel is HBM-bound at 28 bytes per pixel (frame and running sum read, running sum
// and one packed colour written).
#include <cstdint>
#include <string>

#include "../host/tt_host.h"
#include "dev_common.h"
#include "scene_dev.h"
#include "ttmath.h"

namespace tt {
namespace {

// The reference clears its running sum with cudaMemset(buffer, 1, ...): every float
// of a fresh sum has the bit pattern 0x01010101.
constexpr uint32_t kResetBits = 0x01010101u;

__device__ __forceinline__ uint32_t display_channel(float sum, float frames) {
    const float mean = sum / frames;
    // clamp(x, 0, 1) as fmaxf(0, fminf(x, 1)): NaN becomes 1
    const float c = fmaxf(0.0f, fminf(mean, 1.0f));
    return (uint32_t)(uint8_t)(tm_pow(c, 1.0f / 2.2f) * 255.0f);
}

__global__ __launch_bounds__(256) void k_display(const float *__restrict__ frame, float *__restrict__ sum,
                                                 int64_t npix, int32_t frame_number, uint32_t *__restrict__ rgba) {
    const float frames = (float)frame_number;
    const float fresh = __uint_as_float(kResetBits);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += stride) {
        const float *f = frame + 3 * i;
        float *s = sum + 3 * i;
        const float r = (frame_number == 1 ? fresh : s[0]) + f[0];
        const float g = (frame_number == 1 ? fresh : s[1]) + f[1];
        const float b = (frame_number == 1 ? fresh : s[2]) + f[2];
        s[0] = r;
        s[1] = g;
        s[2] = b;
        rgba[i] = display_channel(r, frames) | display_channel(g, frames) << 8 | display_channel(b, frames) << 16 |
                  1u << 24;
    }
}

}  // namespace
}  // namespace tt

using namespace tt;

#define TT_HIPD(call, what)                                                                      \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) return fail(TT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e_)); \
    } while (0)

tt_status tt_display_frame(tt_ctx *ctx, const float *frame, float *sum, int64_t npix, int32_t frame_number,
                           uint32_t *rgba, void *stream) {
    if (!ctx || npix < 0 || (npix && (!frame || !sum || !rgba)) || frame_number < 1)
        return fail(TT_ERR_INVALID, "tt_display_frame: bad arguments (frame_number counts from 1)");
    if (!npix) return TT_OK;
    TT_HIPD(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    int64_t blocks = (npix + 255) / 256;
    if (blocks > (int64_t)ctx->num_cus * 16) blocks = (int64_t)ctx->num_cus * 16;
    hipLaunchKernelGGL(k_display, dim3((unsigned)blocks), dim3(256), 0, st, frame, sum, npix, frame_number, rgba);
    TT_HIPD(hipGetLastError(), "display launch");
    return TT_OK;
}
