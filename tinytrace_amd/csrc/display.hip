// Display pass: the reference's post-process shader
// (src/post_process/display.glsl:6-12) run headless.  Per pixel: the
// accumulated RGBA32F radiance, neutral_tonemap when constants.tonemap != 0
// (src/post_process/tonemap.glsl:1-26), then the sRGB transfer function and
// 8-bit quantisation that the reference's sRGB swapchain store applies, into
// an RGBA8 image (alpha 255).  Timed as "post_process", as
// Renderer::prepare_to_present brackets its dispatch (src/lib.rs:180-193).
#include <hip/hip_runtime.h>

#include "abi_guard.h"
#include "device_math.h"
#include "runtime_internal.h"

using namespace ttd;

namespace {

TT_DEV float tonemap_curve(float v) { return 1.0f - ttm::exp(-v); }

// tonemap.glsl:1-26.  rgb_to_ycbcr is color * mat3(...) with the matrix
// given by columns: component k = dot(color, column k).
TT_DEV f3 neutral_tonemap(f3 c) {
    float y = c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f;
    float cb = c.x * -0.1146f + c.y * -0.3854f + c.z * 0.5f;
    float cr = c.x * 0.5f + c.y * -0.4542f + c.z * -0.0458f;
    float bt = tonemap_curve(sqrt_(cb * cb + cr * cr) * 2.4f);
    float desat = fmax_((bt - 0.7f) * 0.8f, 0.0f);
    desat = desat * desat;
    f3 desat_color = mixv(c, f3(y), desat);
    float tm_lum = tonemap_curve(y);
    float lum = 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z;
    f3 tm0 = c * fmax_(0.0f, tm_lum / fmax_(1e-5f, lum));
    f3 tm1(tonemap_curve(desat_color.x), tonemap_curve(desat_color.y), tonemap_curve(desat_color.z));
    return mixv(tm0, tm1, bt * bt) * 0.97f;
}

// Linear [0, 1] -> sRGB code value, rounded to nearest (NaN -> 0).
TT_DEV uint32_t srgb8(float v) {
    v = v > 0.0f ? (v < 1.0f ? v : 1.0f) : 0.0f;
    float s = v <= 0.0031308f ? 12.92f * v : 1.055f * ttm::pow(v, 1.0f / 2.4f) - 0.055f;
    return uint32_t(s * 255.0f + 0.5f);
}

__global__ void display_kernel(const float4* target, uint32_t count, uint32_t tonemap, uint32_t* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float4 p = target[i];
    f3 c(p.x, p.y, p.z);
    if (tonemap != 0u) c = neutral_tonemap(c);
    out[i] = srgb8(c.x) | srgb8(c.y) << 8 | srgb8(c.z) << 16 | 0xff000000u;
}

}  // namespace

extern "C" tt_status tt_target_display(tt_context* c, const tt_target* t, const tt_constants* constants,
                                       uint8_t* rgba8, size_t bytes) try {
    using tt::fail;
    if (!c || !t || !constants || !rgba8) return fail(c, TT_INVALID_ARGUMENT, "null argument");
    if (t->context != c) return fail(c, TT_STALE_HANDLE, "target belongs to another context");
    size_t count = size_t(t->width) * t->height;
    if (bytes != count * 4) return fail(c, TT_INVALID_ARGUMENT, "display buffer size");
    TT_HIP(c, hipSetDevice(c->device));
    if (c->display_capacity < count) {
        if (c->display_memory) TT_HIP(c, hipFree(c->display_memory));
        c->display_memory = nullptr;
        c->display_capacity = 0;
        TT_HIP(c, hipMalloc(&c->display_memory, count * 4));
        c->display_capacity = count;
    }
    if (!c->post_process.begin) {
        TT_HIP(c, hipEventCreate(&c->post_process.begin));
        TT_HIP(c, hipEventCreate(&c->post_process.end));
    }
    TT_HIP(c, hipEventRecord(c->post_process.begin, tt::api_stream(c)));
    display_kernel<<<uint32_t((count + 255) / 256), 256, 0, tt::api_stream(c)>>>(
        t->pixels, uint32_t(count), constants->tonemap, static_cast<uint32_t*>(c->display_memory));
    TT_HIP(c, hipGetLastError());
    TT_HIP(c, hipEventRecord(c->post_process.end, tt::api_stream(c)));
    c->post_process_recorded = true;
    TT_HIP(c, hipMemcpyAsync(rgba8, c->display_memory, count * 4, hipMemcpyDeviceToHost, tt::api_stream(c)));
    TT_HIP(c, hipStreamSynchronize(tt::api_stream(c)));
    return TT_OK;
} catch (...) {
    return tt::exception_status(c ? &c->error : nullptr);
}
