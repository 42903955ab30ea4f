// Capture sharpening without allocation: the variant of rhd_capture for
// callers that cannot call hipMalloc (code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation). It runs the same kernels
// as rhd_capture (capture.hip) on a caller-owned planar RGB buffer, with its
// working planes and device statistics in `scratch`, a fixed array that each
// GPU holds from the moment this code object is loaded until the process
// ends.
//
// Host code never passes `scratch` itself: its name is a host-side stand-in.
// rhd_create (the library's set-up call) resolves the array's device address
// once per device with hipGetSymbolAddress while that device is current, and
// every later use goes through that address. Calls on one GPU share the
// array and leave their contents in it, so overlapping calls on one GPU race;
// the statistics block is cleared by a fill kernel at the start of each call
// and every working plane is written before it is read.
#include "rhd_context.hpp"

namespace rhd {

int gaussian_weights(float sigma, float* w);  // api.cpp

namespace {

constexpr uint64_t MAX_PIXELS = RHD_NOALLOC_MAX_PIXELS;
constexpr size_t STATS_BYTES = 256;
constexpr int PLANES = 5;  // guide, mask, two estimates, temporary
constexpr size_t SCRATCH_BYTES = STATS_BYTES + (size_t)PLANES * MAX_PIXELS * sizeof(float);
static_assert(sizeof(DeviceStats) <= STATS_BYTES, "statistics block outgrew its slot");

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_scratch[MAX_DEVICES] = {};

}  // namespace

// Resolves `scratch` on `device`, which must be current. Idempotent.
rhd_status noalloc_setup(rhd_context* ctx, int device) {
    if (device < 0 || device >= MAX_DEVICES) return fail(ctx, RHD_UNSUPPORTED, "device index");
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch[device]) return RHD_OK;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        if (e == hipSuccess) e = hipErrorInvalidSymbol;
        return hip_fail(ctx, e, "hipGetSymbolAddress(scratch)");
    }
    g_scratch[device] = ptr;
    return RHD_OK;
}

}  // namespace rhd

using namespace rhd;

extern "C" rhd_status rhd_capture_noalloc(rhd_context* ctx, float* planar_rgb, uint32_t width,
                                          uint32_t height, const rhd_detail_desc* dd) {
    if (!ctx || !planar_rgb || !dd) return RHD_INVALID_INPUT;
    rhd_status st = validate_detail(ctx, dd);
    if (st != RHD_OK) return st;
    if (dd->false_colour > 0 || noise_active(dd))
        return fail(ctx, RHD_INVALID_INPUT,
                    "rhd_capture_noalloc runs capture sharpening only (no false colour or noise "
                    "reduction)");
    if (!capture_active(dd)) return RHD_OK;
    if (std::ceil(3.0f * dd->sigma) > 4.0f)
        return fail(ctx, RHD_UNSUPPORTED_EDITS, "capture radius above 4 taps");
    if (width == 0 || height == 0)
        return fail(ctx, RHD_INVALID_DIMENSIONS, "empty image");
    const uint64_t n = (uint64_t)width * height;
    if (n > MAX_PIXELS)
        return fail(ctx, RHD_INVALID_DIMENSIONS,
                    "image of " + std::to_string(n) + " pixels exceeds the " +
                        std::to_string(MAX_PIXELS) + "-pixel scratch of rhd_capture_noalloc");
    RHD_HIP(ctx, hipSetDevice(ctx->device));
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        base = g_scratch[ctx->device];
    }
    if (!base) {
        if ((st = noalloc_setup(ctx, ctx->device)) != RHD_OK) return st;
        std::lock_guard<std::mutex> lock(g_mu);
        base = g_scratch[ctx->device];
    }
    Lane& L = ctx->lanes[0];
    DeviceStats* stats = (DeviceStats*)base;
    float* planes = (float*)((char*)base + STATS_BYTES);
    CaptureArgs a{};
    a.rgb = planar_rgb;
    a.width = (int)width;
    a.height = (int)height;
    a.amount = dd->amount;
    a.noise_protection = dd->noise_protection;
    for (int c = 0; c < 3; ++c) a.ceilings[c] = dd->ceilings[c];
    a.radius = gaussian_weights(dd->sigma, a.weights);
    a.iterations = dd->iterations;
    a.guide = planes;
    a.mask = planes + n;
    a.est_a = planes + 2 * n;
    a.est_b = planes + 3 * n;
    a.tmp = planes + 4 * n;
    a.stats = stats;
    RHD_HIP(ctx, launch_zero(stats, sizeof(DeviceStats), L.stream));
    hipError_t e = launch_capture(a, L.stream, nullptr);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(L.stream);
        return hip_fail(ctx, e, "capture");
    }
    // The lane's pinned readback slot was allocated when the context was made.
    RHD_HIP(ctx, hipMemcpyAsync(L.h_stats, stats, sizeof(DeviceStats), hipMemcpyDeviceToHost,
                                L.stream));
    RHD_HIP(ctx, hipStreamSynchronize(L.stream));
    return stats_to_metrics(ctx, *L.h_stats, nullptr, RHD_HIGHLIGHT_OFF);
}
