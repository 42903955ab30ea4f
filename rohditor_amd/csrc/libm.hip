// Device side of libm_f32.hpp: the log2 exception bitmap, and a self-test
// entry point that evaluates the libm replicas on the device.
//
// The build lists the inputs in [2^-8, 2^8) where the host libm's log2f is
// not the series value's rounding (gen_log2_table.cc); the list is compiled
// in here. rhd_create expands it once per device into a bitmap of one bit per
// f32 in that range (2^27 bits, 16 MiB), so that a lookup is one load.
#include "libm_f32.hpp"
#include "rhd_context.hpp"

namespace rhd {
namespace {

using namespace libm;

static __device__ const uint32_t log2_list[] = {
#include "log2f_exceptions.inc"
};
constexpr uint32_t LIST_COUNT = sizeof(log2_list) / sizeof(log2_list[0]);

static __device__ __align__(256) uint32_t log2_bitmap[LOG2_BITMAP_WORDS];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
const uint32_t* g_bitmap[MAX_DEVICES] = {};

// One thread per bitmap word: the listed inputs in its 32-value span, found
// by binary search in the sorted list. The list is read by name here, in
// device code (which also keeps it in the device code object); only the
// bitmap's address comes from the host.
__global__ __launch_bounds__(256) void libm_fill_bitmap(uint32_t* __restrict__ bitmap) {
    const uint64_t w = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (w >= LOG2_BITMAP_WORDS) return;
    const uint32_t* list = log2_list;
    const uint32_t count = LIST_COUNT;
    const uint32_t lo = LOG2_FIRST + (uint32_t)(w * 32), hi = lo + 32;
    uint32_t a = 0, b = count;
    while (a < b) {
        const uint32_t m = (a + b) >> 1;
        if (list[m] < lo)
            a = m + 1;
        else
            b = m;
    }
    uint32_t word = 0;
    for (uint32_t i = a; i < count && list[i] < hi; ++i) word |= 1u << (list[i] - lo);
    bitmap[w] = word;
}

__global__ __launch_bounds__(256) void libm_eval_kernel(int fn, const float* __restrict__ in,
                                                        float* __restrict__ out, uint64_t n,
                                                        const uint32_t* __restrict__ bitmap) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float x = in[i];
    out[i] = fn == 0 ? cbrt_libm(x) : log2_libm(x, bitmap);
}

}  // namespace

// Builds the log2 bitmap on `device`, which must be current. Idempotent.
rhd_status libm_setup(rhd_context* ctx, int device) {
    if (device < 0 || device >= MAX_DEVICES) return fail(ctx, RHD_UNSUPPORTED, "device index");
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_bitmap[device]) return RHD_OK;
    void* bitmap = nullptr;
    RHD_HIP(ctx, hipGetSymbolAddress(&bitmap, HIP_SYMBOL(log2_bitmap)));
    if (!bitmap) return hip_fail(ctx, hipErrorInvalidSymbol, "libm bitmap");
    const unsigned blocks = (unsigned)((LOG2_BITMAP_WORDS + 255) / 256);
    hipLaunchKernelGGL(libm_fill_bitmap, dim3(blocks), dim3(256), 0, ctx->lanes[0].stream,
                       (uint32_t*)bitmap);
    RHD_HIP(ctx, hipGetLastError());
    RHD_HIP(ctx, hipStreamSynchronize(ctx->lanes[0].stream));
    g_bitmap[device] = (const uint32_t*)bitmap;
    return RHD_OK;
}

const uint32_t* libm_log2_bitmap(int device) {
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_mu);
    return g_bitmap[device];
}

}  // namespace rhd

using namespace rhd;

extern "C" rhd_status rhd_libm_eval(rhd_context* ctx, int32_t fn, const float* in, float* out,
                                    uint64_t count) {
    if (!ctx || (fn != 0 && fn != 1) || (count && (!in || !out))) return RHD_INVALID_INPUT;
    if (count == 0) return RHD_OK;
    RHD_HIP(ctx, hipSetDevice(ctx->device));
    const uint32_t* bitmap = libm_log2_bitmap(ctx->device);
    if (!bitmap) return fail(ctx, RHD_SYNCHRONIZATION, "libm tables are not set up");
    Lane& L = ctx->lanes[0];
    const uint64_t chunk = std::min<uint64_t>(count, 1ull << 26);
    void* d = nullptr;
    rhd_status st = device_alloc(ctx, &L, 2 * chunk * sizeof(float), &d);
    if (st != RHD_OK) return st;
    float* din = (float*)d;
    float* dout = din + chunk;
    hipError_t e = hipSuccess;
    for (uint64_t base = 0; base < count && e == hipSuccess; base += chunk) {
        const uint64_t n = std::min(chunk, count - base);
        e = hipMemcpyAsync(din, in + base, n * sizeof(float), hipMemcpyHostToDevice, L.stream);
        if (e != hipSuccess) break;
        hipLaunchKernelGGL(libm_eval_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           L.stream, fn, din, dout, n, bitmap);
        e = hipGetLastError();
        if (e != hipSuccess) break;
        e = hipMemcpyAsync(out + base, dout, n * sizeof(float), hipMemcpyDeviceToHost, L.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(L.stream);
    }
    (void)hipStreamSynchronize(L.stream);
    device_free(ctx, d, 2 * chunk * sizeof(float));
    if (e != hipSuccess) return hip_fail(ctx, e, "libm self-test");
    return RHD_OK;
}
