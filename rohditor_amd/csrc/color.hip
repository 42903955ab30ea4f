// White balance and the camera matrix to linear Rec.2020 on a resident
// camera source, in place (crates/core/src/cpu/stages.rs:250-308): each
// channel is multiplied by its gain, then every output channel is the matrix
// row's dot product evaluated left to right (crates/color/src/matrix.rs:173),
// with no fused multiply-adds. One thread per pixel, the three planes read
// and written once: 24 bytes per pixel, HBM-bound.
#include "rhd_internal.hpp"

namespace rhd {
namespace {

struct ColorTransform {
    float gains[3];
    float m[9];
};

__global__ __launch_bounds__(256) void color_kernel(float* __restrict__ rgb, size_t n,
                                                    ColorTransform t) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float r = rgb[i] * t.gains[0];
    const float g = rgb[n + i] * t.gains[1];
    const float b = rgb[2 * n + i] * t.gains[2];
    rgb[i] = t.m[0] * r + t.m[1] * g + t.m[2] * b;
    rgb[n + i] = t.m[3] * r + t.m[4] * g + t.m[5] * b;
    rgb[2 * n + i] = t.m[6] * r + t.m[7] * g + t.m[8] * b;
}

}  // namespace

hipError_t launch_color(float* rgb, size_t n, const float gains[3], const float m[9],
                        hipStream_t s, Profiler* prof) {
    ColorTransform t{};
    for (int c = 0; c < 3; ++c) t.gains[c] = gains[c];
    for (int k = 0; k < 9; ++k) t.m[k] = m[k];
    const unsigned blocks = (unsigned)((n + 255) / 256);
    RHD_TIMED(prof, s, "color_kernel", 24 * (uint64_t)n,
              color_kernel<<<blocks, 256, 0, s>>>(rgb, n, t));
    return hipGetLastError();
}

}  // namespace rhd
