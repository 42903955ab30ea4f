// False-colour suppression, the first step of the detail stage
// (crates/detail/src/false_color.rs:45-135): colour differences of the
// balanced image, `steps` 3x3 medians of each (crates/detail/src/
// median.rs:22-110, with the reference's selection network so that ties
// resolve the same way), then red and blue rebuilt from green. Planar RGB in
// place; four scratch planes (two differences, two median outputs).
#include "rhd_internal.hpp"

namespace rhd {
namespace {

struct Fc {
    float balance[3];
    float ceilings[3];
};

__global__ __launch_bounds__(256) void fc_diff(const float* __restrict__ rgb, size_t n, Fc p,
                                               float* __restrict__ c1, float* __restrict__ c2) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float g = rgb[n + i] * p.balance[1];
    const float d1 = rgb[i] * p.balance[0] - g;
    const float d2 = rgb[2 * n + i] * p.balance[2] - g;
    c1[i] = is_finite(d1) ? d1 : 0.0f;
    c2[i] = is_finite(d2) ? d2 : 0.0f;
}

__device__ __forceinline__ float med3(float a, float b, float c) {
    return sel_max(sel_min(a, b), sel_min(sel_max(a, b), c));
}

// Median of the 3x3 window whose columns (top to bottom) are given.
__device__ __forceinline__ float median_of_columns(const float col[3][3]) {
    float low[3], mid[3], high[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        low[k] = sel_min(sel_min(col[k][0], col[k][1]), col[k][2]);
        mid[k] = med3(col[k][0], col[k][1], col[k][2]);
        high[k] = sel_max(sel_max(col[k][0], col[k][1]), col[k][2]);
    }
    return med3(sel_max(sel_max(low[0], low[1]), low[2]), med3(mid[0], mid[1], mid[2]),
                sel_min(sel_min(high[0], high[1]), high[2]));
}

// One median step of both difference planes; half-sample mirrored borders.
__global__ __launch_bounds__(256) void fc_median(const float* __restrict__ a,
                                                 const float* __restrict__ b,
                                                 float* __restrict__ oa, float* __restrict__ ob,
                                                 int width, int height) {
    const uint3 blk = xcd_block();
    const int x = blk.x * 64 + (threadIdx.x & 63);
    const int y = blk.y * 4 + (threadIdx.x >> 6);
    if (x >= width || y >= height) return;
    int rows[3], cols[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        rows[k] = mirror_near(y + k - 1, height) * width;
        cols[k] = mirror_near(x + k - 1, width);
    }
    float ca[3][3], cb[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            ca[k][r] = a[rows[r] + cols[k]];
            cb[k][r] = b[rows[r] + cols[k]];
        }
    const size_t o = (size_t)y * width + x;
    oa[o] = median_of_columns(ca);
    ob[o] = median_of_columns(cb);
}

__global__ __launch_bounds__(256) void fc_write(float* __restrict__ rgb, size_t n, Fc p,
                                                const float* __restrict__ c1,
                                                const float* __restrict__ c2) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float r0 = rgb[i], g0 = rgb[n + i], b0 = rgb[2 * n + i];
    const bool finite = is_finite(r0) && is_finite(g0) && is_finite(b0);
    const bool blown = r0 >= p.ceilings[0] || g0 >= p.ceilings[1] || b0 >= p.ceilings[2];
    if (!finite || blown) return;
    const float g = g0 * p.balance[1];
    const float r = (g + c1[i]) / p.balance[0];
    const float b = (g + c2[i]) / p.balance[2];
    if (is_finite(r) && is_finite(b)) {
        rgb[i] = r;
        rgb[2 * n + i] = b;
    }
}

}  // namespace

hipError_t launch_false_colour(float* rgb, int width, int height, int steps,
                               const float balance[3], const float ceilings[3], float* scratch,
                               hipStream_t s, Profiler* prof) {
    if (steps <= 0) return hipSuccess;
    const size_t n = (size_t)width * height;
    Fc p{};
    for (int c = 0; c < 3; ++c) {
        p.balance[c] = balance[c];
        p.ceilings[c] = ceilings[c];
    }
    float* c1 = scratch;
    float* c2 = scratch + n;
    float* t1 = scratch + 2 * n;
    float* t2 = scratch + 3 * n;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    RHD_TIMED(prof, s, "fc_diff", 20 * (uint64_t)n,
              fc_diff<<<blocks, 256, 0, s>>>(rgb, n, p, c1, c2));
    const dim3 grid((width + 63) / 64, (height + 3) / 4);
    for (int k = 0; k < steps; ++k) {
        RHD_TIMED(prof, s, "fc_median", 16 * (uint64_t)n,
                  fc_median<<<grid, 256, 0, s>>>(c1, c2, t1, t2, width, height));
        float* u = c1;
        c1 = t1;
        t1 = u;
        u = c2;
        c2 = t2;
        t2 = u;
    }
    RHD_TIMED(prof, s, "fc_write", 28 * (uint64_t)n,
              fc_write<<<blocks, 256, 0, s>>>(rgb, n, p, c1, c2));
    return hipGetLastError();
}

}  // namespace rhd
