// Bilinear and Malvar-He-Cutler demosaic, one output pixel per thread,
// planar RGB out.
//
// Bilinear: crates/demosaic/src/bilinear.rs:57-113 (neighbour means in the
// reference's fixed offset order, 0.0 start, divided by the count). MHC:
// crates/demosaic/src/malvar_he_cutler.rs:16-119 (row-major sum of the
// non-zero taps of three integer/16 kernels, times 1/16), bilinear in the
// 2-pixel border. The bilinear path also serves AMaZE and RCD on images too
// small for their tiles; then samples are scaled by the channel balance on
// load and the output carries the inverse gains (lib.rs:342-356).
//
// HBM-bound: 4 B read (neighbours come from cache) and 12 B written per pixel.
#include "rhd_internal.hpp"

namespace rhd {
namespace {

struct Src {
    const float* m;
    int w, h, pattern;
    float fwd[3];
    __device__ __forceinline__ float at(int x, int y) const {
        return m[(size_t)y * w + x] * fwd[color_at(pattern, x, y)];
    }
};

// bilinear.rs:82-113 for one offset list; offsets as (dx, dy).
template <int N>
__device__ __forceinline__ float average(const Src& s, int x, int y, const int (&o)[N][2]) {
    float sum = 0.0f;
    unsigned count = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        int nx = x + o[k][0], ny = y + o[k][1];
        if (nx >= 0 && ny >= 0 && nx < s.w && ny < s.h) {
            sum += s.at(nx, ny);
            ++count;
        }
    }
    return count == 0 ? s.at(x, y) : sum / (float)count;
}

__device__ __forceinline__ void bilinear_pixel(const Src& s, int x, int y, float rgb[3]) {
    const int cross[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
    const int diag[4][2] = {{-1, -1}, {1, -1}, {-1, 1}, {1, 1}};
    const int horiz[2][2] = {{-1, 0}, {1, 0}};
    const int vert[2][2] = {{0, -1}, {0, 1}};
    int c = color_at(s.pattern, x, y);
    if (c == 0) {
        rgb[0] = s.at(x, y);
        rgb[1] = average(s, x, y, cross);
        rgb[2] = average(s, x, y, diag);
    } else if (c == 2) {
        rgb[0] = average(s, x, y, diag);
        rgb[1] = average(s, x, y, cross);
        rgb[2] = s.at(x, y);
    } else {
        bool red_horizontal = color_at(s.pattern, x + 1, y) == 0;
        rgb[0] = red_horizontal ? average(s, x, y, horiz) : average(s, x, y, vert);
        rgb[1] = s.at(x, y);
        rgb[2] = red_horizontal ? average(s, x, y, vert) : average(s, x, y, horiz);
    }
}

__global__ __launch_bounds__(256) void bilinear_kernel(Src s, float g0, float g1, float g2,
                                                       float* __restrict__ out,
                                                       DeviceStats* stats) {
    const uint3 blk = xcd_block();
    int x = blk.x * 64 + (threadIdx.x & 63);
    int y = blk.y * 4 + (threadIdx.x >> 6);
    if (x >= s.w || y >= s.h) return;
    float rgb[3];
    bilinear_pixel(s, x, y, rgb);
    rgb[0] *= g0;
    rgb[1] *= g1;
    rgb[2] *= g2;
    size_t n = (size_t)s.w * s.h, i = (size_t)y * s.w + x;
    out[i] = rgb[0];
    out[n + i] = rgb[1];
    out[2 * n + i] = rgb[2];
    if (!(is_finite(rgb[0]) && is_finite(rgb[1]) && is_finite(rgb[2])))
        atomicOr(&stats->nonfinite, 2u);
}

// malvar_he_cutler.rs:16-38, stored x16 as small integers.
__constant__ signed char MHC_KERNELS[3][5][5] = {
    {{0, 0, -2, 0, 0}, {0, 0, 4, 0, 0}, {-2, 4, 8, 4, -2}, {0, 0, 4, 0, 0}, {0, 0, -2, 0, 0}},
    {{0, 0, 1, 0, 0}, {0, -2, 0, -2, 0}, {-2, 8, 10, 8, -2}, {0, -2, 0, -2, 0}, {0, 0, 1, 0, 0}},
    {{0, 0, -3, 0, 0}, {0, 4, 0, 4, 0}, {-3, 0, 12, 0, -3}, {0, 4, 0, 4, 0}, {0, 0, -3, 0, 0}}};

// malvar_he_cutler.rs:97-119; taps visited row-major over the (possibly
// transposed) kernel, skipping zero coefficients.
__device__ __forceinline__ float mhc_apply(const Src& s, int x, int y, int which, bool transpose) {
    float sum = 0.0f;
#pragma unroll
    for (int ky = 0; ky < 5; ++ky)
#pragma unroll
        for (int kx = 0; kx < 5; ++kx) {
            int c = transpose ? MHC_KERNELS[which][kx][ky] : MHC_KERNELS[which][ky][kx];
            if (c != 0) sum += (float)c * s.m[(size_t)(y + ky - 2) * s.w + (x + kx - 2)];
        }
    return sum * (1.0f / 16.0f);
}

__global__ __launch_bounds__(256) void mhc_kernel(Src s, float* __restrict__ out,
                                                  DeviceStats* stats) {
    const uint3 blk = xcd_block();
    int x = blk.x * 64 + (threadIdx.x & 63);
    int y = blk.y * 4 + (threadIdx.x >> 6);
    if (x >= s.w || y >= s.h) return;
    float rgb[3];
    bool border = x < 2 || y < 2 || x + 2 >= s.w || y + 2 >= s.h;
    if (border) {
        bilinear_pixel(s, x, y, rgb);
    } else {
        float measured = s.m[(size_t)y * s.w + x];
        int c = color_at(s.pattern, x, y);
        if (c == 0) {
            rgb[0] = measured;
            rgb[1] = mhc_apply(s, x, y, 0, false);
            rgb[2] = mhc_apply(s, x, y, 2, false);
        } else if (c == 2) {
            rgb[0] = mhc_apply(s, x, y, 2, false);
            rgb[1] = mhc_apply(s, x, y, 0, false);
            rgb[2] = measured;
        } else {
            bool red_horizontal = color_at(s.pattern, x + 1, y) == 0;
            rgb[0] = mhc_apply(s, x, y, 1, !red_horizontal);
            rgb[1] = measured;
            rgb[2] = mhc_apply(s, x, y, 1, red_horizontal);
        }
    }
    size_t n = (size_t)s.w * s.h, i = (size_t)y * s.w + x;
    out[i] = rgb[0];
    out[n + i] = rgb[1];
    out[2 * n + i] = rgb[2];
    if (!(is_finite(rgb[0]) && is_finite(rgb[1]) && is_finite(rgb[2])))
        atomicOr(&stats->nonfinite, 2u);
}

}  // namespace

hipError_t launch_bilinear(const float* mosaic, int width, int height, int pattern,
                           const float forward[3], const float gains[3], float* out,
                           DeviceStats* stats, hipStream_t s, Profiler* prof) {
    Src src{mosaic, width, height, pattern, {forward[0], forward[1], forward[2]}};
    dim3 grid((width + 63) / 64, (height + 3) / 4);
    RHD_TIMED(prof, s, "bilinear", 16ull * width * height,
              hipLaunchKernelGGL(bilinear_kernel, grid, dim3(256), 0, s, src, gains[0], gains[1],
                                 gains[2], out, stats));
    return hipGetLastError();
}

hipError_t launch_mhc(const float* mosaic, int width, int height, int pattern, float* out,
                      DeviceStats* stats, hipStream_t s, Profiler* prof) {
    Src src{mosaic, width, height, pattern, {1.0f, 1.0f, 1.0f}};
    dim3 grid((width + 63) / 64, (height + 3) / 4);
    RHD_TIMED(prof, s, "mhc", 16ull * width * height,
              hipLaunchKernelGGL(mhc_kernel, grid, dim3(256), 0, s, src, out, stats));
    return hipGetLastError();
}

}  // namespace rhd
