// Normalization fused with Clip or Opposed detection.
//
// crates/core/src/pipeline/sensor.rs:221-277: out = (f32(u16) - black) /
// (white - black), unclamped. crates/highlight/src/clip.rs:24-72 with the
// predicates of detect.rs:5-16: affected if s >= limit, over-white if s > 1,
// changed (and clipped to limit) if s > limit. Opposed detection counts
// suspected sites (s >= level) per channel (opposed.rs:182-215).
//
// HBM-bound: 2 B read and 4 B written per sample. Each thread handles 4
// consecutive samples of a row; counters are reduced per wavefront and then
// per block before one 64-bit atomic per counter.
#include "rhd_internal.hpp"

namespace rhd {
namespace {

constexpr int BLOCK = 256;
constexpr int PER_THREAD = 4;

struct Counters {
    unsigned c[6];
};

__device__ void flush_counters(Counters& k, int n, unsigned long long* dst) {
    __shared__ unsigned partial[BLOCK / 64][6];
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int i = 0; i < n; ++i) {
        unsigned v = wave_sum(k.c[i]);
        if (lane == 0) partial[wave][i] = v;
    }
    __syncthreads();
    if (threadIdx.x < n) {
        unsigned long long total = 0;
        for (int w = 0; w < BLOCK / 64; ++w) total += partial[w][threadIdx.x];
        if (total) atomicAdd(dst + threadIdx.x, total);
    }
}

__device__ __forceinline__ void classify(float& v, int c, int mode, const float levels[3],
                                         Counters& k) {
    float limit = levels[c];
    if (mode == 1) {
        if (v >= limit) {
            k.c[0]++;
            k.c[3 + c]++;
        }
        if (v > 1.0f) k.c[2]++;
        if (v > limit) {
            k.c[1]++;
            v = limit;
        }
    } else if (mode == 2) {
        if (v >= limit) {
            k.c[0]++;
            k.c[3 + c]++;
        }
    }
}

__global__ __launch_bounds__(BLOCK) void normalize_kernel(NormalizeArgs a, float* __restrict__ out,
                                                          DeviceStats* stats) {
    const int y = blockIdx.y;
    const int x0 = (blockIdx.x * BLOCK + threadIdx.x) * PER_THREAD;
    Counters k = {{0, 0, 0, 0, 0, 0}};
    const int sy = a.crop_y + y;
    const int ty = sy % a.table.period_y;
    const uint16_t* src = a.mosaic + (uint64_t)sy * a.row_stride + a.crop_x;
    float* dst = out + (size_t)y * a.width;
#pragma unroll
    for (int i = 0; i < PER_THREAD; ++i) {
        int x = x0 + i;
        if (x < a.width) {
            int sx = a.crop_x + x;
            int t = ty * a.table.period_x + sx % a.table.period_x;
            float v = ((float)src[x] - a.table.black[t]) / a.table.range[t];
            classify(v, color_at(a.crop_pattern, x, y), a.mode, a.levels, k);
            dst[x] = v;
        }
    }
    if (a.mode == 1) flush_counters(k, 6, stats->clip);
    if (a.mode == 2) flush_counters(k, 6, stats->detect);
}

// Clip or detection on an existing f32 mosaic (the staged apply_highlight).
__global__ __launch_bounds__(BLOCK) void classify_kernel(float* __restrict__ plane, int width,
                                                         int height, int pattern, int mode,
                                                         float l0, float l1, float l2,
                                                         DeviceStats* stats) {
    const int y = blockIdx.y;
    const int x0 = (blockIdx.x * BLOCK + threadIdx.x) * PER_THREAD;
    const float levels[3] = {l0, l1, l2};
    Counters k = {{0, 0, 0, 0, 0, 0}};
    unsigned bad = 0;
    float* row = plane + (size_t)y * width;
#pragma unroll
    for (int i = 0; i < PER_THREAD; ++i) {
        int x = x0 + i;
        if (x < width) {
            float v = row[x];
            if (!is_finite(v)) {
                bad = 1;
                continue;
            }
            float before = v;
            classify(v, color_at(pattern, x, y), mode, levels, k);
            if (v != before) row[x] = v;
        }
    }
    if (bad) atomicOr(&stats->nonfinite, 1u);
    flush_counters(k, 6, mode == 1 ? stats->clip : stats->detect);
}

}  // namespace

hipError_t launch_normalize(const NormalizeArgs& a, float* out, DeviceStats* stats,
                            hipStream_t s, Profiler* prof) {
    dim3 grid((a.width + BLOCK * PER_THREAD - 1) / (BLOCK * PER_THREAD), a.height);
    RHD_TIMED(prof, s, "normalize", 6ull * a.width * a.height,
              hipLaunchKernelGGL(normalize_kernel, grid, dim3(BLOCK), 0, s, a, out, stats));
    return hipGetLastError();
}

hipError_t launch_clip(float* plane, int width, int height, int pattern, const float limits[3],
                       DeviceStats* stats, hipStream_t s) {
    dim3 grid((width + BLOCK * PER_THREAD - 1) / (BLOCK * PER_THREAD), height);
    hipLaunchKernelGGL(classify_kernel, grid, dim3(BLOCK), 0, s, plane, width, height, pattern, 1,
                       limits[0], limits[1], limits[2], stats);
    return hipGetLastError();
}

hipError_t launch_detect(const float* plane, int width, int height, int pattern,
                         const float levels[3], DeviceStats* stats, hipStream_t s) {
    dim3 grid((width + BLOCK * PER_THREAD - 1) / (BLOCK * PER_THREAD), height);
    hipLaunchKernelGGL(classify_kernel, grid, dim3(BLOCK), 0, s, const_cast<float*>(plane), width,
                       height, pattern, 2, levels[0], levels[1], levels[2], stats);
    return hipGetLastError();
}

}  // namespace rhd
