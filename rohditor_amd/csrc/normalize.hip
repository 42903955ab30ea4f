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
constexpr int PER_THREAD = 4, ROWS_IN_FLIGHT = 4;
// Row blocks per launch: each block strides over height / ROW_BLOCKS rows.
constexpr int ROW_BLOCKS = 256;

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

// One thread per 4 consecutive samples of a row, striding over rows so that
// each block flushes its counters once. The level table is staged in LDS and
// indexed by a phase that is stepped, not recomputed; rows whose u16 source
// is 8-byte aligned load 4 samples at once, and rows of a width that is a
// multiple of 4 store 4 floats at once.
__global__ __launch_bounds__(BLOCK) void normalize_kernel(NormalizeArgs a, float* __restrict__ out,
                                                          DeviceStats* stats) {
    __shared__ float s_black[256], s_range[256];
    const int px = a.table.period_x, entries = px * a.table.period_y;
    if ((int)threadIdx.x < entries) {
        s_black[threadIdx.x] = a.table.black[threadIdx.x];
        s_range[threadIdx.x] = a.table.range[threadIdx.x];
    }
    __syncthreads();
    const int x0 = (blockIdx.x * BLOCK + threadIdx.x) * PER_THREAD;
    Counters k = {{0, 0, 0, 0, 0, 0}};
    const bool full = x0 + PER_THREAD <= a.width;
    // The samples of ROWS_IN_FLIGHT rows are requested before the first of
    // them is converted, so each thread keeps that many loads in flight.
    auto load = [&](int y, uint16_t raw[PER_THREAD]) {
        const uint64_t first = (uint64_t)(a.crop_y + y) * a.row_stride + a.crop_x;
        const uint16_t* src = a.mosaic + first;
        if (full && (first & 3) == 0) {
            const ushort4 q = *reinterpret_cast<const ushort4*>(src + x0);
            raw[0] = q.x;
            raw[1] = q.y;
            raw[2] = q.z;
            raw[3] = q.w;
        } else {
#pragma unroll
            for (int i = 0; i < PER_THREAD; ++i) raw[i] = x0 + i < a.width ? src[x0 + i] : 0;
        }
    };
    auto convert = [&](int y, const uint16_t raw[PER_THREAD]) {
        const int sy = a.crop_y + y;
        const int row_t = (sy % a.table.period_y) * px;
        int phase = (a.crop_x + x0) % px;
        float* dst = out + (size_t)y * a.width;
        float v[PER_THREAD];
#pragma unroll
        for (int i = 0; i < PER_THREAD; ++i) {
            const int t = row_t + phase;
            v[i] = ((float)raw[i] - s_black[t]) / s_range[t];
            if (x0 + i < a.width) classify(v[i], color_at(a.crop_pattern, x0 + i, y), a.mode, a.levels, k);
            phase = phase + 1 == px ? 0 : phase + 1;
        }
        if (full && (a.width & 3) == 0) {
            *reinterpret_cast<float4*>(dst + x0) = make_float4(v[0], v[1], v[2], v[3]);
        } else {
#pragma unroll
            for (int i = 0; i < PER_THREAD; ++i)
                if (x0 + i < a.width) dst[x0 + i] = v[i];
        }
    };
    const int step = (int)gridDim.y;
    for (int y0 = blockIdx.y; y0 < a.height && x0 < a.width; y0 += ROWS_IN_FLIGHT * step) {
        uint16_t raw[ROWS_IN_FLIGHT][PER_THREAD];
#pragma unroll
        for (int u = 0; u < ROWS_IN_FLIGHT; ++u)
            if (y0 + u * step < a.height) load(y0 + u * step, raw[u]);
#pragma unroll
        for (int u = 0; u < ROWS_IN_FLIGHT; ++u)
            if (y0 + u * step < a.height) convert(y0 + u * step, raw[u]);
    }
    if (a.mode == 1) flush_counters(k, 6, stats->clip);
    if (a.mode == 2) flush_counters(k, 6, stats->detect);
}

// Clip or detection on an existing f32 mosaic (the staged apply_highlight).
__global__ __launch_bounds__(BLOCK) void classify_kernel(float* __restrict__ plane, int width,
                                                         int height, int pattern, int mode,
                                                         float l0, float l1, float l2,
                                                         DeviceStats* stats) {
    const int x0 = (blockIdx.x * BLOCK + threadIdx.x) * PER_THREAD;
    const float levels[3] = {l0, l1, l2};
    Counters k = {{0, 0, 0, 0, 0, 0}};
    unsigned bad = 0;
    for (int y = blockIdx.y; y < height; y += gridDim.y) {
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
    }
    if (bad) atomicOr(&stats->nonfinite, 1u);
    flush_counters(k, 6, mode == 1 ? stats->clip : stats->detect);
}

}  // namespace

namespace {
__global__ __launch_bounds__(BLOCK) void zero_kernel(unsigned* __restrict__ p, size_t words) {
    for (size_t i = blockIdx.x * (size_t)BLOCK + threadIdx.x; i < words; i += (size_t)gridDim.x * BLOCK)
        p[i] = 0u;
}
}  // namespace

hipError_t launch_zero(void* p, size_t bytes, hipStream_t s) {
    const size_t words = bytes / 4;
    const unsigned blocks = (unsigned)std::min<size_t>((words + BLOCK - 1) / BLOCK, 1024);
    hipLaunchKernelGGL(zero_kernel, dim3(blocks ? blocks : 1), dim3(BLOCK), 0, s, (unsigned*)p, words);
    return hipGetLastError();
}

hipError_t launch_normalize(const NormalizeArgs& a, float* out, DeviceStats* stats,
                            hipStream_t s, Profiler* prof) {
    dim3 grid((a.width + BLOCK * PER_THREAD - 1) / (BLOCK * PER_THREAD), std::min(a.height, ROW_BLOCKS));
    RHD_TIMED(prof, s, "normalize", 6ull * a.width * a.height,
              hipLaunchKernelGGL(normalize_kernel, grid, dim3(BLOCK), 0, s, a, out, stats));
    return hipGetLastError();
}

hipError_t launch_clip(float* plane, int width, int height, int pattern, const float limits[3],
                       DeviceStats* stats, hipStream_t s) {
    dim3 grid((width + BLOCK * PER_THREAD - 1) / (BLOCK * PER_THREAD), std::min(height, ROW_BLOCKS));
    hipLaunchKernelGGL(classify_kernel, grid, dim3(BLOCK), 0, s, plane, width, height, pattern, 1,
                       limits[0], limits[1], limits[2], stats);
    return hipGetLastError();
}

hipError_t launch_detect(const float* plane, int width, int height, int pattern,
                         const float levels[3], DeviceStats* stats, hipStream_t s) {
    dim3 grid((width + BLOCK * PER_THREAD - 1) / (BLOCK * PER_THREAD), std::min(height, ROW_BLOCKS));
    hipLaunchKernelGGL(classify_kernel, grid, dim3(BLOCK), 0, s, const_cast<float*>(plane), width,
                       height, pattern, 2, levels[0], levels[1], levels[2], stats);
    return hipGetLastError();
}

}  // namespace rhd
