// Optics correction and exact area reduction of a resident camera source:
// the spatial step after the sensor stage (crates/optics/src/correction.rs,
// crates/core/src/resample.rs; the reference GPU executor is
// crates/gpu/src/spatial/spatial.wgsl + reduction.rs).
//
// Kernels, in launch order:
//   opt_vig_rows   PA vignetting's incremental radius (correction.rs:213-262)
//                  is a sequential f32 sum along each row. One thread per row
//                  runs it and keeps r^2 at every 16th column, so any pixel's
//                  exact r^2 is at most 15 additions away.
//   opt_vignette   (only when vignetting precedes a remap) vignetted copy of
//                  the three planes, which the cubic remap then samples.
//   opt_reduce_h   the corrected image is never stored: each workgroup
//                  computes the corrected pixels of one row segment into LDS
//                  (vignetting, inverse distortion and TCA in f64, 4x4
//                  Catmull-Rom sampling) and reduces them horizontally with
//                  the area weights, in tap order, for up to 256 output
//                  columns. With `direct` (full-resolution output) it writes
//                  the corrected pixels instead.
//   opt_reduce_v   vertical area pass over the horizontal results.
// Failures (an invalid vignetting gain, a non-finite value or coordinate, a
// cubic footprint outside the image) set bits in DeviceStats::nonfinite; the
// host turns them into an error and publishes nothing.
#include "optics_math.hpp"
#include "rhd_internal.hpp"

namespace rhd {

namespace {

constexpr int CK = 16;         // checkpoint spacing of the vignetting radius
constexpr int SEG = 256;       // corrected pixels per workgroup pass
constexpr unsigned BAD_VALUE = 8u;      // invalid gain or non-finite value
constexpr unsigned BAD_FOOTPRINT = 16u; // cubic footprint outside the image

struct Vig {
    const float* steps;  // per column: 2 ns x_k + ns^2 (f32, built on the host)
    const float* rows;   // r^2 at columns 0, 16, 32, ... per row: [ckpt][height]
    float k1, k2, k3;
};

__device__ __forceinline__ void flag(DeviceStats* s, unsigned bit) {
    atomicOr(&s->nonfinite, bit);
}

// r^2 of pixel (x, y) as the reference's row recurrence leaves it.
__device__ __forceinline__ float vig_radius(const Vig& v, int x, int y, int height) {
    float r2 = v.rows[(size_t)(x / CK) * height + y];
    for (int k = x - x % CK; k < x; ++k) r2 += v.steps[k];
    return r2;
}

// 1 / gain, or 0 with `bad` set for an invalid gain.
__device__ __forceinline__ float vig_multiplier(const Vig& v, float r2, bool& bad) {
    const float r4 = r2 * r2;
    const float r6 = r4 * r2;
    const float gain = 1.0f + v.k1 * r2 + v.k2 * r4 + v.k3 * r6;
    if (!optics::finite32(gain) || gain == 0.0f) {
        bad = true;
        return 0.0f;
    }
    return 1.0f / gain;
}

// One thread per row; the row-independent increments pass through LDS in
// chunks so that every lane reads the same word (a broadcast).
constexpr int ROW_CHUNK = 1024;
__global__ __launch_bounds__(64) void opt_vig_rows(const float* __restrict__ steps, float start_x,
                                                   float start_y, float ns, int width, int height,
                                                   float* __restrict__ rows) {
    __shared__ __align__(16) float chunk[ROW_CHUNK];
    const int y = blockIdx.x * 64 + threadIdx.x;
    const bool live = y < height;
    const float yy = start_y + ns * (float)(live ? y : 0);
    float r2 = start_x * start_x + yy * yy;
    for (int c0 = 0; c0 < width; c0 += ROW_CHUNK) {
        const int n = min(ROW_CHUNK, width - c0);
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += 64) chunk[i] = steps[c0 + i];
        __syncthreads();
        int k0 = 0;
        for (; k0 + CK <= n; k0 += CK) {
            if (live) rows[(size_t)((c0 + k0) / CK) * height + y] = r2;
            // 16 increments read ahead of the dependent additions.
            const float4* q = reinterpret_cast<const float4*>(chunk + k0);
            const float4 a = q[0], b = q[1], c = q[2], d = q[3];
            r2 += a.x; r2 += a.y; r2 += a.z; r2 += a.w;
            r2 += b.x; r2 += b.y; r2 += b.z; r2 += b.w;
            r2 += c.x; r2 += c.y; r2 += c.z; r2 += c.w;
            r2 += d.x; r2 += d.y; r2 += d.z; r2 += d.w;
        }
        if (k0 < n) {
            if (live) rows[(size_t)((c0 + k0) / CK) * height + y] = r2;
            for (int k = k0; k < n; ++k) r2 += chunk[k];
        }
    }
}

// Vignetted copy of the three planes. A workgroup takes 4096 columns of one
// row: each thread first runs the radius recurrence over one 16-column
// segment from its checkpoint (no walk) into LDS, then the workgroup streams
// the pixels through with coalesced loads and stores.
constexpr int VIG_COLS = 4096;
constexpr int VIG_PITCH = CK + 1;  // pads the per-segment rows against bank conflicts
__global__ __launch_bounds__(256) void opt_vignette(const float* __restrict__ src,
                                                    float* __restrict__ dst, Vig v, int width,
                                                    int height, DeviceStats* stats) {
    __shared__ float mult[(VIG_COLS / CK) * VIG_PITCH];
    const uint3 b = xcd_block();
    const int y = b.y;
    const int x0 = b.x * VIG_COLS;
    const int t = threadIdx.x;
    bool bad = false;
    {
        const int seg_x = x0 + t * CK;
        if (seg_x < width) {
            float r2 = v.rows[(size_t)(seg_x / CK) * height + y];
            const int n = min(CK, width - seg_x);
            for (int j = 0; j < n; ++j) {
                mult[t * VIG_PITCH + j] = vig_multiplier(v, r2, bad);
                r2 += v.steps[seg_x + j];
            }
        }
    }
    __syncthreads();
    const size_t plane = (size_t)width * height;
    const int n = min(VIG_COLS, width - x0);
    for (int i = t; i < n; i += 256) {
        const float m = mult[(i / CK) * VIG_PITCH + i % CK];
        const size_t o = (size_t)y * width + x0 + i;
        for (int c = 0; c < 3; ++c) {
            const float value = src[c * plane + o] * m;
            bad = bad || !optics::finite32(value);
            dst[c * plane + o] = value;
        }
    }
    if (bad) flag(stats, BAD_VALUE);
}

struct ReduceArgs {
    const float* src;     // planar camera RGB, width x height
    const float* sample;  // what the cubic remap samples: vignetted copy or src
    int width, height;
    int mode;             // 0 copy, 1 vignetting only, 2 remap
    Vig vig;
    optics::Model model;
    const float* in_x;    // input (lens) coordinate per column
    const float* in_y;    // per row
    // horizontal area axis
    const uint32_t* first;
    const uint32_t* offset;
    const uint32_t* count;
    const float* weights;
    int target_width;
    int group;            // output columns per workgroup
    int direct;           // write corrected pixels (full resolution), no reduction
    float* out;           // planar: target_width x height (or the corrected image)
    DeviceStats* stats;
};

// Catmull-Rom weights of the four taps x0 .. x0+3 around x = x0 + 1 + f,
// 0 <= f < 1. Each tap's distance lies in a fixed branch of cubic_weight
// (|d| in [1,2), [0,1), (0,1], (1,2]); at the shared end points both
// polynomials give +0, so the values equal cubic_weight's exactly.
__device__ __forceinline__ float weight_inner(float a) {
    return 1.5f * a * a * a - 2.5f * a * a + 1.0f;
}
__device__ __forceinline__ float weight_outer(float a) {
    return -0.5f * a * a * a + 2.5f * a * a - 4.0f * a + 2.0f;
}
__device__ __forceinline__ void tap_weights(float x, int x0, float w[4]) {
    w[0] = weight_outer(x - (float)x0);
    w[1] = weight_inner(x - (float)(x0 + 1));
    w[2] = weight_inner(fabsf(x - (float)(x0 + 2)));
    w[3] = weight_outer(fabsf(x - (float)(x0 + 3)));
}

// Cubic sample of one plane at (x, y) (crates/optics/src/resample.rs:3-42).
__device__ __forceinline__ float cubic(const float* __restrict__ p, int width, int height, float x,
                                       float y, unsigned& bad) {
    const float fx = floorf(x), fy = floorf(y);
    if (!(fx >= 1.0f && fy >= 1.0f && fx < (float)(width - 2) && fy < (float)(height - 2))) {
        bad |= BAD_FOOTPRINT;
        return 0.0f;
    }
    const int x0 = (int)fx - 1, y0 = (int)fy - 1;
    float wx[4], wy[4];
    tap_weights(x, x0, wx);
    tap_weights(y, y0, wy);
    float value = 0.0f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const float* row = p + (size_t)(y0 + r) * width + x0;
#pragma unroll
        for (int k = 0; k < 4; ++k) value += row[k] * (wy[r] * wx[k]);
    }
    if (!optics::finite32(value)) {
        bad |= BAD_VALUE;
        return 0.0f;
    }
    return value;
}

// Corrected RGB of pixel (x, y).
__device__ __forceinline__ void corrected(const ReduceArgs& a, int x, int y, float v[3],
                                          unsigned& bad) {
    const size_t plane = (size_t)a.width * a.height;
    const size_t i = (size_t)y * a.width + x;
    if (a.mode == 0) {
        for (int c = 0; c < 3; ++c) v[c] = a.src[c * plane + i];
        return;
    }
    if (a.mode == 1) {
        bool b = false;
        const float m = vig_multiplier(a.vig, vig_radius(a.vig, x, y, a.height), b);
        for (int c = 0; c < 3; ++c) {
            v[c] = a.src[c * plane + i] * m;
            b = b || !optics::finite32(v[c]);
        }
        if (b) bad |= BAD_VALUE;
        return;
    }
    float s[6];
    if (!optics::source_coordinates(a.model, a.in_x[x], a.in_y[y], s)) {
        bad |= BAD_VALUE;
        v[0] = v[1] = v[2] = 0.0f;
        return;
    }
    for (int c = 0; c < 3; ++c)
        v[c] = cubic(a.sample + c * plane, a.width, a.height, s[2 * c], s[2 * c + 1], bad);
}

__global__ __launch_bounds__(256) void opt_reduce_h(ReduceArgs a) {
    __shared__ float seg[3][SEG];
    const uint3 b = xcd_block();
    const int y = b.y;
    const int t = threadIdx.x;
    unsigned bad = 0;
    if (a.direct) {
        const int x = b.x * SEG + t;
        if (x < a.width) {
            float v[3];
            corrected(a, x, y, v, bad);
            const size_t plane = (size_t)a.width * a.height;
            for (int c = 0; c < 3; ++c) a.out[c * plane + (size_t)y * a.width + x] = v[c];
        }
        if (bad) flag(a.stats, bad);
        return;
    }
    const int ox0 = b.x * a.group;
    const int oxn = min(a.group, a.target_width - ox0);
    const int last = ox0 + oxn - 1;
    const int xa = (int)a.first[ox0];
    const int xb = (int)(a.first[last] + a.count[last]);
    const bool owner = t < oxn;
    int tap = 0, taps = 0, x_first = 0;
    uint32_t woff = 0;
    float acc[3] = {0.0f, 0.0f, 0.0f};
    if (owner) {
        x_first = (int)a.first[ox0 + t];
        taps = (int)a.count[ox0 + t];
        woff = a.offset[ox0 + t];
    }
    for (int cs = xa; cs < xb; cs += SEG) {
        const int x = cs + t;
        if (x < xb) {
            float v[3];
            corrected(a, x, y, v, bad);
            seg[0][t] = v[0];
            seg[1][t] = v[1];
            seg[2][t] = v[2];
        }
        __syncthreads();
        if (owner) {
            for (; tap < taps && x_first + tap < cs + SEG; ++tap) {
                const int j = x_first + tap - cs;
                const float w = a.weights[woff + tap];
                acc[0] += seg[0][j] * w;
                acc[1] += seg[1][j] * w;
                acc[2] += seg[2][j] * w;
            }
        }
        __syncthreads();
    }
    if (owner) {
        const size_t plane = (size_t)a.target_width * a.height;
        const size_t o = (size_t)y * a.target_width + ox0 + t;
        a.out[o] = acc[0];
        a.out[plane + o] = acc[1];
        a.out[2 * plane + o] = acc[2];
    }
    if (bad) flag(a.stats, bad);
}

__global__ __launch_bounds__(256) void opt_reduce_v(const float* __restrict__ mid,
                                                    const uint32_t* __restrict__ first,
                                                    const uint32_t* __restrict__ offset,
                                                    const uint32_t* __restrict__ count,
                                                    const float* __restrict__ weights, int width,
                                                    int source_height, int target_height,
                                                    float* __restrict__ out, DeviceStats* stats) {
    const uint3 b = xcd_block();
    const int x = b.x * 256 + threadIdx.x;
    const int y = b.y;
    if (x >= width) return;
    const size_t in_plane = (size_t)width * source_height;
    const size_t out_plane = (size_t)width * target_height;
    const uint32_t f = first[y], o = offset[y], n = count[y];
    float acc[3] = {0.0f, 0.0f, 0.0f};
    for (uint32_t k = 0; k < n; ++k) {
        const float w = weights[o + k];
        const size_t i = (size_t)(f + k) * width + x;
        acc[0] += mid[i] * w;
        acc[1] += mid[in_plane + i] * w;
        acc[2] += mid[2 * in_plane + i] * w;
    }
    bool bad = false;
    for (int c = 0; c < 3; ++c) {
        bad = bad || !optics::finite32(acc[c]);
        out[c * out_plane + (size_t)y * width + x] = acc[c];
    }
    if (bad) flag(stats, BAD_VALUE);
}

}  // namespace

size_t optics_checkpoints(int width) { return (size_t)(width + CK - 1) / CK; }

hipError_t launch_optics(const OpticsLaunch& L, hipStream_t s, Profiler* prof) {
    const int W = L.width, H = L.height;
    const size_t plane = (size_t)W * H;
    Vig vig{L.vig_steps, L.vig_rows, L.vig_k[0], L.vig_k[1], L.vig_k[2]};
    const int mode = L.remap ? 2 : (L.vignetting ? 1 : 0);
    if (L.vignetting && !L.rows_ready) {
        RHD_TIMED(prof, s, "opt_vig_rows", (uint64_t)optics_checkpoints(W) * H * 4,
                  opt_vig_rows<<<(H + 63) / 64, 64, 0, s>>>(L.vig_steps, L.vig_start_x,
                                                            L.vig_start_y, L.vig_ns, W, H,
                                                            L.vig_rows));
    }
    if (L.vignetting) {
        if (L.remap) {
            RHD_TIMED(prof, s, "opt_vignette", plane * 24,
                      opt_vignette<<<dim3((W + VIG_COLS - 1) / VIG_COLS, H), 256, 0, s>>>(
                          L.src, L.vignetted, vig, W, H, L.stats));
        }
    }
    ReduceArgs a{};
    a.src = L.src;
    a.sample = (L.remap && L.vignetting) ? L.vignetted : L.src;
    a.width = W;
    a.height = H;
    a.mode = mode;
    a.vig = vig;
    a.model = L.model;
    a.in_x = L.in_x;
    a.in_y = L.in_y;
    a.stats = L.stats;
    if (L.direct) {
        a.direct = 1;
        a.target_width = W;
        a.out = L.out;
        RHD_TIMED(prof, s, "opt_correct", plane * 24,
                  opt_reduce_h<<<dim3((W + SEG - 1) / SEG, H), 256, 0, s>>>(a));
        return hipGetLastError();
    }
    a.first = L.h_first;
    a.offset = L.h_offset;
    a.count = L.h_count;
    a.weights = L.h_weights;
    a.target_width = L.target_width;
    a.group = L.group;
    a.out = L.mid;
    const int groups = (L.target_width + L.group - 1) / L.group;
    RHD_TIMED(prof, s, "opt_reduce_h", plane * 12 + (uint64_t)L.target_width * H * 12,
              opt_reduce_h<<<dim3(groups, H), 256, 0, s>>>(a));
    RHD_TIMED(prof, s, "opt_reduce_v",
              (uint64_t)L.target_width * (H + L.target_height) * 12,
              opt_reduce_v<<<dim3((L.target_width + 255) / 256, L.target_height), 256, 0, s>>>(
                  L.mid, L.v_first, L.v_offset, L.v_count, L.v_weights, L.target_width, H,
                  L.target_height, L.out, L.stats));
    return hipGetLastError();
}

}  // namespace rhd
