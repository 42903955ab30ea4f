// Masked Richardson-Lucy capture sharpening on planar camera RGB.
//
// Reference: crates/core/src/sharpening.rs:87-127 (apply), :180-310 (solve)
// and :312-384 (smoothstep, mirror, blur); Gaussian weights from
// crates/detail/src/gaussian.rs:9-17 (computed on the host with the platform
// expf, as Rust's f32::exp does on Linux). Results are bit-identical to the
// reference: every blurred value is computed at its image coordinate with the
// reference's tap order (0.0 start, taps in ascending order, -0.0 start in
// the mirrored horizontal border) and half-sample mirroring.
//
// Kernels: init (guide, highlight mask), two single blurs with the mask
// epilogues, one fused kernel per RL iteration (blur estimate -> ratio ->
// blur ratio -> clamped update, all in LDS for a 64x32 output tile with a
// 2r halo), and the gain apply. Iterations ping-pong between two estimate
// planes. HBM-bound: per iteration the estimate (with halo) and the guide are
// read and the estimate written.
#include "rhd_internal.hpp"

namespace rhd {
namespace {

constexpr int TW = 64, TH = 32, MAXR = 4, THREADS = 256;
constexpr float FLOOR = 1.0e-6f;

struct Weights {
    float w[2 * MAXR + 1];
};

__device__ __forceinline__ float smoothstep(float low, float high, float value) {
    float t = (value - low) / (high - low);
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    return t * t * (3.0f - 2.0f * t);
}

// sharpening.rs:194-218
__global__ __launch_bounds__(THREADS) void cap_init(const float* __restrict__ rgb, int n,
                                                    float l0, float l1, float l2,
                                                    float* __restrict__ guide,
                                                    float* __restrict__ mask,
                                                    DeviceStats* stats) {
    int i = blockIdx.x * THREADS + threadIdx.x;
    if (i >= n) return;
    float r = rgb[i], g = rgb[n + i], b = rgb[2 * (size_t)n + i];
    if (!(is_finite(r) && is_finite(g) && is_finite(b))) atomicOr(&stats->nonfinite, 4u);
    float s = -0.0f;
    s += num_max(r, 0.0f) / 3.0f;
    s += num_max(g, 0.0f) / 3.0f;
    s += num_max(b, 0.0f) / 3.0f;
    guide[i] = num_max(s, FLOOR);
    float peak = 0.0f;
    peak = num_max(peak, r / l0);
    peak = num_max(peak, g / l1);
    peak = num_max(peak, b / l2);
    mask[i] = 1.0f - smoothstep(0.85f, 0.98f, peak);
}

// Horizontal taps at image column x of an LDS row whose first element is
// image column `col0`; `border` selects the reference's mirrored `.sum()`.
__device__ __forceinline__ float hblur(const float* row, int col0, int x, int width, int r,
                                       const Weights& k, bool fast) {
    bool border = width <= 2 * r || x < r || x >= width - r;
    float v = border ? -0.0f : 0.0f;
    if (fast) {
        for (int t = 0; t <= 2 * r; ++t) v += row[x + t - r - col0] * k.w[t];
    } else {
        for (int t = 0; t <= 2 * r; ++t) v += row[mirror_half(x + t - r, width) - col0] * k.w[t];
    }
    return v;
}

// Vertical taps at image row y over an LDS plane whose first row is image row
// `row0` and pitch `pitch`, at LDS column `c`.
__device__ __forceinline__ float vblur(const float* plane, int pitch, int row0, int y, int c,
                                       int height, int r, const Weights& k, bool fast) {
    float v = 0.0f;
    if (fast) {
        for (int t = 0; t <= 2 * r; ++t) v += plane[(y + t - r - row0) * pitch + c] * k.w[t];
    } else {
        for (int t = 0; t <= 2 * r; ++t)
            v += plane[(mirror_half(y + t - r, height) - row0) * pitch + c] * k.w[t];
    }
    return v;
}

// One blur of `in` with an epilogue:
//   mode 0: out = min(aux, b^4)                          (sharpening.rs:227-237)
//   mode 1: out = aux * detail gate(guide = in, b)       (sharpening.rs:238-262)
__global__ __launch_bounds__(THREADS) void cap_blur(const float* __restrict__ in,
                                                    const float* __restrict__ aux,
                                                    float* __restrict__ out, int width,
                                                    int height, int r, Weights k, int mode,
                                                    float noise_protection) {
    __shared__ float s_in[(TH + 2 * MAXR) * (TW + 2 * MAXR)];
    __shared__ float s_h[(TH + 2 * MAXR) * TW];
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;
    const int cx0 = max(0, x0 - r), cx1 = min(width, x0 + TW + r);
    const int ry0 = max(0, y0 - r), ry1 = min(height, y0 + TH + r);
    const int pw = cx1 - cx0, ph = ry1 - ry0;
    const int ox1 = min(width, x0 + TW), oy1 = min(height, y0 + TH);
    const bool fast = x0 - r >= 0 && x0 + TW + r <= width && y0 - r >= 0 && y0 + TH + r <= height;
    for (int e = threadIdx.x; e < pw * ph; e += THREADS) {
        int rr = e / pw, cc = e % pw;
        s_in[rr * pw + cc] = in[(size_t)(ry0 + rr) * width + cx0 + cc];
    }
    __syncthreads();
    const int ow = ox1 - x0;
    for (int e = threadIdx.x; e < ow * ph; e += THREADS) {
        int rr = e / ow, cc = e % ow;
        s_h[rr * TW + cc] = hblur(s_in + rr * pw, cx0, x0 + cc, width, r, k, fast);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < ow * (oy1 - y0); e += THREADS) {
        int rr = e / ow, cc = e % ow;
        int y = y0 + rr, x = x0 + cc;
        float b = vblur(s_h, TW, ry0, y, cc, height, r, k, fast);
        size_t i = (size_t)y * width + x;
        if (mode == 0) {
            out[i] = num_min(aux[i], (b * b) * (b * b));
        } else {
            float g = in[i];
            float threshold = 0.0005f + noise_protection * 0.008f + b * noise_protection * 0.02f;
            float detail = fabsf(g - b);
            out[i] = aux[i] * (smoothstep(threshold, threshold * 3.0f, detail) *
                               smoothstep(0.001f, 0.01f, g));
        }
    }
}

// One RL iteration (sharpening.rs:268-304) for a 64x32 tile.
constexpr int E_W = TW + 4 * MAXR, E_H = TH + 4 * MAXR;
constexpr int H1_W = TW + 2 * MAXR;
constexpr int R_W = TW + 2 * MAXR, R_H = TH + 2 * MAXR;
__global__ __launch_bounds__(THREADS) void cap_iterate(const float* __restrict__ est_in,
                                                       const float* __restrict__ guide,
                                                       float* __restrict__ est_out, int width,
                                                       int height, int r, Weights k) {
    __shared__ float s_e[E_H * E_W];
    __shared__ float s_h1[E_H * H1_W];
    __shared__ float s_r[R_H * R_W];
    __shared__ float s_h2[R_H * TW];
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;
    const int ox1 = min(width, x0 + TW), oy1 = min(height, y0 + TH);
    // Estimate: image window widened by 2r; ratio window by r.
    const int ex0 = max(0, x0 - 2 * r), ex1 = min(width, x0 + TW + 2 * r);
    const int ey0 = max(0, y0 - 2 * r), ey1 = min(height, y0 + TH + 2 * r);
    const int rx0 = max(0, x0 - r), rx1 = min(width, x0 + TW + r);
    const int ry0 = max(0, y0 - r), ry1 = min(height, y0 + TH + r);
    const int ew = ex1 - ex0, eh = ey1 - ey0, rw = rx1 - rx0, rh = ry1 - ry0, ow = ox1 - x0;
    const bool fast =
        x0 - 2 * r >= 0 && x0 + TW + 2 * r <= width && y0 - 2 * r >= 0 && y0 + TH + 2 * r <= height;
    for (int e = threadIdx.x; e < ew * eh; e += THREADS) {
        int rr = e / ew, cc = e % ew;
        s_e[rr * E_W + cc] = est_in[(size_t)(ey0 + rr) * width + ex0 + cc];
    }
    __syncthreads();
    // Horizontal blur of the estimate over the ratio columns, all estimate rows.
    for (int e = threadIdx.x; e < rw * eh; e += THREADS) {
        int rr = e / rw, cc = e % rw;
        s_h1[rr * H1_W + cc] = hblur(s_e + rr * E_W, ex0, rx0 + cc, width, r, k, fast);
    }
    __syncthreads();
    // Vertical blur -> ratio over the ratio window.
    for (int e = threadIdx.x; e < rw * rh; e += THREADS) {
        int rr = e / rw, cc = e % rw;
        int y = ry0 + rr, x = rx0 + cc;
        float b = vblur(s_h1, H1_W, ey0, y, cc, height, r, k, fast);
        s_r[rr * R_W + cc] = guide[(size_t)y * width + x] / num_max(b, FLOOR);
    }
    __syncthreads();
    // Horizontal blur of the ratio over the output columns, all ratio rows.
    for (int e = threadIdx.x; e < ow * rh; e += THREADS) {
        int rr = e / ow, cc = e % ow;
        s_h2[rr * TW + cc] = hblur(s_r + rr * R_W, rx0, x0 + cc, width, r, k, fast);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < ow * (oy1 - y0); e += THREADS) {
        int rr = e / ow, cc = e % ow;
        int y = y0 + rr, x = x0 + cc;
        float b = vblur(s_h2, TW, ry0, y, cc, height, r, k, fast);
        float g = guide[(size_t)y * width + x];
        float v = s_e[(y - ey0) * E_W + (x - ex0)] * b;
        float lo = g * 0.5f, hi = g * 2.0f;
        est_out[(size_t)y * width + x] = v < lo ? lo : (v > hi ? hi : v);
    }
}

// sharpening.rs:105-126
__global__ __launch_bounds__(THREADS) void cap_apply(float* __restrict__ rgb, int n, float amount,
                                                     const float* __restrict__ guide,
                                                     const float* __restrict__ mask,
                                                     const float* __restrict__ est) {
    int i = blockIdx.x * THREADS + threadIdx.x;
    if (i >= n) return;
    float gain = 1.0f + amount * mask[i] * (est[i] / guide[i] - 1.0f);
    if (gain == 1.0f) return;
    float r = rgb[i], g = rgb[n + i], b = rgb[2 * (size_t)n + i];
    if (is_finite(r * gain) && is_finite(g * gain) && is_finite(b * gain)) {
        rgb[i] = r * gain;
        rgb[n + i] = g * gain;
        rgb[2 * (size_t)n + i] = b * gain;
    }
}

}  // namespace

size_t capture_scratch_bytes(int width, int height) {
    return (size_t)width * height * sizeof(float) * 5;
}

hipError_t launch_capture(const CaptureArgs& a, hipStream_t s, Profiler* prof) {
    if (a.radius < 0 || a.radius > MAXR) return hipErrorInvalidValue;
    const int n = a.width * a.height;
    const uint64_t px = (uint64_t)n;
    Weights k{};
    for (int t = 0; t <= 2 * a.radius; ++t) k.w[t] = a.weights[t];
    const int blocks = (n + THREADS - 1) / THREADS;
    RHD_TIMED(prof, s, "cap_init", 20 * px,
              hipLaunchKernelGGL(cap_init, dim3(blocks), dim3(THREADS), 0, s, a.rgb, n,
                                 a.ceilings[0], a.ceilings[1], a.ceilings[2], a.guide, a.mask,
                                 a.stats));
    const dim3 tiles((a.width + TW - 1) / TW, (a.height + TH - 1) / TH);
    // mask = min(mask, blur(mask)^4), then the detail gate on blur(guide).
    RHD_TIMED(prof, s, "cap_blur", 12 * px,
              hipLaunchKernelGGL(cap_blur, tiles, dim3(THREADS), 0, s, a.mask, a.mask, a.tmp,
                                 a.width, a.height, a.radius, k, 0, a.noise_protection));
    RHD_TIMED(prof, s, "cap_blur", 12 * px,
              hipLaunchKernelGGL(cap_blur, tiles, dim3(THREADS), 0, s, a.guide, a.tmp, a.mask,
                                 a.width, a.height, a.radius, k, 1, a.noise_protection));
    hipMemcpyAsync(a.est_a, a.guide, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, s);
    float* cur = a.est_a;
    float* next = a.est_b;
    for (int it = 0; it < a.iterations; ++it) {
        // Reads the estimate and the guide, writes the estimate: 12 B per pixel.
        RHD_TIMED(prof, s, "cap_iterate", 12 * px,
                  hipLaunchKernelGGL(cap_iterate, tiles, dim3(THREADS), 0, s, cur, a.guide, next,
                                     a.width, a.height, a.radius, k));
        float* t = cur;
        cur = next;
        next = t;
    }
    RHD_TIMED(prof, s, "cap_apply", 36 * px,
              hipLaunchKernelGGL(cap_apply, dim3(blocks), dim3(THREADS), 0, s, a.rgb, n, a.amount,
                                 a.guide, a.mask, cur));
    return hipGetLastError();
}

}  // namespace rhd
