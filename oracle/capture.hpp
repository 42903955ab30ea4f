// Oracle (test infrastructure only): masked Richardson-Lucy capture
// sharpening restated from crates/core/src/sharpening.rs:87-127 (apply),
// :180-310 (solve), :312-384 (smoothstep, mirror, blur) and the Gaussian
// weights of crates/detail/src/gaussian.rs:9-17. The iteration count is a
// parameter; the reference fixes it at 8 (sharpening.rs:13).
#pragma once

#include <cmath>
#include <vector>

#include "common.hpp"
#include "demosaic.hpp"

namespace rhdo {
namespace capture {

// gaussian.rs:9-17. f32 weights use the platform libm expf, which is what
// Rust's f32::exp calls on Linux; the running sum starts at -0.0 like Rust's
// f32 `Sum`.
template <class T> std::vector<T> gaussian_kernel_t(T sigma) {
    int radius = (int)std::ceil(T(3) * sigma);
    std::vector<T> k;
    for (int x = -radius; x <= radius; ++x) {
        T t = (T)x / sigma;
        k.push_back(std::exp(RHDO_LIT(T, -0.5) * (t * t)));
    }
    T sum = T(-0.0);
    for (T v : k) sum += v;
    for (T& v : k) v /= sum;
    return k;
}
inline std::vector<float> gaussian_kernel(float sigma) { return gaussian_kernel_t<float>(sigma); }

// sharpening.rs:312-315
template <class T> inline T smoothstep(T low, T high, T value) {
    T t = (value - low) / (high - low);
    if (t < T(0)) t = T(0);
    else if (t > T(1)) t = T(1);
    return t * t * (T(3) - T(2) * t);
}

// sharpening.rs:328-384: horizontal pass into `temporary`, then vertical.
template <class T>
void blur(const std::vector<T>& input, std::vector<T>& temporary, std::vector<T>& output,
          size_t width, size_t height, const std::vector<T>& kernel, int threads) {
    const size_t r = kernel.size() / 2;
    parallel_for(height, threads, [&](size_t y) {
        T* row = &temporary[y * width];
        for (size_t x = 0; x < width; ++x) row[x] = T(0);
        if (width > 2 * r) {
            for (size_t k = 0; k < kernel.size(); ++k) {
                const T* src = &input[y * width + k];
                for (size_t x = r; x < width - r; ++x) row[x] += src[x - r] * kernel[k];
            }
        }
        for (size_t x = 0; x < width; ++x) {
            if (width <= 2 * r || x < r || x >= width - r) {
                T v = T(-0.0);
                for (size_t k = 0; k < kernel.size(); ++k)
                    v += input[y * width +
                               mirror_half((ptrdiff_t)x + (ptrdiff_t)k - (ptrdiff_t)r, width)] *
                         kernel[k];
                row[x] = v;
            }
        }
    });
    parallel_for(height, threads, [&](size_t y) {
        T* row = &output[y * width];
        for (size_t x = 0; x < width; ++x) row[x] = T(0);
        for (size_t k = 0; k < kernel.size(); ++k) {
            size_t sy = mirror_half((ptrdiff_t)y + (ptrdiff_t)k - (ptrdiff_t)r, height);
            const T* src = &temporary[sy * width];
            for (size_t x = 0; x < width; ++x) row[x] += src[x] * kernel[k];
        }
    });
}

// sharpening.rs:87-127 + :180-310 on interleaved RGB (stride 3 * width).
// Returns 0, or 8 when a non-finite input sample is found.
template <class T>
int apply(T* rgb, size_t width, size_t height, T amount, T sigma, T noise_protection,
          const T ceilings[3], int iterations, float* guide_out, float* mask_out,
          float* estimate_out, int threads) {
    if (!(amount > T(0))) return 0;
    const T FLOOR = RHDO_LIT(T, 1.0e-6);
    const size_t n = width * height;
    std::vector<T> kernel = gaussian_kernel_t<T>(sigma);
    std::vector<T> guide(n), mask(n);
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c)
            if (!finite(rgb[i * 3 + c])) return 8;
    parallel_for(height, threads, [&](size_t y) {
        for (size_t x = 0; x < width; ++x) {
            size_t i = y * width + x;
            const T* p = rgb + i * 3;
            T g = T(-0.0);
            for (int c = 0; c < 3; ++c) g += num_max(p[c], T(0)) / T(3);
            guide[i] = num_max(g, FLOOR);
            T peak = T(0);
            for (int c = 0; c < 3; ++c) peak = num_max(peak, p[c] / ceilings[c]);
            // 6: upper highlight edge 0.95, 8: lower edge 0.80
            mask[i] = T(1) - smoothstep(RHDO_MUTANT(8) ? RHDO_LIT(T, 0.80) : RHDO_LIT(T, 0.85),
                                        RHDO_MUTANT(6) ? RHDO_LIT(T, 0.95) : RHDO_LIT(T, 0.98),
                                        peak);
        }
    });
    std::vector<T> estimate = guide, temporary(n), blurred(n), ratio(n);
    blur(mask, temporary, blurred, width, height, kernel, threads);
    for (size_t i = 0; i < n; ++i) {
        T b = blurred[i];
        mask[i] = num_min(mask[i], RHDO_MUTANT(5) ? b * b : (b * b) * (b * b));  // 5: power 2
    }
    blur(guide, temporary, blurred, width, height, kernel, threads);
    parallel_for(height, threads, [&](size_t y) {
        for (size_t x = 0; x < width; ++x) {
            size_t i = y * width + x;
            T threshold = RHDO_LIT(T, 0.0005) + noise_protection * RHDO_LIT(T, 0.008) +
                          blurred[i] * noise_protection * RHDO_LIT(T, 0.02);
            T detail = std::fabs(guide[i] - blurred[i]);
            mask[i] *= smoothstep(threshold, threshold * T(3), detail) *
                       smoothstep(RHDO_LIT(T, 0.001), RHDO_LIT(T, 0.01), guide[i]);
        }
    });
    for (int it = 0; it < iterations; ++it) {
        blur(estimate, temporary, blurred, width, height, kernel, threads);
        for (size_t i = 0; i < n; ++i) ratio[i] = guide[i] / num_max(blurred[i], FLOOR);
        blur(ratio, temporary, blurred, width, height, kernel, threads);
        for (size_t i = 0; i < n; ++i) {
            T e = estimate[i] * blurred[i];
            // 7: clamp to [0.25 g, 4 g]
            T lo = guide[i] * (RHDO_MUTANT(7) ? RHDO_LIT(T, 0.25) : RHDO_LIT(T, 0.5)),
              hi = guide[i] * (RHDO_MUTANT(7) ? T(4) : T(2));
            estimate[i] = e < lo ? lo : (e > hi ? hi : e);
        }
    }
    parallel_for(height, threads, [&](size_t y) {
        for (size_t x = 0; x < width; ++x) {
            size_t i = y * width + x;
            T gain = T(1) + amount * mask[i] * (estimate[i] / guide[i] - T(1));
            T* p = rgb + i * 3;
            if (gain != T(1) && finite(p[0] * gain) && finite(p[1] * gain) && finite(p[2] * gain))
                for (int c = 0; c < 3; ++c) p[c] *= gain;
        }
    });
    if (guide_out)
        for (size_t i = 0; i < n; ++i) guide_out[i] = (float)guide[i];
    if (mask_out)
        for (size_t i = 0; i < n; ++i) mask_out[i] = (float)mask[i];
    if (estimate_out)
        for (size_t i = 0; i < n; ++i) estimate_out[i] = (float)estimate[i];
    return 0;
}

}  // namespace capture
}  // namespace rhdo
