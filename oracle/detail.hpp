// Oracle (test infrastructure only): false-colour suppression, the first step
// of the detail stage (crates/detail/src/false_color.rs, with median3 from
// crates/detail/src/median.rs:5-110 and mirror from crates/detail/src/
// lib.rs:95-103). Colour differences C1 = r*b0 - g*b1 and C2 = b*b2 - g*b1
// (non-finite -> 0), `steps` 3x3 medians of each, then r = (g + C1) / b0 and
// b = (g + C2) / b2 where the pixel is finite, below its ceilings, and the
// rebuilt values are finite. Green is never written.
#pragma once

#include <cmath>
#include <cstddef>
#include <vector>

namespace rhdo {
namespace detail {

// Half-sample reflection (lib.rs:95-103).
inline size_t mirror(long index, size_t length) {
    const long period = (long)(length * 2);
    long folded = index % period;
    if (folded < 0) folded += period;
    return (size_t)folded < length ? (size_t)folded : length * 2 - 1 - (size_t)folded;
}

// Selection with `<` / `>` (median.rs:5-20): ties keep the reference's pick.
inline float mn(float a, float b) { return a < b ? a : b; }
inline float mx(float a, float b) { return a > b ? a : b; }
inline float med3(float a, float b, float c) { return mx(mn(a, b), mn(mx(a, b), c)); }

// Median of a 3x3 window from its three columns, each top to bottom
// (median.rs:22-37).
inline float median_of_columns(const float col[3][3]) {
    float low[3], mid[3], high[3];
    for (int k = 0; k < 3; ++k) {
        low[k] = mn(mn(col[k][0], col[k][1]), col[k][2]);
        mid[k] = med3(col[k][0], col[k][1], col[k][2]);
        high[k] = mx(mx(col[k][0], col[k][1]), col[k][2]);
    }
    return med3(mx(mx(low[0], low[1]), low[2]), med3(mid[0], mid[1], mid[2]),
                mn(mn(high[0], high[1]), high[2]));
}

inline void median3(const std::vector<float>& src, std::vector<float>& out, size_t w, size_t h) {
    for (size_t y = 0; y < h; ++y) {
        const size_t rows[3] = {mirror((long)y - 1, h), y, mirror((long)y + 1, h)};
        for (size_t x = 0; x < w; ++x) {
            float col[3][3];
            for (int k = 0; k < 3; ++k) {
                const size_t sx = mirror((long)x + k - 1, w);
                for (int r = 0; r < 3; ++r) col[k][r] = src[rows[r] * w + sx];
            }
            out[y * w + x] = median_of_columns(col);
        }
    }
}

// In place on tight interleaved RGB. Returns 1 for an invalid balance.
inline int false_color(float* rgb, size_t w, size_t h, int steps, const float balance[3],
                       const float ceilings[3]) {
    if (steps == 0) return 0;
    for (int c = 0; c < 3; ++c)
        if (!std::isfinite(balance[c]) || !(balance[c] > 0.0f)) return 1;
    const size_t n = w * h;
    std::vector<float> c1(n), c2(n), t1(n), t2(n);
    for (size_t i = 0; i < n; ++i) {
        const float* p = rgb + 3 * i;
        const float g = p[1] * balance[1];
        const float d1 = p[0] * balance[0] - g, d2 = p[2] * balance[2] - g;
        c1[i] = std::isfinite(d1) ? d1 : 0.0f;
        c2[i] = std::isfinite(d2) ? d2 : 0.0f;
    }
    for (int s = 0; s < steps; ++s) {
        median3(c1, t1, w, h);
        median3(c2, t2, w, h);
        c1.swap(t1);
        c2.swap(t2);
    }
    for (size_t i = 0; i < n; ++i) {
        float* p = rgb + 3 * i;
        bool finite = true, blown = false;
        for (int c = 0; c < 3; ++c) {
            finite = finite && std::isfinite(p[c]);
            blown = blown || p[c] >= ceilings[c];
        }
        if (!finite || blown) continue;
        const float g = p[1] * balance[1];
        const float r = (g + c1[i]) / balance[0], b = (g + c2[i]) / balance[2];
        if (std::isfinite(r) && std::isfinite(b)) {
            p[0] = r;
            p[2] = b;
        }
    }
    return 0;
}

}  // namespace detail
}  // namespace rhdo
