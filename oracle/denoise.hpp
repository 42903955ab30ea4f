// Oracle (test infrastructure only): noise reduction of the detail stage.
//
//   * the noise profile sigma^2(x) = a x + b of the balanced luma and of the
//     two colour differences (crates/detail/src/noise.rs:48-170, sampled rows
//     and opponent split from crates/detail/src/denoise/mod.rs:60-150 and
//     crates/detail/src/planes.rs:96-152);
//   * chroma noise reduction, a five-level luma-guided a-trous shrinkage of
//     C1 = R-G and C2 = B-G (crates/detail/src/denoise/chroma.rs:88-425);
//   * luminance noise reduction, a five-level B3 a-trous BayesShrink of the
//     balanced luma applied back as a common RGB gain
//     (crates/detail/src/denoise/luma.rs:128-366, B3 smoothing from
//     crates/detail/src/atrous.rs:63-120).
//
// Everything is IEEE f32 (f64 where the reference accumulates in f64) in the
// reference's operation order; no libm beyond sqrt, so results are exact.
// Interleaved tight RGB, in place.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

#include "detail.hpp"

namespace rhdo {
namespace denoise {

using detail::mirror;

// noise.rs:24-47.
struct Profile {
    float a = 0.0f, b = 0.0f;
    uint64_t bins = 0;
    float variance(float x) const { return a * std::fmax(x, 0.0f) + b; }
    float sigma(float x) const { return std::sqrt(variance(x)); }
};

// f32::total_cmp as an unsigned key: ascending key = ascending total order.
inline uint32_t order_key(float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

inline bool finite_positive(const float b[3]) {
    for (int c = 0; c < 3; ++c)
        if (!std::isfinite(b[c]) || b[c] <= 0.0f) return false;
    return true;
}

// atrous.rs:9, 63-120: separable B3 [1 4 6 4 1]/16 with taps `step` apart,
// mirrored; horizontal into `tmp`, then vertical into `out`.
constexpr float B3[5] = {1.0f / 16.0f, 4.0f / 16.0f, 6.0f / 16.0f, 4.0f / 16.0f, 1.0f / 16.0f};

inline void smooth(const float* src, size_t w, size_t h, size_t step, std::vector<float>& tmp,
                   float* out) {
    tmp.resize(w * h);
    const long s = (long)step;
    for (size_t y = 0; y < h; ++y)
        for (size_t x = 0; x < w; ++x) {
            const float* line = src + y * w;
            float t[5];
            for (int k = 0; k < 5; ++k) t[k] = line[mirror((long)x + (k - 2) * s, w)];
            tmp[y * w + x] = B3[0] * t[0] + B3[1] * t[1] + B3[2] * t[2] + B3[3] * t[3] + B3[4] * t[4];
        }
    for (size_t y = 0; y < h; ++y) {
        size_t rows[5];
        for (int k = 0; k < 5; ++k) rows[k] = mirror((long)y + (k - 2) * s, h) * w;
        for (size_t x = 0; x < w; ++x)
            out[y * w + x] = B3[0] * tmp[rows[0] + x] + B3[1] * tmp[rows[1] + x] +
                             B3[2] * tmp[rows[2] + x] + B3[3] * tmp[rows[3] + x] +
                             B3[4] * tmp[rows[4] + x];
    }
}

// noise.rs:127-139: median absolute deviation about the median, both
// selections at index len/2 in total order.
inline float mad(std::vector<float>& v) {
    auto less = [](float a, float b) { return order_key(a) < order_key(b); };
    const size_t k = v.size() / 2;
    std::nth_element(v.begin(), v.begin() + k, v.end(), less);
    const float median = v[k];
    for (float& x : v) x = std::fabs(x - median);
    std::nth_element(v.begin(), v.begin() + k, v.end(), less);
    return v[k];
}

// noise.rs:142-170.
inline Profile fit_line(const std::vector<std::pair<float, float>>& points) {
    const double n = (double)points.size();
    double sx = 0.0, sy = 0.0, sxx = 0.0, sxy = 0.0;
    for (auto& p : points) {
        const double x = p.first, y = p.second;
        sx += x;
        sy += y;
        sxx += x * x;
        sxy += x * y;
    }
    const double den = n * sxx - sx * sx;
    const double slope = std::fabs(den) < 1e-18 ? 0.0 : (n * sxy - sx * sy) / den;
    const double a = std::fmax(slope, 0.0);
    const double b = std::fmax((sy - a * sx) / n, 0.0);
    Profile p;
    p.a = (float)a;
    p.b = (float)b;
    p.bins = points.size();
    return p;
}

constexpr size_t TILE = 16;
constexpr float DETAIL_NOISE_GAIN = 0.8908f;
constexpr float MAD_TO_SIGMA = 1.4826f;

// noise.rs:48-125 on one plane with its brightness plane. Returns 1 when the
// plane holds no whole 16x16 tile (InvalidDimensions).
inline int estimate(const std::vector<float>& plane, const std::vector<float>& bright, size_t w,
                    size_t h, Profile& out) {
    std::vector<float> tmp, band(w * h);
    smooth(plane.data(), w, h, 1, tmp, band.data());
    for (size_t i = 0; i < w * h; ++i) band[i] = plane[i] - band[i];
    struct Tile {
        float energy, brightness;
        size_t x, y;
    };
    std::vector<Tile> tiles;
    for (size_t ty = 0; ty + TILE <= (h / TILE) * TILE; ty += TILE)
        for (size_t tx = 0; tx + TILE <= (w / TILE) * TILE; tx += TILE) {
            double energy = 0.0, mean = 0.0;
            for (size_t y = ty; y < ty + TILE; ++y)
                for (size_t x = tx; x < tx + TILE; ++x) {
                    const double v = band[y * w + x];
                    energy += v * v;
                    mean += (double)bright[y * w + x];
                }
            const double n = (double)(TILE * TILE);
            tiles.push_back({(float)(energy / n), (float)(mean / n), tx, ty});
        }
    if (tiles.empty()) return 1;
    std::stable_sort(tiles.begin(), tiles.end(), [](const Tile& a, const Tile& b) {
        return order_key(a.brightness) < order_key(b.brightness);
    });
    const size_t bin_count = std::min<size_t>(tiles.size() * 1 / 4 / 8, 16);
    auto bin_estimate = [&](Tile* group, size_t len) {
        std::stable_sort(group, group + len, [](const Tile& a, const Tile& b) {
            return order_key(a.energy) < order_key(b.energy);
        });
        const size_t flat = std::max<size_t>(len * 1 / 4, 1);
        std::vector<float> samples;
        samples.reserve(flat * TILE * TILE);
        for (size_t t = 0; t < flat; ++t)
            for (size_t y = group[t].y; y < group[t].y + TILE; ++y)
                for (size_t x = group[t].x; x < group[t].x + TILE; ++x)
                    samples.push_back(band[y * w + x]);
        const float sigma = MAD_TO_SIGMA * mad(samples) / DETAIL_NOISE_GAIN;
        // f32 Sum folds from -0.0.
        float sum = -0.0f;
        for (size_t t = 0; t < flat; ++t) sum += group[t].brightness;
        const float x = sum / (float)flat;
        return std::make_pair(x, sigma * sigma);
    };
    if (bin_count < 8) {
        out = Profile{0.0f, bin_estimate(tiles.data(), tiles.size()).second, 0};
        return 0;
    }
    const size_t len = tiles.size();
    std::vector<std::pair<float, float>> points;
    for (size_t i = 0; i < bin_count; ++i) {
        const size_t a = i * len / bin_count, b = (i + 1) * len / bin_count;
        points.push_back(bin_estimate(tiles.data() + a, b - a));
    }
    out = fit_line(points);
    return 0;
}

// mod.rs:112-122: every fourth 64-row band; images under 1024 rows whole.
inline std::vector<size_t> sample_rows(size_t h) {
    std::vector<size_t> rows;
    if (h < 4 * 256) {
        for (size_t y = 0; y < h; ++y) rows.push_back(y);
        return rows;
    }
    for (size_t start = 0; start < h; start += 256)
        for (size_t y = start; y < std::min(start + 64, h); ++y) rows.push_back(y);
    return rows;
}

// planes.rs:125-152 on the given rows: balanced Y, C1, C2.
inline void split(const float* rgb, size_t w, const std::vector<size_t>& rows, const float bal[3],
                  std::vector<float>& y, std::vector<float>& c1, std::vector<float>& c2) {
    const size_t n = rows.size() * w;
    y.assign(n, 0.0f);
    c1.assign(n, 0.0f);
    c2.assign(n, 0.0f);
    for (size_t r = 0; r < rows.size(); ++r)
        for (size_t x = 0; x < w; ++x) {
            const float* px = rgb + (rows[r] * w + x) * 3;
            const float R = px[0] * bal[0], G = px[1] * bal[1], B = px[2] * bal[2];
            const size_t i = r * w + x;
            y[i] = (R + 2.0f * G + B) * 0.25f;
            c1[i] = R - G;
            c2[i] = B - G;
        }
}

inline void zero_non_finite(std::vector<float>& p) {
    for (float& v : p)
        if (!std::isfinite(v)) v = 0.0f;
}

// mod.rs:60-110 (estimate_noise_profiles over noise_sample_rows). Returns 0
// with out = {luma, c1, c2}, 1 when the image is too small to profile, 2 for
// an invalid balance.
inline int profiles(const float* rgb, size_t w, size_t h, const float bal[3], Profile out[3]) {
    if (!finite_positive(bal)) return 2;
    const std::vector<size_t> rows = sample_rows(h);
    std::vector<float> y, c1, c2;
    split(rgb, w, rows, bal, y, c1, c2);
    zero_non_finite(y);
    zero_non_finite(c1);
    zero_non_finite(c2);
    const size_t hs = rows.size();
    if (estimate(y, y, w, hs, out[0])) return 1;
    if (estimate(c1, y, w, hs, out[1])) return 1;
    if (estimate(c2, y, w, hs, out[2])) return 1;
    return 0;
}

// chroma.rs:376-378 and luma.rs:274-277.
inline float soft_threshold(float v, float t) {
    const float sign = std::isnan(v) ? v : (std::signbit(v) ? -1.0f : 1.0f);
    return sign * std::fmax(std::fabs(v) - t, 0.0f);
}

// chroma.rs:184-192.
inline float fast_exp_neg(float t) {
    float x = std::fmax(1.0f - t * 0.125f, 0.0f);
    x = x * x;
    x = x * x;
    return x * x;
}

constexpr float CHROMA_LEVEL_WEIGHTS[5] = {1.0f, 0.8f, 0.6f, 0.4f, 0.3f};
constexpr float KERNEL[3] = {0.25f, 0.5f, 0.25f};

// chroma.rs:88-148, 198-425. profiles = {luma, c1, c2}. Returns 1 for an
// invalid strength or balance.
inline int chroma(float* rgb, size_t w, size_t h, float strength, const float bal[3],
                  const float ceilings[3], const Profile prof[3]) {
    if (!std::isfinite(strength) || !(strength >= 0.0f && strength <= 1.0f)) return 1;
    if (strength == 0.0f) return 0;
    if (!finite_positive(bal)) return 1;
    std::vector<size_t> all(h);
    for (size_t y = 0; y < h; ++y) all[y] = y;
    std::vector<float> luma, cur1, cur2;
    split(rgb, w, all, bal, luma, cur1, cur2);
    zero_non_finite(luma);
    zero_non_finite(cur1);
    zero_non_finite(cur2);
    const size_t n = w * h;
    std::vector<float> acc1(n, 0.0f), acc2(n, 0.0f), s1(n, 0.0f), s2(n, 0.0f);
    for (int level = 0; level < 5; ++level) {
        const long step = 1L << level;
        const float lw = CHROMA_LEVEL_WEIGHTS[level];
        for (size_t y = 0; y < h; ++y) {
            size_t rows[3];
            for (int k = 0; k < 3; ++k) rows[k] = mirror((long)y + (k - 1) * step, h) * w;
            for (size_t x = 0; x < w; ++x) {
                size_t cols[3];
                for (int k = 0; k < 3; ++k) cols[k] = mirror((long)x + (k - 1) * step, w);
                const size_t i = y * w + x;
                const float centre = luma[i];
                const float sg = std::fmax(prof[0].sigma(centre), 1e-6f) * 4.0f;
                const float inv = 1.0f / (sg * sg);
                float sum1 = 0.0f, sum2 = 0.0f, wsum = 0.0f;
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k) {
                        const size_t q = rows[j] + cols[k];
                        const float d = luma[q] - centre;
                        const float wt = KERNEL[j] * KERNEL[k] * fast_exp_neg(d * d * inv);
                        sum1 += wt * cur1[q];
                        sum2 += wt * cur2[q];
                        wsum += wt;
                    }
                const float o1 = sum1 / wsum, o2 = sum2 / wsum;
                s1[i] = o1;
                s2[i] = o2;
                const float scale = lw * strength * 5.0f;
                const float t1 = scale * prof[1].sigma(centre), t2 = scale * prof[2].sigma(centre);
                acc1[i] += soft_threshold(cur1[i] - o1, t1);
                acc2[i] += soft_threshold(cur2[i] - o2, t2);
            }
        }
        std::swap(cur1, s1);
        std::swap(cur2, s2);
    }
    for (size_t i = 0; i < n; ++i) {
        acc1[i] += cur1[i];
        acc2[i] += cur2[i];
    }
    for (size_t i = 0; i < n; ++i) {
        float* px = rgb + i * 3;
        bool finite = true, blown = false;
        for (int c = 0; c < 3; ++c) {
            finite = finite && std::isfinite(px[c]);
            blown = blown || px[c] >= ceilings[c];
        }
        if (!finite || blown) continue;
        const float g = luma[i] - (acc1[i] + acc2[i]) * 0.25f;
        const float o[3] = {(g + acc1[i]) / bal[0], g / bal[1], (g + acc2[i]) / bal[2]};
        if (std::isfinite(o[0]) && std::isfinite(o[1]) && std::isfinite(o[2]))
            for (int c = 0; c < 3; ++c) px[c] = o[c];
    }
    return 0;
}

constexpr float BAND_NOISE_GAIN[5] = {0.8908f, 0.2006f, 0.0856f, 0.0413f, 0.0203f};
constexpr float SIGNAL_FLOOR = 0.4f;
constexpr float DETAIL_RELIEF = 0.8f;
constexpr float MAX_GAIN = 4.0f;
constexpr float LUMA_FLOOR = 1.0e-6f;
constexpr size_t WINDOW_RADIUS = 3;

// luma.rs:237-250.
inline float luma_threshold(float scale, float gain_sq, const Profile& p, float brightness,
                            float energy) {
    const float nv = p.variance(brightness) * gain_sq;
    const float floor = SIGNAL_FLOOR * SIGNAL_FLOOR * nv;
    const float sigma_signal = std::sqrt(std::fmax(energy - nv, floor));
    return sigma_signal > 0.0f ? scale * nv / sigma_signal : 0.0f;
}

// luma.rs:279-321: mean of band^2 over the 7x7 mirrored window.
inline void local_mean_square(const std::vector<float>& band, size_t w, size_t h,
                              std::vector<float>& tmp, std::vector<float>& out) {
    const size_t win = 2 * WINDOW_RADIUS + 1;
    const float norm = 1.0f / (float)(win * win);
    tmp.assign(w * h, 0.0f);
    out.assign(w * h, 0.0f);
    std::vector<float> padded(w + 2 * WINDOW_RADIUS);
    for (size_t y = 0; y < h; ++y) {
        for (size_t i = 0; i < padded.size(); ++i) {
            const float v = band[y * w + mirror((long)i - (long)WINDOW_RADIUS, w)];
            padded[i] = v * v;
        }
        for (size_t k = 0; k < win; ++k)
            for (size_t x = 0; x < w; ++x) tmp[y * w + x] += padded[k + x];
    }
    for (size_t y = 0; y < h; ++y) {
        for (size_t k = 0; k < win; ++k) {
            const size_t sy = mirror((long)y + (long)k - (long)WINDOW_RADIUS, h);
            for (size_t x = 0; x < w; ++x) out[y * w + x] += tmp[sy * w + x];
        }
        for (size_t x = 0; x < w; ++x) out[y * w + x] *= norm;
    }
}

// luma.rs:128-176, 179-277, 323-366. Returns 1 for invalid parameters.
inline int luma(float* rgb, size_t w, size_t h, float strength, float detail_, const float bal[3],
                const float ceilings[3], const Profile& prof) {
    for (float v : {strength, detail_})
        if (!(v >= 0.0f && v <= 1.0f)) return 1;
    if (strength == 0.0f) return 0;
    if (!finite_positive(bal)) return 1;
    const size_t n = w * h;
    // luma_plane (planes.rs:108-121).
    std::vector<float> cur(n), sm(n, 0.0f), acc(n, 0.0f), energy, tmp, tmp2;
    for (size_t i = 0; i < n; ++i) {
        const float* px = rgb + i * 3;
        const float v = (px[0] * bal[0] + 2.0f * px[1] * bal[1] + px[2] * bal[2]) * 0.25f;
        cur[i] = std::isfinite(v) ? v : 0.0f;
    }
    for (int level = 0; level < 5; ++level) {
        smooth(cur.data(), w, h, (size_t)1 << level, tmp, sm.data());
        for (size_t i = 0; i < n; ++i) cur[i] -= sm[i];
        local_mean_square(cur, w, h, tmp2, energy);
        const float relief = level < 2 ? 1.0f - DETAIL_RELIEF * detail_ : 1.0f;
        const float scale = strength * 1.0f * relief;
        const float gain_sq = BAND_NOISE_GAIN[level] * BAND_NOISE_GAIN[level];
        for (size_t i = 0; i < n; ++i)
            acc[i] += soft_threshold(cur[i], luma_threshold(scale, gain_sq, prof, sm[i], energy[i]));
        std::swap(cur, sm);
    }
    for (size_t i = 0; i < n; ++i) acc[i] += cur[i];
    for (size_t i = 0; i < n; ++i) {
        float* px = rgb + i * 3;
        bool finite = true, blown = false;
        for (int c = 0; c < 3; ++c) {
            finite = finite && std::isfinite(px[c]);
            blown = blown || px[c] >= ceilings[c];
        }
        if (!finite || blown) continue;
        const float old = (px[0] * bal[0] + 2.0f * px[1] * bal[1] + px[2] * bal[2]) * 0.25f;
        if (old <= LUMA_FLOOR) continue;
        float gain = acc[i] / old;
        if (gain < 0.0f) gain = 0.0f;
        if (gain > MAX_GAIN) gain = MAX_GAIN;
        const float o[3] = {px[0] * gain, px[1] * gain, px[2] * gain};
        if (std::isfinite(o[0]) && std::isfinite(o[1]) && std::isfinite(o[2]))
            for (int c = 0; c < 3; ++c) px[c] = o[c];
    }
    return 0;
}

}  // namespace denoise
}  // namespace rhdo
