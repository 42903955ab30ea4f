// Oracle (test infrastructure only): restatement of the reference's optics
// correction and exact area reduction, the step after the sensor stage.
//
//   plan.rs      crates/optics/src/plan.rs (OpticsExecution :36-48, mapping
//                :465-496, model application :498-519, find_scale :521-559)
//   correction   crates/optics/src/correction.rs (authority remap :157-185,
//                vignetting :213-262)
//   resample     crates/optics/src/resample.rs (cubic_sample :3-42, weight :44-55)
//   area         crates/core/src/resample.rs (axis tables :46-112, passes :235-316)
//   preview      crates/core/src/cpu/stages.rs:112-147 (preview_dimensions)
//
// The inverse lens models come from the third-party crate lensfun 0.7.0
// (Cargo.lock:3234-3235), which is not vendored in the reference. They are
// restated from lensfun's published Newton-Raphson inverses
// (ModifyCoord_UnDist_Poly3/Poly5/PTLens, ModifyCoord_TCA_Poly3 reverse): start
// at ru = rd, stop when |F(ru)| < 1e-5, give up after six updates. The operand
// types follow what crates/optics/src/radial.rs:317-352 fixes for the first
// residual (rd = sqrt(f64(x*x + y*y)) with the sum in f32; Poly3 truncates
// rd/k1 to f32); the Newton updates run in f64. Non-convergence returns NaN for
// Poly3 and the input unchanged for the others (radial.rs:21-26). The bits of
// the crate itself are not reachable here: the inverse kernels are pinned by
// their defining equations (residual below 1e-5, forward map, direction), not
// by lensfun's outputs (parity unpinned at that boundary).
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace rhdo {
namespace optics {

enum Distortion { DIST_NONE = 0, DIST_POLY3 = 1, DIST_POLY5 = 2, DIST_PTLENS = 3 };
enum Tca { TCA_NONE = 0, TCA_LINEAR = 1, TCA_POLY3 = 2 };
enum Vignetting { VIG_NONE = 0, VIG_PA = 1 };
enum Status { OK = 0, ERR_CORRECTION = 1, ERR_FOOTPRINT = 2, ERR_DIMENSIONS = 3 };

// OpticsExecution (plan.rs:36-48). Linear TCA keeps kr in red[0], kb in
// blue[0]; Poly3 TCA keeps [v, c, b] per channel.
struct Execution {
    size_t width, height;
    double norm_scale, norm_unscale, center_x, center_y;
    float scale;
    int distortion;
    float dist[3];
    int tca;
    float red[3], blue[3];
    int vignetting;
    float vig[3];
};

constexpr double NEWTON_EPS = 0.00001;
constexpr int NEWTON_UPDATES = 6;

inline bool converged(double f) { return f >= -NEWTON_EPS && f < NEWTON_EPS; }

// rd of a float point: the sum of squares in f32, widened, f64 sqrt.
inline double radius_of(float x, float y) { return std::sqrt((double)(x * x + y * y)); }

inline void scale_point(float x, float y, double ru, double rd, float* ox, float* oy) {
    const double k = ru / rd;
    *ox = (float)((double)x * k);
    *oy = (float)((double)y * k);
}

// Poly3 inverse: ru^3 + ru/k1 - rd/k1 = 0 (1/k1 in f32, rd/k1 truncated to f32).
inline void undist_poly3(float x, float y, float k1, float* ox, float* oy) {
    const double rd = radius_of(x, y);
    if (rd == 0.0) { *ox = x; *oy = y; return; }
    const double inv_k1 = (double)(1.0f / k1);
    const double rd_div_k1 = (double)(float)(rd * inv_k1);
    double ru = rd;
    for (int step = 0;; ++step) {
        const double f = ru * ru * ru + ru * inv_k1 - rd_div_k1;
        if (converged(f)) break;
        if (step >= NEWTON_UPDATES) { *ox = *oy = NAN; return; }
        ru -= f / (3.0 * ru * ru + inv_k1);
    }
    if (ru < 0.0) { *ox = *oy = NAN; return; }
    scale_point(x, y, ru, rd, ox, oy);
}

// Poly5 inverse: ru (1 + k1 ru^2 + k2 ru^4) = rd.
inline void undist_poly5(float x, float y, float k1, float k2, float* ox, float* oy) {
    *ox = x; *oy = y;
    const double rd = radius_of(x, y);
    if (rd == 0.0) return;
    const double a = k1, b = k2;
    double ru = rd;
    for (int step = 0;; ++step) {
        const double ru2 = ru * ru;
        const double f = ru * (1.0 + a * ru2 + b * ru2 * ru2) - rd;
        if (converged(f)) break;
        if (step >= NEWTON_UPDATES) return;
        ru -= f / (1.0 + 3.0 * a * ru2 + 5.0 * b * ru2 * ru2);
    }
    if (ru < 0.0) return;
    scale_point(x, y, ru, rd, ox, oy);
}

// PTLens inverse: ru (a ru^3 + b ru^2 + c ru + 1) = rd.
inline void undist_ptlens(float x, float y, float a_, float b_, float c_, float* ox, float* oy) {
    *ox = x; *oy = y;
    const double rd = radius_of(x, y);
    if (rd == 0.0) return;
    const double a = a_, b = b_, c = c_;
    double ru = rd;
    for (int step = 0;; ++step) {
        const double f = ru * (a * ru * ru * ru + b * ru * ru + c * ru + 1.0) - rd;
        if (converged(f)) break;
        if (step >= NEWTON_UPDATES) return;
        ru -= f / (4.0 * a * ru * ru * ru + 3.0 * b * ru * ru + 2.0 * c * ru + 1.0);
    }
    if (ru < 0.0) return;
    scale_point(x, y, ru, rd, ox, oy);
}

// TCA Poly3 reverse for one channel with coefficients [v, c, b]:
// b ru^3 + c ru^2 + v ru = rd.
inline void tca_poly3_channel(float x, float y, const float k[3], float* ox, float* oy) {
    *ox = x; *oy = y;
    const double rd = radius_of(x, y);
    if (rd == 0.0) return;
    const double v = k[0], c = k[1], b = k[2];
    double ru = rd;
    for (int step = 0;; ++step) {
        const double f = b * ru * ru * ru + c * ru * ru + v * ru - rd;
        if (converged(f)) break;
        if (step >= NEWTON_UPDATES) return;
        ru -= f / (3.0 * b * ru * ru + 2.0 * c * ru + v);
    }
    if (ru <= 0.0) return;
    scale_point(x, y, ru, rd, ox, oy);
}

inline void apply_distortion(const Execution& e, float x, float y, float* ox, float* oy) {
    switch (e.distortion) {
        case DIST_POLY3: undist_poly3(x, y, e.dist[0], ox, oy); return;
        case DIST_POLY5: undist_poly5(x, y, e.dist[0], e.dist[1], ox, oy); return;
        case DIST_PTLENS: undist_ptlens(x, y, e.dist[0], e.dist[1], e.dist[2], ox, oy); return;
        default: *ox = x; *oy = y; return;
    }
}

// Source coordinates (x, y per channel) of output pixel (x, y) at the plan's
// scale (plan.rs:465-496).
inline int map_coordinates(const Execution& e, float scale, size_t x, size_t y, float out[6]) {
    const double cpx = e.center_x * e.norm_unscale;
    const double cpy = e.center_y * e.norm_unscale;
    const double mx = cpx + ((double)x - cpx) / (double)scale;
    const double my = cpy + ((double)y - cpy) / (double)scale;
    const double ix = mx * e.norm_scale - e.center_x;
    const double iy = my * e.norm_scale - e.center_y;
    float gx, gy;
    apply_distortion(e, (float)ix, (float)iy, &gx, &gy);
    if (!std::isfinite(gx) || !std::isfinite(gy)) return ERR_CORRECTION;
    float n[6] = {gx, gy, gx, gy, gx, gy};
    if (e.tca == TCA_LINEAR) {
        n[0] = gx * e.red[0]; n[1] = gy * e.red[0];
        n[4] = gx * e.blue[0]; n[5] = gy * e.blue[0];
    } else if (e.tca == TCA_POLY3) {
        tca_poly3_channel(gx, gy, e.red, &n[0], &n[1]);
        tca_poly3_channel(gx, gy, e.blue, &n[4], &n[5]);
    }
    for (int c = 0; c < 3; ++c) {
        out[2 * c] = (float)(((double)n[2 * c] + e.center_x) * e.norm_unscale);
        out[2 * c + 1] = (float)(((double)n[2 * c + 1] + e.center_y) * e.norm_unscale);
        if (!std::isfinite(out[2 * c]) || !std::isfinite(out[2 * c + 1])) return ERR_CORRECTION;
    }
    return OK;
}

// Catmull-Rom weight, tension -0.5 (resample.rs:44-55), in the reference's
// left-to-right evaluation order.
inline float cubic_weight(float distance) {
    const float a = std::fabs(distance);
    if (a <= 1.0f) return 1.5f * a * a * a - 2.5f * a * a + 1.0f;
    if (a < 2.0f) return -0.5f * a * a * a + 2.5f * a * a - 4.0f * a + 2.0f;
    return 0.0f;
}

// 4x4 cubic sample of one channel of interleaved RGB (resample.rs:3-42).
inline int cubic_sample(const float* data, size_t row_stride, size_t width, size_t height,
                        int channel, float x, float y, float* out) {
    if (!std::isfinite(x) || !std::isfinite(y)) return ERR_CORRECTION;
    const long fx = (long)std::floor(x);
    const long fy = (long)std::floor(y);
    if (fx - 1 < 0 || fy - 1 < 0 || fx + 2 >= (long)width || fy + 2 >= (long)height)
        return ERR_FOOTPRINT;
    float value = 0.0f;
    for (long oy = -1; oy <= 2; ++oy) {
        const float wy = cubic_weight(y - (float)(fy + oy));
        const size_t sy = (size_t)(fy + oy);
        for (long ox = -1; ox <= 2; ++ox) {
            const float w = wy * cubic_weight(x - (float)(fx + ox));
            const size_t sx = (size_t)(fx + ox);
            value += data[sy * row_stride + sx * 3 + channel] * w;
        }
    }
    if (!std::isfinite(value)) return ERR_CORRECTION;
    *out = value;
    return OK;
}

// PA vignetting in place, with the reference's incremental radius: x and
// r^2 advance along each row by f32 additions (correction.rs:213-262).
inline int vignette(float* rgb, size_t row_stride, const Execution& e) {
    if (e.vignetting != VIG_PA) return OK;
    const float k1 = e.vig[0], k2 = e.vig[1], k3 = e.vig[2];
    const float ns = (float)e.norm_scale;
    const float step = 2.0f * ns;
    const float step_sq = ns * ns;
    const float start_x = (float)(-e.center_x);
    const float start_y = (float)(-e.center_y);
    for (size_t row = 0; row < e.height; ++row) {
        float* p = rgb + row * row_stride;
        const float y = start_y + ns * (float)row;
        float x = start_x;
        float r2 = x * x + y * y;
        for (size_t col = 0; col < e.width; ++col) {
            const float r4 = r2 * r2;
            const float r6 = r4 * r2;
            const float gain = 1.0f + k1 * r2 + k2 * r4 + k3 * r6;
            if (!std::isfinite(gain) || gain == 0.0f) return ERR_CORRECTION;
            const float m = 1.0f / gain;
            for (int c = 0; c < 3; ++c) {
                p[col * 3 + c] *= m;
                if (!std::isfinite(p[col * 3 + c])) return ERR_CORRECTION;
            }
            r2 += step * x + step_sq;
            x += ns;
        }
    }
    return OK;
}

inline bool remaps(const Execution& e) { return e.distortion != DIST_NONE || e.tca != TCA_NONE; }

// The authority path of apply_inner (correction.rs:69-185): vignetting in
// place, then the per-pixel mapping and cubic remap into `out`. `rgb` is
// tight interleaved RGB (row stride 3 * width); on return `out` holds the
// corrected image (a copy of the vignetted one when nothing remaps).
inline int apply(float* rgb, const Execution& e, float* out) {
    if (e.width < 4 || e.height < 4) return ERR_DIMENSIONS;
    const size_t stride = e.width * 3;
    int st = vignette(rgb, stride, e);
    if (st != OK) return st;
    if (!remaps(e)) {
        for (size_t i = 0; i < e.width * e.height * 3; ++i) out[i] = rgb[i];
        return OK;
    }
    for (size_t y = 0; y < e.height; ++y) {
        for (size_t x = 0; x < e.width; ++x) {
            float coords[6];
            if ((st = map_coordinates(e, e.scale, x, y, coords)) != OK) return st;
            for (int c = 0; c < 3; ++c) {
                st = cubic_sample(rgb, stride, e.width, e.height, c, coords[2 * c],
                                  coords[2 * c + 1], &out[(y * e.width + x) * 3 + c]);
                if (st != OK) return st;
            }
        }
    }
    return OK;
}

// Automatic scale: grow by 1.025 until every perimeter probe maps inside the
// cubic footprint (plan.rs:521-559).
inline int find_scale(const Execution& e, float* out) {
    const size_t max_x = e.width - 1, max_y = e.height - 1;
    float scale = 1.0f;
    for (int attempt = 0; attempt < 160; ++attempt) {
        bool safe = true;
        for (int index = 0; index < 64 && safe; ++index) {
            const float t = (float)index / 63.0f;
            const float px[4] = {t * (float)max_x, t * (float)max_x, 0.0f, (float)max_x};
            const float py[4] = {0.0f, (float)max_y, t * (float)max_y, t * (float)max_y};
            for (int k = 0; k < 4 && safe; ++k) {
                float c[6];
                if (map_coordinates(e, scale, (size_t)px[k], (size_t)py[k], c) != OK) {
                    safe = false;
                    break;
                }
                for (int ch = 0; ch < 3; ++ch) {
                    if (!(c[2 * ch] >= 1.0f && c[2 * ch + 1] >= 1.0f &&
                          c[2 * ch] <= (float)(e.width - 2) &&
                          c[2 * ch + 1] <= (float)(e.height - 2)))
                        safe = false;
                }
            }
        }
        if (safe) {
            *out = scale;
            return OK;
        }
        scale *= 1.025f;
    }
    return ERR_FOOTPRINT;
}

// One axis of exact area reduction (core resample.rs:46-112): overlap of
// each source cell with the destination cell, in f64, divided by the scale
// and narrowed to f32.
struct Axis {
    std::vector<size_t> first, offset, count;
    std::vector<float> weights;
};

inline bool make_axis(size_t source, size_t target, Axis* a) {
    if (source == 0 || target == 0 || target > source) return false;
    const double scale = (double)source / (double)target;
    a->first.clear(); a->offset.clear(); a->count.clear(); a->weights.clear();
    for (size_t t = 0; t < target; ++t) {
        const double left = (double)t * scale;
        const double right = (double)(t + 1) * scale;
        const size_t first = (size_t)std::floor(left);
        size_t end = (size_t)std::ceil(right);
        if (end > source) end = source;
        a->first.push_back(first);
        a->offset.push_back(a->weights.size());
        for (size_t s = first; s < end; ++s) {
            const double hi = right < (double)(s + 1) ? right : (double)(s + 1);
            const double lo = left > (double)s ? left : (double)s;
            a->weights.push_back((float)((hi - lo) / scale));
        }
        a->count.push_back(end - first);
    }
    return true;
}

// Horizontal pass into a source-height intermediate, then the vertical pass;
// every destination accumulates from 0.0 in tap order (resample.rs:235-316).
// Equal dimensions return the input unchanged.
inline int resize_area(const float* rgb, size_t w, size_t h, float* out, size_t tw, size_t th) {
    if (tw == w && th == h) {
        for (size_t i = 0; i < w * h * 3; ++i) out[i] = rgb[i];
        return OK;
    }
    Axis ha, va;
    if (!make_axis(w, tw, &ha) || !make_axis(h, th, &va)) return ERR_DIMENSIONS;
    std::vector<float> mid(tw * 3 * h, 0.0f);
    for (size_t y = 0; y < h; ++y) {
        for (size_t tx = 0; tx < tw; ++tx) {
            float* d = &mid[(y * tw + tx) * 3];
            for (size_t k = 0; k < ha.count[tx]; ++k) {
                const float wgt = ha.weights[ha.offset[tx] + k];
                const float* s = &rgb[(y * w + ha.first[tx] + k) * 3];
                for (int c = 0; c < 3; ++c) d[c] += s[c] * wgt;
            }
        }
    }
    for (size_t i = 0; i < tw * th * 3; ++i) out[i] = 0.0f;
    for (size_t ty = 0; ty < th; ++ty) {
        for (size_t k = 0; k < va.count[ty]; ++k) {
            const float wgt = va.weights[va.offset[ty] + k];
            const float* s = &mid[(va.first[ty] + k) * tw * 3];
            float* d = &out[ty * tw * 3];
            for (size_t i = 0; i < tw * 3; ++i) d[i] += s[i] * wgt;
        }
    }
    return OK;
}

// Preview dimensions for a long-edge limit (stages.rs:112-147).
inline int preview_dimensions(size_t w, size_t h, size_t max_long, size_t* tw, size_t* th) {
    if (w < 2 || h < 2 || max_long < 2) return ERR_DIMENSIONS;
    const size_t long_edge = w > h ? w : h;
    if (long_edge <= max_long) { *tw = w; *th = h; return OK; }
    auto scale = [&](size_t d) {
        size_t v = (d * max_long + long_edge / 2) / long_edge;
        if (v < 2) v = 2;
        if (v > d) v = d;
        return v;
    };
    *tw = scale(w);
    *th = scale(h);
    return OK;
}

}  // namespace optics
}  // namespace rhdo
