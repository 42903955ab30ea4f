// Lens-correction coordinate mapping shared by the optics kernels and the
// host-side automatic-scale search. Compiled for host and device alike, with
// the same IEEE operations in the same order on both sides (no contraction,
// correctly rounded f64 division and square root), so a host search and a
// device remap of one plan agree bit for bit.
//
// Mapping of one output pixel (crates/optics/src/plan.rs:465-496): the input
// (lens) coordinate comes from per-column and per-row tables built on the
// host in f64 and narrowed to f32; the inverse distortion model runs a
// Newton-Raphson solve in f64 from ru = rd, stopping at |F(ru)| < 1e-5 and
// giving up after six updates (lensfun 0.7.0's inverse kernels; Poly3 gives
// NaN on failure, the others the input unchanged); transverse chromatic
// aberration scales red and blue; the result is shifted back to pixels in f64
// and narrowed once.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace rhd {
namespace optics {

enum { DIST_NONE = 0, DIST_POLY3 = 1, DIST_POLY5 = 2, DIST_PTLENS = 3 };
enum { TCA_NONE = 0, TCA_LINEAR = 1, TCA_POLY3 = 2 };
enum { VIG_NONE = 0, VIG_PA = 1 };

// Plan constants the mapping needs after the input tables.
struct Model {
    double center_x, center_y, norm_unscale;
    int distortion;
    float dist[3];
    int tca;
    float red[3], blue[3];
};

constexpr double NEWTON_EPS = 0.00001;
constexpr int NEWTON_UPDATES = 6;

__host__ __device__ inline bool newton_done(double f) {
    return f >= -NEWTON_EPS && f < NEWTON_EPS;
}

__host__ __device__ inline bool finite32(float v) { return v == v && fabsf(v) <= 3.40282347e38f; }

// Radius of a point whose squared length is summed in f32.
__host__ __device__ inline double point_radius(float x, float y) {
    return sqrt((double)(x * x + y * y));
}

// Scales (x, y) by ru / rd in f64 and narrows once.
__host__ __device__ inline void rescale(float& x, float& y, double ru, double rd) {
    const double k = ru / rd;
    x = (float)((double)x * k);
    y = (float)((double)y * k);
}

// Inverse radial distortion in place; false when the model gives no finite
// point (Poly3 without a solution).
__host__ __device__ inline bool undistort(const Model& m, float& x, float& y) {
    if (m.distortion == DIST_NONE) return true;
    const double rd = point_radius(x, y);
    if (rd == 0.0) return true;
    double ru = rd;
    if (m.distortion == DIST_POLY3) {
        const double inv = (double)(1.0f / m.dist[0]);
        const double target = (double)(float)(rd * inv);
        for (int step = 0;; ++step) {
            const double f = ru * ru * ru + ru * inv - target;
            if (newton_done(f)) break;
            if (step >= NEWTON_UPDATES) return false;
            ru -= f / (3.0 * ru * ru + inv);
        }
        if (ru < 0.0) return false;
    } else if (m.distortion == DIST_POLY5) {
        const double k1 = m.dist[0], k2 = m.dist[1];
        for (int step = 0;; ++step) {
            const double ru2 = ru * ru;
            const double f = ru * (1.0 + k1 * ru2 + k2 * ru2 * ru2) - rd;
            if (newton_done(f)) break;
            if (step >= NEWTON_UPDATES) return true;
            ru -= f / (1.0 + 3.0 * k1 * ru2 + 5.0 * k2 * ru2 * ru2);
        }
        if (ru < 0.0) return true;
    } else {
        const double a = m.dist[0], b = m.dist[1], c = m.dist[2];
        for (int step = 0;; ++step) {
            const double f = ru * (a * ru * ru * ru + b * ru * ru + c * ru + 1.0) - rd;
            if (newton_done(f)) break;
            if (step >= NEWTON_UPDATES) return true;
            ru -= f / (4.0 * a * ru * ru * ru + 3.0 * b * ru * ru + 2.0 * c * ru + 1.0);
        }
        if (ru < 0.0) return true;
    }
    rescale(x, y, ru, rd);
    return true;
}

// Reverse Poly3 TCA of one channel, coefficients [v, c, b].
__host__ __device__ inline void tca_poly3(const float k[3], float& x, float& y) {
    const double rd = point_radius(x, y);
    if (rd == 0.0) return;
    const double v = k[0], c = k[1], b = k[2];
    double ru = rd;
    for (int step = 0;; ++step) {
        const double f = b * ru * ru * ru + c * ru * ru + v * ru - rd;
        if (newton_done(f)) break;
        if (step >= NEWTON_UPDATES) return;
        ru -= f / (3.0 * b * ru * ru + 2.0 * c * ru + v);
    }
    if (ru <= 0.0) return;
    rescale(x, y, ru, rd);
}

// Source pixel coordinates (x, y for red, green, blue) of the input (lens)
// coordinate (ix, iy). False on a non-finite result.
__host__ __device__ inline bool source_coordinates(const Model& m, float ix, float iy,
                                                   float out[6]) {
    float gx = ix, gy = iy;
    if (!undistort(m, gx, gy) || !finite32(gx) || !finite32(gy)) return false;
    float rx = gx, ry = gy, bx = gx, by = gy;
    if (m.tca == TCA_LINEAR) {
        rx = gx * m.red[0];
        ry = gy * m.red[0];
        bx = gx * m.blue[0];
        by = gy * m.blue[0];
    } else if (m.tca == TCA_POLY3) {
        tca_poly3(m.red, rx, ry);
        tca_poly3(m.blue, bx, by);
    }
    const float n[6] = {rx, ry, gx, gy, bx, by};
    bool ok = true;
    for (int c = 0; c < 3; ++c) {
        out[2 * c] = (float)(((double)n[2 * c] + m.center_x) * m.norm_unscale);
        out[2 * c + 1] = (float)(((double)n[2 * c + 1] + m.center_y) * m.norm_unscale);
        ok = ok && finite32(out[2 * c]) && finite32(out[2 * c + 1]);
    }
    return ok;
}

// Input (lens) coordinate of output column or row `i` at `scale`, narrowed
// to f32 (plan.rs:470-475).
__host__ __device__ inline float input_coordinate(uint32_t i, double center, double norm_scale,
                                                  double norm_unscale, float scale) {
    const double center_px = center * norm_unscale;
    const double mapped = center_px + ((double)i - center_px) / (double)scale;
    return (float)(mapped * norm_scale - center);
}

// Catmull-Rom weight with tension -0.5 (crates/optics/src/resample.rs:44-55),
// evaluated left to right as the reference writes it.
__host__ __device__ inline float cubic_weight(float distance) {
    const float a = fabsf(distance);
    if (a <= 1.0f) return 1.5f * a * a * a - 2.5f * a * a + 1.0f;
    if (a < 2.0f) return -0.5f * a * a * a + 2.5f * a * a - 4.0f * a + 2.0f;
    return 0.0f;
}

}  // namespace optics
}  // namespace rhd
