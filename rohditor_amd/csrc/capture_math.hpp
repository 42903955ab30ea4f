// Capture sharpening's per-pixel inputs: the guide and the highlight mask
// (crates/core/src/sharpening.rs:194-218, smoothstep :312-316).
//
// capture_inputs is the reference's arithmetic as written (IEEE division).
// capture_inputs_fast gives the same bits with each division replaced by a
// reciprocal product and one FMA correction; the AMaZE finish uses it, where
// seven correctly rounded divisions per pixel would lengthen the slowest
// stage of its row pipeline. Plain IEEE f32 operations with no contraction,
// so the host and the device compute the same bits (tests/capture_check.cc
// checks the two forms against each other on the host).
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define RHD_CAP_HD __host__ __device__ __forceinline__
#else
#define RHD_CAP_HD inline
#endif

namespace rhd {

constexpr float CAPTURE_FLOOR = 1.0e-6f;

struct CaptureInputs {
    float guide, mask;
};

// Hermite smoothstep clamped to [0, 1].
RHD_CAP_HD float smoothstep(float low, float high, float value) {
    float t = (value - low) / (high - low);
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    return t * t * (3.0f - 2.0f * t);
}

// The guide is the mean of the non-negative channels (a sum from -0.0 of each
// channel divided by 3), floored at 1e-6; the mask falls from 1 to 0 as the
// largest channel-to-ceiling ratio rises from 0.85 to 0.98. Maxima are IEEE
// maxNum, as Rust's f32::max.
RHD_CAP_HD CaptureInputs capture_inputs(float r, float g, float b, float l0, float l1, float l2) {
    float s = -0.0f;
    s += fmaxf(r, 0.0f) / 3.0f;
    s += fmaxf(g, 0.0f) / 3.0f;
    s += fmaxf(b, 0.0f) / 3.0f;
    float peak = 0.0f;
    peak = fmaxf(peak, r / l0);
    peak = fmaxf(peak, g / l1);
    peak = fmaxf(peak, b / l2);
    return {fmaxf(s, CAPTURE_FLOOR), 1.0f - smoothstep(0.85f, 0.98f, peak)};
}

// x / y from c = RN(1 / y) and one FMA correction (Markstein's theorem):
// q0 = RN(x * c) lies within an ulp of x / y, the remainder x - q0 * y is
// exact in an FMA, and RN(q0 + r * c) is the correctly rounded quotient while
// the quotient and the remainder stay in the normal range.
RHD_CAP_HD float div_by(float x, float y, float c) {
    const float q0 = x * c;
    const float r = __builtin_fmaf(-q0, y, x);
    return __builtin_fmaf(r, c, q0);
}

// capture_inputs with each division taken by div_by (`inv` holds RN(1 / l)).
// Equal to capture_inputs for finite inputs when every ceiling lies in
// [2^-20, 2^20]: the channel means divide by 3, which div_by rounds correctly
// for every float but -0 (whose sign the sum and the floor discard); a
// channel-to-ceiling ratio only matters between 0.85 and 0.98, where the
// numerator and the quotient are normal, and below or above that band any
// value on the same side gives the same mask. Channels above 2^100 are taken
// as 2^100 for the ratio (still far above the band) so that x * RN(1 / l)
// cannot overflow.
RHD_CAP_HD CaptureInputs capture_inputs_fast(float r, float g, float b, const float l[3],
                                             const float inv[3]) {
    constexpr float THIRD = 1.0f / 3.0f;
    constexpr float LOW = 0.85f, SPAN = 0.98f - 0.85f, INV_SPAN = 1.0f / SPAN;
    float s = -0.0f;
    s += div_by(fmaxf(r, 0.0f), 3.0f, THIRD);
    s += div_by(fmaxf(g, 0.0f), 3.0f, THIRD);
    s += div_by(fmaxf(b, 0.0f), 3.0f, THIRD);
    float peak = 0.0f;
    constexpr float BIG = 0x1p100f;
    peak = fmaxf(peak, div_by(fminf(r, BIG), l[0], inv[0]));
    peak = fmaxf(peak, div_by(fminf(g, BIG), l[1], inv[1]));
    peak = fmaxf(peak, div_by(fminf(b, BIG), l[2], inv[2]));
    float t = div_by(peak - LOW, SPAN, INV_SPAN);
    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
    return {fmaxf(s, CAPTURE_FLOOR), 1.0f - t * t * (3.0f - 2.0f * t)};
}

}  // namespace rhd

#undef RHD_CAP_HD
