// Oracle (test infrastructure only): white balance and the camera matrix to
// linear Rec.2020, the colour step after optics.
//   apply_white_balance_cancellable   crates/core/src/cpu/stages.rs:286-308
//   apply_camera_color_transform      crates/core/src/cpu/stages.rs:250-284
//   Matrix3::transform / dot          crates/color/src/matrix.rs:46-48, 173-175
//   WhiteBalanceGains::validate       crates/demosaic/src/lib.rs:269-280
// The DCP profile look applied after the matrix (ProfileLook) is not part of
// this step's contract here.
#pragma once

#include <cmath>
#include <cstddef>

namespace rhdo {
namespace color {

inline bool valid_gains(const float g[3]) {
    for (int c = 0; c < 3; ++c)
        if (!std::isfinite(g[c]) || !(g[c] > 0.0f)) return false;
    return true;
}

// In place on `n` interleaved RGB pixels: per-channel gains, then each output
// channel is the row's dot product evaluated left to right.
inline int apply(float* rgb, size_t n, const float gains[3], const float m[9]) {
    if (!valid_gains(gains)) return 1;
    for (size_t i = 0; i < n; ++i) {
        float* p = rgb + 3 * i;
        p[0] *= gains[0];
        p[1] *= gains[1];
        p[2] *= gains[2];
        const float r = p[0], g = p[1], b = p[2];
        for (int c = 0; c < 3; ++c) p[c] = m[3 * c] * r + m[3 * c + 1] * g + m[3 * c + 2] * b;
    }
    return 0;
}

}  // namespace color
}  // namespace rhdo
