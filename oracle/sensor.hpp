// Oracle (test infrastructure only): normalization and Clip restated from
// crates/core/src/pipeline/sensor.rs:221-277,598-610 and
// crates/highlight/src/{clip,detect}.rs.
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace rhdo {

struct LevelPattern {
    const float* black;  // repeat_w * repeat_h values
    size_t repeat_w, repeat_h;
    const float* white;  // 1, 3 or repeat_w * repeat_h values
    size_t white_count;
};

// sensor.rs:598-602 (level_index) and :604-610 (white_level)
inline void levels_at(const LevelPattern& lv, int sensor_pattern, size_t sx, size_t sy,
                      float& black, float& white) {
    size_t index = (sy % lv.repeat_h) * lv.repeat_w + (sx % lv.repeat_w);
    black = lv.black[index];
    if (lv.white_count == 1)
        white = lv.white[0];
    else if (lv.white_count == 3)
        white = lv.white[color_at(sensor_pattern, sx, sy)];
    else
        white = lv.white[index];
}

// sensor.rs:221-277 at full crop size (phase_preserving_sample is the
// identity there): out = (f32(u16) - black) / (white - black).
inline void normalize(const uint16_t* mosaic, size_t row_stride, int sensor_pattern,
                      const LevelPattern& lv, size_t crop_x, size_t crop_y, size_t crop_w,
                      size_t crop_h, float* out) {
    for (size_t y = 0; y < crop_h; ++y)
        for (size_t x = 0; x < crop_w; ++x) {
            size_t sx = crop_x + x, sy = crop_y + y;
            float black, white;
            levels_at(lv, sensor_pattern, sx, sy, black, white);
            out[y * crop_w + x] = ((float)mosaic[sy * row_stride + sx] - black) / (white - black);
        }
}

// crates/highlight/src/lib.rs:253-258 (ClipStats), widened to fixed-size
// counters for the C ABI.
struct ClipStats {
    uint64_t affected_sites, changed_sites, nominal_over_white_sites, affected_by_channel[3];
};

// clip.rs:24-72 with detect.rs:5-16 predicates.
inline int clip(float* mosaic, size_t width, size_t height, size_t stride, int pattern,
                const float limits[3], ClipStats* stats) {
    ClipStats st{};
    for (size_t y = 0; y < height; ++y)
        for (size_t x = 0; x < width; ++x) {
            float& s = mosaic[y * stride + x];
            if (!finite(s)) return 8;
            int c = color_at(pattern, x, y);
            float limit = limits[c];
            if (s >= limit) {
                st.affected_sites++;
                st.affected_by_channel[c]++;
            }
            if (s > 1.0f) st.nominal_over_white_sites++;
            if (s > limit) {
                st.changed_sites++;
                s = limit;
            }
        }
    if (stats) *stats = st;
    return 0;
}

}  // namespace rhdo
