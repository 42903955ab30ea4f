// Oracle (test infrastructure only): a CPU restatement of the rohditor
// sensor-development stage, used by tests/, the smoke check and the CPU
// baseline leg of bench.py as the checker. It is never the product path.
//
// Shared helpers: Bayer colour lookup, parity-preserving mirroring, and the
// literal macro that keeps every decimal constant rounded straight to the
// working type (so f32 constants equal Rust's `f32` literals bit for bit).
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <type_traits>

namespace rhdo {

// Planted mistakes for the tests that show the parity pins catch them: the
// oracle built with -DRHDO_MUTATION=k (k > 0) carries mistake k, which the
// headers select with RHDO_MUTANT(k). The oracle used as the checker is
// always built with 0.
#ifndef RHDO_MUTATION
#define RHDO_MUTATION 0
#endif
#define RHDO_MUTANT(k) (RHDO_MUTATION == (k))

// Decimal literal rounded directly to T (no double->float double rounding).
#define RHDO_LIT(T, x) (std::is_same<T, float>::value ? (T)(x##f) : (T)(x))

enum Pattern : int { RGGB = 0, BGGR = 1, GRBG = 2, GBRG = 3 };

// crates/image/src/lib.rs:90-117 (BayerPattern::color_at): channel index
// 0 = red, 1 = green, 2 = blue.
inline int color_at(int pattern, size_t x, size_t y) {
    static const int table[4][4] = {
        {0, 1, 1, 2},  // RGGB
        {2, 1, 1, 0},  // BGGR
        {1, 0, 2, 1},  // GRBG
        {1, 2, 0, 1},  // GBRG
    };
    return table[pattern][((y & 1) << 1) | (x & 1)];
}

// crates/image/src/lib.rs:119-135 (BayerPattern::shifted). A shift by (x, y)
// re-labels which colour sits at the origin.
inline int shifted(int pattern, size_t x, size_t y) {
    for (int candidate = 0; candidate < 4; ++candidate) {
        bool same = true;
        for (size_t yy = 0; yy < 2 && same; ++yy)
            for (size_t xx = 0; xx < 2 && same; ++xx)
                same = color_at(candidate, xx, yy) == color_at(pattern, xx + x, yy + y);
        if (same) return candidate;
    }
    return pattern;
}

// crates/demosaic/src/amaze_geometry.rs:72-80 (mirror_coordinate): whole-sample
// reflection about the first and last sample, parity preserving.
inline size_t mirror_coordinate(ptrdiff_t index, size_t length) {
    if (length <= 1) return 0;
    ptrdiff_t period = 2 * (ptrdiff_t)(length - 1);
    ptrdiff_t wrapped = index % period;
    if (wrapped < 0) wrapped += period;
    ptrdiff_t other = period - wrapped;
    return (size_t)(wrapped < other ? wrapped : other);
}

// crates/core/src/sharpening.rs:318-326 (mirror): half-sample reflection about
// the outer pixel boundary.
inline size_t mirror_half(ptrdiff_t index, size_t length) {
    ptrdiff_t period = (ptrdiff_t)(length * 2);
    ptrdiff_t folded = index % period;
    if (folded < 0) folded += period;
    return (size_t)folded < length ? (size_t)folded : length * 2 - 1 - (size_t)folded;
}

// Compare-select min/max of crates/demosaic/src/amaze/math.rs:95-105.
template <class T> inline T sel_min(T a, T b) { return a < b ? a : b; }
template <class T> inline T sel_max(T a, T b) { return a > b ? a : b; }

// Rust f32::max / f32::min (IEEE maxNum/minNum: a NaN operand loses).
template <class T> inline T num_max(T a, T b) {
    if (a != a) return b;
    if (b != b) return a;
    return a > b ? a : b;
}
template <class T> inline T num_min(T a, T b) {
    if (a != a) return b;
    if (b != b) return a;
    return a < b ? a : b;
}

template <class T> inline bool finite(T v) { return std::isfinite(v); }

}  // namespace rhdo
