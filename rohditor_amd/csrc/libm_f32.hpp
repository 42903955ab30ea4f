// f32 cube root and base-2 logarithm that return exactly what the platform
// libm (glibc) returns, for the Opposed estimator.
//
// The reference calls Rust's f32::cbrt and f32::log2, which are glibc's
// cbrtf and log2f on Linux (crates/highlight/src/support.rs:65,106). Their
// results are not always the correctly rounded ones, so a correctly rounded
// device function would differ from the reference in the last bit on a few
// percent of the cube roots and a fraction of a percent of the logarithms.
//
// cbrt_libm: one Halley step for y^3 = m (y <- y (y^3 + 2m) / (2y^3 + m))
// from a quadratic minimax seed on m in [0.5, 1), then scaled by
// 2^(e/3) (e = the binary exponent), with the seed and y^3 in f32 and the
// Halley quotient in f64 rounded once to f32. This is the scheme and
// operand typing of glibc's flt-32 cbrtf (sysdeps/ieee754/flt-32/s_cbrtf.c);
// evaluated this way it returns glibc's result on every positive f32
// (checked exhaustively against the host libm).
//
// log2_libm: log2(m) for m in [sqrt(1/2), sqrt(2)) from the atanh series
// ln m = 2 (s + s^3/3 + ... + s^23/23), s = (m - 1)/(m + 1), in f64, plus
// the binary exponent, rounded once to f32. That is the correctly rounded
// value except very near a rounding midpoint. glibc's log2f (a table-driven
// f64 polynomial) lands on the other side of such a midpoint for about 0.2 %
// of inputs; those inputs are listed, for x in [2^-8, 2^8), by a table that
// the build generates by calling the host libm (gen_log2_table.cc), and for
// them the other neighbour of the rounding midpoint is returned. With equal
// channel detection levels every ratio Opposed takes the logarithm of lies
// in (0.2, 5), well inside the table's range.
//
// Everything here is IEEE f32/f64 add, subtract, multiply and divide with no
// contraction, so the host and the device compute the same bits.
#pragma once

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define RHD_HD __host__ __device__
#else
#define RHD_HD
#endif

namespace rhd {
namespace libm {

// Table range of log2 exceptions: binary exponents [LOG2_EMIN, LOG2_EMAX).
constexpr int LOG2_EMIN = -8, LOG2_EMAX = 8;
constexpr uint32_t LOG2_FIRST = (uint32_t)(127 + LOG2_EMIN) << 23;  // bits of 2^-8
constexpr uint32_t LOG2_END = (uint32_t)(127 + LOG2_EMAX) << 23;    // bits of 2^8
constexpr uint64_t LOG2_BITMAP_WORDS = (uint64_t)(LOG2_END - LOG2_FIRST) / 32;

RHD_HD inline uint32_t f2u(float v) {
    uint32_t u;
    __builtin_memcpy(&u, &v, 4);
    return u;
}
RHD_HD inline float u2f(uint32_t u) {
    float v;
    __builtin_memcpy(&v, &u, 4);
    return v;
}

// Positive finite x (normal or subnormal) = m * 2^e with m in [0.5, 1).
RHD_HD inline float split_half(float x, int& e) {
    uint32_t b = f2u(x);
    int bias = 126;
    if ((b >> 23) == 0) {  // subnormal: scale into the normal range first
        b = f2u(x * 33554432.0f);  // 2^25, exact
        bias += 25;
    }
    e = (int)(b >> 23) - bias;
    return u2f((b & 0x007fffffu) | 0x3f000000u);
}

// 2^k as an f32 for k in [-126, 127].
RHD_HD inline float pow2f(int k) { return u2f((uint32_t)(k + 127) << 23); }

// glibc's cbrtf for every f32 (zero, infinity and NaN return x + x).
RHD_HD inline float cbrt_libm(float x) {
    const float ax = x < 0.0f ? -x : x;
    if (!(ax > 0.0f) || ax == __builtin_huge_valf()) return x + x;
    int e;
    const float xm = split_half(ax, e);
    const double c0 = 0.492659620528969547, c1 = 0.697570460207922770,
                 c2 = 0.191502161678719066;
    const double factor[5] = {1.0 / 1.5874010519681994748, 1.0 / 1.2599210498948731648, 1.0,
                              1.2599210498948731648, 1.5874010519681994748};
    const double dm = (double)xm;
    const float u = (float)(c0 + (c1 - c2 * dm) * dm);
    const float t2 = u * u * u;
    const double q = (double)u * ((double)t2 + 2.0 * dm) / (2.0 * (double)t2 + dm);
    const float ym = (float)(q * factor[2 + e % 3]);
    // e / 3 lies in [-50, 43], so the scaling is one exact multiply.
    const float y = ym * pow2f(e / 3);
    return x > 0.0f ? y : -y;
}

// The series value of log2(x) in f64, for positive finite x.
RHD_HD inline double log2_series(float x) {
    uint32_t b = f2u(x);
    int k = -127;
    if ((b >> 23) == 0) {
        b = f2u(x * 33554432.0f);
        k -= 25;
    }
    k += (int)(b >> 23);
    double m = (double)u2f((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.4142135623730951) {
        m *= 0.5;
        k += 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = 1.0 / 23.0;
    p = p * s2 + 1.0 / 21.0;
    p = p * s2 + 1.0 / 19.0;
    p = p * s2 + 1.0 / 17.0;
    p = p * s2 + 1.0 / 15.0;
    p = p * s2 + 1.0 / 13.0;
    p = p * s2 + 1.0 / 11.0;
    p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0;
    p = p * s2 + 1.0 / 5.0;
    p = p * s2 + 1.0 / 3.0;
    const double ln = 2.0 * s + 2.0 * s * s2 * p;
    return (double)k + ln * 1.4426950408889634;
}

// The f32 on the other side of the rounding midpoint nearest to `r` from
// f = (float)r: one step away from f towards r.
RHD_HD inline float across_midpoint(float f, double r) {
    const uint32_t b = f2u(f);
    const bool away_from_zero = (r > (double)f) == (f > 0.0f);
    if (f == 0.0f) return f;  // log2 is zero only at x = 1, where it is exact
    return u2f(away_from_zero ? b + 1 : b - 1);
}

// True when the bit for `x` is set in the exception bitmap (x in range).
RHD_HD inline bool log2_listed(const uint32_t* bitmap, float x) {
    const uint32_t b = f2u(x);
    if (b < LOG2_FIRST || b >= LOG2_END) return false;
    const uint32_t i = b - LOG2_FIRST;
    return (bitmap[i >> 5] >> (i & 31)) & 1u;
}

// glibc's log2f for positive finite x in [2^-8, 2^8); elsewhere (and for a
// null bitmap) the series value rounded to f32.
RHD_HD inline float log2_libm(float x, const uint32_t* bitmap) {
    const double r = log2_series(x);
    const float f = (float)r;
    return bitmap && log2_listed(bitmap, x) ? across_midpoint(f, r) : f;
}

}  // namespace libm
}  // namespace rhd
