// Host check of rohditor_amd/csrc/capture_math.hpp: capture_inputs_fast (the
// divisions as reciprocal products with one FMA correction) against
// capture_inputs (IEEE division, the reference's arithmetic).
// Compiled and run by tests/test_capture_math.py; prints "name count" lines.
//
//   third_mismatch     f32 x >= +0 (every finite one) where div_by(x, 3) != x / 3
//   third_uncorrected  stride-31 sample: the bare product x * RN(1/3) differs (why the
//                      FMA step exists)
//   mask_mismatch      f32 peak in [0.25, 4) (every one) and in [0, 0.25) (stride 101)
//                      where the fast mask != the reference mask
//   ratio_mismatch     f32 x in [0.8 l, 1.05 l] where div_by(x, l) != x / l, for
//                      ceilings l across [2^-20, 2^20]
//   ratio_mask_sample  stride-97 sample of x in [0, 8 l): fast guide or mask != reference
//   random_mismatch    random (r, g, b, ceilings): guide or mask bits differ
//   checked            how many cases the above covered
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "capture_math.hpp"

using namespace rhd;

static float bits_f(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}
static uint32_t f_bits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

static std::atomic<uint64_t> checked{0};

// Counts i in [lo, hi) (stride) where bad(i) holds, on every host thread.
template <class F>
static uint64_t count_over(uint64_t lo, uint64_t hi, uint64_t stride, F bad) {
    const unsigned threads = std::max(1u, std::thread::hardware_concurrency());
    std::vector<std::thread> pool;
    std::atomic<uint64_t> total{0};
    for (unsigned t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            uint64_t n = 0, seen = 0;
            for (uint64_t i = lo + t * stride; i < hi; i += threads * stride) {
                n += bad(i) ? 1 : 0;
                ++seen;
            }
            total += n;
            checked += seen;
        });
    for (auto& th : pool) th.join();
    return total.load();
}

static float ref_mask(float peak) { return 1.0f - smoothstep(0.85f, 0.98f, peak); }

static float fast_mask(float peak) {
    const float l[3] = {1.0f, 1.0f, 1.0f}, inv[3] = {1.0f, 1.0f, 1.0f};
    return capture_inputs_fast(peak, 0.0f, 0.0f, l, inv).mask;
}

int main() {
    const float third = 1.0f / 3.0f;
    const uint64_t finite_end = 0x7f800000ull;
    printf("third_mismatch %llu\n",
           (unsigned long long)count_over(0, finite_end, 1, [&](uint64_t i) {
               const float x = bits_f((uint32_t)i);
               return f_bits(div_by(x, 3.0f, third)) != f_bits(x / 3.0f);
           }));
    printf("third_uncorrected %llu\n",
           (unsigned long long)count_over(0, finite_end, 31, [&](uint64_t i) {
               const float x = bits_f((uint32_t)i);
               return f_bits(x * third) != f_bits(x / 3.0f);
           }));
    auto mask_differs = [&](uint64_t i) {
        const float p = bits_f((uint32_t)i);
        return f_bits(fast_mask(p)) != f_bits(ref_mask(p));
    };
    printf("mask_mismatch %llu\n",
           (unsigned long long)(count_over(f_bits(0.25f), f_bits(4.0f), 1, mask_differs) +
                                count_over(0, f_bits(0.25f), 101, mask_differs)));
    const float ceilings[] = {0x1p-20f, 0x1.5ep-20f, 1.0e-3f,  0.1234567f, 0.5f, 0.70710677f,
                              0.93f,    1.0f,        1.618034f, 3.0f,       1000.5f, 0x1p20f};
    uint64_t ratio = 0, ratio_mask = 0;
    for (float l : ceilings) {
        const float inv = 1.0f / l;
        ratio += count_over(f_bits(0.8f * l), f_bits(1.05f * l) + 1, 1, [&](uint64_t i) {
            const float x = bits_f((uint32_t)i);
            return f_bits(div_by(x, l, inv)) != f_bits(x / l);
        });
        const float ls[3] = {l, l, l}, invs[3] = {inv, inv, inv};
        ratio_mask += count_over(0, f_bits(8.0f * l), 97, [&](uint64_t i) {
            const float x = bits_f((uint32_t)i);
            const CaptureInputs a = capture_inputs(0.0f, x, 0.0f, l, l, l);
            const CaptureInputs b = capture_inputs_fast(0.0f, x, 0.0f, ls, invs);
            return f_bits(a.mask) != f_bits(b.mask) || f_bits(a.guide) != f_bits(b.guide);
        });
    }
    printf("ratio_mismatch %llu\n", (unsigned long long)ratio);
    printf("ratio_mask_sample %llu\n", (unsigned long long)ratio_mask);
    // Random pixels: per case, ceilings log-uniform in [2^-20, 2^20] and
    // channels either near a ceiling, anywhere below 8 ceilings, negative, or
    // of a random exponent.
    printf("random_mismatch %llu\n",
           (unsigned long long)count_over(0, 1ull << 26, 1, [&](uint64_t i) {
               uint64_t s = (i + 1) * 0x9e3779b97f4a7c15ull;
               auto next = [&]() {
                   s ^= s << 13;
                   s ^= s >> 7;
                   s ^= s << 17;
                   return s;
               };
               auto unit = [&]() { return (float)(next() >> 40) * 0x1p-24f; };
               float l[3], inv[3], c[3];
               for (int k = 0; k < 3; ++k) {
                   l[k] = ldexpf(1.0f + unit(), (int)(next() % 40) - 20);
                   if (l[k] > 0x1p20f) l[k] = 0x1p20f;
                   inv[k] = 1.0f / l[k];
               }
               for (int k = 0; k < 3; ++k) {
                   switch (next() % 4) {
                       case 0: c[k] = l[k] * (0.84f + 0.15f * unit()); break;
                       case 1: c[k] = l[k] * 8.0f * unit(); break;
                       case 2: c[k] = -l[k] * unit(); break;
                       default: {
                           const uint32_t sign = (uint32_t)(next() & 1) << 31;
                           c[k] = bits_f(sign | (uint32_t)(next() % finite_end));
                           break;
                       }
                   }
               }
               const CaptureInputs a = capture_inputs(c[0], c[1], c[2], l[0], l[1], l[2]);
               const CaptureInputs b = capture_inputs_fast(c[0], c[1], c[2], l, inv);
               return f_bits(a.mask) != f_bits(b.mask) || f_bits(a.guide) != f_bits(b.guide);
           }));
    printf("checked %llu\n", (unsigned long long)checked.load());
    return 0;
}
