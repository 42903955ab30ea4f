// Exhaustive host check of rohditor_amd/csrc/libm_f32.hpp against the host
// libm (the functions the reference's f32::cbrt and f32::log2 call).
// Compiled and run by tests/test_libm.py; prints "name count" lines.
//
//   cbrt_mismatch       positive f32 (normal and subnormal) where cbrt_libm != cbrtf
//   cbrt_negative       the same over a stride-7 sample of negative f32
//   cbrt_not_rounded    positive f32 where cbrtf != (float)cbrt((double)x), i.e. where a
//                       correctly rounded replica would be wrong
//   log2_mismatch       f32 in [2^-8, 2^8) where log2_libm != log2f
//   log2_unlisted       the same with no exception table (the series value alone)
//   log2_outside_far    f32 outside the table range (stride-5 sample of positive
//                       finite values) where log2_libm is more than 1 ulp from log2f
#include <math.h>
#include <stdio.h>

#include <atomic>
#include <thread>
#include <vector>

#include "libm_f32.hpp"

using namespace rhd::libm;

static const uint32_t kList[] = {
#include "log2f_exceptions.inc"
};

template <class F>
static uint64_t count_over(uint64_t lo, uint64_t hi, uint64_t stride, F bad) {
    const unsigned threads = std::max(1u, std::thread::hardware_concurrency());
    std::atomic<uint64_t> total{0};
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            uint64_t n = 0;
            for (uint64_t b = lo + (uint64_t)t * stride; b < hi; b += (uint64_t)threads * stride)
                n += bad((uint32_t)b) ? 1 : 0;
            total += n;
        });
    for (auto& th : pool) th.join();
    return total.load();
}

int main() {
    std::vector<uint32_t> bitmap(LOG2_BITMAP_WORDS, 0u);
    for (uint32_t b : kList) {
        const uint32_t i = b - LOG2_FIRST;
        bitmap[i >> 5] |= 1u << (i & 31);
    }
    const uint32_t* bm = bitmap.data();
    const uint64_t inf = 0x7f800000u;

    printf("cbrt_mismatch %llu\n", (unsigned long long)count_over(1, inf, 1, [](uint32_t b) {
               const float x = u2f(b);
               return f2u(cbrt_libm(x)) != f2u(cbrtf(x));
           }));
    printf("cbrt_negative %llu\n",
           (unsigned long long)count_over(0x80000001u, 0xff800000u, 7, [](uint32_t b) {
               const float x = u2f(b);
               return f2u(cbrt_libm(x)) != f2u(cbrtf(x));
           }));
    printf("cbrt_not_rounded %llu\n", (unsigned long long)count_over(1, inf, 1, [](uint32_t b) {
               const float x = u2f(b);
               return f2u(cbrtf(x)) != f2u((float)cbrt((double)x));
           }));
    printf("log2_mismatch %llu\n",
           (unsigned long long)count_over(LOG2_FIRST, LOG2_END, 1, [bm](uint32_t b) {
               const float x = u2f(b);
               return f2u(log2_libm(x, bm)) != f2u(log2f(x));
           }));
    printf("log2_unlisted %llu\n",
           (unsigned long long)count_over(LOG2_FIRST, LOG2_END, 1, [](uint32_t b) {
               const float x = u2f(b);
               return f2u(log2_libm(x, nullptr)) != f2u(log2f(x));
           }));
    printf("log2_outside_far %llu\n", (unsigned long long)count_over(1, inf, 5, [bm](uint32_t b) {
               if (b >= LOG2_FIRST && b < LOG2_END) return false;
               const float x = u2f(b);
               const uint32_t a = f2u(log2_libm(x, bm)), g = f2u(log2f(x));
               return (a > g ? a - g : g - a) > 1u;
           }));
    return 0;
}
