// Build step: lists the f32 inputs x in [2^-8, 2^8) for which the host
// libm's log2f returns the other neighbour of the rounding midpoint than the
// series value of libm_f32.hpp rounds to (see log2_libm there). Writes the
// sorted input bits as the body of a C initializer list, which opposed.hip
// turns into a bitmap on the device. Fails if the libm result is ever anything other than
// the series value or that neighbour.
//
// Usage: gen_log2_table OUT.inc
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "libm_f32.hpp"

using namespace rhd::libm;

int main(int argc, char** argv) {
    if (argc != 2) {
        fprintf(stderr, "usage: %s OUT.inc\n", argv[0]);
        return 2;
    }
    std::vector<uint32_t> listed;
    for (uint32_t b = LOG2_FIRST; b < LOG2_END; ++b) {
        const float x = u2f(b);
        const double r = log2_series(x);
        const float f = (float)r;
        const float g = log2f(x);
        if (f2u(g) == f2u(f)) continue;
        if (f2u(g) != f2u(across_midpoint(f, r))) {
            fprintf(stderr, "log2f(%a) = %a is neither %a nor its midpoint neighbour\n", x, g, f);
            return 1;
        }
        listed.push_back(b);
    }
    FILE* out = fopen(argv[1], "w");
    if (!out) {
        perror(argv[1]);
        return 1;
    }
    fprintf(out, "// Generated by gen_log2_table from the host libm's log2f: %zu inputs.\n",
            listed.size());
    for (size_t i = 0; i < listed.size(); ++i)
        fprintf(out, "0x%08xu,%s", listed[i], (i % 8 == 7 || i + 1 == listed.size()) ? "\n" : " ");
    if (fclose(out) != 0) {
        perror(argv[1]);
        return 1;
    }
    return 0;
}
