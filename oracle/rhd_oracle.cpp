// Oracle (test infrastructure only): C entry points over the restatement of
// the reference's sensor-development stage. Only tests/, the smoke check and
// bench.py's CPU-baseline leg load this library, and only as the checker.
//
// Suffix `_f32` runs the bit-faithful f32 restatement; `_f64` runs the same
// algorithm in double precision (the FP64 twin).
#include <cstring>

#include "capture.hpp"
#include "demosaic.hpp"
#include "opposed.hpp"
#include "rcd.hpp"
#include "sensor.hpp"

using namespace rhdo;

extern "C" {

int rhdo_normalize(const uint16_t* mosaic, size_t row_stride, int sensor_pattern,
                   const float* black, size_t repeat_w, size_t repeat_h, const float* white,
                   size_t white_count, size_t crop_x, size_t crop_y, size_t crop_w, size_t crop_h,
                   float* out) {
    LevelPattern lv{black, repeat_w, repeat_h, white, white_count};
    normalize(mosaic, row_stride, sensor_pattern, lv, crop_x, crop_y, crop_w, crop_h, out);
    return 0;
}

int rhdo_clip(float* mosaic, size_t width, size_t height, int pattern, const float limits[3],
              uint64_t stats_out[6]) {
    ClipStats st{};
    int status = clip(mosaic, width, height, width, pattern, limits, &st);
    if (stats_out) {
        stats_out[0] = st.affected_sites;
        stats_out[1] = st.changed_sites;
        stats_out[2] = st.nominal_over_white_sites;
        for (int c = 0; c < 3; ++c) stats_out[3 + c] = st.affected_by_channel[c];
    }
    return status;
}

int rhdo_demosaic_f32(const float* mosaic, size_t width, size_t height, int pattern,
                      const float balance[3], int algorithm, float* out, int threads, int poison) {
    return demosaic<float>(mosaic, width, height, width, pattern, balance, algorithm, out, threads,
                           poison != 0);
}

int rhdo_demosaic_f64(const double* mosaic, size_t width, size_t height, int pattern,
                      const double balance[3], int algorithm, double* out, int threads) {
    return demosaic<double>(mosaic, width, height, width, pattern, balance, algorithm, out,
                            threads, false);
}

// One AMaZE tile at a given origin, exporting every scratch plane for the
// reference's stage probes (amaze/mod.rs:355-430). `planes` receives 29
// planes of 65536 floats in the order of `rhdo_amaze_plane_names`.
const char* rhdo_amaze_plane_names() {
    return "cfa,green,delhvsqsum,dirwts0,dirwts1,vcd,hcd,vcdalt,hcdalt,cddiffsq,dgintv,dginth,"
           "hvwt,dgrb0,dgrb1,dgrb2_h,dgrb2_v,delp,delm,sq1m,sq1p,nyqutest,pmwt,rbm,rbp,rbint,"
           "nyquist,nyquist2,red";
}

int rhdo_amaze_tile_planes(const float* mosaic, size_t width, size_t height, int pattern,
                           size_t origin_x, size_t origin_y, float clip_point, float* planes) {
    using namespace amaze;
    Scratch<float> s;
    s.clear(true);
    Mosaic<float> m{mosaic, width, height, width, pattern};
    amaze_populate(s, m, origin_x, origin_y);
    reconstruct_tile(s, pattern, clip_point);
    const std::vector<float>* full[] = {&s.cfa,     &s.green,  &s.delhvsqsum, &s.dirwts0,
                                        &s.dirwts1, &s.vcd,    &s.hcd,        &s.vcdalt,
                                        &s.hcdalt,  &s.cddiffsq, &s.dgintv,   &s.dginth};
    const std::vector<float>* half[] = {&s.hvwt,  &s.dgrb0, &s.dgrb1,    &s.dgrb2_h, &s.dgrb2_v,
                                        &s.delp,  &s.delm,  &s.sq1m,     &s.sq1p,    &s.nyqutest,
                                        &s.pmwt,  &s.rbm,   &s.rbp,      &s.rbint};
    size_t k = 0;
    for (auto* p : full) std::memcpy(planes + (k++) * FULL, p->data(), FULL * sizeof(float));
    for (auto* p : half) {
        std::memset(planes + k * FULL, 0, FULL * sizeof(float));
        std::memcpy(planes + (k++) * FULL, p->data(), HALF * sizeof(float));
    }
    for (auto* p : {&s.nyquist, &s.nyquist2}) {
        float* dst = planes + (k++) * FULL;
        std::memset(dst, 0, FULL * sizeof(float));
        for (int i = 0; i < HALF; ++i) dst[i] = (float)(*p)[i];
    }
    std::memcpy(planes + (k++) * FULL, s.red.data(), FULL * sizeof(float));
    return 0;
}

int rhdo_capture_f32(float* rgb, size_t width, size_t height, float amount, float sigma,
                     float noise_protection, const float ceilings[3], int iterations,
                     float* guide_out, float* mask_out, float* estimate_out, int threads) {
    return capture::apply<float>(rgb, width, height, amount, sigma, noise_protection, ceilings,
                                 iterations, guide_out, mask_out, estimate_out, threads);
}

int rhdo_capture_f64(double* rgb, size_t width, size_t height, double amount, double sigma,
                     double noise_protection, const double ceilings[3], int iterations,
                     int threads) {
    return capture::apply<double>(rgb, width, height, amount, sigma, noise_protection, ceilings,
                                  iterations, nullptr, nullptr, nullptr, threads);
}

int rhdo_gaussian_kernel(float sigma, float* out, int capacity) {
    std::vector<float> k = capture::gaussian_kernel(sigma);
    if ((int)k.size() > capacity) return -(int)k.size();
    for (size_t i = 0; i < k.size(); ++i) out[i] = k[i];
    return (int)k.size();
}

int rhdo_opposed(float* mosaic, size_t width, size_t height, int pattern,
                 const float detection[3], const float* neutral_ceiling, uint64_t stats_out[9],
                 uint32_t* labels_out, uint32_t* region_count_out) {
    opposed::Stats st{};
    int status = opposed::reconstruct(mosaic, width, height, pattern, detection, neutral_ceiling,
                                      &st, labels_out, region_count_out);
    if (stats_out) {
        stats_out[0] = st.suspected;
        stats_out[1] = st.reconstructed;
        stats_out[2] = st.changed;
        stats_out[3] = st.fallback;
        stats_out[4] = st.fully_unsupported;
        stats_out[5] = st.neutralised;
        for (int c = 0; c < 3; ++c) stats_out[6 + c] = st.suspected_by_channel[c];
    }
    return status;
}

int rhdo_rcd_probe(const float* mosaic, size_t width, size_t height, int pattern, size_t red,
                   size_t green, float out[9]) {
    Mosaic<float> m{mosaic, width, height, width, pattern};
    rcd::probe_first_tile(m, red, green, out);
    return 0;
}

int rhdo_label_components(uint32_t* labels, size_t width, size_t height, uint32_t* count) {
    *count = opposed::label_components(labels, width, height);
    return 0;
}

}  // extern "C"
