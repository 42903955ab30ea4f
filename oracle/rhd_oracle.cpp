// Oracle (test infrastructure only): C entry points over the restatement of
// the reference's sensor-development stage. Only tests/, the smoke check and
// bench.py's CPU-baseline leg load this library, and only as the checker.
//
// Suffix `_f32` runs the bit-faithful f32 restatement; `_f64` runs the same
// algorithm in double precision (the FP64 twin).
#include <cmath>
#include <cstring>

#include "arw.hpp"
#include "capture.hpp"
#include "color.hpp"
#include "demosaic.hpp"
#include "denoise.hpp"
#include "detail.hpp"
#include "opposed.hpp"
#include "optics.hpp"
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

// The FP64 twin of rhdo_normalize: (u16 - black) / (white - black) in double.
int rhdo_normalize_f64(const uint16_t* mosaic, size_t row_stride, int sensor_pattern,
                       const float* black, size_t repeat_w, size_t repeat_h, const float* white,
                       size_t white_count, size_t crop_x, size_t crop_y, size_t crop_w,
                       size_t crop_h, double* out) {
    LevelPattern lv{black, repeat_w, repeat_h, white, white_count};
    for (size_t y = 0; y < crop_h; ++y)
        for (size_t x = 0; x < crop_w; ++x) {
            size_t sx = crop_x + x, sy = crop_y + y;
            float b, w;
            levels_at(lv, sensor_pattern, sx, sy, b, w);
            out[y * crop_w + x] =
                ((double)mosaic[sy * row_stride + sx] - (double)b) / ((double)w - (double)b);
        }
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

// sharpening.rs:328-384 with the kernel of gaussian.rs:9-17 for `sigma`.
int rhdo_blur_f32(const float* in, size_t width, size_t height, float sigma, float* out) {
    const size_t n = width * height;
    std::vector<float> input(in, in + n), temporary(n), output(n);
    capture::blur(input, temporary, output, width, height, capture::gaussian_kernel(sigma), 1);
    for (size_t i = 0; i < n; ++i) out[i] = output[i];
    return 0;
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

// The FP64 twin of rhdo_opposed: the same algorithm in double precision.
int rhdo_opposed_f64(double* mosaic, size_t width, size_t height, int pattern,
                     const double detection[3], const double* neutral_ceiling,
                     uint64_t stats_out[9]) {
    opposed::Stats st{};
    int status = opposed::reconstruct(mosaic, width, height, pattern, detection, neutral_ceiling,
                                      &st, nullptr, nullptr);
    if (stats_out) {
        const uint64_t v[9] = {st.suspected, st.reconstructed, st.changed, st.fallback,
                               st.fully_unsupported, st.neutralised, st.suspected_by_channel[0],
                               st.suspected_by_channel[1], st.suspected_by_channel[2]};
        std::memcpy(stats_out, v, sizeof(v));
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

// Same field order and types as rhd_optics_desc in include/rhd.h, so one
// ctypes structure serves both libraries.
struct OpticsDesc {
    uint32_t width, height;
    double norm_scale, norm_unscale, center_x, center_y;
    float scale;
    int32_t distortion;
    float distortion_k[3];
    int32_t tca;
    float tca_red[3], tca_blue[3];
    int32_t vignetting;
    float vignetting_k[3];
};

static optics::Execution execution_of(const OpticsDesc* d) {
    optics::Execution e{};
    e.width = d->width;
    e.height = d->height;
    e.norm_scale = d->norm_scale;
    e.norm_unscale = d->norm_unscale;
    e.center_x = d->center_x;
    e.center_y = d->center_y;
    e.scale = d->scale;
    e.distortion = d->distortion;
    e.tca = d->tca;
    e.vignetting = d->vignetting;
    for (int i = 0; i < 3; ++i) {
        e.dist[i] = d->distortion_k[i];
        e.red[i] = d->tca_red[i];
        e.blue[i] = d->tca_blue[i];
        e.vig[i] = d->vignetting_k[i];
    }
    return e;
}

// rgb: tight interleaved RGB, vignetted in place; out: the corrected image.
int rhdo_optics_apply(float* rgb, const OpticsDesc* d, float* out) {
    return optics::apply(rgb, execution_of(d), out);
}

int rhdo_optics_map(const OpticsDesc* d, float scale, size_t x, size_t y, float out[6]) {
    return optics::map_coordinates(execution_of(d), scale, x, y, out);
}

int rhdo_optics_find_scale(const OpticsDesc* d, float* out) {
    return optics::find_scale(execution_of(d), out);
}

// model: 1 Poly3, 2 Poly5, 3 PTLens; 4 = TCA Poly3 reverse with k = [v, c, b].
int rhdo_inverse_model(int model, const float k[3], float x, float y, float out[2]) {
    switch (model) {
        case 1: optics::undist_poly3(x, y, k[0], &out[0], &out[1]); return 0;
        case 2: optics::undist_poly5(x, y, k[0], k[1], &out[0], &out[1]); return 0;
        case 3: optics::undist_ptlens(x, y, k[0], k[1], k[2], &out[0], &out[1]); return 0;
        case 4: optics::tca_poly3_channel(x, y, k, &out[0], &out[1]); return 0;
        default: return -1;
    }
}

int rhdo_cubic_sample(const float* rgb, size_t width, size_t height, int channel, float x,
                      float y, float* out) {
    return optics::cubic_sample(rgb, width * 3, width, height, channel, x, y, out);
}

float rhdo_cubic_weight(float distance) { return optics::cubic_weight(distance); }

// Writes first/offset/count per target (capacity `target`) and the weights
// (capacity `weight_capacity`); returns the weight count, or -1.
long rhdo_area_axis(size_t source, size_t target, uint64_t* first, uint64_t* offset,
                    uint64_t* count, float* weights, size_t weight_capacity) {
    optics::Axis a;
    if (!optics::make_axis(source, target, &a)) return -1;
    if (a.weights.size() > weight_capacity) return -1;
    for (size_t t = 0; t < target; ++t) {
        first[t] = a.first[t];
        offset[t] = a.offset[t];
        count[t] = a.count[t];
    }
    for (size_t i = 0; i < a.weights.size(); ++i) weights[i] = a.weights[i];
    return (long)a.weights.size();
}

int rhdo_resize_area(const float* rgb, size_t w, size_t h, float* out, size_t tw, size_t th) {
    return optics::resize_area(rgb, w, h, out, tw, th);
}

// rgb: n interleaved pixels, in place; m: row-major camera -> Rec.2020.
int rhdo_color(float* rgb, size_t n, const float gains[3], const float m[9]) {
    return color::apply(rgb, n, gains, m);
}

// False-colour suppression in place on tight interleaved RGB.
int rhdo_false_color(float* rgb, size_t width, size_t height, int steps, const float balance[3],
                     const float ceilings[3]) {
    return detail::false_color(rgb, width, height, steps, balance, ceilings);
}

int rhdo_preview_dimensions(size_t w, size_t h, size_t max_long, size_t* tw, size_t* th) {
    return optics::preview_dimensions(w, h, max_long, tw, th);
}

// Noise profiles {luma, c1, c2} of tight interleaved RGB as {a, b, bins}
// triples. Returns 0, 1 when the image is too small to profile, 2 for an
// invalid balance.
int rhdo_noise_profiles(const float* rgb, size_t width, size_t height, const float balance[3],
                        double out[9]) {
    denoise::Profile p[3];
    const int st = denoise::profiles(rgb, width, height, balance, p);
    if (st == 0)
        for (int c = 0; c < 3; ++c) {
            out[3 * c] = p[c].a;
            out[3 * c + 1] = p[c].b;
            out[3 * c + 2] = (double)p[c].bins;
        }
    return st;
}

// One plane's profile with its own brightness plane (noise.rs:48-125).
int rhdo_estimate_noise(const float* plane, const float* brightness, size_t width, size_t height,
                        double out[3]) {
    std::vector<float> p(plane, plane + width * height), b(brightness, brightness + width * height);
    denoise::Profile r;
    if (denoise::estimate(p, b, width, height, r)) return 1;
    out[0] = r.a;
    out[1] = r.b;
    out[2] = (double)r.bins;
    return 0;
}

static void unpack_profiles(const float prof[6], denoise::Profile p[3]) {
    for (int c = 0; c < 3; ++c) {
        p[c].a = prof[2 * c];
        p[c].b = prof[2 * c + 1];
    }
}

// prof = {a, b} of luma, c1, c2.
int rhdo_chroma_denoise(float* rgb, size_t width, size_t height, float strength,
                        const float balance[3], const float ceilings[3], const float prof[6]) {
    denoise::Profile p[3];
    unpack_profiles(prof, p);
    return denoise::chroma(rgb, width, height, strength, balance, ceilings, p);
}

int rhdo_luma_denoise(float* rgb, size_t width, size_t height, float strength, float detail_,
                      const float balance[3], const float ceilings[3], const float prof[2]) {
    denoise::Profile p;
    p.a = prof[0];
    p.b = prof[1];
    return denoise::luma(rgb, width, height, strength, detail_, balance, ceilings, p);
}

// Band-limited helpers for the pinning tests.
int rhdo_b3_smooth(const float* src, size_t width, size_t height, size_t step, float* out) {
    std::vector<float> tmp;
    denoise::smooth(src, width, height, step, tmp, out);
    return 0;
}

int rhdo_local_mean_square(const float* band, size_t width, size_t height, float* out) {
    std::vector<float> b(band, band + width * height), tmp, o;
    denoise::local_mean_square(b, width, height, tmp, o);
    std::memcpy(out, o.data(), o.size() * sizeof(float));
    return 0;
}

float rhdo_mad(float* values, size_t n) {
    std::vector<float> v(values, values + n);
    return denoise::mad(v);
}

// The detail stage's steps in order (crates/core/src/pipeline/detail.rs:
// 210-307): noise profiles of the stage input, false colour, chroma NR,
// luminance NR, capture sharpening. prof_out (may be null) receives the
// profiles as {a, b, bins} x {luma, c1, c2}; *profiled is 1 when they
// exist. Returns the first failing step's status.
int rhdo_detail_stage(float* rgb, size_t width, size_t height, int false_colour, float chroma,
                      float luminance, float luminance_detail, const float balance[3],
                      const float ceilings[3], int capture_enabled, float amount, float sigma,
                      float noise_protection, int iterations, double prof_out[9],
                      int* profiled) {
    denoise::Profile p[3];
    int have = 0;
    if (chroma > 0.0f || luminance > 0.0f) {
        const int st = denoise::profiles(rgb, width, height, balance, p);
        if (st == 2) return 2;
        have = st == 0;
    }
    if (profiled) *profiled = have;
    if (prof_out && have)
        for (int c = 0; c < 3; ++c) {
            prof_out[3 * c] = p[c].a;
            prof_out[3 * c + 1] = p[c].b;
            prof_out[3 * c + 2] = (double)p[c].bins;
        }
    int st = detail::false_color(rgb, width, height, false_colour, balance, ceilings);
    if (st) return 10 + st;
    if (have && chroma > 0.0f && (st = denoise::chroma(rgb, width, height, chroma, balance,
                                                       ceilings, p)))
        return 20 + st;
    if (have && luminance > 0.0f && (st = denoise::luma(rgb, width, height, luminance,
                                                        luminance_detail, balance, ceilings, p[0])))
        return 30 + st;
    if (capture_enabled && amount > 0.0f &&
        (st = capture::apply<float>(rgb, width, height, amount, sigma, noise_protection, ceilings,
                                    iterations, nullptr, nullptr, nullptr, 1)))
        return 40 + st;
    return 0;
}

// Sony compressed ARW strip -> u16 mosaic (width x height).
int rhdo_decode_arw2(const uint8_t* strip, size_t bytes, size_t width, size_t height,
                     const uint16_t curve[4], uint16_t* out) {
    return arw::decode_arw2(strip, bytes, width, height, curve, out);
}

// The tone curve table (4096 entries) and its dithering lookup.
int rhdo_arw_curve(const uint16_t curve[4], uint16_t table[4096], uint16_t base[4096],
                   uint16_t delta[4096]) {
    const std::vector<uint16_t> t = arw::curve_table(curve);
    const arw::Lookup l = arw::lookup(t);
    for (size_t i = 0; i < 4096; ++i) {
        table[i] = t[i];
        base[i] = l.base[i];
        delta[i] = l.delta[i];
    }
    return 0;
}

// out = {width, height, bits_per_sample, compression, offset, bytes, curve[4],
// has_curve, kind}; returns 0, 1 (not TIFF), 2 (corrupt), 3 (no raw IFD).
int rhdo_arw_probe(const uint8_t* file, size_t size, uint64_t out[12]) {
    arw::Layout l;
    const int st = arw::probe(file, size, l);
    if (st == 0) {
        const uint64_t v[12] = {l.width, l.height, l.bits_per_sample, l.compression, l.offset,
                                l.bytes, l.curve[0], l.curve[1], l.curve[2], l.curve[3],
                                (uint64_t)l.has_curve, (uint64_t)l.kind};
        std::memcpy(out, v, sizeof(v));
    }
    return st;
}

// The platform libm's cbrtf (fn 0), log2f (fn 1) or cosf (fn 2) over n
// floats: what the reference's f32::cbrt, f32::log2 and f32::cos return on
// this host.
int rhdo_libm(int fn, const float* in, float* out, size_t n) {
    if (fn < 0 || fn > 2) return 1;
    for (size_t i = 0; i < n; ++i)
        out[i] = fn == 0 ? std::cbrt(in[i]) : fn == 1 ? std::log2(in[i]) : std::cos(in[i]);
    return 0;
}

}  // extern "C"
