// Spatial completion on the host side: plan validation, the f64 tables the
// optics and area-reduction kernels read (built here once per call and
// uploaded together), the automatic-scale search, and the C ABI entry points
// rhd_reduce_preview / rhd_optics_full / rhd_preview_dimensions /
// rhd_optics_find_scale.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "rhd_context.hpp"

namespace rhd {
namespace {

// One axis of exact area reduction (crates/core/src/resample.rs:46-112):
// destination t covers [t*s, (t+1)*s) of the source with s = source/target;
// each overlapped source cell gets weight overlap/s, computed in f64 and
// stored as f32.
struct Axis {
    std::vector<uint32_t> first, offset, count;
    std::vector<float> weights;
};

bool build_axis(uint32_t source, uint32_t target, Axis& a) {
    if (source == 0 || target == 0 || target > source) return false;
    const double scale = (double)source / (double)target;
    a.first.resize(target);
    a.offset.resize(target);
    a.count.resize(target);
    a.weights.clear();
    a.weights.reserve((size_t)source + 2 * (size_t)target);
    for (uint32_t t = 0; t < target; ++t) {
        const double left = (double)t * scale;
        const double right = (double)(t + 1) * scale;
        const uint64_t first = (uint64_t)std::floor(left);
        uint64_t end = (uint64_t)std::ceil(right);
        if (end > source) end = source;
        a.first[t] = (uint32_t)first;
        a.offset[t] = (uint32_t)a.weights.size();
        a.count[t] = (uint32_t)(end - first);
        for (uint64_t s = first; s < end; ++s) {
            const double lo = std::max(left, (double)s);
            const double hi = std::min(right, (double)(s + 1));
            a.weights.push_back((float)((hi - lo) / scale));
        }
    }
    return true;
}

optics::Model model_of(const rhd_optics_desc& d) {
    optics::Model m{};
    m.center_x = d.center_x;
    m.center_y = d.center_y;
    m.norm_unscale = d.norm_unscale;
    m.distortion = d.distortion;
    m.tca = d.tca;
    for (int i = 0; i < 3; ++i) {
        m.dist[i] = d.distortion_k[i];
        m.red[i] = d.tca_red[i];
        m.blue[i] = d.tca_blue[i];
    }
    return m;
}

bool finite_positive(double v) { return std::isfinite(v) && v > 0.0; }

// Structural checks of a resolved plan (plan.rs:173-356 guarantees these
// for plans it builds).
rhd_status validate_optics(rhd_context* ctx, const rhd_optics_desc& d) {
    if (d.distortion < RHD_DISTORTION_NONE || d.distortion > RHD_DISTORTION_PTLENS ||
        d.tca < RHD_TCA_NONE || d.tca > RHD_TCA_POLY3 || d.vignetting < RHD_VIGNETTING_NONE ||
        d.vignetting > RHD_VIGNETTING_PA)
        return fail(ctx, RHD_INVALID_INPUT, "unknown optics model");
    if (d.width < 4 || d.height < 4)
        return fail(ctx, RHD_INVALID_DIMENSIONS,
                    "cubic correction requires at least four pixels per axis");
    if (!finite_positive(d.norm_scale) || !finite_positive(d.norm_unscale) ||
        !std::isfinite(d.center_x) || !std::isfinite(d.center_y) ||
        !finite_positive((double)d.scale))
        return fail(ctx, RHD_INVALID_INPUT, "optics normalisation or scale is not finite");
    for (int i = 0; i < 3; ++i) {
        if (!std::isfinite(d.distortion_k[i]) || !std::isfinite(d.tca_red[i]) ||
            !std::isfinite(d.tca_blue[i]) || !std::isfinite(d.vignetting_k[i]))
            return fail(ctx, RHD_INVALID_INPUT, "optics coefficient is not finite");
    }
    if (d.distortion == RHD_DISTORTION_POLY3 && d.distortion_k[0] == 0.0f)
        return fail(ctx, RHD_INVALID_INPUT, "Poly3 distortion with k1 = 0");
    return RHD_OK;
}

// Whether the perimeter of the output maps inside the cubic footprint at
// `scale` (plan.rs:537-559).
bool perimeter_safe(const rhd_optics_desc& d, const optics::Model& m, float scale) {
    const uint32_t max_x = d.width - 1, max_y = d.height - 1;
    for (int index = 0; index < 64; ++index) {
        const float t = (float)index / 63.0f;
        const float px[4] = {t * (float)max_x, t * (float)max_x, 0.0f, (float)max_x};
        const float py[4] = {0.0f, (float)max_y, t * (float)max_y, t * (float)max_y};
        for (int k = 0; k < 4; ++k) {
            const uint32_t x = (uint32_t)px[k], y = (uint32_t)py[k];
            const float ix =
                optics::input_coordinate(x, d.center_x, d.norm_scale, d.norm_unscale, scale);
            const float iy =
                optics::input_coordinate(y, d.center_y, d.norm_scale, d.norm_unscale, scale);
            float s[6];
            if (!optics::source_coordinates(m, ix, iy, s)) return false;
            for (int c = 0; c < 3; ++c) {
                if (!(s[2 * c] >= 1.0f && s[2 * c + 1] >= 1.0f &&
                      s[2 * c] <= (float)(d.width - 2) && s[2 * c + 1] <= (float)(d.height - 2)))
                    return false;
            }
        }
    }
    return true;
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Byte layout of one call's tables and scratch inside the lane's optics
// buffer. Tables come first (one host-to-device copy), scratch after.
struct Layout {
    size_t steps, in_x, in_y, h_first, h_offset, h_count, h_weights, v_first, v_offset,
        v_count, v_weights, tables_end, rows, tables_begin, mid, vignetted, total;
};

}  // namespace

rhd_status spatial_on(rhd_context* ctx, Lane& L, const rhd_source* src,
                      const rhd_optics_desc* optics_desc, uint32_t target_width,
                      uint32_t target_height, bool full, const volatile uint8_t* cancel,
                      rhd_source** out, float* gpu_ms) {
    if (!ctx || !src || !out) return RHD_INVALID_INPUT;
    *out = nullptr;
    if (gpu_ms) *gpu_ms = 0.0f;
    rhd_status st = check_handle_ctx(ctx, src->ctx);
    if (st != RHD_OK) return st;
    const uint32_t W = src->width, H = src->height;
    if (full) {
        target_width = W;
        target_height = H;
    }
    if (W < 2 || H < 2)
        return fail(ctx, RHD_INVALID_DIMENSIONS, "spatial completion needs at least 2x2 pixels");
    // Rows index the launch grid's y dimension; columns stay exact in f32.
    if (H > 65535 || W > (1u << 24))
        return fail(ctx, RHD_INVALID_DIMENSIONS, "source too large for spatial completion");
    if (target_width == 0 || target_height == 0 || target_width > W || target_height > H)
        return fail(ctx, RHD_INVALID_DIMENSIONS,
                    "area reduction target must be non-zero and no larger than the source");
    rhd_optics_desc d{};
    bool vignetting = false, remap = false;
    if (optics_desc) {
        d = *optics_desc;
        if (d.width != W || d.height != H)
            return fail(ctx, RHD_BASE_MISMATCH, "optics plan dimensions differ from the source");
        if ((st = validate_optics(ctx, d)) != RHD_OK) return st;
        vignetting = d.vignetting == RHD_VIGNETTING_PA;
        remap = d.distortion != RHD_DISTORTION_NONE || d.tca != RHD_TCA_NONE;
    }
    const bool reduce = target_width != W || target_height != H;
    if (cancelled(cancel)) return fail(ctx, RHD_CANCELLED, "cancelled");

    // Host tables.
    Axis ha, va;
    if (reduce && (!build_axis(W, target_width, ha) || !build_axis(H, target_height, va)))
        return fail(ctx, RHD_INVALID_DIMENSIONS, "invalid area reduction axis");
    std::vector<float> steps, in_x, in_y;
    float start_x = 0.0f, start_y = 0.0f, ns = 0.0f;
    if (vignetting) {
        // correction.rs:220-257: x and r^2 advance by f32 additions; the
        // per-column increment 2 ns x_k + ns^2 does not depend on the row.
        ns = (float)d.norm_scale;
        const float step = 2.0f * ns;
        const float step_sq = ns * ns;
        start_x = (float)(-d.center_x);
        start_y = (float)(-d.center_y);
        steps.resize(W);
        float x = start_x;
        for (uint32_t k = 0; k < W; ++k) {
            const float product = step * x;
            steps[k] = product + step_sq;
            x += ns;
        }
    }
    if (remap) {
        in_x.resize(W);
        in_y.resize(H);
        for (uint32_t x = 0; x < W; ++x)
            in_x[x] = optics::input_coordinate(x, d.center_x, d.norm_scale, d.norm_unscale, d.scale);
        for (uint32_t y = 0; y < H; ++y)
            in_y[y] = optics::input_coordinate(y, d.center_y, d.norm_scale, d.norm_unscale, d.scale);
    }

    Layout lay{};
    size_t at = 0;
    auto place = [&](size_t bytes) {
        const size_t here = at;
        at = align256(at + bytes);
        return here;
    };
    // The checkpoints first, at a fixed offset, so a later call with the same
    // geometry finds them in place.
    lay.rows = place(vignetting ? optics_checkpoints((int)W) * H * 4 : 0);
    lay.tables_begin = at;
    lay.steps = place(steps.size() * 4);
    lay.in_x = place(in_x.size() * 4);
    lay.in_y = place(in_y.size() * 4);
    lay.h_first = place(ha.first.size() * 4);
    lay.h_offset = place(ha.offset.size() * 4);
    lay.h_count = place(ha.count.size() * 4);
    lay.h_weights = place(ha.weights.size() * 4);
    lay.v_first = place(va.first.size() * 4);
    lay.v_offset = place(va.offset.size() * 4);
    lay.v_count = place(va.count.size() * 4);
    lay.v_weights = place(va.weights.size() * 4);
    lay.tables_end = at;
    lay.mid = place(reduce ? (size_t)target_width * H * 12 : 0);
    lay.vignetted = place(vignetting && remap ? (size_t)W * H * 12 : 0);
    lay.total = at;

    std::vector<uint8_t> host(lay.tables_end - lay.tables_begin);
    auto put = [&](size_t off, const void* p, size_t bytes) {
        if (bytes) std::memcpy(host.data() + (off - lay.tables_begin), p, bytes);
    };
    put(lay.steps, steps.data(), steps.size() * 4);
    put(lay.in_x, in_x.data(), in_x.size() * 4);
    put(lay.in_y, in_y.data(), in_y.size() * 4);
    put(lay.h_first, ha.first.data(), ha.first.size() * 4);
    put(lay.h_offset, ha.offset.data(), ha.offset.size() * 4);
    put(lay.h_count, ha.count.data(), ha.count.size() * 4);
    put(lay.h_weights, ha.weights.data(), ha.weights.size() * 4);
    put(lay.v_first, va.first.data(), va.first.size() * 4);
    put(lay.v_offset, va.offset.data(), va.offset.size() * 4);
    put(lay.v_count, va.count.data(), va.count.size() * 4);
    put(lay.v_weights, va.weights.data(), va.weights.size() * 4);

    (void)hipSetDevice(ctx->device);
    void* previous = L.optics.ptr;
    if ((st = ensure_scratch(ctx, L, L.optics, std::max<size_t>(lay.total, 256))) != RHD_OK)
        return st;
    // Without vignetting the tables start at offset 0, over the checkpoints.
    if (L.optics.ptr != previous || !vignetting) L.optics_rows = Lane::RowsKey{};
    const bool rows_ready = vignetting && L.optics_rows.buffer == L.optics.ptr &&
                            L.optics_rows.width == W && L.optics_rows.height == H &&
                            L.optics_rows.ns == ns && L.optics_rows.start_x == start_x &&
                            L.optics_rows.start_y == start_y;
    rhd_source* result = nullptr;
    if ((st = new_source(ctx, &L, target_width, target_height, &result)) != RHD_OK) return st;
    uint8_t* base = (uint8_t*)L.optics.ptr;
    auto fptr = [&](size_t off) { return (float*)(base + off); };
    auto uptr = [&](size_t off) { return (const uint32_t*)(base + off); };

    auto abandon = [&](rhd_status s) {
        (void)hipStreamSynchronize(L.stream);
        rhd_source_release(result);
        return s;
    };
    if ((st = reset_stats(ctx, L)) != RHD_OK) return abandon(st);
    hipError_t e = hipEventRecord(L.ev[3], L.stream);
    if (e == hipSuccess && !host.empty())
        e = hipMemcpyAsync(base + lay.tables_begin, host.data(), host.size(),
                           hipMemcpyHostToDevice, L.stream);
    if (e != hipSuccess) return abandon(hip_fail(ctx, e, "optics tables"));

    if (!reduce && !vignetting && !remap) {
        // Nothing to do: a copy of the source (resample.rs:253-255).
        e = hipMemcpyAsync(result->d, src->d, src->bytes, hipMemcpyDeviceToDevice, L.stream);
    } else {
        OpticsLaunch a{};
        a.src = src->d;
        a.vignetted = fptr(lay.vignetted);
        a.width = (int)W;
        a.height = (int)H;
        a.vignetting = vignetting;
        a.remap = remap;
        a.direct = !reduce;
        a.rows_ready = rows_ready;
        a.vig_steps = fptr(lay.steps);
        a.vig_rows = fptr(lay.rows);
        a.vig_start_x = start_x;
        a.vig_start_y = start_y;
        a.vig_ns = ns;
        for (int i = 0; i < 3; ++i) a.vig_k[i] = d.vignetting_k[i];
        a.model = model_of(d);
        a.in_x = fptr(lay.in_x);
        a.in_y = fptr(lay.in_y);
        a.h_first = uptr(lay.h_first);
        a.h_offset = uptr(lay.h_offset);
        a.h_count = uptr(lay.h_count);
        a.h_weights = fptr(lay.h_weights);
        a.v_first = uptr(lay.v_first);
        a.v_offset = uptr(lay.v_offset);
        a.v_count = uptr(lay.v_count);
        a.v_weights = fptr(lay.v_weights);
        a.target_width = (int)target_width;
        a.target_height = (int)target_height;
        // As many output columns per workgroup as keep its corrected span
        // within one 256-pixel pass.
        const double s = (double)W / (double)target_width;
        a.group = (int)std::max(1.0, std::min(256.0, std::floor(254.0 / s)));
        a.mid = fptr(lay.mid);
        a.out = result->d;
        a.stats = L.d_stats;
        e = launch_optics(a, L.stream, PROF(L));
    }
    if (e == hipSuccess) e = hipEventRecord(L.ev[4], L.stream);
    if (e != hipSuccess) {
        L.optics_rows = Lane::RowsKey{};
        return abandon(hip_fail(ctx, e, "optics launch"));
    }
    if (vignetting) L.optics_rows = Lane::RowsKey{W, H, ns, start_x, start_y, L.optics.ptr};
    e = hipMemcpyAsync(L.h_stats, L.d_stats, sizeof(DeviceStats), hipMemcpyDeviceToHost,
                       L.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(L.stream);
    if (e != hipSuccess) return abandon(hip_fail(ctx, e, "optics"));
    const unsigned bad = L.h_stats->nonfinite;
    if (bad & 16u)
        return abandon(fail(ctx, RHD_INVALID_INPUT,
                            "optics sampled outside the image (invalid cubic footprint)"));
    if (bad & 8u)
        return abandon(fail(ctx, RHD_INVALID_INPUT,
                            "optics produced an invalid gain or a non-finite value"));
    if (cancelled(cancel)) return abandon(fail(ctx, RHD_CANCELLED, "cancelled"));
    if (gpu_ms) *gpu_ms = elapsed(L.ev[3], L.ev[4]);
    *out = result;
    return RHD_OK;
}

}  // namespace rhd

using namespace rhd;

extern "C" {

rhd_status rhd_preview_dimensions(uint32_t width, uint32_t height, uint32_t max_long_edge,
                                  uint32_t* target_width, uint32_t* target_height) {
    if (!target_width || !target_height) return RHD_INVALID_INPUT;
    if (width < 2 || height < 2 || max_long_edge < 2) return RHD_INVALID_DIMENSIONS;
    const uint64_t long_edge = std::max(width, height);
    if (long_edge <= max_long_edge) {
        *target_width = width;
        *target_height = height;
        return RHD_OK;
    }
    auto scale = [&](uint64_t dim) {
        uint64_t v = (dim * max_long_edge + long_edge / 2) / long_edge;
        return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(v, 2), dim);
    };
    *target_width = scale(width);
    *target_height = scale(height);
    return RHD_OK;
}

rhd_status rhd_optics_find_scale(const rhd_optics_desc* optics_desc, float* scale) {
    if (!optics_desc || !scale) return RHD_INVALID_INPUT;
    rhd_optics_desc d = *optics_desc;
    d.scale = 1.0f;
    if (validate_optics(nullptr, d) != RHD_OK) return RHD_INVALID_INPUT;
    const optics::Model m = model_of(d);
    float s = 1.0f;
    for (int attempt = 0; attempt < 160; ++attempt) {
        if (perimeter_safe(d, m, s)) {
            *scale = s;
            return RHD_OK;
        }
        s *= 1.025f;
    }
    return RHD_INVALID_INPUT;
}

rhd_status rhd_reduce_preview(rhd_context* ctx, const rhd_source* source,
                              const rhd_optics_desc* optics, uint32_t target_width,
                              uint32_t target_height, const volatile uint8_t* cancel,
                              rhd_source** out, float* gpu_ms) {
    if (!ctx) return RHD_INVALID_INPUT;
    return spatial_on(ctx, ctx->lanes[0], source, optics, target_width, target_height, false,
                      cancel, out, gpu_ms);
}

rhd_status rhd_color_transform(rhd_context* ctx, rhd_source* source, const float gains[3],
                               const float camera_to_rec2020[9], const volatile uint8_t* cancel,
                               float* gpu_ms) {
    if (!ctx || !source || !gains || !camera_to_rec2020) return RHD_INVALID_INPUT;
    if (gpu_ms) *gpu_ms = 0.0f;
    rhd_status st = check_handle_ctx(ctx, source->ctx);
    if (st != RHD_OK) return st;
    for (int c = 0; c < 3; ++c)
        if (!std::isfinite(gains[c]) || !(gains[c] > 0.0f))
            return fail(ctx, RHD_INVALID_INPUT, "white-balance gains must be finite and positive");
    for (int k = 0; k < 9; ++k)
        if (!std::isfinite(camera_to_rec2020[k]))
            return fail(ctx, RHD_INVALID_INPUT, "camera matrix is not finite");
    if (cancelled(cancel)) return fail(ctx, RHD_CANCELLED, "cancelled");
    Lane& L = ctx->lanes[0];
    (void)hipSetDevice(ctx->device);
    RHD_HIP(ctx, hipEventRecord(L.ev[3], L.stream));
    RHD_HIP(ctx, launch_color(source->d, (size_t)source->width * source->height, gains,
                              camera_to_rec2020, L.stream, PROF(L)));
    RHD_HIP(ctx, hipEventRecord(L.ev[4], L.stream));
    RHD_HIP(ctx, hipStreamSynchronize(L.stream));
    if (gpu_ms) *gpu_ms = elapsed(L.ev[3], L.ev[4]);
    return RHD_OK;
}

rhd_status rhd_optics_full(rhd_context* ctx, const rhd_source* source,
                           const rhd_optics_desc* optics, const volatile uint8_t* cancel,
                           rhd_source** out, float* gpu_ms) {
    if (!ctx) return RHD_INVALID_INPUT;
    return spatial_on(ctx, ctx->lanes[0], source, optics, 0, 0, true, cancel, out, gpu_ms);
}

}  // extern "C"
