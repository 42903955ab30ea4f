// librhd C ABI: contexts, resident handles, validation and stage sequencing.
//
// Mirrors crates/gpu/src/sensor/{develop,normalize,highlight,demosaic}.rs:
// the same stage order, the same routing errors (UnsupportedEdits and
// OverBudget keep the device; everything else is a device or input error),
// a logical memory budget held by every live handle and scratch block
// (crates/gpu/src/memory.rs:42-81), and a cancellation flag polled between
// stages (crates/core/src/cancel.rs:13-42).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "rhd_context.hpp"
#include "rhd_internal.hpp"

using namespace rhd;

namespace rhd {
void context_unref(rhd_context* ctx) {
    if (ctx->refs.fetch_sub(1) != 1) return;
    (void)hipSetDevice(ctx->device);
    for (Lane& L : ctx->lanes) {
        if (!L.open) continue;
        for (auto& ev : L.ev)
            if (ev) (void)hipEventDestroy(ev);
        for (hipEvent_t ev : {L.fork, L.join})
            if (ev) (void)hipEventDestroy(ev);
        if (L.d_stats) (void)hipFree(L.d_stats);
        if (L.h_stats) (void)hipHostFree(L.h_stats);
        if (L.aux) (void)hipStreamDestroy(L.aux);
        if (L.stream) (void)hipStreamDestroy(L.stream);
    }
    for (const PooledBlock& b : ctx->pool) (void)hipFree(b.ptr);
    delete ctx;
}
}  // namespace rhd


namespace rhd {

hipEvent_t Profiler::take_event() {
    if (!free_.empty()) {
        hipEvent_t e = free_.back();
        free_.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void Profiler::begin(hipStream_t s, const char* name, uint64_t bytes) {
    if (!only.empty() && only != name) {
        skipped_ = true;
        return;
    }
    size_t k = 0;
    while (k < totals.size() && totals[k].name != name) ++k;
    if (k == totals.size()) totals.push_back(Total{name, 0.0, 0, 0});
    open_ = take_event();
    open_total_ = k;
    open_bytes_ = bytes;
    if (open_) (void)hipEventRecord(open_, s);
}

void Profiler::end(hipStream_t s) {
    if (skipped_) {
        skipped_ = false;
        return;
    }
    if (!open_) return;
    hipEvent_t stop = take_event();
    if (!stop) {
        free_.push_back(open_);
        open_ = nullptr;
        return;
    }
    (void)hipEventRecord(stop, s);
    pending_.push_back(Pending{open_total_, open_, stop, open_bytes_});
    open_ = nullptr;
    if (pending_.size() > 4096) collect();
}

void Profiler::collect() {
    for (const Pending& p : pending_) {
        float ms = 0.0f;
        if (hipEventSynchronize(p.stop) == hipSuccess &&
            hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
            totals[p.total].ms += ms;
            totals[p.total].launches += 1;
            totals[p.total].bytes = p.bytes;
        }
        free_.push_back(p.start);
        free_.push_back(p.stop);
    }
    pending_.clear();
}

void Profiler::reset() {
    collect();
    totals.clear();
}

Profiler::~Profiler() {
    for (const Pending& p : pending_) {
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    for (hipEvent_t e : free_) (void)hipEventDestroy(e);
}

}  // namespace rhd

namespace rhd {

rhd_status fail(rhd_context* ctx, rhd_status status, const std::string& msg) {
    if (ctx) {
        std::lock_guard<std::mutex> g(ctx->mu);
        ctx->last_error = msg;
    }
    return status;
}

rhd_status hip_fail(rhd_context* ctx, hipError_t e, const char* what) {
    return fail(ctx, RHD_SYNCHRONIZATION, std::string(what) + ": " + hipGetErrorString(e));
}


bool cancelled(const volatile uint8_t* flag) {
    return flag && __atomic_load_n(flag, __ATOMIC_ACQUIRE) != 0;
}

rhd_status reserve(rhd_context* ctx, uint64_t bytes) {
    uint64_t cur = ctx->reserved.load();
    while (true) {
        if (cur + bytes > ctx->budget.load())
            return fail(ctx, RHD_OVER_BUDGET,
                        "GPU work does not fit: " + std::to_string(cur + bytes) +
                            " bytes exceed the budget of " + std::to_string(ctx->budget.load()));
        if (ctx->reserved.compare_exchange_weak(cur, cur + bytes)) return RHD_OK;
    }
}
void unreserve(rhd_context* ctx, uint64_t bytes) { ctx->reserved.fetch_sub(bytes); }

// Frees pooled blocks (largest first) until the pool plus live bytes plus
// `need` fit the budget, or all of them when `all`.
void pool_trim(rhd_context* ctx, uint64_t need, bool all) {
    std::vector<PooledBlock> victims;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        std::sort(ctx->pool.begin(), ctx->pool.end(),
                  [](const PooledBlock& a, const PooledBlock& b) { return a.bytes < b.bytes; });
        while (!ctx->pool.empty() &&
               (all || ctx->reserved.load() + ctx->pooled + need > ctx->budget.load())) {
            victims.push_back(ctx->pool.back());
            ctx->pooled -= ctx->pool.back().bytes;
            ctx->pool.pop_back();
        }
    }
    for (const PooledBlock& b : victims) (void)hipFree(b.ptr);
}

void* pool_take(rhd_context* ctx, uint64_t bytes) {
    std::lock_guard<std::mutex> g(ctx->mu);
    for (size_t i = 0; i < ctx->pool.size(); ++i)
        if (ctx->pool[i].bytes == bytes) {
            void* p = ctx->pool[i].ptr;
            ctx->pool[i] = ctx->pool.back();
            ctx->pool.pop_back();
            ctx->pooled -= bytes;
            return p;
        }
    return nullptr;
}

// Returns a block to the pool. Every entry point finishes its own work
// before returning, so a released block has no work pending on any lane.
void device_free(rhd_context* ctx, void* p, uint64_t bytes) {
    if (!p) return;
    unreserve(ctx, bytes);
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->pool.push_back(PooledBlock{p, bytes});
    ctx->pooled += bytes;
}

void free_buffer(rhd_context* ctx, DeviceBuffer& b) {
    device_free(ctx, b.ptr, b.bytes);
    b.ptr = nullptr;
    b.bytes = 0;
}

void drop_scratch(rhd_context* ctx, Lane& L) {
    if (!L.open) return;
    (void)hipStreamSynchronize(L.stream);
    free_buffer(ctx, L.amaze);
    free_buffer(ctx, L.amaze_areas);
    free_buffer(ctx, L.rcd);
    free_buffer(ctx, L.capture);
    free_buffer(ctx, L.opposed);
    free_buffer(ctx, L.optics);
    free_buffer(ctx, L.raw);
    L.optics_rows = Lane::RowsKey{};
}

// Reserve and allocate: a pooled block of the same size if there is one,
// else a new one. When the budget is short, the lane's cached scratch is
// dropped once; pooled blocks are freed as needed to keep the device's
// footprint within the budget.
rhd_status device_alloc(rhd_context* ctx, Lane* L, uint64_t bytes, void** out) {
    rhd_status st = reserve(ctx, bytes);
    if (st == RHD_OVER_BUDGET && L) {
        drop_scratch(ctx, *L);
        st = reserve(ctx, bytes);
    }
    if (st != RHD_OK) return st;
    if ((*out = pool_take(ctx, bytes)) != nullptr) return RHD_OK;
    pool_trim(ctx, bytes, false);
    hipError_t e = hipMalloc(out, bytes ? bytes : 1);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        pool_trim(ctx, 0, true);
        e = hipMalloc(out, bytes ? bytes : 1);
    }
    if (e != hipSuccess) {
        unreserve(ctx, bytes);
        (void)hipGetLastError();
        return fail(ctx, RHD_OVER_BUDGET, std::string("device allocation failed: ") +
                                              hipGetErrorString(e));
    }
    return RHD_OK;
}

rhd_status ensure_scratch(rhd_context* ctx, Lane& L, DeviceBuffer& b, uint64_t bytes) {
    if (b.bytes >= bytes) return RHD_OK;
    if (b.ptr) {
        (void)hipStreamSynchronize(L.stream);
        free_buffer(ctx, b);
    }
    void* p = nullptr;
    rhd_status st = device_alloc(ctx, &L, bytes, &p);
    if (st != RHD_OK) return st;
    b.ptr = p;
    b.bytes = bytes;
    return RHD_OK;
}

// Creates lane `i` on first use.
rhd_status lane_open(rhd_context* ctx, int i, Lane** out) {
    if (i < 0 || i >= MAX_LANES) return fail(ctx, RHD_INVALID_INPUT, "no such lane");
    Lane& L = ctx->lanes[i];
    std::lock_guard<std::mutex> g(ctx->mu);
    if (!L.open) {
        L.index = i;
        bool ok = hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) == hipSuccess &&
                  hipStreamCreateWithFlags(&L.aux, hipStreamNonBlocking) == hipSuccess &&
                  hipEventCreateWithFlags(&L.fork, hipEventDisableTiming) == hipSuccess &&
                  hipEventCreateWithFlags(&L.join, hipEventDisableTiming) == hipSuccess &&
                  hipMalloc(&L.d_stats, sizeof(DeviceStats)) == hipSuccess &&
                  hipHostMalloc((void**)&L.h_stats, sizeof(DeviceStats), hipHostMallocDefault) ==
                      hipSuccess;
        for (auto& ev : L.ev) ok = ok && hipEventCreate(&ev) == hipSuccess;
        if (!ok) {
            for (auto& ev : L.ev)
                if (ev) (void)hipEventDestroy(ev);
            for (hipEvent_t ev : {L.fork, L.join})
                if (ev) (void)hipEventDestroy(ev);
            if (L.d_stats) (void)hipFree(L.d_stats);
            if (L.h_stats) (void)hipHostFree(L.h_stats);
            if (L.aux) (void)hipStreamDestroy(L.aux);
            if (L.stream) (void)hipStreamDestroy(L.stream);
            L = Lane{};
            ctx->last_error = "could not create a stream, events or statistics for a lane";
            return RHD_UNSUPPORTED;
        }
        L.open = true;
    }
    *out = &L;
    return RHD_OK;
}

bool finite_positive3(const float v[3]) {
    for (int c = 0; c < 3; ++c)
        if (!std::isfinite(v[c]) || v[c] <= 0.0f) return false;
    return true;
}

// crates/core/src/pipeline/sensor.rs:508-528 (development_geometry),
// :544-596 (validate_levels), :180-202 (validate_frame) and the
// normalization output bounds of :227-244.
rhd_status validate_desc(rhd_context* ctx, const rhd_sensor_desc* d) {
    if (!d) return fail(ctx, RHD_INVALID_INPUT, "missing sensor description");
    if (d->sensor_width == 0 || d->sensor_height == 0 || d->row_stride < d->sensor_width)
        return fail(ctx, RHD_INVALID_DIMENSIONS, "decoded row stride is shorter than the sensor width");
    if (d->sensor_pattern < 0 || d->sensor_pattern > 3)
        return fail(ctx, RHD_INVALID_INPUT, "unsupported CFA pattern");
    if (d->crop_width == 0 || d->crop_height == 0 ||
        (uint64_t)d->crop_x + d->crop_width > d->sensor_width ||
        (uint64_t)d->crop_y + d->crop_height > d->sensor_height)
        return fail(ctx, RHD_INVALID_INPUT, "crop region is outside the sensor");
    if (d->crop_width < 2 || d->crop_height < 2)
        return fail(ctx, RHD_INVALID_DIMENSIONS,
                    "normalization output must be at least 2x2 and no larger than the resolved crop");
    if ((uint64_t)d->crop_width * d->crop_height > (1ull << 31))
        return fail(ctx, RHD_INVALID_DIMENSIONS, "crop exceeds 2^31 samples");
    uint32_t rw = d->black_repeat_width, rh = d->black_repeat_height;
    if (!d->black_levels || rw == 0 || rh == 0)
        return fail(ctx, RHD_INVALID_INPUT, "a non-empty one-component black repeat pattern is required");
    uint32_t expected = rw * rh;
    for (uint32_t k = 0; k < expected; ++k)
        if (!std::isfinite(d->black_levels[k]))
            return fail(ctx, RHD_INVALID_INPUT, "black levels must be finite");
    if (!d->white_levels || !(d->white_count == 1 || d->white_count == 3 || d->white_count == expected))
        return fail(ctx, RHD_INVALID_INPUT, "expected 1, 3, or repeat-pattern white levels");
    for (uint32_t k = 0; k < d->white_count; ++k)
        if (!std::isfinite(d->white_levels[k]))
            return fail(ctx, RHD_INVALID_INPUT, "all white levels must be finite");
    for (uint32_t y = 0; y < rh; ++y)
        for (uint32_t x = 0; x < rw; ++x) {
            uint32_t index = y * rw + x;
            float black = d->black_levels[index];
            float white = d->white_count == 1   ? d->white_levels[0]
                          : d->white_count == 3 ? d->white_levels[color_at(d->sensor_pattern, x, y)]
                                                : d->white_levels[index];
            if (white <= black)
                return fail(ctx, RHD_INVALID_INPUT, "white level must exceed black level");
        }
    switch (d->highlight) {
        case RHD_HIGHLIGHT_OFF:
            break;
        case RHD_HIGHLIGHT_CLIP:
            if (!finite_positive3(d->clip_levels))
                return fail(ctx, RHD_INVALID_INPUT, "highlight levels must be finite and positive");
            break;
        case RHD_HIGHLIGHT_OPPOSED:
            if (!finite_positive3(d->detection_levels))
                return fail(ctx, RHD_INVALID_INPUT, "highlight levels must be finite and positive");
            if (d->has_neutral_ceiling && !finite_positive3(d->neutral_ceiling))
                return fail(ctx, RHD_INVALID_INPUT, "neutral ceiling must be finite and positive");
            break;
        default:
            return fail(ctx, RHD_INVALID_INPUT, "unknown highlight method");
    }
    if (d->demosaic < RHD_DEMOSAIC_BILINEAR || d->demosaic > RHD_DEMOSAIC_AMAZE)
        return fail(ctx, RHD_INVALID_INPUT, "unknown demosaic algorithm");
    if (!finite_positive3(d->demosaic_balance))
        return fail(ctx, RHD_INVALID_INPUT, "effective R, G, and B gains must be finite and positive");
    return RHD_OK;
}

bool capture_active(const rhd_detail_desc* dd) {
    return dd && dd->capture_enabled && dd->amount > 0.0f;
}

bool noise_active(const rhd_detail_desc* dd) {
    return dd && (dd->nr_chroma > 0.0f || dd->nr_luminance > 0.0f);
}

bool detail_active(const rhd_detail_desc* dd) {
    return capture_active(dd) || (dd && dd->false_colour > 0) || noise_active(dd);
}

// Bytes of lane scratch the detail stage works in, then a small area at its
// end that holds the noise profiles between the estimate and the filters.
static uint64_t detail_work_bytes(int w, int h, const rhd_detail_desc* dd) {
    uint64_t bytes = capture_scratch_bytes(w, h);
    if (noise_active(dd)) {
        bytes = std::max<uint64_t>(bytes, noise_filter_bytes(w, h));
        const NoiseGeometry g = noise_geometry(w, h);
        if (g.tiles > 0) bytes = std::max<uint64_t>(bytes, g.bytes);
    }
    return (bytes + 255) & ~(uint64_t)255;
}

uint64_t detail_scratch_bytes(int w, int h, const rhd_detail_desc* dd) {
    return detail_work_bytes(w, h, dd) + 256;
}

// crates/edit/src/sharpening.rs:50-81 and crates/core/src/sharpening.rs:26-42.
rhd_status validate_detail(rhd_context* ctx, const rhd_detail_desc* dd) {
    if (!dd) return RHD_OK;
    auto in = [](float v, float lo, float hi) { return std::isfinite(v) && v >= lo && v <= hi; };
    if (!in(dd->amount, 0.0f, 1.0f) || !in(dd->sigma, 0.3f, 1.2f) ||
        !in(dd->noise_protection, 0.0f, 1.0f))
        return fail(ctx, RHD_INVALID_INPUT, "capture sharpening settings are out of range");
    if (dd->false_colour < 0 || dd->false_colour > 5)
        return fail(ctx, RHD_INVALID_INPUT, "false-colour steps must be in 0..=5");
    // crates/edit/src/noise_reduction.rs:63-91.
    if (!in(dd->nr_chroma, 0.0f, 1.0f) || !in(dd->nr_luminance, 0.0f, 1.0f) ||
        !in(dd->nr_luminance_detail, 0.0f, 1.0f))
        return fail(ctx, RHD_INVALID_INPUT, "noise reduction settings must be finite in 0..=1");
    if (dd->false_colour > 0 || noise_active(dd)) {
        if (!finite_positive3(dd->balance))
            return fail(ctx, RHD_INVALID_INPUT, "channel balance must be finite and positive");
        for (int c = 0; c < 3; ++c)
            if (std::isnan(dd->ceilings[c]))
                return fail(ctx, RHD_INVALID_INPUT, "detail ceilings must not be NaN");
    }
    if (capture_active(dd)) {
        if (!finite_positive3(dd->ceilings))
            return fail(ctx, RHD_INVALID_INPUT, "capture ceilings must be finite and positive");
        if (dd->iterations < 0 || dd->iterations > 64)
            return fail(ctx, RHD_INVALID_INPUT, "capture iterations must be in 0..=64");
    }
    return RHD_OK;
}

// crates/core/src/pipeline/sensor.rs:598-610, tabulated over the combined
// period of the black pattern and the 2x2 CFA.
rhd_status level_table(rhd_context* ctx, const rhd_sensor_desc* d, LevelTable& t) {
    uint32_t rw = d->black_repeat_width, rh = d->black_repeat_height;
    uint32_t px = rw % 2 ? 2 * rw : rw, py = rh % 2 ? 2 * rh : rh;
    if (px > 16 || py > 16)
        return fail(ctx, RHD_UNSUPPORTED_EDITS, "black level repeat pattern larger than 8x8");
    t.period_x = (int)px;
    t.period_y = (int)py;
    for (uint32_t y = 0; y < py; ++y)
        for (uint32_t x = 0; x < px; ++x) {
            uint32_t index = (y % rh) * rw + (x % rw);
            float black = d->black_levels[index];
            float white = d->white_count == 1   ? d->white_levels[0]
                          : d->white_count == 3 ? d->white_levels[color_at(d->sensor_pattern, x, y)]
                                                : d->white_levels[index];
            t.black[y * px + x] = black;
            t.range[y * px + x] = white - black;
        }
    return RHD_OK;
}

int crop_pattern(const rhd_sensor_desc* d) {
    // BayerPattern::shifted (crates/image/src/lib.rs:119-135).
    for (int candidate = 0; candidate < 4; ++candidate) {
        bool same = true;
        for (int y = 0; y < 2 && same; ++y)
            for (int x = 0; x < 2 && same; ++x)
                same = color_at(candidate, x, y) ==
                       color_at(d->sensor_pattern, x + (int)d->crop_x, y + (int)d->crop_y);
        if (same) return candidate;
    }
    return d->sensor_pattern;
}

rhd_status new_plane(rhd_context* ctx, Lane* L, uint32_t w, uint32_t h, int32_t pattern,
                     rhd_plane** out) {
    uint64_t bytes = (uint64_t)w * h * sizeof(float);
    void* p = nullptr;
    rhd_status st = device_alloc(ctx, L, bytes, &p);
    if (st != RHD_OK) return st;
    ctx->refs.fetch_add(1);
    *out = new rhd_plane{ctx, (float*)p, w, h, pattern, bytes};
    return RHD_OK;
}

rhd_status new_source(rhd_context* ctx, Lane* L, uint32_t w, uint32_t h, rhd_source** out) {
    uint64_t bytes = (uint64_t)w * h * 3 * sizeof(float);
    void* p = nullptr;
    rhd_status st = device_alloc(ctx, L, bytes, &p);
    if (st != RHD_OK) return st;
    ctx->refs.fetch_add(1);
    *out = new rhd_source{ctx, (float*)p, w, h, bytes};
    return RHD_OK;
}

// A fill kernel, not hipMemsetAsync, so that the sequence can be captured
// into a HIP graph.
rhd_status reset_stats(rhd_context* ctx, Lane& L) {
    RHD_HIP(ctx, launch_zero(L.d_stats, sizeof(DeviceStats), L.stream));
    return RHD_OK;
}

// Reads the device counters back and turns the non-finite flags into the
// reference's errors.
rhd_status finish_stats(rhd_context* ctx, Lane& L, rhd_metrics* m, int highlight) {
    RHD_HIP(ctx, hipMemcpyAsync(L.h_stats, L.d_stats, sizeof(DeviceStats),
                                hipMemcpyDeviceToHost, L.stream));
    RHD_HIP(ctx, hipStreamSynchronize(L.stream));
    return stats_to_metrics(ctx, *L.h_stats, m, highlight);
}

rhd_status stats_to_metrics(rhd_context* ctx, const DeviceStats& s, rhd_metrics* m, int highlight) {
    if (m) {
        m->readback_bytes += sizeof(DeviceStats);
        m->highlight_method = highlight;
        if (highlight == RHD_HIGHLIGHT_CLIP) {
            m->clip.affected_sites = s.clip[0];
            m->clip.changed_sites = s.clip[1];
            m->clip.nominal_over_white_sites = s.clip[2];
            for (int c = 0; c < 3; ++c) m->clip.affected_by_channel[c] = s.clip[3 + c];
        } else if (highlight == RHD_HIGHLIGHT_OPPOSED) {
            m->opposed.suspected_clipped_sites = s.detect[0];
            for (int c = 0; c < 3; ++c) m->opposed.suspected_by_channel[c] = s.detect[3 + c];
            m->opposed.reconstructed_sites = s.opposed[0];
            m->opposed.changed_sites = s.opposed[1];
            m->opposed.fallback_sites = s.opposed[2];
            m->opposed.fully_unsupported_sites = s.opposed[3];
            m->opposed.neutralised_sites = s.opposed[4];
            m->opposed.regions = s.regions;
        }
    }
    if (s.nonfinite & 1u)
        return fail(ctx, RHD_INVALID_INPUT, "highlight processing received a non-finite sample");
    if (s.nonfinite & 2u)
        return fail(ctx, RHD_INVALID_INPUT, "demosaicing produced non-finite image data");
    if (s.nonfinite & 4u)
        return fail(ctx, RHD_INVALID_INPUT, "capture sharpening received non-finite image data");
    return RHD_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();
        return 0.0f;
    }
    return ms;
}

// Highlight on a plane already on the stream (normalize may have fused Clip
// and the detection counts). Opposed always runs here.
rhd_status run_opposed(rhd_context* ctx, Lane& L, rhd_plane* plane, const rhd_sensor_desc* d) {
    size_t need = opposed_scratch_bytes((int)plane->width, (int)plane->height);
    rhd_status st = ensure_scratch(ctx, L, L.opposed, need);
    if (st != RHD_OK) return st;
    OpposedArgs a{};
    a.mosaic = plane->d;
    a.width = (int)plane->width;
    a.height = (int)plane->height;
    a.pattern = plane->pattern;
    for (int c = 0; c < 3; ++c) {
        a.levels[c] = d->detection_levels[c];
        a.ceiling[c] = d->neutral_ceiling[c];
    }
    a.has_ceiling = d->has_neutral_ceiling;
    a.scratch = L.opposed.ptr;
    a.scratch_bytes = L.opposed.bytes;
    a.stats = L.d_stats;
    a.log2_bitmap = libm_log2_bitmap(ctx->device);
    if (!a.log2_bitmap) return fail(ctx, RHD_SYNCHRONIZATION, "libm tables are not set up");
    a.aux = L.aux;
    a.fork = L.fork;
    a.join = L.join;
    hipError_t e = launch_opposed(a, L.stream, PROF(L));
    if (e != hipSuccess) return hip_fail(ctx, e, "opposed");
    return RHD_OK;
}

struct Balance {
    float forward[3], inverse[3], clip_point;
};

// crates/demosaic/src/balance.rs:33-66 and lib.rs:321-356.
Balance resolve_balance(const rhd_sensor_desc* d) {
    Balance b{{1, 1, 1}, {1, 1, 1}, 1.0f};
    if (d->demosaic != RHD_DEMOSAIC_AMAZE && d->demosaic != RHD_DEMOSAIC_RCD) return b;
    const float* g = d->demosaic_balance;
    float largest = std::fmax(std::fmax(g[0], g[1]), g[2]);
    float f[3] = {g[0] / largest, g[1] / largest, g[2] / largest};
    if (f[0] == 1.0f && f[1] == 1.0f && f[2] == 1.0f) return b;
    for (int c = 0; c < 3; ++c) {
        b.forward[c] = f[c];
        b.inverse[c] = 1.0f / f[c];
    }
    b.clip_point = std::fmin(std::fmin(f[0], f[1]), f[2]);
    return b;
}

// Capture sharpening is the only detail step: nothing changes the demosaiced
// RGB before it, so its guide and mask can be made as the RGB is stored. The
// fused form divides by the ceilings through their reciprocals, exact for
// ceilings in [2^-20, 2^20].
static bool capture_follows_demosaic(const rhd_detail_desc* dd) {
    if (!capture_active(dd) || dd->false_colour != 0 || noise_active(dd)) return false;
    for (int c = 0; c < 3; ++c)
        if (!(dd->ceilings[c] >= 0x1p-20f && dd->ceilings[c] <= 0x1p20f)) return false;
    return true;
}

// `dd` (may be null) is the detail stage that follows on `src`. When the
// demosaic also writes capture sharpening's guide and mask into the lane's
// capture scratch, *capture_ready is set, and run_detail takes it.
rhd_status run_demosaic(rhd_context* ctx, Lane& L, const rhd_plane* plane,
                        const rhd_sensor_desc* d, rhd_source* src,
                        const rhd_detail_desc* dd = nullptr, bool* capture_ready = nullptr) {
    const int w = (int)plane->width, h = (int)plane->height, pattern = plane->pattern;
    const Balance b = resolve_balance(d);
    hipError_t e = hipSuccess;
    if (capture_ready) *capture_ready = false;
    switch (d->demosaic) {
        case RHD_DEMOSAIC_BILINEAR:
            e = launch_bilinear(plane->d, w, h, pattern, b.forward, b.inverse, src->d, L.d_stats,
                                L.stream, PROF(L));
            break;
        case RHD_DEMOSAIC_MHC:
            e = launch_mhc(plane->d, w, h, pattern, src->d, L.d_stats, L.stream, PROF(L));
            break;
        case RHD_DEMOSAIC_AMAZE: {
            if (w <= 32 || h <= 32) {
                e = launch_bilinear(plane->d, w, h, pattern, b.forward, b.inverse, src->d,
                                    L.d_stats, L.stream, PROF(L));
                break;
            }
            const int tx = (w + 223) / 224, ty = (h + 223) / 224, tiles = tx * ty;
            const int per_launch = std::min(tiles, MAX_TILES_PER_LAUNCH);
            rhd_status st = ensure_scratch(ctx, L, L.amaze,
                                           (uint64_t)per_launch * amaze_scratch_bytes_per_tile());
            if (st != RHD_OK) return st;
            st = ensure_scratch(ctx, L, L.amaze_areas, (uint64_t)per_launch * 4 * sizeof(int));
            if (st != RHD_OK) return st;
            AmazeArgs a{};
            a.mosaic = plane->d;
            a.width = w;
            a.height = h;
            a.pattern = pattern;
            a.tiles_x = tx;
            a.tiles_y = ty;
            for (int c = 0; c < 3; ++c) {
                a.forward[c] = b.forward[c];
                a.inverse[c] = b.inverse[c];
            }
            a.clip_point = b.clip_point;
            a.scratch = (float*)L.amaze.ptr;
            a.areas = (int*)L.amaze_areas.ptr;
            a.out = src->d;
            a.stats = L.d_stats;
            if (capture_ready && capture_follows_demosaic(dd)) {
                // The same size run_detail asks for, so the planes survive.
                st = ensure_scratch(ctx, L, L.capture, detail_scratch_bytes(w, h, dd));
                if (st != RHD_OK) return st;
                a.guide = (float*)L.capture.ptr;
                a.mask = a.guide + (size_t)w * h;
                for (int c = 0; c < 3; ++c) {
                    a.ceilings[c] = dd->ceilings[c];
                    a.ceilings_inv[c] = 1.0f / dd->ceilings[c];
                }
                *capture_ready = true;
            }
            for (int base = 0; base < tiles && e == hipSuccess; base += per_launch) {
                a.tile_base = base;
                e = launch_amaze(a, std::min(per_launch, tiles - base), L.stream, PROF(L));
            }
            break;
        }
        case RHD_DEMOSAIC_RCD: {
            if (w <= 24 || h <= 24) {
                e = launch_bilinear(plane->d, w, h, pattern, b.forward, b.inverse, src->d,
                                    L.d_stats, L.stream, PROF(L));
                break;
            }
            const int tx = (w + 173) / 174, ty = (h + 173) / 174, tiles = tx * ty;
            const int per_launch = std::min(tiles, MAX_TILES_PER_LAUNCH);
            rhd_status st =
                ensure_scratch(ctx, L, L.rcd, (uint64_t)per_launch * rcd_scratch_bytes_per_tile());
            if (st != RHD_OK) return st;
            for (int base = 0; base < tiles && e == hipSuccess; base += per_launch)
                e = launch_rcd(plane->d, w, h, pattern, b.forward, b.inverse, (float*)L.rcd.ptr,
                               base, std::min(per_launch, tiles - base), src->d, L.d_stats,
                               L.stream, PROF(L));
            break;
        }
        default:
            return fail(ctx, RHD_INVALID_INPUT, "unknown demosaic algorithm");
    }
    if (e != hipSuccess) return hip_fail(ctx, e, "demosaic");
    return RHD_OK;
}

// Gaussian weights of crates/detail/src/gaussian.rs:9-17, with the platform
// expf and a sum that starts at -0.0 like Rust's f32 Sum.
int gaussian_weights(float sigma, float* w) {
    int radius = (int)std::ceil(3.0f * sigma);
    if (radius > 4) radius = 4;
    float sum = -0.0f;
    for (int x = -radius; x <= radius; ++x) {
        float t = (float)x / sigma;
        w[x + radius] = std::exp(-0.5f * (t * t));
    }
    for (int k = 0; k <= 2 * radius; ++k) sum += w[k];
    for (int k = 0; k <= 2 * radius; ++k) w[k] /= sum;
    return radius;
}

rhd_status run_capture(rhd_context* ctx, Lane& L, rhd_source* src, const rhd_detail_desc* dd,
                       bool inputs_ready) {
    const int w = (int)src->width, h = (int)src->height;
    rhd_status st = ensure_scratch(ctx, L, L.capture, capture_scratch_bytes(w, h));
    if (st != RHD_OK) return st;
    CaptureArgs a{};
    a.rgb = src->d;
    a.width = w;
    a.height = h;
    a.amount = dd->amount;
    a.noise_protection = dd->noise_protection;
    for (int c = 0; c < 3; ++c) a.ceilings[c] = dd->ceilings[c];
    a.radius = gaussian_weights(dd->sigma, a.weights);
    a.iterations = dd->iterations;
    float* base = (float*)L.capture.ptr;
    size_t n = (size_t)w * h;
    a.guide = base;
    a.mask = base + n;
    a.est_a = base + 2 * n;
    a.est_b = base + 3 * n;
    a.tmp = base + 4 * n;
    a.stats = L.d_stats;
    a.inputs_ready = inputs_ready;
    hipError_t e = launch_capture(a, L.stream, PROF(L));
    if (e != hipSuccess) return hip_fail(ctx, e, "capture");
    return RHD_OK;
}

// The detail stage's steps in the reference's order (crates/core/src/
// pipeline/detail.rs:210-307): noise profiles of the stage input, false
// colour, chroma NR, luminance NR, capture sharpening. Everything is queued
// on the lane's stream; the profiles never leave the device.
rhd_status run_detail(rhd_context* ctx, Lane& L, rhd_source* src, const rhd_detail_desc* dd,
                      bool capture_ready) {
    if (!detail_active(dd)) return RHD_OK;
    const int w = (int)src->width, h = (int)src->height;
    const uint64_t work = detail_work_bytes(w, h, dd);
    rhd_status st = ensure_scratch(ctx, L, L.capture, work + 256);
    if (st != RHD_OK) return st;
    float* scratch = (float*)L.capture.ptr;
    float* prof = (float*)((char*)L.capture.ptr + work);
    bool denoise = false;
    if (noise_active(dd)) {
        const NoiseGeometry g = noise_geometry(w, h);
        if (g.tiles > MAX_NOISE_TILES || h > 65535)
            return fail(ctx, RHD_INVALID_DIMENSIONS, "image too large for the noise profile");
        if (g.tiles > 0) {
            hipError_t e = launch_noise_profile(src->d, w, h, dd->balance, scratch, prof,
                                                L.stream, PROF(L));
            if (e != hipSuccess) return hip_fail(ctx, e, "noise profile");
            denoise = true;
        }
    }
    if (dd->false_colour > 0) {
        hipError_t e = launch_false_colour(src->d, w, h, dd->false_colour, dd->balance,
                                           dd->ceilings, scratch, L.stream, PROF(L));
        if (e != hipSuccess) return hip_fail(ctx, e, "false colour");
    }
    if (denoise && dd->nr_chroma > 0.0f) {
        hipError_t e = launch_chroma_nr(src->d, w, h, dd->nr_chroma, dd->balance, dd->ceilings,
                                        prof, scratch, L.stream, PROF(L));
        if (e != hipSuccess) return hip_fail(ctx, e, "chroma noise reduction");
    }
    if (denoise && dd->nr_luminance > 0.0f) {
        hipError_t e = launch_luma_nr(src->d, w, h, dd->nr_luminance, dd->nr_luminance_detail,
                                      dd->balance, dd->ceilings, prof, scratch, L.stream, PROF(L));
        if (e != hipSuccess) return hip_fail(ctx, e, "luminance noise reduction");
    }
    if (capture_active(dd)) return run_capture(ctx, L, src, dd, capture_ready);
    return RHD_OK;
}

rhd_status check_handle_ctx(rhd_context* ctx, const void* handle_ctx) {
    if (!ctx) return RHD_INVALID_INPUT;
    if (handle_ctx != ctx) return fail(ctx, RHD_INVALID_INPUT, "handle belongs to another context");
    return RHD_OK;
}

rhd_status mosaic_upload(rhd_context* ctx, Lane& L, const uint16_t* host, uint64_t row_stride,
                         uint32_t width, uint32_t height, rhd_mosaic** out) {
    *out = nullptr;
    if (!host || width == 0 || height == 0 || row_stride < width)
        return fail(ctx, RHD_INVALID_DIMENSIONS, "decoded row stride is shorter than the sensor width");
    (void)hipSetDevice(ctx->device);
    uint64_t bytes = row_stride * height * sizeof(uint16_t);
    void* p = nullptr;
    rhd_status st = device_alloc(ctx, &L, bytes, &p);
    if (st != RHD_OK) return st;
    hipError_t e = hipMemcpyAsync(p, host, bytes, hipMemcpyHostToDevice, L.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(L.stream);
    if (e != hipSuccess) {
        device_free(ctx, p, bytes);
        return hip_fail(ctx, e, "mosaic upload");
    }
    ctx->refs.fetch_add(1);
    *out = new rhd_mosaic{ctx, (uint16_t*)p, row_stride, width, height, bytes};
    return RHD_OK;
}

// Grows lane L's scratch to what the stage needs for a frame of `d`.
rhd_status stage_scratch(rhd_context* ctx, Lane& L, const rhd_sensor_desc* d,
                         const rhd_detail_desc* dd) {
    const int w = (int)d->crop_width, h = (int)d->crop_height;
    rhd_status st = RHD_OK;
    if (d->highlight == RHD_HIGHLIGHT_OPPOSED)
        st = ensure_scratch(ctx, L, L.opposed, opposed_scratch_bytes(w, h));
    if (st == RHD_OK && d->demosaic == RHD_DEMOSAIC_AMAZE && w > 32 && h > 32) {
        const uint64_t tiles = std::min<uint64_t>(
            (uint64_t)((w + 223) / 224) * ((h + 223) / 224), MAX_TILES_PER_LAUNCH);
        st = ensure_scratch(ctx, L, L.amaze, tiles * amaze_scratch_bytes_per_tile());
        if (st == RHD_OK) st = ensure_scratch(ctx, L, L.amaze_areas, tiles * 4 * sizeof(int));
    }
    if (st == RHD_OK && d->demosaic == RHD_DEMOSAIC_RCD && w > 24 && h > 24) {
        const uint64_t tiles = std::min<uint64_t>(
            (uint64_t)((w + 173) / 174) * ((h + 173) / 174), MAX_TILES_PER_LAUNCH);
        st = ensure_scratch(ctx, L, L.rcd, tiles * rcd_scratch_bytes_per_tile());
    }
    if (st == RHD_OK && detail_active(dd))
        st = ensure_scratch(ctx, L, L.capture, detail_scratch_bytes(w, h, dd));
    return st;
}

// Everything develop_on launches for one frame, without allocating or
// waiting: `plane` and `src` are given, the lane's scratch must already fit
// (stage_scratch), the statistics are reset first and copied to `stats_out`
// (pinned host memory) last. A HIP graph may capture the sequence: it holds
// kernels and copies only.
rhd_status enqueue_develop(rhd_context* ctx, Lane& L, const rhd_mosaic* mosaic,
                           const rhd_sensor_desc* d, const rhd_detail_desc* dd, rhd_plane* plane,
                           rhd_source* src, DeviceStats* stats_out) {
    NormalizeArgs a{};
    rhd_status st = level_table(ctx, d, a.table);
    if (st != RHD_OK) return st;
    a.mosaic = mosaic->d;
    a.row_stride = mosaic->row_stride;
    a.crop_x = (int)d->crop_x;
    a.crop_y = (int)d->crop_y;
    a.width = (int)d->crop_width;
    a.height = (int)d->crop_height;
    a.crop_pattern = plane->pattern;
    a.mode = d->highlight == RHD_HIGHLIGHT_CLIP ? 1 : (d->highlight == RHD_HIGHLIGHT_OPPOSED ? 2 : 0);
    const float* lv = d->highlight == RHD_HIGHLIGHT_CLIP ? d->clip_levels : d->detection_levels;
    for (int c = 0; c < 3; ++c) a.levels[c] = lv[c];
    if ((st = reset_stats(ctx, L)) != RHD_OK) return st;
    hipError_t e = launch_normalize(a, plane->d, L.d_stats, L.stream, PROF(L));
    if (e != hipSuccess) return hip_fail(ctx, e, "normalize");
    if (d->highlight == RHD_HIGHLIGHT_OPPOSED && (st = run_opposed(ctx, L, plane, d)) != RHD_OK)
        return st;
    bool ready = false;
    if ((st = run_demosaic(ctx, L, plane, d, src, dd, &ready)) != RHD_OK) return st;
    if (detail_active(dd) && (st = run_detail(ctx, L, src, dd, ready)) != RHD_OK) return st;
    RHD_HIP(ctx, hipMemcpyAsync(stats_out, L.d_stats, sizeof(DeviceStats), hipMemcpyDeviceToHost,
                                L.stream));
    return RHD_OK;
}

// The whole stage on one lane: normalize (with Clip or the Opposed detection
// fused), Opposed, demosaic and capture, one statistics readback at the end
// (crates/gpu/src/sensor/develop.rs:68-142).
rhd_status develop_on(rhd_context* ctx, Lane& L, const rhd_mosaic* mosaic,
                      const rhd_sensor_desc* d, const rhd_detail_desc* dd,
                      const volatile uint8_t* cancel, rhd_source** out, rhd_metrics* m,
                      rhd_source** base) {
    *out = nullptr;
    if (base) *base = nullptr;
    rhd_status st = check_handle_ctx(ctx, mosaic->ctx);
    if (st == RHD_OK && cancelled(cancel)) st = fail(ctx, RHD_CANCELLED, "cancelled");
    if (st == RHD_OK) st = validate_desc(ctx, d);
    if (st == RHD_OK) st = validate_detail(ctx, dd);
    if (st == RHD_OK && !rhd_supports(d, dd))
        st = fail(ctx, RHD_UNSUPPORTED_EDITS, "GPU sensor development does not support these edits");
    if (st != RHD_OK) return st;
    if (mosaic->width != d->sensor_width || mosaic->height != d->sensor_height ||
        mosaic->row_stride != d->row_stride)
        return fail(ctx, RHD_INVALID_INPUT, "decoded frame does not match its normalization contract");
    (void)hipSetDevice(ctx->device);
    NormalizeArgs a{};
    if ((st = level_table(ctx, d, a.table)) != RHD_OK) return st;
    rhd_plane* plane = nullptr;
    if ((st = new_plane(ctx, &L, d->crop_width, d->crop_height, crop_pattern(d), &plane)) != RHD_OK)
        return st;
    rhd_source* src = nullptr;
    if ((st = new_source(ctx, &L, plane->width, plane->height, &src)) != RHD_OK) {
        rhd_plane_release(plane);
        return st;
    }
    a.mosaic = mosaic->d;
    a.row_stride = mosaic->row_stride;
    a.crop_x = (int)d->crop_x;
    a.crop_y = (int)d->crop_y;
    a.width = (int)d->crop_width;
    a.height = (int)d->crop_height;
    a.crop_pattern = plane->pattern;
    a.mode = d->highlight == RHD_HIGHLIGHT_CLIP ? 1 : (d->highlight == RHD_HIGHLIGHT_OPPOSED ? 2 : 0);
    const float* lv = d->highlight == RHD_HIGHLIGHT_CLIP ? d->clip_levels : d->detection_levels;
    for (int c = 0; c < 3; ++c) a.levels[c] = lv[c];
    auto step = [&](bool ok, rhd_status err_status, const char* what) {
        if (!ok && st == RHD_OK) st = fail(ctx, err_status, what);
    };
    st = reset_stats(ctx, L);
    step(hipEventRecord(L.ev[0], L.stream) == hipSuccess, RHD_SYNCHRONIZATION, "event");
    if (st == RHD_OK) {
        hipError_t e = launch_normalize(a, plane->d, L.d_stats, L.stream, PROF(L));
        if (e != hipSuccess) st = hip_fail(ctx, e, "normalize");
    }
    step(hipEventRecord(L.ev[1], L.stream) == hipSuccess, RHD_SYNCHRONIZATION, "event");
    if (st == RHD_OK && cancelled(cancel)) st = fail(ctx, RHD_CANCELLED, "cancelled");
    if (st == RHD_OK && d->highlight == RHD_HIGHLIGHT_OPPOSED) st = run_opposed(ctx, L, plane, d);
    step(hipEventRecord(L.ev[2], L.stream) == hipSuccess, RHD_SYNCHRONIZATION, "event");
    if (st == RHD_OK && cancelled(cancel)) st = fail(ctx, RHD_CANCELLED, "cancelled");
    bool ready = false;
    if (st == RHD_OK) st = run_demosaic(ctx, L, plane, d, src, dd, &ready);
    step(hipEventRecord(L.ev[3], L.stream) == hipSuccess, RHD_SYNCHRONIZATION, "event");
    if (st == RHD_OK && base) {
        st = new_source(ctx, &L, src->width, src->height, base);
        if (st == RHD_OK) {
            hipError_t e = hipMemcpyAsync((*base)->d, src->d, src->bytes, hipMemcpyDeviceToDevice,
                                          L.stream);
            if (e != hipSuccess) st = hip_fail(ctx, e, "base copy");
        }
    }
    if (st == RHD_OK && detail_active(dd)) {
        if (cancelled(cancel)) st = fail(ctx, RHD_CANCELLED, "cancelled");
        if (st == RHD_OK) st = run_detail(ctx, L, src, dd, ready);
    }
    step(hipEventRecord(L.ev[4], L.stream) == hipSuccess, RHD_SYNCHRONIZATION, "event");
    if (st == RHD_OK) st = finish_stats(ctx, L, m, d->highlight);
    if (st != RHD_OK) {
        (void)hipStreamSynchronize(L.stream);
        rhd_plane_release(plane);
        rhd_source_release(src);
        if (base && *base) {
            rhd_source_release(*base);
            *base = nullptr;
        }
        return st;
    }
    if (m) {
        m->normalize_ms = elapsed(L.ev[0], L.ev[1]);
        m->highlight_ms = elapsed(L.ev[1], L.ev[2]);
        m->demosaic_ms = elapsed(L.ev[2], L.ev[3]);
        m->capture_ms = elapsed(L.ev[3], L.ev[4]);
        m->total_ms = elapsed(L.ev[0], L.ev[4]);
        m->launches += 4;
        m->estimated_gpu_bytes = std::max<uint64_t>(m->estimated_gpu_bytes, ctx->reserved.load());
    }
    rhd_plane_release(plane);
    *out = src;
    return RHD_OK;
}

}  // namespace rhd

// Lane 0 scratch with `work` bytes for the filters and the profile area
// after them.
static rhd_status noise_scratch(rhd_context* ctx, Lane& L, uint64_t work, float** prof) {
    work = (work + 255) & ~(uint64_t)255;
    rhd_status st = ensure_scratch(ctx, L, L.capture, work + 256);
    if (st == RHD_OK) *prof = (float*)((char*)L.capture.ptr + work);
    return st;
}

// Copies host profiles next to the filters' scratch and runs `launch`.
template <class Launch>
static rhd_status run_filter(rhd_context* ctx, rhd_source* src, const rhd_noise_profile* p,
                             int count, float* gpu_ms, const char* what, Launch launch) {
    Lane& L = ctx->lanes[0];
    (void)hipSetDevice(ctx->device);
    const int w = (int)src->width, h = (int)src->height;
    float* prof = nullptr;
    rhd_status st = noise_scratch(ctx, L, noise_filter_bytes(w, h), &prof);
    if (st != RHD_OK) return st;
    float host[6] = {};
    for (int c = 0; c < count; ++c) {
        host[2 * c] = p[c].a;
        host[2 * c + 1] = p[c].b;
    }
    RHD_HIP(ctx, hipMemcpyAsync(prof, host, sizeof(host), hipMemcpyHostToDevice, L.stream));
    RHD_HIP(ctx, hipEventRecord(L.ev[3], L.stream));
    hipError_t e = launch(L, prof, (float*)L.capture.ptr);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(L.stream);
        return hip_fail(ctx, e, what);
    }
    RHD_HIP(ctx, hipEventRecord(L.ev[4], L.stream));
    RHD_HIP(ctx, hipStreamSynchronize(L.stream));
    if (gpu_ms) *gpu_ms = elapsed(L.ev[3], L.ev[4]);
    return RHD_OK;
}

extern "C" {

rhd_status rhd_create(int device, uint64_t budget_bytes, rhd_context** out) {
    if (!out) return RHD_INVALID_INPUT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
        (void)hipGetLastError();
        return RHD_UNSUPPORTED;
    }
    if (hipSetDevice(device) != hipSuccess) return RHD_UNSUPPORTED;
    rhd_context* ctx = new rhd_context();
    ctx->device = device;
    if (budget_bytes) ctx->budget = budget_bytes;
    Lane* L = nullptr;
    if (lane_open(ctx, 0, &L) != RHD_OK || libm_setup(ctx, device) != RHD_OK ||
        noalloc_setup(ctx, device) != RHD_OK) {
        context_unref(ctx);
        return RHD_UNSUPPORTED;
    }
    *out = ctx;
    return RHD_OK;
}

// The context outlives rhd_destroy while handles made from it are alive: each
// handle holds a reference, and the last release frees the lanes and the pool.
void rhd_destroy(rhd_context* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    for (Lane& L : ctx->lanes) drop_scratch(ctx, L);
    context_unref(ctx);
}

// Not to be called while other calls on the context are in flight.
rhd_status rhd_set_budget(rhd_context* ctx, uint64_t budget_bytes) {
    if (!ctx) return RHD_INVALID_INPUT;
    ctx->budget = budget_bytes ? budget_bytes : DEFAULT_BUDGET;
    if (ctx->reserved.load() > ctx->budget.load())
        for (Lane& L : ctx->lanes) drop_scratch(ctx, L);
    pool_trim(ctx, 0, false);
    return RHD_OK;
}

uint64_t rhd_reserved_bytes(const rhd_context* ctx) { return ctx ? ctx->reserved.load() : 0; }

const char* rhd_last_error(const rhd_context* ctx) {
    return ctx ? ctx->last_error.c_str() : "no context";
}

const char* rhd_status_name(rhd_status s) {
    switch (s) {
        case RHD_OK: return "Ok";
        case RHD_CANCELLED: return "Cancelled";
        case RHD_UNSUPPORTED: return "Unsupported";
        case RHD_OVER_BUDGET: return "OverBudget";
        case RHD_INVALID_DIMENSIONS: return "InvalidDimensions";
        case RHD_BASE_MISMATCH: return "BaseMismatch";
        case RHD_UNSUPPORTED_EDITS: return "UnsupportedEdits";
        case RHD_SYNCHRONIZATION: return "Synchronization";
        case RHD_INVALID_INPUT: return "InvalidInput";
        case RHD_READBACK: return "Readback";
        case RHD_RAW_UNSUPPORTED: return "RawUnsupported";
        case RHD_RAW_CORRUPT: return "RawCorrupt";
    }
    return "Unknown";
}

int rhd_status_is_routing(rhd_status s) {
    return s == RHD_UNSUPPORTED_EDITS || s == RHD_OVER_BUDGET;
}

int rhd_supports(const rhd_sensor_desc* d, const rhd_detail_desc* dd) {
    if (!d) return 0;
    bool highlight = d->highlight == RHD_HIGHLIGHT_OFF || d->highlight == RHD_HIGHLIGHT_CLIP ||
                     d->highlight == RHD_HIGHLIGHT_OPPOSED;
    bool demosaic = d->demosaic >= RHD_DEMOSAIC_BILINEAR && d->demosaic <= RHD_DEMOSAIC_AMAZE;
    bool detail = !dd || (dd->false_colour >= 0 && dd->false_colour <= 5);
    bool capture = !capture_active(dd) || std::ceil(3.0f * dd->sigma) <= 4.0f;
    return highlight && demosaic && detail && capture ? 1 : 0;
}

rhd_status rhd_mosaic_upload(rhd_context* ctx, const uint16_t* host, uint64_t row_stride,
                             uint32_t width, uint32_t height, rhd_mosaic** out) {
    if (!ctx || !out) return RHD_INVALID_INPUT;
    return mosaic_upload(ctx, ctx->lanes[0], host, row_stride, width, height, out);
}

void rhd_mosaic_release(rhd_mosaic* m) {
    if (!m) return;
    rhd_context* ctx = m->ctx;
    device_free(ctx, m->d, m->bytes);
    delete m;
    context_unref(ctx);
}

rhd_status rhd_normalize(rhd_context* ctx, const rhd_mosaic* mosaic, const rhd_sensor_desc* d,
                         const volatile uint8_t* cancel, rhd_plane** out, rhd_metrics* m) {
    if (!ctx || !out || !mosaic) return RHD_INVALID_INPUT;
    *out = nullptr;
    Lane& L = ctx->lanes[0];
    rhd_status st = check_handle_ctx(ctx, mosaic->ctx);
    if (st != RHD_OK) return st;
    if ((st = validate_desc(ctx, d)) != RHD_OK) return st;
    if (mosaic->width != d->sensor_width || mosaic->height != d->sensor_height ||
        mosaic->row_stride != d->row_stride)
        return fail(ctx, RHD_INVALID_INPUT, "decoded frame does not match its normalization contract");
    if (cancelled(cancel)) return fail(ctx, RHD_CANCELLED, "cancelled");
    (void)hipSetDevice(ctx->device);
    NormalizeArgs a{};
    if ((st = level_table(ctx, d, a.table)) != RHD_OK) return st;
    rhd_plane* plane = nullptr;
    if ((st = new_plane(ctx, &L, d->crop_width, d->crop_height, crop_pattern(d), &plane)) != RHD_OK)
        return st;
    a.mosaic = mosaic->d;
    a.row_stride = mosaic->row_stride;
    a.crop_x = (int)d->crop_x;
    a.crop_y = (int)d->crop_y;
    a.width = (int)d->crop_width;
    a.height = (int)d->crop_height;
    a.crop_pattern = plane->pattern;
    a.mode = 0;
    RHD_HIP(ctx, hipEventRecord(L.ev[0], L.stream));
    hipError_t e = launch_normalize(a, plane->d, L.d_stats, L.stream, PROF(L));
    if (e != hipSuccess) {
        rhd_plane_release(plane);
        return hip_fail(ctx, e, "normalize");
    }
    RHD_HIP(ctx, hipEventRecord(L.ev[1], L.stream));
    e = hipEventSynchronize(L.ev[1]);
    if (e != hipSuccess) {
        rhd_plane_release(plane);
        return hip_fail(ctx, e, "normalize");
    }
    if (m) {
        m->normalize_ms = elapsed(L.ev[0], L.ev[1]);
        m->launches += 1;
        m->estimated_gpu_bytes = std::max<uint64_t>(m->estimated_gpu_bytes, plane->bytes);
    }
    *out = plane;
    return RHD_OK;
}

rhd_status rhd_highlight(rhd_context* ctx, rhd_plane* plane, const rhd_sensor_desc* d,
                         const volatile uint8_t* cancel, rhd_metrics* m) {
    if (!ctx || !plane) return RHD_INVALID_INPUT;
    Lane& L = ctx->lanes[0];
    rhd_status st = check_handle_ctx(ctx, plane->ctx);
    if (st != RHD_OK) return st;
    if ((st = validate_desc(ctx, d)) != RHD_OK) return st;
    if (cancelled(cancel)) return fail(ctx, RHD_CANCELLED, "cancelled");
    (void)hipSetDevice(ctx->device);
    if ((st = reset_stats(ctx, L)) != RHD_OK) return st;
    RHD_HIP(ctx, hipEventRecord(L.ev[0], L.stream));
    hipError_t e = hipSuccess;
    const int w = (int)plane->width, h = (int)plane->height;
    if (d->highlight == RHD_HIGHLIGHT_CLIP) {
        e = launch_clip(plane->d, w, h, plane->pattern, d->clip_levels, L.d_stats, L.stream);
    } else {
        // Off and Opposed validate every sample (opposed.rs:182-215).
        float levels[3] = {INFINITY, INFINITY, INFINITY};
        e = launch_detect(plane->d, w, h, plane->pattern,
                          d->highlight == RHD_HIGHLIGHT_OPPOSED ? d->detection_levels : levels,
                          L.d_stats, L.stream);
    }
    if (e != hipSuccess) return hip_fail(ctx, e, "highlight");
    if (d->highlight == RHD_HIGHLIGHT_OPPOSED) {
        // Samples must be finite before reconstruction reads them.
        if ((st = finish_stats(ctx, L, nullptr, d->highlight)) != RHD_OK) return st;
        if ((st = run_opposed(ctx, L, plane, d)) != RHD_OK) return st;
    }
    RHD_HIP(ctx, hipEventRecord(L.ev[1], L.stream));
    if ((st = finish_stats(ctx, L, m, d->highlight)) != RHD_OK) return st;
    if (m) {
        m->highlight_ms = elapsed(L.ev[0], L.ev[1]);
        m->launches += 1;
        m->readback_bytes += sizeof(DeviceStats);
    }
    return RHD_OK;
}

rhd_status rhd_demosaic(rhd_context* ctx, rhd_plane* plane, const rhd_sensor_desc* d,
                        const rhd_detail_desc* dd, const volatile uint8_t* cancel,
                        rhd_source** out, rhd_metrics* m) {
    if (!ctx || !out || !plane) {
        rhd_plane_release(plane);
        return RHD_INVALID_INPUT;
    }
    *out = nullptr;
    Lane& L = ctx->lanes[0];
    rhd_status st = check_handle_ctx(ctx, plane->ctx);
    if (st == RHD_OK) st = validate_desc(ctx, d);
    if (st == RHD_OK) st = validate_detail(ctx, dd);
    if (st == RHD_OK && !rhd_supports(d, dd))
        st = fail(ctx, RHD_UNSUPPORTED_EDITS, "GPU sensor development does not support these edits");
    if (st == RHD_OK && cancelled(cancel)) st = fail(ctx, RHD_CANCELLED, "cancelled");
    if (st != RHD_OK) {
        rhd_plane_release(plane);
        return st;
    }
    (void)hipSetDevice(ctx->device);
    rhd_source* src = nullptr;
    st = new_source(ctx, &L, plane->width, plane->height, &src);
    if (st == RHD_OK) st = reset_stats(ctx, L);
    if (st == RHD_OK && hipEventRecord(L.ev[2], L.stream) != hipSuccess)
        st = fail(ctx, RHD_SYNCHRONIZATION, "event");
    bool ready = false;
    if (st == RHD_OK) st = run_demosaic(ctx, L, plane, d, src, dd, &ready);
    if (st == RHD_OK && hipEventRecord(L.ev[3], L.stream) != hipSuccess)
        st = fail(ctx, RHD_SYNCHRONIZATION, "event");
    if (st == RHD_OK && detail_active(dd)) {
        if (cancelled(cancel)) st = fail(ctx, RHD_CANCELLED, "cancelled");
        if (st == RHD_OK) st = run_detail(ctx, L, src, dd, ready);
    }
    if (st == RHD_OK && hipEventRecord(L.ev[4], L.stream) != hipSuccess)
        st = fail(ctx, RHD_SYNCHRONIZATION, "event");
    if (st == RHD_OK) st = finish_stats(ctx, L, nullptr, RHD_HIGHLIGHT_OFF);
    if (st != RHD_OK) (void)hipStreamSynchronize(L.stream);
    rhd_plane_release(plane);
    if (st != RHD_OK) {
        rhd_source_release(src);
        return st;
    }
    if (m) {
        m->demosaic_ms = elapsed(L.ev[2], L.ev[3]);
        m->capture_ms = elapsed(L.ev[3], L.ev[4]);
        m->launches += 2;
        m->estimated_gpu_bytes = std::max<uint64_t>(m->estimated_gpu_bytes, ctx->reserved.load());
    }
    *out = src;
    return RHD_OK;
}

rhd_status rhd_capture(rhd_context* ctx, rhd_source* src, const rhd_detail_desc* dd,
                       const volatile uint8_t* cancel, rhd_metrics* m) {
    if (!ctx || !src) return RHD_INVALID_INPUT;
    Lane& L = ctx->lanes[0];
    rhd_status st = check_handle_ctx(ctx, src->ctx);
    if (st != RHD_OK) return st;
    if ((st = validate_detail(ctx, dd)) != RHD_OK) return st;
    if (!detail_active(dd)) return RHD_OK;
    if (std::ceil(3.0f * dd->sigma) > 4.0f)
        return fail(ctx, RHD_UNSUPPORTED_EDITS, "capture radius above 4 taps");
    if (cancelled(cancel)) return fail(ctx, RHD_CANCELLED, "cancelled");
    (void)hipSetDevice(ctx->device);
    if ((st = reset_stats(ctx, L)) != RHD_OK) return st;
    RHD_HIP(ctx, hipEventRecord(L.ev[3], L.stream));
    if ((st = run_detail(ctx, L, src, dd)) != RHD_OK) {
        (void)hipStreamSynchronize(L.stream);
        return st;
    }
    RHD_HIP(ctx, hipEventRecord(L.ev[4], L.stream));
    if ((st = finish_stats(ctx, L, nullptr, RHD_HIGHLIGHT_OFF)) != RHD_OK) return st;
    if (m) m->capture_ms = elapsed(L.ev[3], L.ev[4]);
    return RHD_OK;
}

rhd_status rhd_noise_profiles(rhd_context* ctx, const rhd_source* src, const float balance[3],
                              rhd_noise_profile out[3], int32_t* profiled, float* gpu_ms) {
    if (!ctx || !src || !balance || !out || !profiled) return RHD_INVALID_INPUT;
    *profiled = 0;
    if (gpu_ms) *gpu_ms = 0.0f;
    rhd_status st = check_handle_ctx(ctx, src->ctx);
    if (st != RHD_OK) return st;
    if (!finite_positive3(balance))
        return fail(ctx, RHD_INVALID_INPUT, "channel balance must be finite and positive");
    const int w = (int)src->width, h = (int)src->height;
    const NoiseGeometry g = noise_geometry(w, h);
    if (g.tiles == 0) return RHD_OK;
    if (g.tiles > MAX_NOISE_TILES || h > 65535)
        return fail(ctx, RHD_INVALID_DIMENSIONS, "image too large for the noise profile");
    Lane& L = ctx->lanes[0];
    (void)hipSetDevice(ctx->device);
    float* prof = nullptr;
    if ((st = noise_scratch(ctx, L, g.bytes, &prof)) != RHD_OK) return st;
    float host[9];
    RHD_HIP(ctx, hipEventRecord(L.ev[3], L.stream));
    RHD_HIP(ctx, launch_noise_profile(src->d, w, h, balance, L.capture.ptr, prof, L.stream,
                                      PROF(L)));
    RHD_HIP(ctx, hipEventRecord(L.ev[4], L.stream));
    RHD_HIP(ctx, hipMemcpyAsync(host, prof, sizeof(host), hipMemcpyDeviceToHost, L.stream));
    RHD_HIP(ctx, hipStreamSynchronize(L.stream));
    for (int c = 0; c < 3; ++c) out[c] = {host[2 * c], host[2 * c + 1], (uint32_t)host[6 + c]};
    *profiled = 1;
    if (gpu_ms) *gpu_ms = elapsed(L.ev[3], L.ev[4]);
    return RHD_OK;
}

rhd_status rhd_denoise_chroma(rhd_context* ctx, rhd_source* src, float strength,
                              const float balance[3], const float ceilings[3],
                              const rhd_noise_profile profiles[3], float* gpu_ms) {
    if (!ctx || !src || !balance || !ceilings || !profiles) return RHD_INVALID_INPUT;
    if (gpu_ms) *gpu_ms = 0.0f;
    rhd_status st = check_handle_ctx(ctx, src->ctx);
    if (st != RHD_OK) return st;
    if (!std::isfinite(strength) || !(strength >= 0.0f && strength <= 1.0f))
        return fail(ctx, RHD_INVALID_INPUT, "chroma strength must be finite in 0..=1");
    if (strength == 0.0f) return RHD_OK;
    if (!finite_positive3(balance))
        return fail(ctx, RHD_INVALID_INPUT, "channel balance must be finite and positive");
    const int w = (int)src->width, h = (int)src->height;
    return run_filter(ctx, src, profiles, 3, gpu_ms, "chroma noise reduction",
                      [&](Lane& L, float* prof, float* scratch) {
                          return launch_chroma_nr(src->d, w, h, strength, balance, ceilings,
                                                  prof, scratch, L.stream, PROF(L));
                      });
}

rhd_status rhd_denoise_luma(rhd_context* ctx, rhd_source* src, float strength, float detail,
                            const float balance[3], const float ceilings[3],
                            const rhd_noise_profile* luma, float* gpu_ms) {
    if (!ctx || !src || !balance || !ceilings || !luma) return RHD_INVALID_INPUT;
    if (gpu_ms) *gpu_ms = 0.0f;
    rhd_status st = check_handle_ctx(ctx, src->ctx);
    if (st != RHD_OK) return st;
    if (!(strength >= 0.0f && strength <= 1.0f) || !(detail >= 0.0f && detail <= 1.0f))
        return fail(ctx, RHD_INVALID_INPUT, "luminance strength and detail must be in 0..=1");
    if (strength == 0.0f) return RHD_OK;
    if (!finite_positive3(balance))
        return fail(ctx, RHD_INVALID_INPUT, "channel balance must be finite and positive");
    const int w = (int)src->width, h = (int)src->height;
    if (h > 65535) return fail(ctx, RHD_INVALID_DIMENSIONS, "image taller than 65535 rows");
    return run_filter(ctx, src, luma, 1, gpu_ms, "luminance noise reduction",
                      [&](Lane& L, float* prof, float* scratch) {
                          return launch_luma_nr(src->d, w, h, strength, detail, balance,
                                                ceilings, prof, scratch, L.stream, PROF(L));
                      });
}

rhd_status rhd_develop_resident(rhd_context* ctx, const rhd_mosaic* mosaic,
                                const rhd_sensor_desc* d, const rhd_detail_desc* dd,
                                const volatile uint8_t* cancel, rhd_source** out,
                                rhd_metrics* m) {
    if (!ctx || !out || !mosaic) return RHD_INVALID_INPUT;
    return develop_on(ctx, ctx->lanes[0], mosaic, d, dd, cancel, out, m, nullptr);
}

rhd_status rhd_develop(rhd_context* ctx, const uint16_t* host_mosaic, uint64_t row_stride,
                       const rhd_sensor_desc* d, const rhd_detail_desc* dd,
                       const volatile uint8_t* cancel, rhd_source** out, rhd_metrics* m) {
    if (!ctx || !out || !d) return RHD_INVALID_INPUT;
    *out = nullptr;
    rhd_mosaic* mosaic = nullptr;
    rhd_status st =
        rhd_mosaic_upload(ctx, host_mosaic, row_stride, d->sensor_width, d->sensor_height, &mosaic);
    if (st != RHD_OK) return st;
    if (m) m->uploaded_bytes += mosaic->bytes;
    st = rhd_develop_resident(ctx, mosaic, d, dd, cancel, out, m);
    rhd_mosaic_release(mosaic);
    return st;
}

rhd_status rhd_plane_read(rhd_context* ctx, const rhd_plane* plane, float* dst, uint64_t count) {
    if (!ctx || !plane || !dst) return RHD_INVALID_INPUT;
    Lane& L = ctx->lanes[0];
    if (count < (uint64_t)plane->width * plane->height)
        return fail(ctx, RHD_READBACK, "destination too small");
    RHD_HIP(ctx, hipMemcpyAsync(dst, plane->d, (size_t)plane->width * plane->height * sizeof(float),
                                hipMemcpyDeviceToHost, L.stream));
    RHD_HIP(ctx, hipStreamSynchronize(L.stream));
    return RHD_OK;
}

rhd_status rhd_plane_write(rhd_context* ctx, rhd_plane* plane, const float* src, uint64_t count) {
    if (!ctx || !plane || !src) return RHD_INVALID_INPUT;
    Lane& L = ctx->lanes[0];
    if (count < (uint64_t)plane->width * plane->height)
        return fail(ctx, RHD_INVALID_INPUT, "source too small");
    RHD_HIP(ctx, hipMemcpyAsync(plane->d, src, (size_t)plane->width * plane->height * sizeof(float),
                                hipMemcpyHostToDevice, L.stream));
    RHD_HIP(ctx, hipStreamSynchronize(L.stream));
    return RHD_OK;
}

rhd_status rhd_plane_upload(rhd_context* ctx, const float* host, uint32_t width, uint32_t height,
                            int32_t pattern, rhd_plane** out) {
    if (!ctx || !out || !host) return RHD_INVALID_INPUT;
    *out = nullptr;
    if (width < 2 || height < 2)
        return fail(ctx, RHD_INVALID_DIMENSIONS, "Bayer demosaicing requires at least 2x2 samples");
    if (pattern < 0 || pattern > 3) return fail(ctx, RHD_INVALID_INPUT, "unsupported CFA pattern");
    (void)hipSetDevice(ctx->device);
    rhd_plane* plane = nullptr;
    rhd_status st = new_plane(ctx, &ctx->lanes[0], width, height, pattern, &plane);
    if (st != RHD_OK) return st;
    st = rhd_plane_write(ctx, plane, host, (uint64_t)width * height);
    if (st != RHD_OK) {
        rhd_plane_release(plane);
        return st;
    }
    *out = plane;
    return RHD_OK;
}

void rhd_plane_dims(const rhd_plane* p, uint32_t* w, uint32_t* h, int32_t* pattern) {
    if (!p) return;
    if (w) *w = p->width;
    if (h) *h = p->height;
    if (pattern) *pattern = p->pattern;
}

void rhd_plane_release(rhd_plane* p) {
    if (!p) return;
    rhd_context* ctx = p->ctx;
    device_free(ctx, p->d, p->bytes);
    delete p;
    context_unref(ctx);
}

rhd_status rhd_source_read_rgb(rhd_context* ctx, const rhd_source* s, float* dst, uint64_t count) {
    if (!ctx || !s || !dst) return RHD_INVALID_INPUT;
    Lane& L = ctx->lanes[0];
    const size_t n = (size_t)s->width * s->height;
    if (count < 3 * (uint64_t)n) return fail(ctx, RHD_READBACK, "destination too small");
    std::vector<float> planar(3 * n);
    RHD_HIP(ctx, hipMemcpyAsync(planar.data(), s->d, 3 * n * sizeof(float), hipMemcpyDeviceToHost,
                                L.stream));
    RHD_HIP(ctx, hipStreamSynchronize(L.stream));
    for (size_t i = 0; i < n; ++i) {
        dst[3 * i] = planar[i];
        dst[3 * i + 1] = planar[n + i];
        dst[3 * i + 2] = planar[2 * n + i];
    }
    return RHD_OK;
}

rhd_status rhd_source_upload(rhd_context* ctx, const float* rgb, uint32_t width, uint32_t height,
                             rhd_source** out) {
    if (!ctx || !out || !rgb) return RHD_INVALID_INPUT;
    *out = nullptr;
    if (width == 0 || height == 0)
        return fail(ctx, RHD_INVALID_DIMENSIONS, "camera source must be non-empty");
    (void)hipSetDevice(ctx->device);
    Lane& L = ctx->lanes[0];
    rhd_source* src = nullptr;
    rhd_status st = new_source(ctx, &L, width, height, &src);
    if (st != RHD_OK) return st;
    const size_t n = (size_t)width * height;
    std::vector<float> planar(3 * n);
    for (size_t i = 0; i < n; ++i) {
        planar[i] = rgb[3 * i];
        planar[n + i] = rgb[3 * i + 1];
        planar[2 * n + i] = rgb[3 * i + 2];
    }
    hipError_t e = hipMemcpyAsync(src->d, planar.data(), 3 * n * sizeof(float),
                                  hipMemcpyHostToDevice, L.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(L.stream);
    if (e != hipSuccess) {
        rhd_source_release(src);
        return hip_fail(ctx, e, "camera source upload");
    }
    *out = src;
    return RHD_OK;
}

rhd_status rhd_source_planes(const rhd_source* s, float** r, float** g, float** b) {
    if (!s) return RHD_INVALID_INPUT;
    const size_t n = (size_t)s->width * s->height;
    if (r) *r = s->d;
    if (g) *g = s->d + n;
    if (b) *b = s->d + 2 * n;
    return RHD_OK;
}

void rhd_source_dims(const rhd_source* s, uint32_t* w, uint32_t* h) {
    if (!s) return;
    if (w) *w = s->width;
    if (h) *h = s->height;
}

void rhd_source_release(rhd_source* s) {
    if (!s) return;
    rhd_context* ctx = s->ctx;
    device_free(ctx, s->d, s->bytes);
    delete s;
    context_unref(ctx);
}

rhd_status rhd_synchronize(rhd_context* ctx) {
    if (!ctx) return RHD_INVALID_INPUT;
    for (Lane& L : ctx->lanes)
        if (L.open) RHD_HIP(ctx, hipStreamSynchronize(L.stream));
    return RHD_OK;
}

rhd_status rhd_profile_enable(rhd_context* ctx, int enable) {
    if (!ctx) return RHD_INVALID_INPUT;
    ctx->lanes[0].profiling = enable != 0;
    return RHD_OK;
}

rhd_status rhd_profile_only(rhd_context* ctx, const char* kernel) {
    if (!ctx) return RHD_INVALID_INPUT;
    ctx->lanes[0].profiler.only = kernel ? kernel : "";
    return RHD_OK;
}

rhd_status rhd_profile_reset(rhd_context* ctx) {
    if (!ctx) return RHD_INVALID_INPUT;
    ctx->lanes[0].profiler.reset();
    return RHD_OK;
}

uint32_t rhd_profile_count(rhd_context* ctx) {
    if (!ctx) return 0;
    ctx->lanes[0].profiler.collect();
    return (uint32_t)ctx->lanes[0].profiler.totals.size();
}

rhd_status rhd_profile_entry(rhd_context* ctx, uint32_t index, const char** name, double* total_ms,
                             uint64_t* launches, uint64_t* bytes_per_launch) {
    if (!ctx) return RHD_INVALID_INPUT;
    Profiler& prof = ctx->lanes[0].profiler;
    prof.collect();
    if (index >= prof.totals.size()) return fail(ctx, RHD_INVALID_INPUT, "no such entry");
    const Profiler::Total& t = prof.totals[index];
    if (name) *name = t.name.c_str();
    if (total_ms) *total_ms = t.ms;
    if (launches) *launches = t.launches;
    if (bytes_per_launch) *bytes_per_launch = t.bytes;
    return RHD_OK;
}

}  // extern "C"
