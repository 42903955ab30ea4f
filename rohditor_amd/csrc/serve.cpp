// Device-resident serving loop for sensor develops.
//
// Replaces the engine worker's one-at-a-time preview develop
// (crates/engine/src/coordinator/worker.rs:723-800) and its one-slot mailbox
// (crates/engine/src/coordinator/scheduler.rs:44-173: a newer request for a
// document replaces the pending one and cancels the active one) with
// continuous batching on one GPU:
//   - requests are admitted in arrival order, skipping ahead only when a later
//     request fits the device bytes left by the running ones, onto free lanes
//     (each lane is its own HIP stream with its own scratch), so their work
//     overlaps on the device;
//   - newest-wins per frame id, as should_replace_preview (scheduler.rs:
//     169-173) with submission order as the revision: a newer request drops
//     the frame's queued ones and cancels its running one, which stops at the
//     next stage boundary;
//   - decoded u16 mosaics stay resident per frame id, and demosaiced bases
//     (camera RGB before capture sharpening) are kept on the device in an
//     LRU cache keyed by frame id and the sensor recipe (the role of
//     SpatialCache::prepare_sensor's cached source, crates/engine/src/
//     preview_cache/spatial.rs:475-611), so a re-develop with the same sensor
//     settings only runs capture sharpening;
//   - latency (submit -> RGB ready on the device) is recorded per request and
//     summarised as p50 / p99 beside throughput.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "rhd_context.hpp"

using namespace rhd;

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

constexpr int FIRST_WORKER_LANE = 1;
constexpr int UPLOAD_LANE = MAX_LANES - 1;
constexpr int MAX_WORKERS = UPLOAD_LANE - FIRST_WORKER_LANE;

enum class State { Queued, Running, Done };

struct Request {
    uint64_t id = 0, frame = 0;
    rhd_sensor_desc desc{};
    std::vector<float> black, white;  // owned copies of the desc's level arrays
    rhd_detail_desc detail{};
    bool has_detail = false;
    std::string base_key;
    uint64_t footprint = 0;
    double megapixels = 0.0;
    volatile uint8_t cancel = 0;
    State state = State::Queued;
    rhd_status status = RHD_OK;
    rhd_source* out = nullptr;
    int lane = -1;
    bool base_hit = false;
    // Spatial completion after the stage (optics + area reduction).
    bool spatial = false, has_optics = false;
    rhd_optics_desc optics{};
    uint32_t target_width = 0, target_height = 0;
    rhd_metrics metrics{};
    Clock::time_point submitted, started, finished;
};

struct ResidentMosaic {
    rhd_mosaic* mosaic = nullptr;
    int users = 0;
    bool dropped = false;
};

struct BaseEntry {
    std::string key;  // frame id + sensor recipe
    uint64_t frame = 0;
    uint64_t generation = 0;  // the frame's content generation it was developed from
    rhd_source* source = nullptr;
    int users = 0;
};

// The sensor settings that determine a demosaiced base, as bytes.
std::string recipe_key(uint64_t frame, const rhd_sensor_desc& d, const std::vector<float>& black,
                       const std::vector<float>& white) {
    std::string k;
    auto put = [&](const void* p, size_t n) { k.append((const char*)p, n); };
    put(&frame, sizeof frame);
    rhd_sensor_desc c = d;
    c.black_levels = nullptr;
    c.white_levels = nullptr;
    put(&c, sizeof c);
    put(black.data(), black.size() * sizeof(float));
    put(white.data(), white.size() * sizeof(float));
    return k;
}

uint64_t tiles_of(uint64_t w, uint64_t h, uint64_t core) {
    return ((w + core - 1) / core) * ((h + core - 1) / core);
}

// Device bytes a request holds while it runs: its mosaic plane, RGB source,
// base copy and the scratch of its stages.
uint64_t footprint_of(const rhd_sensor_desc& d, const rhd_detail_desc* dd) {
    const uint64_t w = d.crop_width, h = d.crop_height, px = w * h;
    uint64_t bytes = px * (4 + 12 + 12);
    if (d.demosaic == RHD_DEMOSAIC_AMAZE) bytes += tiles_of(w, h, 224) * amaze_scratch_bytes_per_tile();
    if (d.demosaic == RHD_DEMOSAIC_RCD) bytes += tiles_of(w, h, 174) * rcd_scratch_bytes_per_tile();
    if (d.highlight == RHD_HIGHLIGHT_OPPOSED) bytes += opposed_scratch_bytes((int)w, (int)h);
    if (detail_active(dd)) bytes += detail_scratch_bytes((int)w, (int)h, dd);
    return bytes;
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t lo = (size_t)pos;
    const size_t hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
}

}  // namespace

struct rhd_server {
    rhd_context* ctx = nullptr;
    rhd_server_config cfg{};
    std::mutex mu;
    std::mutex upload_mu;
    std::condition_variable work_cv, done_cv;
    std::deque<Request*> queue;
    std::unordered_map<uint64_t, Request*> requests;  // not yet collected by rhd_server_wait
    std::unordered_map<uint64_t, ResidentMosaic> mosaics;
    std::list<BaseEntry> bases;  // most recently used first
    // Content generation per frame id, bumped whenever new samples are
    // installed under the id (never reset, also across drops): a base is
    // valid only for the generation it was developed from.
    std::unordered_map<uint64_t, uint64_t> generations;
    uint64_t base_bytes = 0;
    uint64_t running_bytes = 0;
    uint64_t next_id = 1;
    bool stop = false;
    std::vector<std::thread> workers;
    // statistics since the last reset
    std::vector<double> latencies;
    uint64_t submitted = 0, completed = 0, cancelled = 0, failed = 0, base_hits = 0, coalesced = 0;
    double megapixels = 0.0;
    bool have_span = false;
    Clock::time_point span_start, span_end;
};

namespace {

// Called with the server lock held.
void finish_locked(rhd_server* s, Request* r) {
    r->state = State::Done;
    r->finished = Clock::now();
    if (r->status == RHD_OK) {
        s->completed += 1;
        s->latencies.push_back(ms_between(r->submitted, r->finished));
        s->megapixels += r->megapixels;
        if (r->base_hit) s->base_hits += 1;
    } else if (r->status == RHD_CANCELLED) {
        s->cancelled += 1;
    } else {
        s->failed += 1;
    }
    if (!s->have_span || r->finished > s->span_end) s->span_end = r->finished;
    s->done_cv.notify_all();
}

// Evicts least recently used unused bases until `need` more bytes fit.
void evict_bases_locked(rhd_server* s, uint64_t need) {
    for (auto it = s->bases.end(); it != s->bases.begin() && s->base_bytes + need > s->cfg.base_cache_bytes;) {
        --it;
        if (it->users != 0) continue;
        s->base_bytes -= it->source->bytes;
        rhd_source_release(it->source);
        it = s->bases.erase(it);
    }
}

// Runs one admitted request on lane L (no server lock held).
void run_request(rhd_server* s, Lane& L, Request* r) {
    rhd_context* ctx = s->ctx;
    (void)hipSetDevice(ctx->device);
    rhd_mosaic* mosaic = nullptr;
    BaseEntry* base = nullptr;
    uint64_t generation = 0;
    {
        std::lock_guard<std::mutex> g(s->mu);
        auto m = s->mosaics.find(r->frame);
        if (m != s->mosaics.end() && !m->second.dropped) {
            mosaic = m->second.mosaic;
            m->second.users += 1;
        }
        auto gen = s->generations.find(r->frame);
        generation = gen == s->generations.end() ? 0 : gen->second;
        if (s->cfg.base_cache_bytes)
            for (auto it = s->bases.begin(); it != s->bases.end(); ++it)
                if (it->key == r->base_key && it->generation == generation) {
                    s->bases.splice(s->bases.begin(), s->bases, it);
                    base = &s->bases.front();
                    base->users += 1;
                    break;
                }
    }
    const rhd_detail_desc* dd = r->has_detail ? &r->detail : nullptr;
    if (cancelled(&r->cancel)) {
        r->status = RHD_CANCELLED;
    } else if (base) {
        // Base hit: copy the cached demosaiced RGB, then capture sharpening.
        r->base_hit = true;
        rhd_source* out = nullptr;
        r->status = new_source(ctx, &L, base->source->width, base->source->height, &out);
        if (r->status == RHD_OK) {
            hipError_t e = hipMemcpyAsync(out->d, base->source->d, out->bytes,
                                          hipMemcpyDeviceToDevice, L.stream);
            r->status = e == hipSuccess ? RHD_OK : fail(ctx, RHD_SYNCHRONIZATION, "base copy");
        }
        if (r->status == RHD_OK) r->status = reset_stats(ctx, L);
        if (r->status == RHD_OK && detail_active(dd)) {
            if (cancelled(&r->cancel)) r->status = fail(ctx, RHD_CANCELLED, "cancelled");
            if (r->status == RHD_OK) r->status = run_detail(ctx, L, out, dd);
        }
        if (r->status == RHD_OK) r->status = finish_stats(ctx, L, nullptr, RHD_HIGHLIGHT_OFF);
        if (r->status != RHD_OK) {
            (void)hipStreamSynchronize(L.stream);
            rhd_source_release(out);
        } else {
            r->out = out;
        }
    } else if (!mosaic) {
        r->status = fail(ctx, RHD_INVALID_INPUT, "frame is not resident on the server");
    } else {
        rhd_source* fresh = nullptr;
        r->status = develop_on(ctx, L, mosaic, &r->desc, dd, &r->cancel, &r->out, &r->metrics,
                               s->cfg.base_cache_bytes ? &fresh : nullptr);
        if (fresh) {
            std::lock_guard<std::mutex> g(s->mu);
            evict_bases_locked(s, fresh->bytes);
            if (s->base_bytes + fresh->bytes <= s->cfg.base_cache_bytes) {
                s->bases.push_front(BaseEntry{r->base_key, r->frame, generation, fresh, 0});
                s->base_bytes += fresh->bytes;
                fresh = nullptr;
            }
        }
        if (fresh) rhd_source_release(fresh);
    }
    if (r->status == RHD_OK && r->spatial) {
        rhd_source* small = nullptr;
        float ms = 0.0f;
        r->status = spatial_on(ctx, L, r->out, r->has_optics ? &r->optics : nullptr,
                               r->target_width, r->target_height, false, &r->cancel, &small, &ms);
        rhd_source_release(r->out);
        r->out = small;
        r->metrics.total_ms += ms;
    }
    std::lock_guard<std::mutex> g(s->mu);
    if (base) {
        base->users -= 1;
        // A base whose frame got new samples while this request used it is
        // released as soon as it is free.
        auto gen = s->generations.find(base->frame);
        if (base->users == 0 && gen != s->generations.end() && gen->second != base->generation) {
            for (auto it = s->bases.begin(); it != s->bases.end(); ++it)
                if (&*it == base) {
                    s->base_bytes -= it->source->bytes;
                    rhd_source_release(it->source);
                    s->bases.erase(it);
                    break;
                }
        }
    }
    if (mosaic) {
        auto m = s->mosaics.find(r->frame);
        if (m != s->mosaics.end()) {
            m->second.users -= 1;
            if (m->second.dropped && m->second.users == 0) {
                rhd_mosaic_release(m->second.mosaic);
                s->mosaics.erase(m);
            }
        }
    }
    s->running_bytes -= r->footprint;
    finish_locked(s, r);
    s->work_cv.notify_all();
}

void worker(rhd_server* s, int lane_index) {
    Lane* L = nullptr;
    (void)hipSetDevice(s->ctx->device);
    if (lane_open(s->ctx, lane_index, &L) != RHD_OK) return;
    while (true) {
        Request* r = nullptr;
        {
            std::unique_lock<std::mutex> lock(s->mu);
            s->work_cv.wait(lock, [&] {
                if (s->stop) return true;
                for (Request* q : s->queue)
                    if (s->running_bytes == 0 || s->running_bytes + q->footprint <= s->cfg.admission_bytes)
                        return true;
                return false;
            });
            if (s->stop) return;
            for (auto it = s->queue.begin(); it != s->queue.end(); ++it)
                if (s->running_bytes == 0 ||
                    s->running_bytes + (*it)->footprint <= s->cfg.admission_bytes) {
                    r = *it;
                    s->queue.erase(it);
                    break;
                }
            r->state = State::Running;
            r->lane = lane_index;
            r->started = Clock::now();
            s->running_bytes += r->footprint;
        }
        run_request(s, *L, r);
    }
}

}  // namespace

namespace {

rhd_status submit(rhd_server* s, uint64_t frame_id, const rhd_sensor_desc* d,
                  const rhd_detail_desc* dd, bool spatial, const rhd_optics_desc* optics,
                  uint32_t target_width, uint32_t target_height, uint64_t* request_id) {
    if (!s || !d || !request_id) return RHD_INVALID_INPUT;
    if (spatial) {
        if (target_width == 0 || target_height == 0 || target_width > d->crop_width ||
            target_height > d->crop_height)
            return fail(s->ctx, RHD_INVALID_DIMENSIONS,
                        "area reduction target must be non-zero and no larger than the crop");
        if (optics && (optics->width != d->crop_width || optics->height != d->crop_height))
            return fail(s->ctx, RHD_BASE_MISMATCH, "optics plan dimensions differ from the crop");
    }
    rhd_status st = validate_detail(s->ctx, dd);
    if (st != RHD_OK) return st;
    if (!rhd_supports(d, dd))
        return fail(s->ctx, RHD_UNSUPPORTED_EDITS, "GPU sensor development does not support these edits");
    if (!d->black_levels || !d->white_levels || d->black_repeat_width == 0 ||
        d->black_repeat_height == 0 || d->white_count == 0)
        return fail(s->ctx, RHD_INVALID_INPUT, "missing level patterns");
    Request* r = new Request();
    r->frame = frame_id;
    r->desc = *d;
    r->black.assign(d->black_levels, d->black_levels + d->black_repeat_width * d->black_repeat_height);
    r->white.assign(d->white_levels, d->white_levels + d->white_count);
    r->desc.black_levels = r->black.data();
    r->desc.white_levels = r->white.data();
    if (dd) {
        r->detail = *dd;
        r->has_detail = true;
    }
    r->base_key = recipe_key(frame_id, r->desc, r->black, r->white);
    r->footprint = footprint_of(r->desc, dd);
    if (spatial) {
        r->spatial = true;
        r->has_optics = optics != nullptr;
        if (optics) r->optics = *optics;
        r->target_width = target_width;
        r->target_height = target_height;
        // Vignetted copy, radius checkpoints, horizontal results, output.
        const uint64_t px = (uint64_t)d->crop_width * d->crop_height;
        r->footprint += px * 12 + px / 4 +
                        (uint64_t)target_width * (d->crop_height + target_height) * 12;
    }
    r->megapixels = (double)d->crop_width * d->crop_height / 1e6;
    r->submitted = Clock::now();
    {
        std::lock_guard<std::mutex> g(s->mu);
        r->id = s->next_id++;
        s->submitted += 1;
        if (!s->have_span) {
            s->have_span = true;
            s->span_start = r->submitted;
            s->span_end = r->submitted;
        }
        // Newest wins (scheduler.rs:43-87): the frame's queued requests are
        // coalesced into this one and its running request is cancelled.
        for (auto it = s->queue.begin(); it != s->queue.end();) {
            if ((*it)->frame == frame_id) {
                (*it)->status = RHD_CANCELLED;
                s->coalesced += 1;
                finish_locked(s, *it);
                it = s->queue.erase(it);
            } else {
                ++it;
            }
        }
        for (auto& kv : s->requests)
            if (kv.second->frame == frame_id && kv.second->state == State::Running)
                kv.second->cancel = 1;
        s->requests[r->id] = r;
        s->queue.push_back(r);
        *request_id = r->id;
    }
    s->work_cv.notify_all();
    return RHD_OK;
}

}  // namespace

namespace rhd {
rhd_status install_mosaic(rhd_server* s, uint64_t frame_id, rhd_mosaic* m);
}  // namespace rhd

extern "C" {

rhd_status rhd_server_create(rhd_context* ctx, const rhd_server_config* cfg, rhd_server** out) {
    if (!ctx || !out) return RHD_INVALID_INPUT;
    *out = nullptr;
    rhd_server_config c = cfg ? *cfg : rhd_server_config{};
    if (c.lanes == 0) c.lanes = 3;
    if (c.lanes > (uint32_t)MAX_WORKERS)
        return fail(ctx, RHD_INVALID_INPUT, "too many server lanes");
    if (c.admission_bytes == 0) c.admission_bytes = ctx->budget.load() / 2;
    Lane* upload = nullptr;
    rhd_status st = lane_open(ctx, UPLOAD_LANE, &upload);
    if (st != RHD_OK) return st;
    for (uint32_t i = 0; i < c.lanes; ++i) {
        Lane* L = nullptr;
        if ((st = lane_open(ctx, FIRST_WORKER_LANE + (int)i, &L)) != RHD_OK) return st;
    }
    rhd_server* s = new rhd_server();
    s->ctx = ctx;
    s->cfg = c;
    ctx->refs.fetch_add(1);
    for (uint32_t i = 0; i < c.lanes; ++i) s->workers.emplace_back(worker, s, FIRST_WORKER_LANE + (int)i);
    *out = s;
    return RHD_OK;
}

void rhd_server_destroy(rhd_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        for (Request* r : s->queue) {
            r->status = RHD_CANCELLED;
            finish_locked(s, r);
        }
        s->queue.clear();
        for (auto& kv : s->requests)
            if (kv.second->state == State::Running) kv.second->cancel = 1;
        s->stop = true;
    }
    s->work_cv.notify_all();
    for (std::thread& t : s->workers) t.join();
    for (auto& kv : s->requests) {
        if (kv.second->out) rhd_source_release(kv.second->out);
        delete kv.second;
    }
    for (BaseEntry& b : s->bases) rhd_source_release(b.source);
    for (auto& kv : s->mosaics) rhd_mosaic_release(kv.second.mosaic);
    rhd_context* ctx = s->ctx;
    for (uint32_t i = 0; i < s->cfg.lanes; ++i) drop_scratch(ctx, ctx->lanes[FIRST_WORKER_LANE + i]);
    delete s;
    context_unref(ctx);
}

rhd_status rhd_server_put_mosaic(rhd_server* s, uint64_t frame_id, const uint16_t* host,
                                 uint64_t row_stride, uint32_t width, uint32_t height) {
    if (!s) return RHD_INVALID_INPUT;
    rhd_mosaic* m = nullptr;
    rhd_status st;
    {
        std::lock_guard<std::mutex> g(s->upload_mu);  // one upload at a time on the upload lane
        st = mosaic_upload(s->ctx, s->ctx->lanes[UPLOAD_LANE], host, row_stride, width, height, &m);
    }
    if (st != RHD_OK) return st;
    return install_mosaic(s, frame_id, m);
}

rhd_status rhd_server_put_arw(rhd_server* s, uint64_t frame_id, const uint8_t* file,
                              uint64_t size, const rhd_arw_layout* layout) {
    if (!s || !file || !layout) return RHD_INVALID_INPUT;
    rhd_mosaic* m = nullptr;
    rhd_status st;
    {
        std::lock_guard<std::mutex> g(s->upload_mu);
        st = decode_arw_on(s->ctx, s->ctx->lanes[UPLOAD_LANE], file, size, layout, &m, nullptr);
    }
    if (st != RHD_OK) return st;
    return install_mosaic(s, frame_id, m);
}

}  // extern "C"

namespace rhd {

// Publishes `m` as frame `frame_id`, replacing an idle previous mosaic and
// dropping the bases developed from it.
rhd_status install_mosaic(rhd_server* s, uint64_t frame_id, rhd_mosaic* m) {
    rhd_mosaic* old = nullptr;
    {
        std::lock_guard<std::mutex> g(s->mu);
        ResidentMosaic& slot = s->mosaics[frame_id];
        if (slot.mosaic && slot.users == 0) old = slot.mosaic;
        if (slot.mosaic && slot.users != 0) {
            // Still read by a running request: keep the new frame under a
            // fresh slot only once the old one is released.
            rhd_mosaic_release(m);
            return fail(s->ctx, RHD_BASE_MISMATCH, "frame is in use by a running develop");
        }
        slot.mosaic = m;
        slot.dropped = false;
        // Bases of the earlier samples no longer match the frame: the new
        // generation keeps them from being served; the unused ones go now,
        // the ones still in use when their last request finishes.
        s->generations[frame_id] += 1;
        for (auto it = s->bases.begin(); it != s->bases.end();)
            if (it->frame == frame_id && it->users == 0) {
                s->base_bytes -= it->source->bytes;
                rhd_source_release(it->source);
                it = s->bases.erase(it);
            } else {
                ++it;
            }
    }
    if (old) rhd_mosaic_release(old);
    return RHD_OK;
}

}  // namespace rhd

extern "C" {

rhd_status rhd_server_drop_mosaic(rhd_server* s, uint64_t frame_id) {
    if (!s) return RHD_INVALID_INPUT;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->mosaics.find(frame_id);
    if (it == s->mosaics.end()) return RHD_OK;
    if (it->second.users == 0) {
        rhd_mosaic_release(it->second.mosaic);
        s->mosaics.erase(it);
    } else {
        it->second.dropped = true;
    }
    return RHD_OK;
}

rhd_status rhd_server_submit(rhd_server* s, uint64_t frame_id, const rhd_sensor_desc* d,
                             const rhd_detail_desc* dd, uint64_t* request_id) {
    return submit(s, frame_id, d, dd, false, nullptr, 0, 0, request_id);
}

rhd_status rhd_server_submit_spatial(rhd_server* s, uint64_t frame_id, const rhd_sensor_desc* d,
                                     const rhd_detail_desc* dd, const rhd_optics_desc* optics,
                                     uint32_t target_width, uint32_t target_height,
                                     uint64_t* request_id) {
    return submit(s, frame_id, d, dd, true, optics, target_width, target_height, request_id);
}

rhd_status rhd_server_wait(rhd_server* s, uint64_t request_id, int32_t timeout_ms, rhd_source** out,
                           rhd_request_info* info) {
    if (!s) return RHD_INVALID_INPUT;
    if (out) *out = nullptr;
    std::unique_lock<std::mutex> lock(s->mu);
    auto it = s->requests.find(request_id);
    if (it == s->requests.end()) return fail(s->ctx, RHD_INVALID_INPUT, "unknown or collected request");
    Request* r = it->second;
    auto done = [&] { return r->state == State::Done; };
    if (timeout_ms < 0) {
        s->done_cv.wait(lock, done);
    } else if (!s->done_cv.wait_for(lock, std::chrono::milliseconds(timeout_ms), done)) {
        return RHD_SYNCHRONIZATION;  // still queued or running; the request stays collectable
    }
    if (info) {
        info->status = r->status;
        info->lane = r->lane < 0 ? 0u : (uint32_t)r->lane;
        info->base_cache_hit = r->base_hit ? 1 : 0;
        info->queued_ms = r->lane < 0 ? 0.0 : ms_between(r->submitted, r->started);
        info->run_ms = r->lane < 0 ? 0.0 : ms_between(r->started, r->finished);
        info->latency_ms = ms_between(r->submitted, r->finished);
        info->megapixels = r->megapixels;
        info->metrics = r->metrics;
    }
    const rhd_status st = r->status;
    if (out) {
        *out = r->out;
    } else if (r->out) {
        rhd_source_release(r->out);
    }
    s->requests.erase(it);
    delete r;
    return st;
}

rhd_status rhd_server_read_stats(rhd_server* s, rhd_server_stats* out) {
    if (!s || !out) return RHD_INVALID_INPUT;
    std::lock_guard<std::mutex> g(s->mu);
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->cancelled = s->cancelled;
    out->coalesced = s->coalesced;
    out->failed = s->failed;
    out->base_hits = s->base_hits;
    out->p50_ms = percentile(s->latencies, 0.50);
    out->p99_ms = percentile(s->latencies, 0.99);
    double sum = 0.0;
    for (double v : s->latencies) sum += v;
    out->mean_ms = s->latencies.empty() ? 0.0 : sum / (double)s->latencies.size();
    out->megapixels = s->megapixels;
    out->span_ms = s->have_span ? ms_between(s->span_start, s->span_end) : 0.0;
    out->base_cache_bytes = s->base_bytes;
    out->running_bytes = s->running_bytes;
    return RHD_OK;
}

rhd_status rhd_server_reset_stats(rhd_server* s) {
    if (!s) return RHD_INVALID_INPUT;
    std::lock_guard<std::mutex> g(s->mu);
    s->latencies.clear();
    s->submitted = s->completed = s->cancelled = s->failed = s->base_hits = s->coalesced = 0;
    s->megapixels = 0.0;
    s->have_span = false;
    return RHD_OK;
}

}  // extern "C"
