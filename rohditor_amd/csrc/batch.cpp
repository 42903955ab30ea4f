// Batched develop: many resident mosaics of uneven sizes and recipes in one
// HIP graph launch.
//
// The reference develops one frame per call (crates/gpu/src/sensor/
// develop.rs:68-142, driven one request at a time by crates/engine/src/
// coordinator/worker.rs:723-800). Here a batch is laid out once:
//   - every frame gets its f32 mosaic plane and RGB source up front;
//   - frames are spread over K lanes (HIP streams with their own scratch),
//     largest first onto the least loaded lane;
//   - each lane's scratch is grown to its largest frame, so nothing allocates
//     while the stream is being captured;
//   - the whole batch is captured into one graph: K independent branches
//     forked from lane 0, each the stage sequence of its frames (kernels and
//     copies only; statistics are cleared by a kernel and copied to a pinned
//     per-frame slot).
// rhd_batch_run launches the graph once and waits for it. K comes from a
// small autotuner: every candidate lane count is captured and timed on the
// batch itself and the fastest is kept (and remembered per batch shape).
#include <algorithm>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "rhd_context.hpp"

using namespace rhd;

namespace {

constexpr int FIRST_BRANCH_LANE = 1;
constexpr int MAX_BRANCHES = MAX_LANES - 2;  // lanes 1..6
const uint32_t LANE_CANDIDATES[] = {1, 2, 3, 4, 6};

struct Job {
    const rhd_mosaic* mosaic = nullptr;
    rhd_sensor_desc desc{};
    std::vector<float> black, white;
    rhd_detail_desc detail{};
    rhd_plane* plane = nullptr;
    rhd_source* src = nullptr;
    int lane = 0;
    double cost = 0.0;
};

// Relative cost of a frame's stage, for balancing lanes.
double cost_of(const rhd_sensor_desc& d, const rhd_detail_desc& dd) {
    const double px = (double)d.crop_width * d.crop_height;
    double per_px = d.demosaic == RHD_DEMOSAIC_AMAZE ? 140.0
                    : d.demosaic == RHD_DEMOSAIC_RCD ? 60.0
                                                      : 8.0;
    if (d.highlight == RHD_HIGHLIGHT_OPPOSED) per_px += 60.0;
    if (capture_active(&dd)) per_px += 20.0 + 6.0 * dd.iterations;
    per_px += 4.0 * dd.false_colour;
    if (dd.nr_chroma > 0.0f) per_px += 12.0;
    if (dd.nr_luminance > 0.0f) per_px += 16.0;
    return px * per_px;
}

std::string signature(const std::vector<Job>& jobs) {
    std::string s;
    for (const Job& j : jobs) {
        s += std::to_string(j.desc.crop_width) + "x" + std::to_string(j.desc.crop_height) + ":" +
             std::to_string(j.desc.highlight) + std::to_string(j.desc.demosaic) + ":" +
             std::to_string(capture_active(&j.detail) ? j.detail.iterations : -1) + ":" +
             std::to_string(j.detail.false_colour) + (noise_active(&j.detail) ? "n" : "") + ";";
    }
    return s;
}

std::mutex tuned_mu;
std::map<std::string, uint32_t> tuned;  // batch shape -> lane count

// Scratch a batch owns for one branch lane. The graph captures these
// pointers, so they must not be the lane's own (which other work may grow
// and free): they are swapped into the lane only while it is captured.
struct BranchScratch {
    DeviceBuffer amaze, amaze_areas, rcd, capture, opposed;
};

void swap_scratch(Lane& L, BranchScratch& s) {
    std::swap(L.amaze, s.amaze);
    std::swap(L.amaze_areas, s.amaze_areas);
    std::swap(L.rcd, s.rcd);
    std::swap(L.capture, s.capture);
    std::swap(L.opposed, s.opposed);
}

}  // namespace

struct rhd_batch {
    rhd_context* ctx = nullptr;
    std::vector<Job> jobs;
    BranchScratch scratch[MAX_LANES];
    DeviceStats* h_stats = nullptr;  // pinned, one slot per frame
    uint32_t lanes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t fork = nullptr, t0 = nullptr, t1 = nullptr;
    hipEvent_t join[MAX_LANES] = {};
};

namespace {

void drop_graph(rhd_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Assigns frames to `lanes` branches (largest first onto the least loaded),
// grows the branch lanes' scratch, and captures and instantiates the graph.
rhd_status build_graph(rhd_batch* b, uint32_t lanes) {
    rhd_context* ctx = b->ctx;
    drop_graph(b);
    b->lanes = lanes;
    std::vector<size_t> order(b->jobs.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return b->jobs[x].cost > b->jobs[y].cost; });
    std::vector<double> load(lanes, 0.0);
    for (size_t i : order) {
        const uint32_t k = (uint32_t)(std::min_element(load.begin(), load.end()) - load.begin());
        load[k] += b->jobs[i].cost;
        b->jobs[i].lane = FIRST_BRANCH_LANE + (int)k;
    }
    for (uint32_t k = 0; k < lanes; ++k) {
        Lane* L = nullptr;
        rhd_status st = lane_open(ctx, FIRST_BRANCH_LANE + (int)k, &L);
        if (st != RHD_OK) return st;
    }
    struct Swapped {
        rhd_batch* b;
        uint32_t lanes;
        Swapped(rhd_batch* b_, uint32_t n) : b(b_), lanes(n) { flip(); }
        ~Swapped() { flip(); }
        void flip() {
            for (uint32_t k = 0; k < lanes; ++k)
                swap_scratch(b->ctx->lanes[FIRST_BRANCH_LANE + k], b->scratch[FIRST_BRANCH_LANE + k]);
        }
    } swapped(b, lanes);
    for (Job& j : b->jobs) {
        rhd_status st = stage_scratch(ctx, ctx->lanes[j.lane], &j.desc, &j.detail);
        if (st != RHD_OK) return st;
    }
    Lane& origin = ctx->lanes[0];
    if (hipStreamBeginCapture(origin.stream, hipStreamCaptureModeRelaxed) != hipSuccess)
        return fail(ctx, RHD_SYNCHRONIZATION, "could not begin graph capture");
    rhd_status st = RHD_OK;
    hipError_t e = hipEventRecord(b->fork, origin.stream);
    for (uint32_t k = 0; k < lanes && e == hipSuccess; ++k)
        e = hipStreamWaitEvent(ctx->lanes[FIRST_BRANCH_LANE + k].stream, b->fork, 0);
    if (e != hipSuccess) st = fail(ctx, RHD_SYNCHRONIZATION, "graph fork");
    for (size_t i = 0; i < b->jobs.size() && st == RHD_OK; ++i) {
        Job& j = b->jobs[i];
        st = enqueue_develop(ctx, ctx->lanes[j.lane], j.mosaic, &j.desc, &j.detail, j.plane, j.src,
                             b->h_stats + i);
    }
    for (uint32_t k = 0; k < lanes && st == RHD_OK; ++k) {
        Lane& L = ctx->lanes[FIRST_BRANCH_LANE + k];
        if (hipEventRecord(b->join[k], L.stream) != hipSuccess ||
            hipStreamWaitEvent(origin.stream, b->join[k], 0) != hipSuccess)
            st = fail(ctx, RHD_SYNCHRONIZATION, "graph join");
    }
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(origin.stream, &graph);
    if (st == RHD_OK && e != hipSuccess) st = fail(ctx, RHD_SYNCHRONIZATION, "graph capture failed");
    if (st != RHD_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        return st;
    }
    b->graph = graph;
    if (hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        drop_graph(b);
        return fail(ctx, RHD_SYNCHRONIZATION, "graph instantiation failed");
    }
    return RHD_OK;
}

rhd_status launch(rhd_batch* b, float* ms) {
    rhd_context* ctx = b->ctx;
    Lane& origin = ctx->lanes[0];
    if (hipEventRecord(b->t0, origin.stream) != hipSuccess ||
        hipGraphLaunch(b->exec, origin.stream) != hipSuccess ||
        hipEventRecord(b->t1, origin.stream) != hipSuccess ||
        hipStreamSynchronize(origin.stream) != hipSuccess)
        return fail(ctx, RHD_SYNCHRONIZATION, "graph launch failed");
    if (ms && hipEventElapsedTime(ms, b->t0, b->t1) != hipSuccess) *ms = 0.0f;
    return RHD_OK;
}

}  // namespace

extern "C" {

void rhd_batch_destroy(rhd_batch* b) {
    if (!b) return;
    rhd_context* ctx = b->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->lanes[0].stream);
    drop_graph(b);
    for (Job& j : b->jobs) {
        rhd_plane_release(j.plane);
        rhd_source_release(j.src);
    }
    for (BranchScratch& s : b->scratch)
        for (DeviceBuffer* d : {&s.amaze, &s.amaze_areas, &s.rcd, &s.capture, &s.opposed})
            device_free(ctx, d->ptr, d->bytes);
    for (hipEvent_t e : {b->fork, b->t0, b->t1})
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->join)
        if (e) (void)hipEventDestroy(e);
    if (b->h_stats) (void)hipHostFree(b->h_stats);
    delete b;
    context_unref(ctx);
}

rhd_status rhd_batch_create(rhd_context* ctx, uint32_t count, const rhd_mosaic* const* mosaics,
                            const rhd_sensor_desc* descs, const rhd_detail_desc* details,
                            uint32_t lanes, rhd_batch** out) {
    if (!ctx || !out || !mosaics || !descs || !details || count == 0) return RHD_INVALID_INPUT;
    *out = nullptr;
    if (lanes > (uint32_t)MAX_BRANCHES) return fail(ctx, RHD_INVALID_INPUT, "too many batch lanes");
    (void)hipSetDevice(ctx->device);
    rhd_batch* b = new rhd_batch();
    b->ctx = ctx;
    ctx->refs.fetch_add(1);
    auto bail = [&](rhd_status st) {
        rhd_batch_destroy(b);
        return st;
    };
    b->jobs.resize(count);
    for (uint32_t i = 0; i < count; ++i) {
        Job& j = b->jobs[i];
        const rhd_sensor_desc& d = descs[i];
        if (!mosaics[i] || mosaics[i]->ctx != ctx)
            return bail(fail(ctx, RHD_INVALID_INPUT, "batch mosaic belongs to another context"));
        rhd_status st = validate_desc(ctx, &d);
        if (st == RHD_OK) st = validate_detail(ctx, &details[i]);
        if (st == RHD_OK && !rhd_supports(&d, &details[i]))
            st = fail(ctx, RHD_UNSUPPORTED_EDITS, "GPU sensor development does not support these edits");
        if (st == RHD_OK && (mosaics[i]->width != d.sensor_width ||
                             mosaics[i]->height != d.sensor_height ||
                             mosaics[i]->row_stride != d.row_stride))
            st = fail(ctx, RHD_INVALID_INPUT, "decoded frame does not match its normalization contract");
        if (st != RHD_OK) return bail(st);
        j.mosaic = mosaics[i];
        j.desc = d;
        j.black.assign(d.black_levels, d.black_levels + d.black_repeat_width * d.black_repeat_height);
        j.white.assign(d.white_levels, d.white_levels + d.white_count);
        j.desc.black_levels = j.black.data();
        j.desc.white_levels = j.white.data();
        j.detail = details[i];
        j.cost = cost_of(j.desc, j.detail);
        if ((st = new_plane(ctx, nullptr, d.crop_width, d.crop_height, crop_pattern(&d), &j.plane)) !=
                RHD_OK ||
            (st = new_source(ctx, nullptr, d.crop_width, d.crop_height, &j.src)) != RHD_OK)
            return bail(st);
    }
    if (hipHostMalloc((void**)&b->h_stats, count * sizeof(DeviceStats), hipHostMallocDefault) !=
            hipSuccess ||
        hipEventCreate(&b->fork) != hipSuccess || hipEventCreate(&b->t0) != hipSuccess ||
        hipEventCreate(&b->t1) != hipSuccess)
        return bail(fail(ctx, RHD_UNSUPPORTED, "could not create batch events or statistics"));
    for (hipEvent_t& e : b->join)
        if (hipEventCreate(&e) != hipSuccess)
            return bail(fail(ctx, RHD_UNSUPPORTED, "could not create batch events"));
    // The profiler would record events inside the capture.
    const bool profiling = ctx->lanes[0].profiling;
    for (Lane& L : ctx->lanes) L.profiling = false;
    rhd_status st = RHD_OK;
    if (lanes != 0) {
        st = build_graph(b, lanes);
    } else {
        const std::string sig = signature(b->jobs);
        uint32_t known = 0;
        {
            std::lock_guard<std::mutex> g(tuned_mu);
            auto it = tuned.find(sig);
            if (it != tuned.end()) known = it->second;
        }
        if (known) {
            st = build_graph(b, known);
        } else {
            // Autotune: capture and time every candidate on the batch itself.
            uint32_t best = 0;
            float best_ms = 0.0f;
            for (uint32_t k : LANE_CANDIDATES) {
                if (k > count) break;
                float ms = 0.0f;
                st = build_graph(b, k);
                if (st == RHD_OK) st = launch(b, nullptr);  // warm-up: first touch of buffers
                if (st == RHD_OK) st = launch(b, &ms);
                if (st != RHD_OK) break;
                if (best == 0 || ms < best_ms) {
                    best = k;
                    best_ms = ms;
                }
            }
            if (st == RHD_OK) {
                std::lock_guard<std::mutex> g(tuned_mu);
                tuned[sig] = best;
            }
            if (st == RHD_OK && b->lanes != best) st = build_graph(b, best);
        }
    }
    ctx->lanes[0].profiling = profiling;
    if (st != RHD_OK) return bail(st);
    *out = b;
    return RHD_OK;
}

rhd_status rhd_batch_run(rhd_batch* b, int32_t* statuses, rhd_metrics* metrics, float* gpu_ms) {
    if (!b) return RHD_INVALID_INPUT;
    (void)hipSetDevice(b->ctx->device);
    float ms = 0.0f;
    rhd_status st = launch(b, &ms);
    if (gpu_ms) *gpu_ms = ms;
    if (st != RHD_OK) return st;
    rhd_status first = RHD_OK;
    for (size_t i = 0; i < b->jobs.size(); ++i) {
        rhd_metrics m{};
        const rhd_status s = stats_to_metrics(b->ctx, b->h_stats[i], &m, b->jobs[i].desc.highlight);
        if (statuses) statuses[i] = s;
        if (metrics) metrics[i] = m;
        if (first == RHD_OK) first = s;
    }
    return first;
}

const rhd_source* rhd_batch_source(const rhd_batch* b, uint32_t index) {
    if (!b || index >= b->jobs.size()) return nullptr;
    return b->jobs[index].src;
}

uint32_t rhd_batch_lanes(const rhd_batch* b) { return b ? b->lanes : 0; }

}  // extern "C"
