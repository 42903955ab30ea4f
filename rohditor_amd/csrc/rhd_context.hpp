// Host-side state shared by the librhd translation units: the context with
// its lanes and device pool, the resident handles, and the internal stage
// sequence. Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "rhd_internal.hpp"

namespace rhd {

constexpr uint64_t DEFAULT_BUDGET = 64ull << 30;  // 64 GiB of the 288 GiB HBM3E
constexpr int MAX_TILES_PER_LAUNCH = 2048;
constexpr int MAX_LANES = 8;

struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
};

// One independent stream of work inside a context: its own HIP stream,
// device statistics (with a pinned readback slot), stage events, scratch and
// profiler. Lane 0 serves the staged API; the others let several develops
// overlap on the device (the serving loop and the batched path).
struct Lane {
    int index = 0;
    bool open = false;
    hipStream_t stream = nullptr;
    // A second stream for work that can run beside the lane's own, forked
    // from and joined back into `stream` with the events.
    hipStream_t aux = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    DeviceStats* d_stats = nullptr;
    DeviceStats* h_stats = nullptr;
    DeviceBuffer amaze, amaze_areas, rcd, capture, opposed, optics, raw;
    hipEvent_t ev[5] = {};
    bool profiling = false;
    Profiler profiler;
    // Plan geometry whose vignetting checkpoints sit at the start of `optics`
    // (they depend on nothing else); cleared whenever that buffer changes.
    struct RowsKey {
        uint32_t width = 0, height = 0;
        float ns = 0.0f, start_x = 0.0f, start_y = 0.0f;
        void* buffer = nullptr;
    } optics_rows;
};

struct PooledBlock {
    void* ptr;
    uint64_t bytes;
};


}  // namespace rhd

struct rhd_context {
    int device = 0;
    std::atomic<uint64_t> budget{rhd::DEFAULT_BUDGET};
    std::atomic<uint64_t> reserved{0};  // bytes held by live handles and scratch
    std::mutex mu;                      // last_error, the pool, opening lanes
    std::string last_error;
    // Released device blocks kept for reuse, so that steady-state develops
    // allocate nothing (hipFree would also wait for the whole device).
    std::vector<rhd::PooledBlock> pool;
    uint64_t pooled = 0;
    rhd::Lane lanes[rhd::MAX_LANES];
    std::atomic<int> refs{1};  // rhd_create's reference plus one per live handle
};

struct rhd_mosaic {
    rhd_context* ctx;
    uint16_t* d;
    uint64_t row_stride;
    uint32_t width, height;
    uint64_t bytes;
};

struct rhd_plane {
    rhd_context* ctx;
    float* d;
    uint32_t width, height;
    int32_t pattern;
    uint64_t bytes;
};

struct rhd_source {
    rhd_context* ctx;
    float* d;
    uint32_t width, height;
    uint64_t bytes;
};
#define PROF(L) ((L).profiling ? &(L).profiler : nullptr)

#define RHD_HIP(ctx, call)                                          \
    do {                                                            \
        hipError_t _e = (call);                                     \
        if (_e != hipSuccess) return hip_fail((ctx), _e, #call);    \
    } while (0)

namespace rhd {

rhd_status fail(rhd_context* ctx, rhd_status status, const std::string& msg);
rhd_status hip_fail(rhd_context* ctx, hipError_t e, const char* what);
rhd_status check_handle_ctx(rhd_context* ctx, const void* handle_ctx);
// Grows a lane scratch buffer to at least `bytes` (waits for the lane first).
rhd_status ensure_scratch(rhd_context* ctx, Lane& L, DeviceBuffer& b, uint64_t bytes);
// Milliseconds between two recorded events (0 on error).
float elapsed(hipEvent_t a, hipEvent_t b);
// Drops one reference; the last one frees the lanes and the pool.
void context_unref(rhd_context* ctx);
// Waits for the lane, then returns its scratch to the pool.
void drop_scratch(rhd_context* ctx, Lane& L);
bool cancelled(const volatile uint8_t* flag);
rhd_status lane_open(rhd_context* ctx, int i, Lane** out);
rhd_status device_alloc(rhd_context* ctx, Lane* L, uint64_t bytes, void** out);
void device_free(rhd_context* ctx, void* p, uint64_t bytes);
bool capture_active(const rhd_detail_desc* dd);
// Any detail-stage step to run (noise reduction, false colour or capture).
bool detail_active(const rhd_detail_desc* dd);
bool noise_active(const rhd_detail_desc* dd);
// Lane scratch the detail stage needs for a w x h frame.
uint64_t detail_scratch_bytes(int w, int h, const rhd_detail_desc* dd);
// Uploads an ARW file's raw strip and decodes it on lane L (raw.cpp).
rhd_status decode_arw_on(rhd_context* ctx, Lane& L, const uint8_t* file, uint64_t size,
                         const rhd_arw_layout* l, rhd_mosaic** out, float* gpu_ms);
// False colour then capture sharpening on `src` (in place).
// capture_ready: the demosaic already wrote capture's guide and mask planes.
rhd_status run_detail(rhd_context* ctx, Lane& L, rhd_source* src, const rhd_detail_desc* dd,
                      bool capture_ready = false);
rhd_status validate_detail(rhd_context* ctx, const rhd_detail_desc* dd);
rhd_status new_source(rhd_context* ctx, Lane* L, uint32_t w, uint32_t h, rhd_source** out);
rhd_status reset_stats(rhd_context* ctx, Lane& L);
rhd_status finish_stats(rhd_context* ctx, Lane& L, rhd_metrics* m, int highlight);
rhd_status run_capture(rhd_context* ctx, Lane& L, rhd_source* src, const rhd_detail_desc* dd,
                       bool inputs_ready = false);
rhd_status stats_to_metrics(rhd_context* ctx, const DeviceStats& s, rhd_metrics* m, int highlight);
rhd_status validate_desc(rhd_context* ctx, const rhd_sensor_desc* d);
int crop_pattern(const rhd_sensor_desc* d);
rhd_status new_plane(rhd_context* ctx, Lane* L, uint32_t w, uint32_t h, int32_t pattern,
                     rhd_plane** out);
rhd_status stage_scratch(rhd_context* ctx, Lane& L, const rhd_sensor_desc* d,
                         const rhd_detail_desc* dd);
rhd_status enqueue_develop(rhd_context* ctx, Lane& L, const rhd_mosaic* mosaic,
                           const rhd_sensor_desc* d, const rhd_detail_desc* dd, rhd_plane* plane,
                           rhd_source* src, DeviceStats* stats_out);
rhd_status mosaic_upload(rhd_context* ctx, Lane& L, const uint16_t* host, uint64_t row_stride,
                         uint32_t width, uint32_t height, rhd_mosaic** out);
// The whole stage on one lane; see api.cpp. With `base`, the demosaiced
// camera RGB before capture sharpening is also copied into it.
rhd_status develop_on(rhd_context* ctx, Lane& L, const rhd_mosaic* mosaic,
                      const rhd_sensor_desc* d, const rhd_detail_desc* dd,
                      const volatile uint8_t* cancel, rhd_source** out, rhd_metrics* m,
                      rhd_source** base);
// Optics correction and area reduction of `src` on one lane (spatial.cpp):
// `optics` may be null; target equal to the source size with `full` set
// gives the full-resolution corrected source.
// Resolves the no-allocation entry point's scratch array on `device`, which
// must be current (noalloc.hip); rhd_create calls it once per context.
rhd_status noalloc_setup(rhd_context* ctx, int device);
// Builds the device's log2 exception bitmap (libm.hip) once per device, with
// the device current; rhd_create calls it. libm_log2_bitmap returns it (null
// before set-up).
rhd_status libm_setup(rhd_context* ctx, int device);
const uint32_t* libm_log2_bitmap(int device);
rhd_status spatial_on(rhd_context* ctx, Lane& L, const rhd_source* src,
                      const rhd_optics_desc* optics, uint32_t target_width,
                      uint32_t target_height, bool full, const volatile uint8_t* cancel,
                      rhd_source** out, float* gpu_ms);

}  // namespace rhd
