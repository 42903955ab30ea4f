/*
 * rhd — MI355X-native sensor development for rohditor (C ABI).
 *
 * The stage is normalize -> highlight (Off / Clip / Opposed) -> demosaic
 * (bilinear / MHC / RCD / AMaZE) -> capture sharpening (masked
 * Richardson-Lucy), producing f32 linear camera RGB kept resident on the GPU
 * as three planes. It replaces the reference's wgpu executor
 * `GpuSensorProcessor` (crates/gpu/src/sensor, crates/gpu/src/capture)
 * behind the same stage API; every entry point below cites the reference
 * interface it replaces.
 *
 * Conventions: plain pointers and sizes only. Every call that launches GPU
 * work runs on the context's own HIP stream. Handles are move-only in the
 * reference (each stage consumes its input); here a stage that consumes a
 * handle says so and releases it, even on error. All functions are safe to
 * call from one thread per context; contexts are independent.
 *
 * `cancel` (may be NULL) points at a one-byte flag laid out like the
 * reference's CancellationToken, an Arc<AtomicBool> (crates/core/src/
 * cancel.rs:13-33): nonzero means cancelled. It is read with acquire
 * ordering between stages; a cancelled call publishes no handle and releases
 * what it reserved.
 */
#ifndef RHD_H
#define RHD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RHD_ABI_VERSION 1

/* crates/gpu/src/lib.rs:55-98 (GpuPreviewError). */
typedef enum rhd_status {
    RHD_OK = 0,
    RHD_CANCELLED = 1,
    RHD_UNSUPPORTED = 2,        /* device cannot run the stage */
    RHD_OVER_BUDGET = 3,        /* routing: run on the CPU, keep the device */
    RHD_INVALID_DIMENSIONS = 4,
    RHD_BASE_MISMATCH = 5,
    RHD_UNSUPPORTED_EDITS = 6,  /* routing: run on the CPU, keep the device */
    RHD_SYNCHRONIZATION = 7,    /* a HIP call or kernel failed */
    RHD_INVALID_INPUT = 8,      /* invalid metadata, non-finite samples, bad handles */
    RHD_READBACK = 9,
    RHD_RAW_UNSUPPORTED = 10,   /* RawError::Unsupported: not a raw this decoder reads */
    RHD_RAW_CORRUPT = 11        /* RawError::Corrupt: the container or strip is damaged */
} rhd_status;

/* crates/image/src/lib.rs:52-58 (BayerPattern); codes as in
 * crates/gpu/src/sensor/normalize.rs:695-702. */
typedef enum rhd_cfa { RHD_RGGB = 0, RHD_BGGR = 1, RHD_GRBG = 2, RHD_GBRG = 3 } rhd_cfa;

/* crates/highlight/src/lib.rs:137-141 (HighlightExecution). */
typedef enum rhd_highlight_method {
    RHD_HIGHLIGHT_OFF = 0,
    RHD_HIGHLIGHT_CLIP = 1,
    RHD_HIGHLIGHT_OPPOSED = 2
} rhd_highlight_method;

/* crates/demosaic/src/lib.rs:125-141 (DemosaicAlgorithm; Neural is not offered). */
typedef enum rhd_demosaic_algorithm {
    RHD_DEMOSAIC_BILINEAR = 0,
    RHD_DEMOSAIC_MHC = 1,
    RHD_DEMOSAIC_RCD = 2,
    RHD_DEMOSAIC_AMAZE = 3
} rhd_demosaic_algorithm;

/*
 * Resolved sensor-development description: crates/core/src/pipeline/
 * sensor.rs:282-338 (SensorDevelopmentDescription) restricted to what the
 * stage reads. Everything is resolved on the host before the call, as
 * `from_frame` does: crop, level patterns, highlight levels, demosaic balance.
 */
typedef struct rhd_sensor_desc {
    uint32_t sensor_width, sensor_height; /* RawFileInfo width/height */
    uint64_t row_stride;                  /* RawFrame.row_stride, in samples */
    int32_t sensor_pattern;               /* rhd_cfa of the full sensor */
    uint32_t crop_x, crop_y, crop_width, crop_height; /* resolved crop */
    const float* black_levels;            /* LevelPattern.values (host memory) */
    uint32_t black_repeat_width, black_repeat_height;
    const float* white_levels;            /* 1, 3 or repeat_w * repeat_h values */
    uint32_t white_count;
    int32_t highlight;                    /* rhd_highlight_method */
    float clip_levels[3];                 /* ChannelClipLevels (Clip) */
    float detection_levels[3];            /* ChannelDetectionLevels (Opposed) */
    int32_t has_neutral_ceiling;          /* OpposedOptions.neutral_ceiling is Some */
    float neutral_ceiling[3];
    int32_t demosaic;                     /* rhd_demosaic_algorithm */
    float demosaic_balance[3];            /* as-shot WhiteBalanceGains */
} rhd_sensor_desc;

/*
 * Detail contract (crates/core/src/pipeline/detail_contract.rs:15-142):
 * capture sharpening (crates/core/src/sharpening.rs:19-69,
 * CaptureSharpeningContract; crates/edit/src/sharpening.rs:22-81), false
 * colour suppression (crates/edit/src/false_color.rs, 0..=5 median steps)
 * and noise reduction (crates/edit/src/noise_reduction.rs:27-35: chroma and
 * luminance strengths, luminance detail, each in [0, 1]). The steps run in
 * the order of crates/core/src/pipeline/detail.rs:210-307: noise profiles of
 * the stage input, false colour, chroma NR, luminance NR, capture. Images
 * with fewer than one 16x16 profile tile skip noise reduction, as there.
 * `iterations` is 8 in the reference (sharpening.rs:13); other counts are an
 * extension.
 */
typedef struct rhd_detail_desc {
    int32_t capture_enabled;
    float amount;           /* [0, 1] */
    float sigma;            /* "radius": Gaussian sigma in [0.3, 1.2] */
    float noise_protection; /* [0, 1] */
    float ceilings[3];      /* DetailContract::ceilings (detail.rs:323-337) */
    int32_t iterations;
    int32_t false_colour;   /* FalseColorSuppression::steps, 0 = off */
    float balance[3];       /* DetailContract::balance: the as-shot demosaic balance */
    float nr_chroma;        /* NoiseReduction::chroma, 0 = off */
    float nr_luminance;     /* NoiseReduction::luminance, 0 = off */
    float nr_luminance_detail; /* NoiseReduction::luminance_detail (0.5 neutral) */
} rhd_detail_desc;

/* crates/highlight/src/lib.rs:253-258 (ClipStats). */
typedef struct rhd_clip_stats {
    uint64_t affected_sites, changed_sites, nominal_over_white_sites;
    uint64_t affected_by_channel[3];
} rhd_clip_stats;

/* crates/highlight/src/lib.rs:212-224 (OpposedStats). */
typedef struct rhd_opposed_stats {
    uint64_t suspected_clipped_sites, reconstructed_sites, changed_sites, fallback_sites;
    uint64_t fully_unsupported_sites, neutralised_sites;
    uint64_t suspected_by_channel[3];
    uint32_t regions; /* clipped regions found (0 when nothing was suspected) */
} rhd_opposed_stats;

/* SensorMetrics / HighlightMetrics / DemosaicMetrics / CaptureMetrics
 * (normalize.rs:16-28, highlight.rs:22-35, demosaic.rs:48-58,
 * capture/mod.rs:30-45), with GPU stage times from HIP events. */
typedef struct rhd_metrics {
    float normalize_ms, highlight_ms, demosaic_ms, capture_ms, total_ms;
    uint64_t uploaded_bytes, readback_bytes, estimated_gpu_bytes;
    uint32_t launches;
    int32_t highlight_method;
    rhd_clip_stats clip;
    rhd_opposed_stats opposed;
} rhd_metrics;

typedef struct rhd_context rhd_context; /* GpuSensorProcessor */
typedef struct rhd_mosaic rhd_mosaic;   /* resident u16 RawFrame mosaic */
typedef struct rhd_plane rhd_plane;     /* GpuNormalizedMosaic / GpuHighlightedMosaic */
typedef struct rhd_source rhd_source;   /* GpuSensorCameraSource: 3 resident planes */

/* GpuSensorProcessor::new (crates/gpu/src/sensor/normalize.rs:91) plus
 * set_budget (:157): `budget_bytes` caps the logical bytes held by live
 * handles and scratch (crates/gpu/src/memory.rs:42-81); 0 selects 64 GiB
 * (the reference defaults to 768 MiB, resources.rs:3, sized for desktop GPUs). */
rhd_status rhd_create(int device, uint64_t budget_bytes, rhd_context** out);
void rhd_destroy(rhd_context* ctx);
rhd_status rhd_set_budget(rhd_context* ctx, uint64_t budget_bytes);
uint64_t rhd_reserved_bytes(const rhd_context* ctx);
/* Message of the last failing call on this context (never NULL). */
const char* rhd_last_error(const rhd_context* ctx);
const char* rhd_status_name(rhd_status status);
/* GpuPreviewError::is_routing (crates/gpu/src/lib.rs:104-109). */
int rhd_status_is_routing(rhd_status status);

/* GpuSensorProcessor::supports (crates/gpu/src/sensor/develop.rs:55-66),
 * extended with Opposed; `detail` may be NULL. Returns 1 or 0. */
int rhd_supports(const rhd_sensor_desc* desc, const rhd_detail_desc* detail);

/* Upload of a decoded RawFrame mosaic (normalize.rs:308-351 packs and
 * uploads per tile; here the whole frame goes up once and stays resident so
 * re-develops reuse it). `host` holds row_stride * height samples. */
rhd_status rhd_mosaic_upload(rhd_context* ctx, const uint16_t* host, uint64_t row_stride,
                             uint32_t width, uint32_t height, rhd_mosaic** out);
void rhd_mosaic_release(rhd_mosaic* mosaic);

/* GpuSensorProcessor::normalize (crates/gpu/src/sensor/normalize.rs:170). */
rhd_status rhd_normalize(rhd_context* ctx, const rhd_mosaic* mosaic, const rhd_sensor_desc* desc,
                         const volatile uint8_t* cancel, rhd_plane** out, rhd_metrics* metrics);
/* GpuSensorProcessor::apply_highlight (crates/gpu/src/sensor/highlight.rs:87),
 * in place on `plane`, which stays owned by the caller. */
rhd_status rhd_highlight(rhd_context* ctx, rhd_plane* plane, const rhd_sensor_desc* desc,
                         const volatile uint8_t* cancel, rhd_metrics* metrics);
/* GpuSensorProcessor::demosaic (crates/gpu/src/sensor/demosaic.rs:74),
 * including capture sharpening when `detail` is active. Consumes `plane`. */
rhd_status rhd_demosaic(rhd_context* ctx, rhd_plane* plane, const rhd_sensor_desc* desc,
                        const rhd_detail_desc* detail, const volatile uint8_t* cancel,
                        rhd_source** out, rhd_metrics* metrics);
/* GpuCaptureProcessor::process_resident_from_resident
 * (crates/gpu/src/capture/mod.rs:54), in place on `source`. */
rhd_status rhd_capture(rhd_context* ctx, rhd_source* source, const rhd_detail_desc* detail,
                       const volatile uint8_t* cancel, rhd_metrics* metrics);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): capture sharpening exactly as rhd_capture does it (same
 * kernels, same results), in place on `planar_rgb`, a caller-owned device
 * buffer of 3 * width * height floats (R plane, then G, then B). It calls no
 * allocator. Its working planes live in a fixed device array of
 * RHD_NOALLOC_MAX_PIXELS pixels that each GPU holds until the process ends;
 * larger images return RHD_INVALID_DIMENSIONS. Only capture sharpening is
 * run: false colour or noise reduction in `detail` is RHD_INVALID_INPUT.
 * Calls on one GPU share that array, so calls that overlap on one GPU race.
 * Runs on the context's stream and waits for its result (the non-finite
 * check, as rhd_capture reports it). */
#define RHD_NOALLOC_MAX_PIXELS (2u * 1024u * 1024u)
rhd_status rhd_capture_noalloc(rhd_context* ctx, float* planar_rgb, uint32_t width,
                               uint32_t height, const rhd_detail_desc* detail);

/*
 * RAW ingest: the Sony ARW strip that crates/raw/src/rawler_adapter.rs:
 * 131-170 turns into RawFrame::mosaic through rawler 0.7.2. rhd_arw_probe
 * walks the TIFF container on the host (IFD0, SubIFDs, the IFD chain) to the
 * raw IFD; rhd_mosaic_decode_arw uploads its strip and decodes it on the
 * device into a resident mosaic (row stride = width) that rhd_develop_resident
 * and the batch and serving paths take like an uploaded one. Compressed
 * (TIFF compression 32767, 8 bits per pixel, tone curve from tag 0x7010) is
 * decoded by a kernel; uncompressed 16-bit little-endian strips are uploaded
 * as they are. Metadata (CFA, levels, white balance) still comes from the
 * caller's RawFileInfo.
 */
typedef enum rhd_arw_kind {
    RHD_ARW_NONE = 0,
    RHD_ARW_COMPRESSED = 1,
    RHD_ARW_UNCOMPRESSED16 = 2
} rhd_arw_kind;
typedef struct rhd_arw_layout {
    uint32_t width, height, bits_per_sample, compression;
    uint64_t offset, bytes;  /* the strip inside the file */
    uint16_t curve[4];       /* SonyCurve knots (compressed only) */
    int32_t kind;            /* rhd_arw_kind */
} rhd_arw_layout;
/* RHD_RAW_UNSUPPORTED (not a TIFF, or no raw IFD this decoder reads) or
 * RHD_RAW_CORRUPT (an IFD, value or strip outside the file, an IFD loop). */
rhd_status rhd_arw_probe(const uint8_t* file, uint64_t size, rhd_arw_layout* out);
rhd_status rhd_mosaic_decode_arw(rhd_context* ctx, const uint8_t* file, uint64_t size,
                                 const rhd_arw_layout* layout, rhd_mosaic** out, float* gpu_ms);
/* Copies a resident mosaic back (row_stride * height samples), for tests. */
rhd_status rhd_mosaic_read(rhd_context* ctx, const rhd_mosaic* mosaic, uint16_t* dst,
                           uint64_t count);

/* NoiseProfile (crates/detail/src/noise.rs:24-34): variance a*max(x,0) + b;
 * bins = 0 for the flat fallback. */
typedef struct rhd_noise_profile {
    float a, b;
    uint32_t bins;
} rhd_noise_profile;
/* estimate_noise_profiles (crates/detail/src/denoise/mod.rs:60-110; the GPU
 * executor's estimate_noise_from_resident, crates/gpu/src/capture/noise.rs:
 * 35-129): profiles {luma, C1, C2} of a resident source for `balance`,
 * computed on the device. *profiled = 0 (and `out` untouched) when the image
 * holds no 16x16 tile in its sampled rows. */
rhd_status rhd_noise_profiles(rhd_context* ctx, const rhd_source* source, const float balance[3],
                              rhd_noise_profile out[3], int32_t* profiled, float* gpu_ms);
/* denoise_chroma (crates/detail/src/denoise/chroma.rs:88-148) in place on a
 * resident source with the given profiles {luma, C1, C2}. */
rhd_status rhd_denoise_chroma(rhd_context* ctx, rhd_source* source, float strength,
                              const float balance[3], const float ceilings[3],
                              const rhd_noise_profile profiles[3], float* gpu_ms);
/* denoise_luma (crates/detail/src/denoise/luma.rs:128-176) in place with
 * the luma profile. */
rhd_status rhd_denoise_luma(rhd_context* ctx, rhd_source* source, float strength, float detail,
                            const float balance[3], const float ceilings[3],
                            const rhd_noise_profile* luma, float* gpu_ms);

/* GpuSensorProcessor::develop (crates/gpu/src/sensor/develop.rs:68-142):
 * normalize -> highlight -> demosaic (+ capture) from a resident mosaic. */
rhd_status rhd_develop_resident(rhd_context* ctx, const rhd_mosaic* mosaic,
                                const rhd_sensor_desc* desc, const rhd_detail_desc* detail,
                                const volatile uint8_t* cancel, rhd_source** out,
                                rhd_metrics* metrics);
/* The same from a host RawFrame mosaic (upload included). */
rhd_status rhd_develop(rhd_context* ctx, const uint16_t* host_mosaic, uint64_t row_stride,
                       const rhd_sensor_desc* desc, const rhd_detail_desc* detail,
                       const volatile uint8_t* cancel, rhd_source** out, rhd_metrics* metrics);

/* Test and CPU-handoff readbacks (the reference's test readbacks). */
rhd_status rhd_plane_read(rhd_context* ctx, const rhd_plane* plane, float* dst, uint64_t count);
rhd_status rhd_plane_write(rhd_context* ctx, rhd_plane* plane, const float* src, uint64_t count);
/* A resident f32 mosaic made from host samples, for staged tests. */
rhd_status rhd_plane_upload(rhd_context* ctx, const float* host, uint32_t width, uint32_t height,
                            int32_t pattern, rhd_plane** out);
void rhd_plane_dims(const rhd_plane* plane, uint32_t* width, uint32_t* height, int32_t* pattern);
void rhd_plane_release(rhd_plane* plane);

/* Interleaved W x H x 3 readback matching LinearRgbImage. */
rhd_status rhd_source_read_rgb(rhd_context* ctx, const rhd_source* source, float* dst,
                               uint64_t count);
/* A resident camera source from interleaved W x H x 3 host RGB, the
 * counterpart of GpuSpatialProcessor::upload_captured_source
 * (crates/gpu/src/spatial/mod.rs:110) for a CPU-developed source. */
rhd_status rhd_source_upload(rhd_context* ctx, const float* rgb, uint32_t width, uint32_t height,
                             rhd_source** out);
/* Device pointers of the three planes (R, G, B; W*H floats each). */
rhd_status rhd_source_planes(const rhd_source* source, float** red, float** green, float** blue);
void rhd_source_dims(const rhd_source* source, uint32_t* width, uint32_t* height);
void rhd_source_release(rhd_source* source);

/* Self-test, no counterpart in the reference: evaluates on the device the
 * replicas of the platform libm's cbrtf (fn 0) and log2f (fn 1) that
 * Opposed uses (crates/highlight/src/support.rs:65,106 call them through
 * Rust's f32::cbrt and f32::log2), on `count` host floats into host `out`. */
rhd_status rhd_libm_eval(rhd_context* ctx, int32_t fn, const float* in, float* out,
                         uint64_t count);

/* Blocks until the context's stream is idle. */
rhd_status rhd_synchronize(rhd_context* ctx);

/* Per-kernel GPU timing, the counterpart of the reference's executor metrics
 * (SensorMetrics etc., crates/gpu/src/sensor/normalize.rs:16-28), which carry
 * estimated bytes but no device timestamps: when enabled, HIP events are
 * recorded on the context's stream around every kernel launch, and each entry
 * reports a kernel group's total milliseconds, launch count and algorithmic
 * bytes per launch. For benchmarking; off by default. */
rhd_status rhd_profile_enable(rhd_context* ctx, int enable);
/* Times only the launches of one kernel group (`kernel`, a name as
 * rhd_profile_entry reports it), so that the events leave the rest of the
 * stage undisturbed; NULL or "" times every group again. */
rhd_status rhd_profile_only(rhd_context* ctx, const char* kernel);
rhd_status rhd_profile_reset(rhd_context* ctx);
uint32_t rhd_profile_count(rhd_context* ctx);
rhd_status rhd_profile_entry(rhd_context* ctx, uint32_t index, const char** name, double* total_ms,
                             uint64_t* launches, uint64_t* bytes_per_launch);

/* ---------------------------------------------------------------------------
 * Spatial completion: lens correction and exact area reduction of a resident
 * camera source, the step after the stage (GpuSpatialProcessor,
 * crates/gpu/src/spatial/mod.rs:33-344 and reduction.rs:15-316, over the
 * CPU contract crates/optics/src/correction.rs and crates/core/src/
 * resample.rs).
 * ------------------------------------------------------------------------- */
typedef enum rhd_distortion_model {
    RHD_DISTORTION_NONE = 0,
    RHD_DISTORTION_POLY3 = 1,  /* k1 */
    RHD_DISTORTION_POLY5 = 2,  /* k1, k2 */
    RHD_DISTORTION_PTLENS = 3  /* a, b, c */
} rhd_distortion_model;

typedef enum rhd_tca_model {
    RHD_TCA_NONE = 0,
    RHD_TCA_LINEAR = 1,  /* red[0] = kr, blue[0] = kb (already inverted, plan.rs:391-424) */
    RHD_TCA_POLY3 = 2    /* [v, c, b] per channel */
} rhd_tca_model;

typedef enum rhd_vignetting_model { RHD_VIGNETTING_NONE = 0, RHD_VIGNETTING_PA = 1 } rhd_vignetting_model;

/* A resolved OpticsExecution (crates/optics/src/plan.rs:36-48): dimensions of
 * the corrected image (the source's), the f64 normalisation and lens centre,
 * the automatic scale, and the rescaled model coefficients. Lens matching and
 * calibration interpolation happen before, on the host. */
typedef struct rhd_optics_desc {
    uint32_t width, height;
    double norm_scale, norm_unscale;
    double center_x, center_y;
    float scale;
    int32_t distortion;    /* rhd_distortion_model */
    float distortion_k[3];
    int32_t tca;           /* rhd_tca_model */
    float tca_red[3];
    float tca_blue[3];
    int32_t vignetting;    /* rhd_vignetting_model */
    float vignetting_k[3]; /* PA k1, k2, k3 */
} rhd_optics_desc;

/* Preview dimensions for a long-edge limit (crates/core/src/cpu/stages.rs:
 * 112-147): (d * max_long + long / 2) / long clamped to [2, d]. */
rhd_status rhd_preview_dimensions(uint32_t width, uint32_t height, uint32_t max_long_edge,
                                  uint32_t* target_width, uint32_t* target_height);
/* Automatic scale of a plan (find_scale, crates/optics/src/plan.rs:521-559):
 * the first of 1, 1.025, 1.025^2 ... (160 tries) at which 256 perimeter
 * probes map inside the cubic footprint; the desc's own `scale` is ignored.
 * RHD_INVALID_INPUT when none does. Host only. */
rhd_status rhd_optics_find_scale(const rhd_optics_desc* optics, float* scale);
/* GpuSpatialProcessor::reduce_preview (crates/gpu/src/spatial/reduction.rs:
 * 15): optics (NULL = off) then exact area reduction to target_width x
 * target_height, into a new source; `source` is not modified. Equal
 * dimensions skip the reduction (resample.rs:253-255). An invalid vignetting
 * gain, a non-finite value or a cubic footprint outside the image gives
 * RHD_INVALID_INPUT and no output. `gpu_ms` (nullable): device time. */
rhd_status rhd_reduce_preview(rhd_context* ctx, const rhd_source* source,
                              const rhd_optics_desc* optics, uint32_t target_width,
                              uint32_t target_height, const volatile uint8_t* cancel,
                              rhd_source** out, float* gpu_ms);
/* GpuSpatialProcessor::full_resolution_source (crates/gpu/src/spatial/
 * mod.rs:200): the corrected image at the source's size. */
rhd_status rhd_optics_full(rhd_context* ctx, const rhd_source* source,
                           const rhd_optics_desc* optics, const volatile uint8_t* cancel,
                           rhd_source** out, float* gpu_ms);

/* White balance then the camera matrix to linear Rec.2020, in place on a
 * resident camera source: apply_white_balance_cancellable and
 * apply_camera_color_transform_cancellable (crates/core/src/cpu/stages.rs:
 * 250-308), run after optics (crates/core/src/pipeline/orchestration.rs:
 * 1464-1470). `gains` are the resolved WhiteBalanceGains (finite, positive);
 * `camera_to_rec2020` is CameraColorTransform::camera_to_linear_rec2020,
 * row-major. A DCP profile look (ProfileLook) is not applied: a caller whose
 * transform has one routes the step to the CPU. */
rhd_status rhd_color_transform(rhd_context* ctx, rhd_source* source, const float gains[3],
                               const float camera_to_rec2020[9], const volatile uint8_t* cancel,
                               float* gpu_ms);

/* ---------------------------------------------------------------------------
 * Serving loop: the counterpart of the engine worker's develop loop
 * (crates/engine/src/coordinator/worker.rs:723-800) with its newest-wins
 * mailbox (crates/engine/src/coordinator/scheduler.rs:44-173) and the
 * resident caches of SpatialCache (crates/engine/src/preview_cache/
 * spatial.rs:322-611), as continuous batching on one GPU. Requests name a
 * frame made resident with rhd_server_put_mosaic; up to `lanes` run at once,
 * each on its own HIP stream, admitted in arrival order while the device
 * bytes of the running ones stay within `admission_bytes`. A newer request
 * for a frame coalesces that frame's queued requests (they complete as
 * Cancelled) and cancels its running one at the next stage boundary.
 * Demosaiced bases (camera RGB before capture sharpening) are cached per
 * frame and sensor recipe, so a re-develop with the same sensor settings only
 * runs capture sharpening. The server holds a reference on its context and
 * uses lanes 1..lanes and 7; do not call the staged API on the same context
 * from other threads meanwhile.
 * ------------------------------------------------------------------------- */
typedef struct rhd_server rhd_server;

typedef struct rhd_server_config {
    uint32_t lanes;            /* concurrent develops, 1..6; 0 selects 3 */
    uint64_t admission_bytes;  /* device bytes running requests may hold; 0 = half the budget */
    uint64_t base_cache_bytes; /* device bytes of cached demosaiced bases; 0 disables */
} rhd_server_config;

typedef struct rhd_request_info {
    int32_t status;            /* rhd_status of the request */
    uint32_t lane;             /* lane it ran on (0 if it never ran) */
    int32_t base_cache_hit;    /* 1 if it reused a cached demosaiced base */
    double queued_ms, run_ms;  /* submit -> start, start -> RGB ready */
    double latency_ms;         /* submit -> RGB ready on the device */
    double megapixels;
    rhd_metrics metrics;       /* stage metrics (zero for a base hit) */
} rhd_request_info;

typedef struct rhd_server_stats {
    uint64_t submitted, completed, cancelled, coalesced, failed, base_hits;
    double p50_ms, p99_ms, mean_ms; /* latency of completed requests */
    double megapixels;              /* sensor megapixels of completed requests */
    double span_ms;                 /* first submit -> last completion */
    uint64_t base_cache_bytes, running_bytes;
} rhd_server_stats;

rhd_status rhd_server_create(rhd_context* ctx, const rhd_server_config* config, rhd_server** out);
/* Cancels queued requests, waits for running ones, frees what the server holds. */
void rhd_server_destroy(rhd_server* server);
/* Makes a decoded frame resident under `frame_id` (replacing a previous one). */
rhd_status rhd_server_put_mosaic(rhd_server* server, uint64_t frame_id, const uint16_t* host,
                                 uint64_t row_stride, uint32_t width, uint32_t height);
/* The same from an ARW file: the raw strip found by rhd_arw_probe is uploaded
 * and decoded on the device (rhd_mosaic_decode_arw). */
rhd_status rhd_server_put_arw(rhd_server* server, uint64_t frame_id, const uint8_t* file,
                              uint64_t size, const rhd_arw_layout* layout);
rhd_status rhd_server_drop_mosaic(rhd_server* server, uint64_t frame_id);
/* Queues a develop of a resident frame; the descriptions are copied. */
rhd_status rhd_server_submit(rhd_server* server, uint64_t frame_id, const rhd_sensor_desc* desc,
                             const rhd_detail_desc* detail, uint64_t* request_id);
/* The same, followed on the request's lane by the spatial completion
 * (rhd_reduce_preview: optics, NULL = off, then area reduction to
 * target_width x target_height), as the worker develops and then reduces a
 * preview (crates/engine/src/preview_cache/spatial.rs:475-611); the result
 * is the reduced source. */
rhd_status rhd_server_submit_spatial(rhd_server* server, uint64_t frame_id,
                                     const rhd_sensor_desc* desc, const rhd_detail_desc* detail,
                                     const rhd_optics_desc* optics, uint32_t target_width,
                                     uint32_t target_height, uint64_t* request_id);
/* Waits up to `timeout_ms` (< 0: forever) for a request. Returns
 * Synchronization if it is still pending (it stays collectable); otherwise
 * its status, and on success hands over the developed source (or releases
 * it when `out` is NULL). A request is collected once. */
rhd_status rhd_server_wait(rhd_server* server, uint64_t request_id, int32_t timeout_ms,
                           rhd_source** out, rhd_request_info* info);
rhd_status rhd_server_read_stats(rhd_server* server, rhd_server_stats* out);
rhd_status rhd_server_reset_stats(rhd_server* server);

/* ---------------------------------------------------------------------------
 * Batched develop: many resident mosaics of uneven sizes and recipes in one
 * HIP graph launch. The reference develops one frame per call
 * (crates/gpu/src/sensor/develop.rs:68-142) for one request at a time
 * (crates/engine/src/coordinator/worker.rs:723-800); a batch captures the
 * stage of all its frames once, as `lanes` independent branches of one graph,
 * and rhd_batch_run replays it. `details` has `count` entries
 * (capture_enabled = 0 for none); the descriptions are copied; the mosaics
 * must stay alive while the batch exists. `lanes` = 0 lets the autotuner
 * choose (every candidate is captured and timed on the batch; the choice is
 * remembered per batch shape). The batch uses lanes 0..6 of its context; do
 * not run other work on the context meanwhile.
 * ------------------------------------------------------------------------- */
typedef struct rhd_batch rhd_batch;

rhd_status rhd_batch_create(rhd_context* ctx, uint32_t count, const rhd_mosaic* const* mosaics,
                            const rhd_sensor_desc* descs, const rhd_detail_desc* details,
                            uint32_t lanes, rhd_batch** out);
/* Launches the graph once and waits. `statuses`, `metrics` (statistics only)
 * and `gpu_ms` (device time of the launch) may be NULL. Returns the first
 * failing frame's status, or Ok. */
rhd_status rhd_batch_run(rhd_batch* batch, int32_t* statuses, rhd_metrics* metrics, float* gpu_ms);
/* The i-th frame's RGB source, owned by the batch and rewritten by each run. */
const rhd_source* rhd_batch_source(const rhd_batch* batch, uint32_t index);
uint32_t rhd_batch_lanes(const rhd_batch* batch);
void rhd_batch_destroy(rhd_batch* batch);

#ifdef __cplusplus
}
#endif

#endif /* RHD_H */
