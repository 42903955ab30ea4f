// Host-side internals of librhd: handles, the context, and the launchers
// each kernel file exports to the C ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <string>
#include <vector>

#include "../../include/rhd.h"
#include "optics_math.hpp"
#include "rhd_device.hpp"

namespace rhd {

// Optional per-launch timing: HIP events recorded on the launch stream around
// one kernel, with the kernel's algorithmic bytes per launch. Launchers take a
// nullable pointer; the C ABI enables it for benchmarking only.
class Profiler {
   public:
    void begin(hipStream_t s, const char* name, uint64_t bytes);
    void end(hipStream_t s);
    // Resolves recorded launches into the per-name totals (synchronizes).
    void collect();
    void reset();
    ~Profiler();

    // When set, only launches of this kernel group are timed.
    std::string only;

    struct Total {
        std::string name;
        double ms = 0.0;
        uint64_t launches = 0;
        uint64_t bytes = 0;  // algorithmic bytes of the latest launch
    };
    std::vector<Total> totals;

   private:
    struct Pending {
        size_t total;
        hipEvent_t start, stop;
        uint64_t bytes;
    };
    hipEvent_t take_event();
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> free_;
    hipEvent_t open_ = nullptr;
    bool skipped_ = false;
    size_t open_total_ = 0;
    uint64_t open_bytes_ = 0;
};

#define RHD_TIMED(prof, stream, name, bytes, ...)          \
    do {                                                   \
        if (prof) (prof)->begin((stream), (name), (bytes)); \
        __VA_ARGS__;                                       \
        if (prof) (prof)->end((stream));                   \
    } while (0)

// Normalization level table: black level and (white - black) per position in
// the combined repeat period of the black pattern and the sensor CFA.
struct LevelTable {
    int period_x, period_y;
    float black[256];
    float range[256];
};

struct NormalizeArgs {
    const uint16_t* mosaic;
    uint64_t row_stride;
    int crop_x, crop_y, width, height;
    int crop_pattern;
    int mode;             // 0 = off, 1 = clip, 2 = opposed detection counts
    float levels[3];      // clip limits or detection levels
    LevelTable table;
};

// Geometry and parameters shared by the AMaZE kernels.
struct AmazeArgs {
    const float* mosaic;  // normalized, highlighted mosaic (width * height)
    int width, height, pattern;
    int tiles_x, tiles_y, tile_base;  // tile_base: first tile index of this launch
    float forward[3];     // channel balance applied while loading
    float inverse[3];     // gains applied to the output
    float clip_point;
    float* scratch;       // per-tile planes
    int* areas;           // per tile: start_row, end_row, start_col, end_col
    float* out;           // planar R, G, B (3 * width * height)
    DeviceStats* stats;
    // When set, capture sharpening's guide and mask planes (width * height
    // each) are written from the output as it is stored, with the channel
    // ceilings below (each in [2^-20, 2^20]) and their reciprocals; capture
    // then starts from them.
    float* guide;
    float* mask;
    float ceilings[3];
    float ceilings_inv[3];
};

// Launchers (each returns the first HIP error, or hipSuccess).
hipError_t launch_normalize(const NormalizeArgs& a, float* out, DeviceStats* stats,
                            hipStream_t s, Profiler* prof);
// Zeroes `bytes` (a multiple of 4) of device memory with a kernel, which a
// HIP graph may capture.
hipError_t launch_zero(void* p, size_t bytes, hipStream_t s);
hipError_t launch_clip(float* plane, int width, int height, int pattern, const float limits[3],
                       DeviceStats* stats, hipStream_t s);
hipError_t launch_detect(const float* plane, int width, int height, int pattern,
                         const float levels[3], DeviceStats* stats, hipStream_t s);
hipError_t launch_bilinear(const float* mosaic, int width, int height, int pattern,
                           const float forward[3], const float gains[3], float* out,
                           DeviceStats* stats, hipStream_t s, Profiler* prof);
hipError_t launch_mhc(const float* mosaic, int width, int height, int pattern, float* out,
                      DeviceStats* stats, hipStream_t s, Profiler* prof);
size_t amaze_scratch_bytes_per_tile();
hipError_t launch_amaze(const AmazeArgs& a, int tiles, hipStream_t s, Profiler* prof);
size_t rcd_scratch_bytes_per_tile();
hipError_t launch_rcd(const float* mosaic, int width, int height, int pattern,
                      const float forward[3], const float inverse[3], float* scratch,
                      int tile_base, int tiles, float* out, DeviceStats* stats, hipStream_t s,
                      Profiler* prof);

struct CaptureArgs {
    float* rgb;  // planar, 3 * n
    int width, height;
    float amount, noise_protection;
    float ceilings[3];
    int radius;
    float weights[9];
    int iterations;
    float* guide;
    float* mask;
    float* est_a;
    float* est_b;
    float* tmp;
    DeviceStats* stats;
    bool inputs_ready;  // guide and mask already hold this image's values
};
size_t capture_scratch_bytes(int width, int height);
// False-colour suppression in place on planar RGB; `scratch` holds 4 planes.
hipError_t launch_false_colour(float* rgb, int width, int height, int steps,
                               const float balance[3], const float ceilings[3], float* scratch,
                               hipStream_t s, Profiler* prof);
hipError_t launch_capture(const CaptureArgs& a, hipStream_t s, Profiler* prof);

// Noise profile of a width x height image: the sampled rows (every fourth
// 64-row band from 1024 rows on), the 16x16 tiles over them, the brightness
// bins, the padded sort size, and the byte offsets of the estimate's scratch.
struct NoiseGeometry {
    int rows, tiles_x, bins;
    int64_t tiles, sort_n, flat_tiles, hist_blocks;
    size_t band0, tmp, energy, bright, keys, keys2, points, state, hist, samples, bytes;
};
constexpr int64_t MAX_NOISE_TILES = 1 << 24;
NoiseGeometry noise_geometry(int width, int height);
// Scratch of the chroma filter (7 planes; the luma filter needs 4).
size_t noise_filter_bytes(int width, int height);
// Writes prof = {a, b} of luma, C1, C2, then their bin counts (9 floats),
// computed entirely on the device. Needs noise_geometry(...).tiles in
// 1..MAX_NOISE_TILES.
hipError_t launch_noise_profile(const float* rgb, int width, int height, const float balance[3],
                                void* scratch, float* prof, hipStream_t s, Profiler* pr);
hipError_t launch_chroma_nr(float* rgb, int width, int height, float strength,
                            const float balance[3], const float ceilings[3], const float* prof,
                            float* scratch, hipStream_t s, Profiler* pr);
hipError_t launch_luma_nr(float* rgb, int width, int height, float strength, float detail,
                          const float balance[3], const float ceilings[3], const float* prof,
                          float* scratch, hipStream_t s, Profiler* pr);

struct OpposedArgs {
    float* mosaic;
    int width, height, pattern;
    float levels[3];
    int has_ceiling;
    float ceiling[3];
    void* scratch;
    size_t scratch_bytes;
    DeviceStats* stats;
    const uint32_t* log2_bitmap;  // libm_f32.hpp exception bitmap on this device
    // Optional second stream (with two events) for the work that depends
    // only on the cells; null runs everything on the launch stream.
    hipStream_t aux;
    hipEvent_t fork, join;
};
size_t opposed_scratch_bytes(int width, int height);
hipError_t launch_opposed(const OpposedArgs& a, hipStream_t s, Profiler* prof);

// One optics + area-reduction launch sequence (optics.hip). Planes are
// planar f32 R, G, B of width x height; tables are device copies of what the
// host built for the plan.
struct OpticsLaunch {
    const float* src;
    float* vignetted;            // 3 planes; needed when vignetting precedes a remap
    int width, height;
    bool vignetting, remap, direct;
    bool rows_ready;             // vig_rows already holds this plan's checkpoints
    const float* vig_steps;      // width floats
    float* vig_rows;             // optics_checkpoints(width) * height floats
    float vig_start_x, vig_start_y, vig_ns;
    float vig_k[3];
    optics::Model model;
    const float* in_x;           // width floats
    const float* in_y;           // height floats
    const uint32_t *h_first, *h_offset, *h_count;
    const float* h_weights;
    const uint32_t *v_first, *v_offset, *v_count;
    const float* v_weights;
    int target_width, target_height;
    int group;                   // output columns per horizontal workgroup (1..256)
    float* mid;                  // 3 planes of target_width x height
    float* out;                  // 3 planes of the output
    DeviceStats* stats;
};
size_t optics_checkpoints(int width);
// White balance then the camera matrix, in place on 3 planes of n pixels.
hipError_t launch_color(float* rgb, size_t n, const float gains[3], const float m[9],
                        hipStream_t s, Profiler* prof);
hipError_t launch_optics(const OpticsLaunch& a, hipStream_t s, Profiler* prof);

}  // namespace rhd
