// Device helpers shared by the rhd kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "capture_math.hpp"

namespace rhd {

// Bayer colour at (x, y) for a pattern code (0 RGGB, 1 BGGR, 2 GRBG, 3 GBRG):
// channel 0 = red, 1 = green, 2 = blue. One byte per pattern, two bits per
// 2x2 site, indexed by ((y & 1) << 1) | (x & 1) (crates/image/src/lib.rs:90-117).
__host__ __device__ __forceinline__ int color_at(int pattern, int x, int y) {
    const uint32_t table = 0x49611694u;
    uint32_t byte = (table >> (8 * pattern)) & 0xffu;
    return (int)((byte >> (2 * (((y & 1) << 1) | (x & 1)))) & 3u);
}

// Whole-sample reflection about the first and last sample, parity preserving
// (crates/demosaic/src/amaze_geometry.rs:72-80).
__host__ __device__ __forceinline__ int mirror_whole(int index, int length) {
    if (length <= 1) return 0;
    int period = 2 * (length - 1);
    int wrapped = index % period;
    if (wrapped < 0) wrapped += period;
    int other = period - wrapped;
    return wrapped < other ? wrapped : other;
}

// Half-sample reflection about the outer pixel boundary
// (crates/core/src/sharpening.rs:318-326).
__host__ __device__ __forceinline__ int mirror_half(int index, int length) {
    int period = 2 * length;
    int folded = index % period;
    if (folded < 0) folded += period;
    return folded < length ? folded : 2 * length - 1 - folded;
}

// mirror_half for stencil taps: inside the image a compare, one reflection
// for taps up to one length outside, the modulo form only beyond that
// (taps farther apart than the image is wide).
__host__ __device__ __forceinline__ int mirror_near(int index, int length) {
    if (index >= 0 && index < length) return index;
    if (index < 0 && index >= -length) return -index - 1;
    if (index >= length && index < 2 * length) return 2 * length - 1 - index;
    return mirror_half(index, length);
}

// mirror_whole for tile loads: inside the image a compare, one reflection for
// indices up to length - 1 outside, the modulo form only beyond that.
__host__ __device__ __forceinline__ int mirror_whole_near(int index, int length) {
    if (index >= 0 && index < length) return index;
    if (index < 0 && index > -length) return -index;
    if (index >= length && index < 2 * length - 1) return 2 * (length - 1) - index;
    return mirror_whole(index, length);
}

// Compare-select min/max (crates/demosaic/src/amaze/math.rs:95-105).
__device__ __forceinline__ float sel_min(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float sel_max(float a, float b) { return a > b ? a : b; }
// IEEE maxNum/minNum, as Rust's f32::max / f32::min.
__device__ __forceinline__ float num_max(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ float num_min(float a, float b) { return fminf(a, b); }

__device__ __forceinline__ bool is_finite(float v) { return __builtin_isfinite(v); }

// Sum over the 64 lanes of a wavefront.
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
    for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, 64);
    return v;
}
__device__ __forceinline__ unsigned wave_sum(unsigned v) {
    for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, 64);
    return v;
}

// MI355X dispatches workgroups to its 8 XCDs round-robin by linear workgroup
// id, and each XCD has its own L2. xcd_block() renumbers the grid so that XCD
// k walks the k-th contiguous eighth of it in order: blocks that share
// stencil halos (the rows of one tile, neighbouring tiles) then run on the
// same XCD and find the halo in its L2. A bijection on the grid; the last
// (count mod 8) blocks keep their own id.
constexpr unsigned XCDS = 8;
__device__ __forceinline__ uint3 xcd_block() {
    const unsigned gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
    const unsigned per = gx * gy * gz / XCDS;
    const unsigned b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned v = b < per * XCDS ? (b % XCDS) * per + b / XCDS : b;
    return make_uint3(v % gx, (v / gx) % gy, v / (gx * gy));
}

// Device-side counters of one develop call; read back once at the end.
struct DeviceStats {
    unsigned long long clip[6];     // affected, changed, over-white, affected by channel
    unsigned long long detect[6];   // suspected, 0, 0, suspected by channel
    unsigned long long opposed[6];  // reconstructed, changed, fallback, fully unsupported,
                                    // neutralised, 0
    unsigned int nonfinite;         // bit 0: highlight input, bit 1: demosaic output,
                                    // bit 2: capture input
    unsigned int regions;
    unsigned int capture_range;     // nonzero: a capture guide value is >= 2^64 or not finite
    unsigned int pad;
};

}  // namespace rhd
