// Sony compressed ARW (TIFF compression 32767, 8 bits per pixel) decoded on
// the device into the resident u16 mosaic the stage reads. The reference
// decodes on the CPU with rawler 0.7.2 (crates/raw/src/rawler_adapter.rs:
// 131-170); the format:
//   * `width` bytes per row; every 32 pixels come from two 16-byte blocks
//     (even columns, then odd), each 128 bits read LSB first: an 11-bit max,
//     an 11-bit min, the 4-bit positions of both, then fourteen 7-bit deltas
//     scaled by 2^max(0, bit_length(max - min) - 7) over the min, clamped to
//     0x7ff;
//   * the sample is the tone curve at value * 2, dithered: base +
//     ((delta * (r & 2047) + 1024) >> 12), where r is a multiply-with-carry
//     stream r' = 15700 (r & 0xffff) + (r >> 16) seeded per row with its
//     first 16 bits and advanced once per pixel in block order.
// The stream is sequential along a row; the decoder jumps it instead: with
// m = 15700 * 2^16 - 1, every state after a seed below 2^16 stays below m and
// r_n = 15700^n r_0 mod m, so each 16-pixel block starts from
// jump[block] * r_0 mod m with jump[block] = 15700^(16 block) mod m from the
// host. One thread decodes one 32-pixel group (its two blocks) and writes it
// as 64 contiguous bytes; a row's tail past its last whole group is written
// as 0.
//
// Blocks are 128 bits only when the max and min positions differ: one whose
// positions coincide carries a 15th delta, and every later block of its row
// starts 7 bits further on (the row's stream then runs on into the following
// rows' bytes, zeros past the strip's end). Well-formed files never do this.
// The block decoder marks the first such block of each row; arw2_repair then
// decodes those rows from that block on, one thread per row, bit by bit.
#include "rhd_internal.hpp"

namespace rhd {
namespace {

constexpr uint64_t MWC_MODULUS = 15700ull * 65536ull - 1ull;

__device__ __forceinline__ uint32_t bits_at(uint64_t lo, uint64_t hi, int pos, int count) {
    const uint64_t word = pos < 64 ? (lo >> pos) | (pos ? hi << (64 - pos) : 0ull) : hi >> (pos - 64);
    return (uint32_t)(word & ((1ull << count) - 1ull));
}

__device__ __forceinline__ uint16_t dither(const uint32_t* lut, uint32_t val, uint32_t& rand) {
    const uint32_t e = lut[val << 1];
    const uint16_t px = (uint16_t)((e & 0xffffu) + (((e >> 16) * (rand & 2047u) + 1024u) >> 12));
    rand = 15700u * (rand & 65535u) + (rand >> 16);
    return px;
}

__global__ __launch_bounds__(256) void arw2_first_bad(uint32_t* __restrict__ first_bad,
                                                      int height) {
    const int row = blockIdx.x * 256 + threadIdx.x;
    if (row < height) first_bad[row] = 0xffffffffu;
}

// One 16-byte block at p (any alignment) as two little-endian words.
__device__ __forceinline__ void load_block(const uint8_t* p, uint64_t& lo, uint64_t& hi) {
    if (((uintptr_t)p & 15) == 0) {
        const uint4 v = *(const uint4*)p;
        lo = (uint64_t)v.x | (uint64_t)v.y << 32;
        hi = (uint64_t)v.z | (uint64_t)v.w << 32;
        return;
    }
    lo = hi = 0;
    for (int k = 0; k < 8; ++k) {
        lo |= (uint64_t)p[k] << (8 * k);
        hi |= (uint64_t)p[8 + k] << (8 * k);
    }
}

// The 16 samples of one aligned block, dithered from stream state `rand`.
__device__ __forceinline__ void decode_block(uint64_t lo, uint64_t hi, const uint32_t* lut,
                                             uint32_t rand, uint16_t px[16]) {
    const uint32_t max = bits_at(lo, hi, 0, 11), min = bits_at(lo, hi, 11, 11);
    const uint32_t imax = bits_at(lo, hi, 22, 4), imin = bits_at(lo, hi, 26, 4);
    const uint32_t delta = max - min;
    const int length = delta ? 32 - __builtin_clz(delta) : 0;
    const int shift = length > 7 ? length - 7 : 0;
    int pos = 30;
#pragma unroll
    for (uint32_t i = 0; i < 16; ++i) {
        uint32_t val;
        if (i == imax) {
            val = max;
        } else if (i == imin) {
            val = min;
        } else {
            val = (bits_at(lo, hi, pos, 7) << shift) + min;
            val = val > 0x7ffu ? 0x7ffu : val;
            pos += 7;
        }
        px[i] = dither(lut, val, rand);
    }
}

// One thread per 32-pixel group: its even-column block and its odd-column
// block, written back as 64 contiguous bytes.
__global__ __launch_bounds__(256) void arw2_decode(const uint8_t* __restrict__ strip, int width,
                                                   const uint32_t* __restrict__ lut,
                                                   const uint32_t* __restrict__ jump,
                                                   uint32_t* __restrict__ first_bad,
                                                   uint16_t* __restrict__ out) {
    const int groups = width / 32;
    const int g = blockIdx.x * 256 + threadIdx.x, row = blockIdx.y;
    const uint8_t* line = strip + (size_t)row * width;
    uint16_t* dst = out + (size_t)row * width;
    if (g == groups) {  // the tail past the last whole group
        for (int x = groups * 32; x < width; ++x) dst[x] = 0;
        return;
    }
    if (g > groups) return;
    uint64_t lo0, hi0, lo1, hi1;
    load_block(line + (size_t)g * 32, lo0, hi0);
    load_block(line + (size_t)g * 32 + 16, lo1, hi1);
    if (bits_at(lo0, hi0, 22, 4) == bits_at(lo0, hi0, 26, 4))
        atomicMin(&first_bad[row], (uint32_t)(2 * g));
    else if (bits_at(lo1, hi1, 22, 4) == bits_at(lo1, hi1, 26, 4))
        atomicMin(&first_bad[row], (uint32_t)(2 * g + 1));
    const uint32_t seed = (uint32_t)line[0] | (uint32_t)line[1] << 8;
    uint16_t even[16], odd[16];
    decode_block(lo0, hi0, lut, (uint32_t)(((uint64_t)jump[2 * g] * seed) % MWC_MODULUS), even);
    decode_block(lo1, hi1, lut, (uint32_t)(((uint64_t)jump[2 * g + 1] * seed) % MWC_MODULUS), odd);
    uint16_t* o = dst + (size_t)g * 32;
    if (((uintptr_t)o & 15) == 0) {
        uint32_t w[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) w[k] = (uint32_t)even[k] | (uint32_t)odd[k] << 16;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            ((uint4*)o)[q] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
    } else {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            o[2 * k] = even[k];
            o[2 * k + 1] = odd[k];
        }
    }
}

// Reads `count` bits LSB first at bit `pos` of the strip; zeros past its end.
__device__ __forceinline__ uint32_t stream_bits(const uint8_t* strip, uint64_t bytes, uint64_t pos,
                                                int count) {
    uint32_t v = 0;
    for (int k = 0; k < count; ++k) {
        const uint64_t b = pos + k;
        const uint32_t bit = (b >> 3) < bytes ? (strip[b >> 3] >> (b & 7)) & 1u : 0u;
        v |= bit << k;
    }
    return v;
}

// Rows with a 135-bit block: decoded sequentially from the first one.
__global__ __launch_bounds__(64) void arw2_repair(const uint8_t* __restrict__ strip,
                                                  uint64_t bytes, int width, int height,
                                                  const uint32_t* __restrict__ lut,
                                                  const uint32_t* __restrict__ jump,
                                                  const uint32_t* __restrict__ first_bad,
                                                  uint16_t* __restrict__ out) {
    const int row = blockIdx.x * 64 + threadIdx.x;
    if (row >= height) return;
    const uint32_t blocks = (uint32_t)(width / 32) * 2, start = first_bad[row];
    if (start >= blocks) return;
    const uint64_t row_at = (uint64_t)row * width;
    const uint32_t seed = stream_bits(strip, bytes, row_at * 8, 16);
    uint32_t rand = (uint32_t)(((uint64_t)jump[start] * seed) % MWC_MODULUS);
    uint64_t pos = row_at * 8 + (uint64_t)start * 128;
    uint16_t* dst = out + row_at;
    for (uint32_t b = start; b < blocks; ++b) {
        const uint32_t max = stream_bits(strip, bytes, pos, 11);
        const uint32_t min = stream_bits(strip, bytes, pos + 11, 11);
        const uint32_t imax = stream_bits(strip, bytes, pos + 22, 4);
        const uint32_t imin = stream_bits(strip, bytes, pos + 26, 4);
        pos += 30;
        const uint32_t delta = max - min;
        const int length = delta ? 32 - __builtin_clz(delta) : 0;
        const int shift = length > 7 ? length - 7 : 0;
        const int column = (int)(b >> 1) * 32 + (int)(b & 1);
        for (uint32_t i = 0; i < 16; ++i) {
            uint32_t val;
            if (i == imax) {
                val = max;
            } else if (i == imin) {
                val = min;
            } else {
                val = (stream_bits(strip, bytes, pos, 7) << shift) + min;
                val = val > 0x7ffu ? 0x7ffu : val;
                pos += 7;
            }
            dst[column + 2 * i] = dither(lut, val, rand);
        }
    }
}

}  // namespace

hipError_t launch_arw2(const uint8_t* strip, uint64_t bytes, int width, int height,
                       const uint32_t* lut, const uint32_t* jump, uint32_t* first_bad,
                       uint16_t* out, hipStream_t s, Profiler* prof) {
    const int groups = width / 32;
    const dim3 grid((groups + 1 + 255) / 256, height);
    arw2_first_bad<<<(height + 255) / 256, 256, 0, s>>>(first_bad, height);
    RHD_TIMED(prof, s, "arw2_decode", 3 * (uint64_t)width * height,
              arw2_decode<<<grid, 256, 0, s>>>(strip, width, lut, jump, first_bad, out));
    arw2_repair<<<(height + 63) / 64, 64, 0, s>>>(strip, bytes, width, height, lut, jump,
                                                  first_bad, out);
    return hipGetLastError();
}

}  // namespace rhd
