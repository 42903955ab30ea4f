// RCD demosaic on gfx950.
//
// Reference: crates/demosaic/src/rcd.rs:36-299 (tiles, populate, write) and
// crates/demosaic/src/rcd_stages.rs (stages), on the tile grid of
// crates/demosaic/src/rcd_geometry.rs (194-pixel tiles, 10-pixel mirrored
// halo, 174-pixel cores). Bit-identical to the reference: the same per-tile
// planes and index arithmetic (including the half-plane low-pass reads two
// rows apart, rcd_stages.rs:107-111), and `interpolate` as a fused
// multiply-add (rcd_stages.rs:455-456).
//
// Every stage is a per-site stencil with no in-place chain, so each runs as
// one grid kernel over (tile, row, column) with the tile planes in HBM.
#include "rhd_internal.hpp"

namespace rhd {
namespace {

constexpr int TS = 194, BORDER = 10, STEP = TS - 2 * BORDER;
constexpr int FULL = TS * TS, HALF = FULL / 2;
enum { CFA, RGB0, RGB1, RGB2, VH, NUM_FULL };
enum { PQ, PD, QD, NUM_HALF };
constexpr size_t TILE_FLOATS = (size_t)NUM_FULL * FULL + (size_t)NUM_HALF * HALF;
constexpr float EPS = 1.0e-5f, EPS_SQ = 1.0e-10f;

struct Args {
    const float* mosaic;
    int width, height, pattern, tiles_x, tile_base;
    float forward[3], inverse[3];
    float* scratch;
    float* out;
    DeviceStats* stats;
};

struct Tile {
    float* base;
    int index;
    __device__ __forceinline__ float* full(int p) const { return base + (size_t)p * FULL; }
    __device__ __forceinline__ float* half(int p) const {
        return base + (size_t)NUM_FULL * FULL + (size_t)p * HALF;
    }
    __device__ __forceinline__ float* rgb(int c) const { return full(RGB0 + c); }
};
__device__ __forceinline__ Tile tile_of(const Args& a, int z) {
    return Tile{a.scratch + (size_t)z * TILE_FLOATS, a.tile_base + z};
}

__device__ __forceinline__ float square(float v) { return v * v; }
__device__ __forceinline__ float vhp(const float* v, int i) {
    return (v[i - 3 * TS] - v[i - TS] - v[i + TS] + v[i + 3 * TS]) -
           3.0f * (v[i - 2 * TS] + v[i + 2 * TS]) + 6.0f * v[i];
}
__device__ __forceinline__ float hhp(const float* v, int i) {
    return (v[i - 3] - v[i - 1] - v[i + 1] + v[i + 3]) - 3.0f * (v[i - 2] + v[i + 2]) + 6.0f * v[i];
}
__device__ __forceinline__ float cardinal_gradient(const float* v, int i, int dx, int dy) {
    int d = dy * TS + dx;
    return EPS + fabsf(v[i + d] - v[i - d]) + fabsf(v[i] - v[i + 2 * d]) +
           fabsf(v[i + d] - v[i + 3 * d]) + fabsf(v[i + 2 * d] - v[i + 4 * d]);
}
__device__ __forceinline__ float diagonal_gradient(const float* v, const float* g, int i, int dx,
                                                   int dy) {
    int d = dy * TS + dx;
    return EPS + fabsf(v[i + d] - v[i - d]) + fabsf(v[i + d] - v[i + 3 * d]) +
           fabsf(g[i] - g[i + 2 * d]);
}
__device__ __forceinline__ float nonzero_denominator(float v) {
    if (fabsf(v) >= EPS) return v;
    return signbit(v) ? -EPS : EPS;
}
__device__ __forceinline__ float ratio_estimate(float sample, float lp, float nlp) {
    float den = nonzero_denominator(EPS + lp + nlp);
    return sample * (2.0f * lp) / den;
}
__device__ __forceinline__ float weighted_pair(float w1, float v1, float w2, float v2) {
    return (w1 * v1 + w2 * v2) / nonzero_denominator(w1 + w2);
}
__device__ __forceinline__ float refined_direction(float c, float nw, float ne, float sw, float se) {
    float hood = 0.25f * (nw + ne + sw + se);
    float sel = fabsf(0.5f - c) < fabsf(0.5f - hood) ? hood : c;
    return sel < 0.0f ? 0.0f : (sel > 1.0f ? 1.0f : sel);
}
__device__ __forceinline__ float interpolate(float d, float h, float v) {
    return __builtin_fmaf(d, h, (1.0f - d) * v);
}

#define RCD_SITE                                                   \
    const uint3 blk = xcd_block();                                 \
    const Tile t = tile_of(a, blk.z);                              \
    const int col = blk.x * 64 + (threadIdx.x & 63);               \
    const int row = blk.y * 4 + (threadIdx.x >> 6);                \
    if (col >= TS || row >= TS) return;                            \
    const int i = row * TS + col;                                  \
    const bool rb = (color_at(a.pattern, col, row) & 1) == 0;

// rcd.rs:210-257 (populate) with the zeroed scratch of rcd.rs:196-208.
__global__ __launch_bounds__(256) void rcd_load(Args a) {
    RCD_SITE
    const int tx = t.index % a.tiles_x, ty = t.index / a.tiles_x;
    const int sx = mirror_whole_near(tx * STEP - BORDER + col, a.width);
    const int sy = mirror_whole_near(ty * STEP - BORDER + row, a.height);
    const int c = color_at(a.pattern, col, row);
    const float v = a.mosaic[(size_t)sy * a.width + sx] * a.forward[c];
    t.full(CFA)[i] = v;
    const int c0 = color_at(a.pattern, 0, row), c1 = color_at(a.pattern, 1, row);
    for (int k = 0; k < 3; ++k) t.rgb(k)[i] = (k == c0 || k == c1) ? v : 0.0f;
    t.full(VH)[i] = 0.0f;
    if ((col & 1) == 0) {
        t.half(PQ)[i >> 1] = 0.0f;
        t.half(PD)[i >> 1] = 0.0f;
        t.half(QD)[i >> 1] = 0.0f;
    }
}

// rcd_stages.rs:8-54 (find_directions), :56-83 (calculate_low_pass) and
// :187-221 (calculate_diagonal_high_pass): all read only the CFA.
__global__ __launch_bounds__(256) void rcd_stats(Args a) {
    RCD_SITE
    const float* cfa = t.full(CFA);
    if (row >= 4 && row < TS - 4 && col >= 4 && col < TS - 4) {
        float vs = num_max(square(vhp(cfa, i - TS)) + square(vhp(cfa, i)) + square(vhp(cfa, i + TS)),
                           EPS_SQ);
        float hs = num_max(square(hhp(cfa, i - 1)) + square(hhp(cfa, i)) + square(hhp(cfa, i + 1)),
                           EPS_SQ);
        t.full(VH)[i] = vs / (vs + hs);
    }
    if (rb && row >= 2 && row < TS - 2 && col >= 2 && col < TS - 2) {
        t.half(PQ)[i >> 1] = cfa[i] + 0.5f * (cfa[i - TS] + cfa[i + TS] + cfa[i - 1] + cfa[i + 1]) +
                             0.25f * (cfa[i - TS - 1] + cfa[i - TS + 1] + cfa[i + TS - 1] +
                                      cfa[i + TS + 1]);
    }
    if ((col & 1) == 1 && row >= 3 && row < TS - 3 && col >= 3 && col < TS - 3) {
        t.half(PD)[i >> 1] =
            square((cfa[i - 3 * TS - 3] - cfa[i - TS - 1] - cfa[i + TS + 1] + cfa[i + 3 * TS + 3]) -
                   3.0f * (cfa[i - 2 * TS - 2] + cfa[i + 2 * TS + 2]) + 6.0f * cfa[i]);
        t.half(QD)[i >> 1] =
            square((cfa[i - 3 * TS + 3] - cfa[i - TS + 1] - cfa[i + TS - 1] + cfa[i + 3 * TS - 3]) -
                   3.0f * (cfa[i - 2 * TS + 2] + cfa[i + 2 * TS - 2]) + 6.0f * cfa[i]);
    }
}

// rcd_stages.rs:85-151 (interpolate_green)
__global__ __launch_bounds__(256) void rcd_green(Args a) {
    RCD_SITE
    if (!rb || row < 4 || row >= TS - 4 || col < 4 || col >= TS - 4) return;
    const float* c = t.full(CFA);
    const float* pq = t.half(PQ);
    const float* vh = t.full(VH);
    float ng = cardinal_gradient(c, i, 0, -1), sg = cardinal_gradient(c, i, 0, 1);
    float wg = cardinal_gradient(c, i, -1, 0), eg = cardinal_gradient(c, i, 1, 0);
    int li = i / 2;
    float lp = pq[li];
    float n_est = ratio_estimate(c[i - TS], lp, pq[li - TS]);
    float s_est = ratio_estimate(c[i + TS], lp, pq[li + TS]);
    float w_est = ratio_estimate(c[i - 1], lp, pq[li - 1]);
    float e_est = ratio_estimate(c[i + 1], lp, pq[li + 1]);
    float vert = weighted_pair(sg, n_est, ng, s_est);
    float horiz = weighted_pair(wg, e_est, eg, w_est);
    float d = refined_direction(vh[i], vh[i - TS - 1], vh[i - TS + 1], vh[i + TS - 1], vh[i + TS + 1]);
    t.rgb(1)[i] = interpolate(d, horiz, vert);
}

// rcd_stages.rs:223-251 (calculate_diagonal_directions)
__global__ __launch_bounds__(256) void rcd_directions(Args a) {
    RCD_SITE
    if (!rb || row < 4 || row >= TS - 4 || col < 4 || col >= TS - 4) return;
    const float* p = t.half(PD);
    const float* q = t.half(QD);
    int h = i / 2, prev = (i - TS - 1) / 2, next = (i + TS - 1) / 2;
    float ps = num_max(p[prev] + p[h] + p[next + 1], EPS_SQ);
    float qs = num_max(q[prev + 1] + q[h] + q[next], EPS_SQ);
    t.half(PQ)[h] = ps / (ps + qs);
}

// rcd_stages.rs:253-312 (interpolate_opposite_color)
__global__ __launch_bounds__(256) void rcd_opposite(Args a) {
    RCD_SITE
    if (!rb || row < 4 || row >= TS - 4 || col < 4 || col >= TS - 4) return;
    const int ch = 2 - color_at(a.pattern, col, row);
    const float* pq = t.half(PQ);
    int h = i / 2, prev = (i - TS - 1) / 2, next = (i + TS - 1) / 2;
    float d = refined_direction(pq[h], pq[prev], pq[prev + 1], pq[next], pq[next + 1]);
    float* v = t.rgb(ch);
    const float* g = t.rgb(1);
    float nwg = diagonal_gradient(v, g, i, -1, -1), neg = diagonal_gradient(v, g, i, 1, -1);
    float swg = diagonal_gradient(v, g, i, -1, 1), seg = diagonal_gradient(v, g, i, 1, 1);
    float nwe = v[i - TS - 1] - g[i - TS - 1], nee = v[i - TS + 1] - g[i - TS + 1];
    float swe = v[i + TS - 1] - g[i + TS - 1], see = v[i + TS + 1] - g[i + TS + 1];
    float dp = weighted_pair(nwg, see, seg, nwe);
    float dq = weighted_pair(neg, swe, swg, nee);
    v[i] = g[i] + interpolate(d, dq, dp);
}

// rcd_stages.rs:314-393 (interpolate_at_green)
__global__ __launch_bounds__(256) void rcd_at_green(Args a) {
    RCD_SITE
    if (rb || row < 4 || row >= TS - 4 || col < 4 || col >= TS - 4) return;
    const float* g = t.rgb(1);
    const float* vh = t.full(VH);
    float d = refined_direction(vh[i], vh[i - TS - 1], vh[i - TS + 1], vh[i + TS - 1], vh[i + TS + 1]);
    float green = g[i];
    float n1 = EPS + fabsf(green - g[i - 2 * TS]);
    float s1 = EPS + fabsf(green - g[i + 2 * TS]);
    float w1 = EPS + fabsf(green - g[i - 2]);
    float e1 = EPS + fabsf(green - g[i + 2]);
    float gn = g[i - TS], gs = g[i + TS], gw = g[i - 1], ge = g[i + 1];
    for (int ch = 0; ch <= 2; ch += 2) {
        float* v = t.rgb(ch);
        float ns = fabsf(v[i - TS] - v[i + TS]);
        float ew = fabsf(v[i - 1] - v[i + 1]);
        float ng = n1 + ns + fabsf(v[i - TS] - v[i - 3 * TS]);
        float sg = s1 + ns + fabsf(v[i + TS] - v[i + 3 * TS]);
        float wg = w1 + ew + fabsf(v[i - 1] - v[i - 3]);
        float eg = e1 + ew + fabsf(v[i + 1] - v[i + 3]);
        float ne = v[i - TS] - gn, se = v[i + TS] - gs, we = v[i - 1] - gw, ee = v[i + 1] - ge;
        float vert = weighted_pair(ng, se, sg, ne);
        float horiz = weighted_pair(eg, we, wg, ee);
        v[i] = green + interpolate(d, horiz, vert);
    }
}

// rcd.rs:261-299 (write_tile): measured sample restored, inverse balance.
__global__ __launch_bounds__(256) void rcd_write(Args a) {
    RCD_SITE
    (void)rb;
    if (row < BORDER || row >= TS - BORDER || col < BORDER || col >= TS - BORDER) return;
    const int tx = t.index % a.tiles_x, ty = t.index / a.tiles_x;
    const int gx = tx * STEP + col - BORDER, gy = ty * STEP + row - BORDER;
    if (gx >= a.width || gy >= a.height) return;
    float rgb[3] = {t.rgb(0)[i], t.rgb(1)[i], t.rgb(2)[i]};
    rgb[color_at(a.pattern, col, row)] = t.full(CFA)[i];
    for (int c = 0; c < 3; ++c) rgb[c] *= a.inverse[c];
    const size_t n = (size_t)a.width * a.height, o = (size_t)gy * a.width + gx;
    a.out[o] = rgb[0];
    a.out[n + o] = rgb[1];
    a.out[2 * n + o] = rgb[2];
    if (!(is_finite(rgb[0]) && is_finite(rgb[1]) && is_finite(rgb[2])))
        atomicOr(&a.stats->nonfinite, 2u);
}

}  // namespace

size_t rcd_scratch_bytes_per_tile() { return TILE_FLOATS * sizeof(float); }

hipError_t launch_rcd(const float* mosaic, int width, int height, int pattern,
                      const float forward[3], const float inverse[3], float* scratch,
                      int tile_base, int tiles, float* out, DeviceStats* stats, hipStream_t s,
                      Profiler* prof) {
    Args a{};
    a.mosaic = mosaic;
    a.width = width;
    a.height = height;
    a.pattern = pattern;
    a.tiles_x = (width + STEP - 1) / STEP;
    a.tile_base = tile_base;
    for (int c = 0; c < 3; ++c) {
        a.forward[c] = forward[c];
        a.inverse[c] = inverse[c];
    }
    a.scratch = scratch;
    a.out = out;
    a.stats = stats;
    const dim3 grid((TS + 63) / 64, (TS + 3) / 4, tiles), blk(256);
    const uint64_t f = 4ull * FULL * tiles, h = 4ull * HALF * tiles;
    RHD_TIMED(prof, s, "rcd_load", f + 4 * f + 3 * h, hipLaunchKernelGGL(rcd_load, grid, blk, 0, s, a));
    RHD_TIMED(prof, s, "rcd_stats", f + f + 3 * h, hipLaunchKernelGGL(rcd_stats, grid, blk, 0, s, a));
    RHD_TIMED(prof, s, "rcd_green", f + f + h + f / 2, hipLaunchKernelGGL(rcd_green, grid, blk, 0, s, a));
    RHD_TIMED(prof, s, "rcd_directions", 3 * h, hipLaunchKernelGGL(rcd_directions, grid, blk, 0, s, a));
    RHD_TIMED(prof, s, "rcd_opposite", h + 2 * f, hipLaunchKernelGGL(rcd_opposite, grid, blk, 0, s, a));
    RHD_TIMED(prof, s, "rcd_at_green", 4 * f, hipLaunchKernelGGL(rcd_at_green, grid, blk, 0, s, a));
    RHD_TIMED(prof, s, "rcd_write", 4 * f, hipLaunchKernelGGL(rcd_write, grid, blk, 0, s, a));
    return hipGetLastError();
}

}  // namespace rhd
