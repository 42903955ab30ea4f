// AMaZE demosaic on gfx950.
//
// Reference: crates/demosaic/src/amaze/{mod,math,green,nyquist,diagonal,
// chroma}.rs with the tile geometry of crates/demosaic/src/amaze_geometry.rs.
// The result is bit-identical to the reference: every stage keeps its loop
// bounds, plane addressing and operation order, and the tile grid (224-pixel
// cores, 16-pixel mirrored halo, 256-pixel tile planes) is kept because the
// Nyquist bounding box and the three in-place chains are per tile.
//
// Layout: every tile owns one scratch block in HBM (11 full planes of
// 256x256 f32, 14 half planes of 256x128 f32 indexed by pixel pair, 2 byte
// planes). Green is stored only at red/blue sites (the GRB half plane); at
// green sites it equals the CFA sample. Stages run as grid kernels over
// (tile, row, pixel or pair); the three sequential chains run as
//   - bound_v: one thread per tile column walking down the rows (vcd),
//   - bound_h: one thread per tile row walking right over LDS-staged chunks (hcd),
//   - chains:  one workgroup per tile, one thread per pair, a barrier per row
//              (hvwt in populate_green and pmwt in combine_diagonals).
// The finish kernel writes planar RGB for the tile core, with the measured
// sample restored and the inverse channel balance applied (mod.rs:196-273).
#include <climits>

#include "rhd_internal.hpp"

namespace rhd {
namespace {

constexpr int TS = 256, TSH = 128, BORDER = 16, VALID = TS - 2 * BORDER;
constexpr int FULL = TS * TS, HALF = TS * TSH;

enum FullPlane { CFA, DW0, DW1, DHV, VCD, HCD, VALT, HALT, DGV, DGH, CDS, NUM_FULL };
enum HalfPlane { GRB, HVWT, DGRB0, DGRB1, D2H, D2V, DELP, DELM, SQ1M, SQ1P, PMWT, RBM, RBP, RBINT,
                 NUM_HALF };
// Per-row partial Nyquist boxes: one {min row, max row, min col, max col}
// slot per (row, half-row) of the tile.
constexpr int BOX_SLOTS = TS * 2;
constexpr size_t TILE_FLOATS =
    (size_t)NUM_FULL * FULL + (size_t)NUM_HALF * HALF + 2 * HALF / 4 + 4 * BOX_SLOTS;

constexpr float EPS = 1.0e-5f, EPSSQ = 1.0e-10f, ARTHRESH = 0.75f;
// math.rs:28-53; GAUSSGRAD carries the exact factor NYQTHRESH = 0.5.
constexpr float GAUSSODD[4] = {0.14659727707323927f, 0.103592713382435f, 0.0732036125103057f,
                               0.0365543548389495f};
constexpr float GAUSSGRAD[6] = {0.5f * 0.07384411893421103f, 0.5f * 0.06207511968171489f,
                                0.5f * 0.0521818194747806f,  0.5f * 0.03687419286733595f,
                                0.5f * 0.03099732204057846f, 0.5f * 0.018413194161458882f};
constexpr float GAUSSEVEN[2] = {0.13719494435797422f, 0.05640252782101291f};
constexpr float GQUINC[4] = {0.169917f, 0.108947f, 0.069855f, 0.0287182f};

struct Tile {
    float* base;
    int index;
    __device__ __forceinline__ float* full(int p) const { return base + (size_t)p * FULL; }
    __device__ __forceinline__ float* half(int p) const {
        return base + (size_t)NUM_FULL * FULL + (size_t)p * HALF;
    }
    __device__ __forceinline__ uint8_t* nyq() const {
        return reinterpret_cast<uint8_t*>(base + (size_t)NUM_FULL * FULL + (size_t)NUM_HALF * HALF);
    }
    __device__ __forceinline__ uint8_t* nyq2() const { return nyq() + HALF; }
    __device__ __forceinline__ int* boxes() const {
        return reinterpret_cast<int*>(nyq() + 2 * HALF);
    }
};

__device__ __forceinline__ Tile tile_of(const AmazeArgs& a, int z) {
    int index = a.tile_base + z;
    return Tile{a.scratch + (size_t)z * TILE_FLOATS, index};
}

__device__ __forceinline__ float sqr(float x) { return x * x; }
__device__ __forceinline__ float median3(float a, float b, float c) {
    return sel_max(sel_min(a, b), sel_min(sel_max(b, a), c));
}
__device__ __forceinline__ float intp(float a, float b, float c) { return a * (b - c) + c; }
__device__ __forceinline__ int fc(int pattern, int row, int col) { return color_at(pattern, col, row); }
__device__ __forceinline__ int gae(int pattern, int row) { return fc(pattern, row, 2) & 1; }

// ---------------------------------------------------------------- load
// mod.rs:171-194 (populate) plus the zeroed state of scratch.rs:113-145.
__global__ __launch_bounds__(256) void amz_load(AmazeArgs a) {
    const uint3 blk = xcd_block();
    const Tile t = tile_of(a, blk.z);
    const int cc = blk.x * 64 + (threadIdx.x & 63);
    const int rr = blk.y * 4 + (threadIdx.x >> 6);
    const int tx = t.index % a.tiles_x, ty = t.index / a.tiles_x;
    const int sx = mirror_whole_near(tx * VALID - BORDER + cc, a.width);
    const int sy = mirror_whole_near(ty * VALID - BORDER + rr, a.height);
    const int c = fc(a.pattern, rr, cc);
    const float v = a.mosaic[(size_t)sy * a.width + sx] * a.forward[c];
    const int i = rr * TS + cc;
    t.full(CFA)[i] = v;
    if ((c & 1) == 0) {
        const int p = i >> 1;
        t.half(GRB)[p] = v;
        t.half(D2H)[p] = 0.0f;
        t.half(D2V)[p] = 0.0f;
        t.nyq()[p] = 0;
        t.nyq2()[p] = 0;
    }
    // vcd/hcd and their alternatives are read up to two rows and columns
    // outside the region the stages write; those entries are zero.
    if (rr < 4 || rr >= TS - 4 || cc < 4 || cc >= TS - 4) {
        t.full(VCD)[i] = 0.0f;
        t.full(HCD)[i] = 0.0f;
        t.full(VALT)[i] = 0.0f;
        t.full(HALT)[i] = 0.0f;
    }
}

// ---------------------------------------------------------------- gradients
// green.rs:13-31 and diagonal.rs:15-42.
__global__ __launch_bounds__(256) void amz_gradients(AmazeArgs a) {
    const uint3 blk = xcd_block();
    const Tile t = tile_of(a, blk.z);
    const int cc = blk.x * 64 + (threadIdx.x & 63);
    const int rr = blk.y * 4 + (threadIdx.x >> 6);
    const float* cfa = t.full(CFA);
    if (rr >= 2 && rr < TS - 2 && cc >= 2 && cc < TS - 2) {
        const int i = rr * TS + cc;
        float delh = fabsf(cfa[i + 1] - cfa[i - 1]);
        float delv = fabsf(cfa[i + TS] - cfa[i - TS]);
        t.full(DW0)[i] = EPS + fabsf(cfa[i + 2 * TS] - cfa[i]) + fabsf(cfa[i] - cfa[i - 2 * TS]) + delv;
        t.full(DW1)[i] = EPS + fabsf(cfa[i + 2] - cfa[i]) + fabsf(cfa[i] - cfa[i - 2]) + delh;
        t.full(DHV)[i] = sqr(delh) + sqr(delv);
    }
    if (rr >= 6 && rr < TS - 6 && cc >= 6 && cc < TS - 6 && (cc & 1) == 0) {
        const float* up = cfa + (rr - 1) * TS;
        const float* mid = cfa + rr * TS;
        const float* down = cfa + (rr + 1) * TS;
        const int p = (rr * TS + cc) >> 1;
        float delp, delm, sqp, sqm;
        if (gae(a.pattern, rr) == 0) {
            delp = fabsf(up[cc + 1] - down[cc - 1]);
            delm = fabsf(down[cc + 1] - up[cc - 1]);
            float g = mid[cc + 1];
            sqp = sqr(g - down[cc]) + sqr(g - up[cc + 2]);
            sqm = sqr(g - up[cc]) + sqr(g - down[cc + 2]);
        } else {
            float g = mid[cc];
            sqp = sqr(g - down[cc - 1]) + sqr(g - up[cc + 1]);
            sqm = sqr(g - up[cc - 1]) + sqr(g - down[cc + 1]);
            delp = fabsf(up[cc + 2] - down[cc]);
            delm = fabsf(down[cc + 2] - up[cc]);
        }
        t.half(DELP)[p] = delp;
        t.half(DELM)[p] = delm;
        t.half(SQ1P)[p] = sqp;
        t.half(SQ1M)[p] = sqm;
    }
}

// ---------------------------------------------------------------- colour differences
__device__ __forceinline__ float adaptive_or(float ratio, float centre, float fallback) {
    return fabsf(1.0f - ratio) < ARTHRESH ? centre * ratio : fallback;
}

// green.rs:46-137
__global__ __launch_bounds__(256) void amz_colour_differences(AmazeArgs a, float clip_pt8) {
    const uint3 blk = xcd_block();
    const Tile t = tile_of(a, blk.z);
    const int cc = blk.x * 64 + (threadIdx.x & 63);
    const int rr = blk.y * 4 + (threadIdx.x >> 6);
    if (rr < 4 || rr >= TS - 4 || cc < 4 || cc >= TS - 4) return;
    const float* cfa = t.full(CFA);
    const float* w0 = t.full(DW0);
    const float* w1 = t.full(DW1);
    const int i = rr * TS + cc;
    const bool fcswitch = (fc(a.pattern, rr, cc) & 1) == 1;  // green site
    float cur = cfa[i];
    float up = cfa[i - TS], up2 = cfa[i - 2 * TS], down = cfa[i + TS], down2 = cfa[i + 2 * TS];
    float left = cfa[i - 1], left2 = cfa[i - 2], right = cfa[i + 1], right2 = cfa[i + 2];
    float w0u2 = w0[i - 2 * TS], w0u1 = w0[i - TS], w0c = w0[i], w0d1 = w0[i + TS],
          w0d2 = w0[i + 2 * TS];
    float w1l2 = w1[i - 2], w1l1 = w1[i - 1], w1c = w1[i], w1r1 = w1[i + 1], w1r2 = w1[i + 2];
    float cru = up * (w0u2 + w0c) / (w0u2 * (EPS + cur) + w0c * (EPS + up2));
    float crd = down * (w0d2 + w0c) / (w0d2 * (EPS + cur) + w0c * (EPS + down2));
    float crl = left * (w1l2 + w1c) / (w1l2 * (EPS + cur) + w1c * (EPS + left2));
    float crr = right * (w1r2 + w1c) / (w1r2 * (EPS + cur) + w1c * (EPS + right2));
    float guha = up + 0.5f * (cur - up2);
    float gdha = down + 0.5f * (cur - down2);
    float glha = left + 0.5f * (cur - left2);
    float grha = right + 0.5f * (cur - right2);
    float guar = adaptive_or(cru, cur, guha);
    float gdar = adaptive_or(crd, cur, gdha);
    float glar = adaptive_or(crl, cur, glha);
    float grar = adaptive_or(crr, cur, grha);
    float hwt = w1l1 / (w1l1 + w1r1);
    float vwt = w0u1 / (w0d1 + w0u1);
    float gintvha = vwt * gdha + (1.0f - vwt) * guha;
    float ginthha = hwt * grha + (1.0f - hwt) * glha;
    float vcd, hcd, vcdalt, hcdalt;
    if (fcswitch) {
        vcd = cur - (vwt * gdar + (1.0f - vwt) * guar);
        hcd = cur - (hwt * grar + (1.0f - hwt) * glar);
        vcdalt = cur - gintvha;
        hcdalt = cur - ginthha;
    } else {
        vcd = (vwt * gdar + (1.0f - vwt) * guar) - cur;
        hcd = (hwt * grar + (1.0f - hwt) * glar) - cur;
        vcdalt = gintvha - cur;
        hcdalt = ginthha - cur;
    }
    bool clipped = (cur > clip_pt8) | (gintvha > clip_pt8) | (ginthha > clip_pt8);
    if (clipped) {
        guar = guha;
        gdar = gdha;
        glar = glha;
        grar = grha;
        vcd = vcdalt;
        hcd = hcdalt;
    }
    t.full(DGV)[i] = sel_min(sqr(guha - gdha), sqr(guar - gdar));
    t.full(DGH)[i] = sel_min(sqr(glha - grha), sqr(glar - grar));
    t.full(VCD)[i] = vcd;
    t.full(HCD)[i] = hcd;
    t.full(VALT)[i] = vcdalt;
    t.full(HALT)[i] = hcdalt;
}

// ---------------------------------------------------------------- bound differences
// green.rs:142-174
__device__ __forceinline__ float bound_at_green(float d, float cur, float a, float b, float clip) {
    float g = -d + cur;
    float alt = -median3(g, a, b) + cur;
    float bounded = d;
    if (d > 0.0f) {
        if (3.0f * d > g + cur) {
            bounded = alt;
        } else {
            float weight = 1.0f - 3.0f * d / (EPS + g + cur);
            bounded = weight * d + (1.0f - weight) * alt;
        }
    }
    return g > clip ? alt : bounded;
}
__device__ __forceinline__ float bound_at_red_blue(float d, float cur, float a, float b,
                                                   float clip) {
    float g = d + cur;
    float alt = median3(g, a, b) - cur;
    float bounded = d;
    if (d < 0.0f) {
        if (3.0f * d < -(g + cur)) {
            bounded = alt;
        } else {
            float weight = 1.0f + 3.0f * d / (EPS + g + cur);
            bounded = weight * d + (1.0f - weight) * alt;
        }
    }
    return g > clip ? alt : bounded;
}
__device__ __forceinline__ float spread(float a, float b, float c) {
    return 3.0f * (sqr(a) + sqr(b) + sqr(c)) - sqr(a + b + c);
}

// green.rs:184-226, vertical half: vcd is bounded in place going down each
// column; the value two rows up is already bounded when read. Rows are
// processed four at a time with the next four rows' inputs loaded up front,
// so the loop waits for memory once per four rows.
__global__ __launch_bounds__(64) void amz_bound_v(AmazeArgs a, float clip_pt) {
    const uint3 blk = xcd_block();
    const Tile t = tile_of(a, blk.z);
    const int cc = blk.x * 64 + threadIdx.x;
    if (cc < 4 || cc >= TS - 4) return;
    float* __restrict__ vcd = t.full(VCD);
    const float* __restrict__ alt = t.full(VALT);
    const float* __restrict__ cfa = t.full(CFA);
    float b_m2 = 0.0f, b_m1 = 0.0f;  // bounded vcd two and one rows up (rows 2, 3 are zero)
    float a_m2 = 0.0f, a_m1 = 0.0f;  // vcdalt two and one rows up
    float u0 = vcd[4 * TS + cc], u1 = vcd[5 * TS + cc];  // unbounded vcd at rr, rr + 1
    float a0 = alt[4 * TS + cc], a1 = alt[5 * TS + cc];
    float c_m1 = cfa[3 * TS + cc], c0 = cfa[4 * TS + cc];
    bool green = (fc(a.pattern, 4, cc) & 1) == 1;
    for (int rr = 4; rr < TS - 4; rr += 4) {
        float U[4], A[4], C[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int r2 = rr + 2 + k;
            U[k] = r2 < TS - 4 ? vcd[r2 * TS + cc] : 0.0f;
            A[k] = r2 < TS - 4 ? alt[r2 * TS + cc] : 0.0f;
            C[k] = cfa[(rr + 1 + k) * TS + cc];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float u2 = U[k], a2 = A[k], c1 = C[k];
            float vcdvar = spread(b_m2, u0, u2);
            float vcdaltvar = spread(a_m2, a0, a2);
            float v = vcdaltvar < vcdvar ? a0 : u0;
            float bounded = green ? bound_at_green(v, c0, c_m1, c1, clip_pt)
                                  : bound_at_red_blue(v, c0, c_m1, c1, clip_pt);
            vcd[(rr + k) * TS + cc] = bounded;
            b_m2 = b_m1;
            b_m1 = bounded;
            a_m2 = a_m1;
            a_m1 = a0;
            u0 = u1;
            u1 = u2;
            a0 = a1;
            a1 = a2;
            c_m1 = c0;
            c0 = c1;
            green = !green;
        }
    }
}

// green.rs:184-226, horizontal half: hcd is bounded in place along each row,
// and cddiffsq is formed at red/blue sites from the final vcd and hcd. One
// thread per tile row walks it in 32-column chunks: each chunk's hcd, hcdalt
// and CFA columns (with the 4-column halo on the left, 8 on the right) and
// its vcd arrive as aligned float4 loads straight into registers, the serial
// recurrence runs fully unrolled over them, and hcd and cddiffsq leave as
// float4 stores. cddiffsq is written as 0 at green sites, which nothing
// reads (the Nyquist stencil touches red/blue sites only).
constexpr int BH_ROWS = 64, BH_CHUNK = 32, BH_IN = BH_CHUNK + 8;

__global__ __launch_bounds__(64) void amz_bound_h(AmazeArgs a, float clip_pt) {
    const uint3 blk = xcd_block();
    const Tile t = tile_of(a, blk.z);
    const int rr = 4 + blk.x * BH_ROWS + (int)threadIdx.x;
    if (rr >= TS - 4) return;
    float* __restrict__ hcd_row = t.full(HCD) + rr * TS;
    const float* __restrict__ alt_row = t.full(HALT) + rr * TS;
    const float* __restrict__ cfa_row = t.full(CFA) + rr * TS;
    const float* __restrict__ vcd_row = t.full(VCD) + rr * TS;
    float* __restrict__ cds_row = t.full(CDS) + rr * TS;
    const bool green4 = (fc(a.pattern, rr, 4) & 1) == 1;  // colour at column 4 (and c0)
    float b_m2 = 0.0f, b_m1 = 0.0f;  // bounded hcd two and one columns left (cols 2, 3 are zero)
    for (int c0 = 4; c0 < TS - 4; c0 += BH_CHUNK) {
        // Register k holds column c0 - 4 + k. hcd and hcdalt read as zero
        // outside [4, TS - 4), the CFA as zero beyond the tile.
        float h[BH_IN], al[BH_IN], cf[BH_IN], vc[BH_CHUNK];
#pragma unroll
        for (int q = 0; q < BH_IN / 4; ++q) {
            const int col = c0 - 4 + 4 * q;
            float4 hv = make_float4(0.0f, 0.0f, 0.0f, 0.0f), av = hv, cv = hv;
            if (col < TS) {
                hv = *reinterpret_cast<const float4*>(hcd_row + col);
                av = *reinterpret_cast<const float4*>(alt_row + col);
                cv = *reinterpret_cast<const float4*>(cfa_row + col);
            }
            const float hs[4] = {hv.x, hv.y, hv.z, hv.w}, as[4] = {av.x, av.y, av.z, av.w};
            const float cs[4] = {cv.x, cv.y, cv.z, cv.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const bool inside = col + e >= 4 && col + e < TS - 4;
                h[4 * q + e] = inside ? hs[e] : 0.0f;
                al[4 * q + e] = inside ? as[e] : 0.0f;
                cf[4 * q + e] = col + e < TS ? cs[e] : 0.0f;
            }
        }
#pragma unroll
        for (int q = 0; q < BH_CHUNK / 4; ++q) {
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (c0 + 4 * q < TS - 4) v = *reinterpret_cast<const float4*>(vcd_row + c0 + 4 * q);
            vc[4 * q] = v.x;
            vc[4 * q + 1] = v.y;
            vc[4 * q + 2] = v.z;
            vc[4 * q + 3] = v.w;
        }
        float out[BH_CHUNK], cd[BH_CHUNK];
#pragma unroll
        for (int k = 0; k < BH_CHUNK; ++k) {
            const bool green = green4 != ((k & 1) != 0);
            const float hcdvar = spread(b_m2, h[k + 4], h[k + 6]);
            const float hcdaltvar = spread(al[k + 2], al[k + 4], al[k + 6]);
            const float hv = hcdaltvar < hcdvar ? al[k + 4] : h[k + 4];
            const float bounded = green ? bound_at_green(hv, cf[k + 4], cf[k + 3], cf[k + 5], clip_pt)
                                        : bound_at_red_blue(hv, cf[k + 4], cf[k + 3], cf[k + 5], clip_pt);
            out[k] = bounded;
            cd[k] = green ? 0.0f : sqr(vc[k] - bounded);
            b_m2 = b_m1;
            b_m1 = bounded;
        }
        // The last chunk holds 24 columns (TS - 8 = 248 = 7 * 32 + 24).
#pragma unroll
        for (int q = 0; q < BH_CHUNK / 4; ++q) {
            if (c0 + 4 * q < TS - 4) {
                *reinterpret_cast<float4*>(hcd_row + c0 + 4 * q) =
                    make_float4(out[4 * q], out[4 * q + 1], out[4 * q + 2], out[4 * q + 3]);
                *reinterpret_cast<float4*>(cds_row + c0 + 4 * q) =
                    make_float4(cd[4 * q], cd[4 * q + 1], cd[4 * q + 2], cd[4 * q + 3]);
            }
        }
    }
}

// ---------------------------------------------------------------- weights, Nyquist test, diagonals
// green.rs:229-292 (direction_weights), nyquist.rs:25-89 (test and bounding
// box), diagonal.rs:98-144 (diagonal_estimates).
__device__ __forceinline__ float diagonal_estimate(float centre, float near_, float far_) {
    float ratio = 2.0f * near_ / (EPS + centre + far_);
    return fabsf(1.0f - ratio) < ARTHRESH ? centre * ratio : near_ + 0.5f * (centre - far_);
}
// diagonal.rs:79-94 and :184-199 (bound_diagonal, bound_green).
__device__ __forceinline__ float bound_towards(float estimate, float centre, float a, float b,
                                               float clip) {
    float value = estimate;
    if (value < centre) {
        if (2.0f * value < centre) {
            value = median3(value, a, b);
        } else {
            float weight = 2.0f * (centre - value) / (EPS + value + centre);
            value = weight * value + (1.0f - weight) * median3(value, a, b);
        }
    }
    return value > clip ? median3(value, a, b) : value;
}

// One thread per red/blue site: cc = 2 * pair + green_at_even(row). A
// wavefront holds 64 consecutive pairs of one row, so the horizontal taps of
// a plane row come from one float2 load per lane (columns 2p, 2p + 1) and
// the neighbouring lanes; only lanes at the wavefront's ends load theirs.
// The arithmetic of each site is the reference's, term for term.
struct Lanes {
    int lane, base, par;  // base = 2 * pair
    // Column cc + dx of a plane row of which this lane holds columns
    // (base, base + 1) in v; `fetch` says whether this lane may read the row
    // itself when the neighbour lies outside the wavefront.
    __device__ __forceinline__ float tap(float2 v, const float* row, int dx, bool fetch) const {
        const int o = par + dx;
        const int po = o >> 1;  // pair offset (floor)
        const float mine = (o & 1) ? v.y : v.x;
        if (po == 0) return mine;
        const int src = lane + po;
        float r = __shfl(mine, src, 64);
        if ((src < 0 || src > 63) && fetch) r = row[base + o];
        return r;
    }
    // Half column pair + s of a half-plane row of which this lane holds
    // half column `pair` in h.
    __device__ __forceinline__ float half_tap(float h, const float* row, int s, bool fetch) const {
        if (s == 0) return h;
        const int src = lane + s;
        float r = __shfl(h, src, 64);
        if ((src < 0 || src > 63) && fetch) r = row[(base >> 1) + s];
        return r;
    }
};

__device__ __forceinline__ float2 pair_load(const float* plane, int row, int base) {
    return *reinterpret_cast<const float2*>(plane + row * TS + base);
}

// even_kernel over half-plane rows rr - 2 .. rr + 2 held one half column per
// lane (e[0] .. e[4]); (cc >> 1) is the lane's own half column, ((cc + 1) >> 1)
// is + par, ((cc - 1) >> 1) is + par - 1, ((cc + 2) >> 1) is + 1 and
// ((cc - 2) >> 1) is - 1.
__device__ __forceinline__ float even_lanes(const Lanes& L, const float* plane, int rr,
                                            const float* e, bool fetch) {
    auto at = [&](int m, int s) { return L.half_tap(e[m], plane + (rr + m - 2) * TSH, s, fetch); };
    const int par = L.par;
    return GAUSSEVEN[0] * (e[1] + at(2, par - 1) + at(2, par) + e[3]) +
           GAUSSEVEN[1] * (at(0, par - 1) + at(0, par) + at(1, -1) + at(1, 1) + at(3, -1) +
                           at(3, 1) + at(4, par - 1) + at(4, par));
}

// Reduces a tile's slots into its Nyquist box (min/max over flagged sites).
__global__ __launch_bounds__(256) void amz_box(AmazeArgs a) {
    const uint3 blk = xcd_block();
    const Tile t = tile_of(a, blk.x);
    const int4* slots = reinterpret_cast<const int4*>(t.boxes());
    int4 b = make_int4(INT_MAX, -1, INT_MAX, -1);
    for (int k = threadIdx.x; k < BOX_SLOTS; k += 256) {
        const int4 v = slots[k];
        b = make_int4(min(b.x, v.x), max(b.y, v.y), min(b.z, v.z), max(b.w, v.w));
    }
    for (int off = 32; off > 0; off >>= 1) {
        b.x = min(b.x, __shfl_xor(b.x, off, 64));
        b.y = max(b.y, __shfl_xor(b.y, off, 64));
        b.z = min(b.z, __shfl_xor(b.z, off, 64));
        b.w = max(b.w, __shfl_xor(b.w, off, 64));
    }
    __shared__ int4 part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        int4 r = part[0];
        for (int w = 1; w < 4; ++w)
            r = make_int4(min(r.x, part[w].x), max(r.y, part[w].y), min(r.z, part[w].z),
                          max(r.w, part[w].w));
        int* area = a.areas + 4 * (size_t)blk.x;
        area[0] = r.x;
        area[1] = r.y;
        area[2] = r.z;
        area[3] = r.w;
    }
}

__global__ __launch_bounds__(256) void amz_weights(AmazeArgs a, float clip_pt) {
    const uint3 blk = xcd_block();
    const Tile t = tile_of(a, blk.z);
    const int lane = threadIdx.x & 63;
    const int p = blk.x * 64 + lane;
    const int rr = blk.y * 4 + (threadIdx.x >> 6);
    const int par = gae(a.pattern, rr);
    const int cc = 2 * p + par, i = rr * TS + cc;
    const Lanes L{lane, 2 * p, par};
    bool flagged = false;
    // Whole wavefronts skip the rows outside [6, TS - 6); inside them every
    // lane takes part in the exchanges and only valid sites fetch or store.
    if (rr >= 6 && rr < TS - 6) {
        const bool valid = cc >= 6 && cc < TS - 6;
        // HV weights (green.rs:229-292).
        {
            const float* vcd = t.full(VCD);
            const float* dgv = t.full(DGV);
            const float* w0 = t.full(DW0);
            const float* hcd_row = t.full(HCD) + rr * TS;
            const float* w1_row = t.full(DW1) + rr * TS;
            const float* dgh_row = t.full(DGH) + rr * TS;
            const float2 HC = pair_load(t.full(HCD), rr, L.base);
            const float2 W1 = pair_load(t.full(DW1), rr, L.base);
            const float2 DH = pair_load(t.full(DGH), rr, L.base);
            float v[7], dv[5];
#pragma unroll
            for (int k = 0; k < 7; ++k) v[k] = vcd[i + (k - 3) * TS];
#pragma unroll
            for (int k = 0; k < 5; ++k) dv[k] = dgv[i + (k - 2) * TS];
            const float w0u = w0[i - TS], w0d = w0[i + TS];
            float hcd[7], dgh[5];
#pragma unroll
            for (int d = -3; d <= 3; ++d) hcd[d + 3] = L.tap(HC, hcd_row, d, valid);
#pragma unroll
            for (int d = -2; d <= 2; ++d) dgh[d + 2] = L.tap(DH, dgh_row, d, valid);
            const float w1l = L.tap(W1, w1_row, -1, valid), w1r = L.tap(W1, w1_row, 1, valid);
            if (valid) {
                float uave = v[3] + v[2] + v[1] + v[0];
                float dave = v[3] + v[4] + v[5] + v[6];
                float lave = hcd[3] + hcd[2] + hcd[1] + hcd[0];
                float rave = hcd[3] + hcd[4] + hcd[5] + hcd[6];
                float varu = sqr(v[3] - uave) + sqr(v[2] - uave) + sqr(v[1] - uave) + sqr(v[0] - uave);
                float vard = sqr(v[3] - dave) + sqr(v[4] - dave) + sqr(v[5] - dave) + sqr(v[6] - dave);
                float varl = sqr(hcd[3] - lave) + sqr(hcd[2] - lave) + sqr(hcd[1] - lave) +
                             sqr(hcd[0] - lave);
                float varr = sqr(hcd[3] - rave) + sqr(hcd[4] - rave) + sqr(hcd[5] - rave) +
                             sqr(hcd[6] - rave);
                float hwt = w1l / (w1l + w1r);
                float vwt = w0u / (w0d + w0u);
                float vcdvar = EPSSQ + vwt * vard + (1.0f - vwt) * varu;
                float hcdvar = EPSSQ + hwt * varr + (1.0f - hwt) * varl;
                float fluct_u = dv[2] + dv[1] + dv[0];
                float fluct_d = dv[2] + dv[3] + dv[4];
                float fluct_l = dgh[2] + dgh[1] + dgh[0];
                float fluct_r = dgh[2] + dgh[3] + dgh[4];
                float vcdvar1 = EPSSQ + vwt * fluct_d + (1.0f - vwt) * fluct_u;
                float hcdvar1 = EPSSQ + hwt * fluct_r + (1.0f - hwt) * fluct_l;
                float varwt = hcdvar / (vcdvar + hcdvar);
                float diffwt = hcdvar1 / (vcdvar1 + hcdvar1);
                t.half(HVWT)[i >> 1] =
                    ((0.5f - varwt) * (0.5f - diffwt) > 0.0f &&
                     fabsf(0.5f - diffwt) < fabsf(0.5f - varwt))
                        ? varwt
                        : diffwt;
            }
        }
        // Nyquist test statistic (nyquist.rs:28-57).
        {
            float2 C[5], D[5];
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                C[m] = pair_load(t.full(CDS), rr + m - 2, L.base);
                D[m] = pair_load(t.full(DHV), rr + m - 2, L.base);
            }
            auto c = [&](int dy, int dx) {
                return L.tap(C[dy + 2], t.full(CDS) + (rr + dy) * TS, dx, valid);
            };
            auto d = [&](int dy, int dx) {
                return L.tap(D[dy + 2], t.full(DHV) + (rr + dy) * TS, dx, valid);
            };
            float test =
                (GAUSSODD[0] * c(0, 0) +
                 GAUSSODD[1] * (c(-1, -1) + c(-1, 1) + c(1, -1) + c(1, 1)) +
                 GAUSSODD[2] * (c(-2, 0) + c(0, -2) + c(0, 2) + c(2, 0)) +
                 GAUSSODD[3] * (c(-2, -2) + c(-2, 2) + c(2, -2) + c(2, 2))) -
                (GAUSSGRAD[0] * d(0, 0) + GAUSSGRAD[1] * (d(-1, 0) + d(0, 1) + d(0, -1) + d(1, 0)) +
                 GAUSSGRAD[2] * (d(-1, -1) + d(-1, 1) + d(1, -1) + d(1, 1)) +
                 GAUSSGRAD[3] * (d(-2, 0) + d(0, -2) + d(0, 2) + d(2, 0)) +
                 GAUSSGRAD[4] * (d(-2, -1) + d(-2, 1) + d(-1, -2) + d(-1, 2) + d(1, -2) +
                                 d(1, 2) + d(2, -1) + d(2, 1)) +
                 GAUSSGRAD[5] * (d(-2, -2) + d(-2, 2) + d(2, -2) + d(2, 2)));
            if (valid && test > 0.0f) {
                t.nyq()[i >> 1] = 1;
                flagged = true;
            }
        }
        // Diagonal estimates (diagonal.rs:98-144).
        if (rr >= 8 && rr < TS - 8) {
            const bool valid8 = cc >= 8 && cc < TS - 8;
            float2 F[5];
#pragma unroll
            for (int m = 0; m < 5; ++m) F[m] = pair_load(t.full(CFA), rr + m - 2, L.base);
            auto f = [&](int dy, int dx) {
                return L.tap(F[dy + 2], t.full(CFA) + (rr + dy) * TS, dx, valid8);
            };
            const float centre = f(0, 0);
            const float se1 = f(1, 1), se2 = f(2, 2);
            const float nw1 = f(-1, -1), nw2 = f(-2, -2);
            const float ne1 = f(-1, 1), ne2 = f(-2, 2);
            const float sw1 = f(1, -1), sw2 = f(2, -2);
            const int hp = rr * TSH + p;
            float dm[5], dp[5], sm[5], sp[5];
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                dm[m] = t.half(DELM)[hp + (m - 2) * TSH];
                dp[m] = t.half(DELP)[hp + (m - 2) * TSH];
                sm[m] = t.half(SQ1M)[hp + (m - 2) * TSH];
                sp[m] = t.half(SQ1P)[hp + (m - 2) * TSH];
            }
            // ((cc + 1) >> 1) is half column p + par, ((cc - 1) >> 1) is
            // p + par - 1, ((cc + 2) >> 1) is p + 1, ((cc - 2) >> 1) is p - 1.
            auto hm = [&](int m, int s) {
                return L.half_tap(dm[m], t.half(DELM) + (rr + m - 2) * TSH, s, valid8);
            };
            auto hq = [&](int m, int s) {
                return L.half_tap(dp[m], t.half(DELP) + (rr + m - 2) * TSH, s, valid8);
            };
            const float dm_se1 = hm(3, par), dm_se2 = hm(4, 1);
            const float dm_nw1 = hm(1, par - 1), dm_nw2 = hm(0, -1);
            const float dp_ne1 = hq(1, par), dp_ne2 = hq(0, 1);
            const float dp_sw1 = hq(3, par - 1), dp_sw2 = hq(4, -1);
            const float rbvarm_even = even_lanes(L, t.half(SQ1M), rr, sm, valid8);
            const float rbvarp_even = even_lanes(L, t.half(SQ1P), rr, sp, valid8);
            if (valid8) {
                float rbse = diagonal_estimate(centre, se1, se2);
                float rbnw = diagonal_estimate(centre, nw1, nw2);
                float rbne = diagonal_estimate(centre, ne1, ne2);
                float rbsw = diagonal_estimate(centre, sw1, sw2);
                float wtse = EPS + dm[2] + dm_se1 + dm_se2;
                float wtnw = EPS + dm[2] + dm_nw1 + dm_nw2;
                float wtne = EPS + dp[2] + dp_ne1 + dp_ne2;
                float wtsw = EPS + dp[2] + dp_sw1 + dp_sw2;
                float rbm_est = (wtse * rbnw + wtnw * rbse) / (wtse + wtnw);
                float rbp_est = (wtne * rbsw + wtsw * rbne) / (wtne + wtsw);
                float rbvarm = EPSSQ + rbvarm_even;
                t.half(PMWT)[hp] = rbvarm / ((EPSSQ + rbvarp_even) + rbvarm);
                t.half(RBP)[hp] = bound_towards(rbp_est, centre, sw1, ne1, clip_pt);
                t.half(RBM)[hp] = bound_towards(rbm_est, centre, nw1, se1, clip_pt);
            }
        }
    }
    // The tile's bounding box of flagged sites (nyquist.rs:59-89). A
    // wavefront is one half row of the tile, with cc increasing by lane, so
    // its flagged sites span [cc of the first flagged lane, cc of the last].
    // Every wavefront writes its slot (empty: INT_MAX, -1, INT_MAX, -1);
    // amz_box reduces the slots, so no two waves touch the same address.
    const unsigned long long m = __ballot(flagged);
    int4 box = make_int4(INT_MAX, -1, INT_MAX, -1);
    if (m) {
        const int first = __ffsll((long long)m) - 1, last = 63 - __clzll((long long)m);
        box = make_int4(rr, rr, __shfl(cc, first, 64), __shfl(cc, last, 64));
    }
    if (lane == 0) reinterpret_cast<int4*>(t.boxes())[rr * 2 + blk.x] = box;
}


// Nyquist area of a tile (nyquist.rs:59-89); false when there is none.
struct Area {
    int start_row, end_row, start_col, end_col;
};
__device__ __forceinline__ bool tile_area(const AmazeArgs& a, int z, Area& out) {
    const int* b = a.areas + 4 * (size_t)z;
    int start_row = b[0], end_row = b[1], start_col = b[2], end_col = b[3];
    if (end_row < 0) return false;  // nothing flagged: the reference's (0, 0) rows
    if (start_row == end_row || start_col == end_col) return false;
    end_row += 1;
    end_col += 1;
    start_col -= start_col & 1;
    out.start_row = max(start_row, 8);
    out.end_row = min(end_row, TS - 8);
    out.start_col = max(start_col, 8);
    out.end_col = min(end_col, TS - 8);
    return true;
}
__device__ __forceinline__ bool in_area(const Area& ar, int pattern, int rr, int cc) {
    return rr >= ar.start_row && rr < ar.end_row && cc >= ar.start_col + gae(pattern, rr) &&
           cc < ar.end_col;
}

// nyquist.rs:100-118: neighbourhood vote.
__global__ __launch_bounds__(256) void amz_vote(AmazeArgs a) {
    const uint3 blk = xcd_block();
    Area ar;
    if (!tile_area(a, blk.z, ar)) return;
    const Tile t = tile_of(a, blk.z);
    const int p = blk.x * 64 + (threadIdx.x & 63);
    const int rr = blk.y * 4 + (threadIdx.x >> 6);
    const int cc = 2 * p + gae(a.pattern, rr);
    if (!in_area(ar, a.pattern, rr, cc)) return;
    const int i = rr * TS + cc;
    const uint8_t* n = t.nyq();
    unsigned votes = n[(i - 2 * TS) >> 1] + n[(i - TS - 1) >> 1] + n[(i - TS + 1) >> 1] +
                     n[(i - 2) >> 1] + n[(i + 2) >> 1] + n[(i + TS - 1) >> 1] +
                     n[(i + TS + 1) >> 1] + n[(i + 2 * TS) >> 1];
    t.nyq2()[i >> 1] = votes >= 5 ? 1 : (votes == 4 ? n[i >> 1] : 0);
}

// nyquist.rs:120-152: area interpolation of hvwt at the sites the vote
// kept (about a tenth of them). A workgroup first lists its kept sites in
// LDS, then its lanes take them densely, so the 7x7 gathers run in full
// wavefronts; each site's result is independent of the order.
__global__ __launch_bounds__(256) void amz_area(AmazeArgs a) {
    const uint3 blk = xcd_block();
    Area ar;
    if (!tile_area(a, blk.z, ar)) return;
    const Tile t = tile_of(a, blk.z);
    __shared__ int s_sites[256];
    __shared__ int s_count;
    if (threadIdx.x == 0) s_count = 0;
    __syncthreads();
    {
        const int p = blk.x * 64 + (threadIdx.x & 63);
        const int rr = blk.y * 4 + (threadIdx.x >> 6);
        const int cc = 2 * p + gae(a.pattern, rr);
        if (in_area(ar, a.pattern, rr, cc) && t.nyq2()[(rr * TS + cc) >> 1] != 0)
            s_sites[atomicAdd(&s_count, 1)] = rr * TS + cc;
    }
    __syncthreads();
    const int n = s_count;
    const uint8_t* n2 = t.nyq2();
    const float* cfa = t.full(CFA);
    for (int k = threadIdx.x; k < n; k += 256) {
        const int i = s_sites[k];
        // Row by row of taps: the row's flags, then the flagged taps' loads
        // (all in flight together), then the sums in the reference's order.
        float sumcfa = 0.0f, sumh = 0.0f, sumv = 0.0f, sumsqh = 0.0f, sumsqv = 0.0f, areawt = 0.0f;
        for (int di = -6; di < 7; di += 2) {
            const int row = i + di * TS - 6;
            bool f[7];
            float c[7], l[7], r[7], u[7], d[7];
#pragma unroll
            for (int j = 0; j < 7; ++j) f[j] = n2[(row + 2 * j) >> 1] != 0;
#pragma unroll
            for (int j = 0; j < 7; ++j) {
                const int i1 = row + 2 * j;
                c[j] = l[j] = r[j] = u[j] = d[j] = 0.0f;
                if (f[j]) {
                    c[j] = cfa[i1];
                    l[j] = cfa[i1 - 1];
                    r[j] = cfa[i1 + 1];
                    u[j] = cfa[i1 - TS];
                    d[j] = cfa[i1 + TS];
                }
            }
#pragma unroll
            for (int j = 0; j < 7; ++j) {
                if (f[j]) {
                    sumcfa += c[j];
                    sumh += l[j] + r[j];
                    sumv += u[j] + d[j];
                    sumsqh += sqr(c[j] - l[j]) + sqr(c[j] - r[j]);
                    sumsqv += sqr(c[j] - u[j]) + sqr(c[j] - d[j]);
                    areawt += 1.0f;
                }
            }
        }
        sumh = sumcfa - 0.5f * sumh;
        sumv = sumcfa - 0.5f * sumv;
        areawt *= 0.5f;
        float hcdvar = EPSSQ + fabsf(areawt * sumsqh - sumh * sumh);
        float vcdvar = EPSSQ + fabsf(areawt * sumsqv - sumv * sumv);
        t.half(HVWT)[i >> 1] = hcdvar / (vcdvar + hcdvar);
    }
}

// ---------------------------------------------------------------- row chains
// nyquist.rs:159-201 (populate_green: hvwt) and diagonal.rs:149-168
// (combine_diagonals: pmwt), both updated in place row after row. One
// workgroup per tile; thread p owns pair p; the previous row's final weights
// live in LDS.
__global__ __launch_bounds__(128) void amz_chains(AmazeArgs a) {
    const uint3 blk = xcd_block();
    const Tile t = tile_of(a, blk.x);
    const int p = threadIdx.x;
    __shared__ float hv_prev[TSH];
    __shared__ float pm_prev[TSH];
    const float* cfa = t.full(CFA);
    const float* vcd = t.full(VCD);
    const float* hcd = t.full(HCD);
    const uint8_t* nyq2 = t.nyq2();
    float* hvwt = t.half(HVWT);
    float* pmwt = t.half(PMWT);
    float* dgrb0 = t.half(DGRB0);
    float* grb = t.half(GRB);
    float* d2h = t.half(D2H);
    float* d2v = t.half(D2V);
    const float* rbm = t.half(RBM);
    const float* rbp = t.half(RBP);
    float* rbint = t.half(RBINT);
    hv_prev[p] = hvwt[7 * TSH + p];
    pm_prev[p] = pmwt[9 * TSH + p];
    // Inputs of one row. Every value is final before the kernel starts except
    // the current row's weights, which only this row's pass writes, so later
    // rows' inputs are loaded while the current row is processed.
    struct RowIn {
        float hv, hv_dl, hv_dr, vcd, hcd, cfa, cfa_l, cfa_r, cfa_u, cfa_d;
        float pm, pm_dl, pm_dr, rbm, rbp;
        uint8_t nyq;
    };
    auto load = [&](int rr) {
        RowIn in{};
        const int cc = 2 * p + gae(a.pattern, rr);
        const int h = rr * TSH + p, down = (rr + 1) * TSH;
        const int i = rr * TS + min(max(cc, 1), TS - 2);
        in.hv = hvwt[h];
        in.hv_dl = hvwt[down + ((cc - 1) >> 1)];
        in.hv_dr = hvwt[down + min((cc + 1) >> 1, TSH - 1)];
        in.vcd = vcd[i];
        in.hcd = hcd[i];
        in.cfa = cfa[i];
        in.cfa_l = cfa[i - 1];
        in.cfa_r = cfa[i + 1];
        in.cfa_u = cfa[i - TS];
        in.cfa_d = cfa[i + TS];
        in.nyq = nyq2[h];
        in.pm = pmwt[h];
        in.pm_dl = pmwt[down + ((cc - 1) >> 1)];
        in.pm_dr = pmwt[down + min((cc + 1) >> 1, TSH - 1)];
        in.rbm = rbm[h];
        in.rbp = rbp[h];
        return in;
    };
    // CHAIN_AHEAD rows of inputs are in flight: row rr + CHAIN_AHEAD is
    // requested as row rr is processed (re-reading row rr past the end).
    constexpr int CHAIN_AHEAD = 8;
    static_assert((TS - 16) % CHAIN_AHEAD == 0, "whole groups of rows");
    RowIn buf[CHAIN_AHEAD];
#pragma unroll
    for (int k = 0; k < CHAIN_AHEAD; ++k) buf[k] = load(8 + k);
    __syncthreads();
    for (int r0 = 8; r0 < TS - 8; r0 += CHAIN_AHEAD)
#pragma unroll
    for (int k = 0; k < CHAIN_AHEAD; ++k) {
        const int rr = r0 + k;
        const RowIn cur = buf[k];
        buf[k] = load(rr + CHAIN_AHEAD < TS - 8 ? rr + CHAIN_AHEAD : rr);
        const int cc = 2 * p + gae(a.pattern, rr);
        const int h = rr * TSH + p;
        float hv = cur.hv;
        if (cc >= 8 && cc < TS - 8) {
            float alt = (hv_prev[(cc - 1) >> 1] + hv_prev[(cc + 1) >> 1] + cur.hv_dl + cur.hv_dr) /
                        4.0f;
            if (fabsf(0.5f - hv) < fabsf(0.5f - alt)) {
                hv = alt;
                hvwt[h] = hv;
            }
            float dg = intp(hv, cur.vcd, cur.hcd);
            float green = cur.cfa + dg;
            dgrb0[h] = dg;
            grb[h] = green;
            bool flagged = cur.nyq != 0;
            d2h[h] = flagged ? sqr(green - 0.5f * (cur.cfa_l + cur.cfa_r)) : 0.0f;
            d2v[h] = flagged ? sqr(green - 0.5f * (cur.cfa_u + cur.cfa_d)) : 0.0f;
        }
        float pm = 0.0f;
        const bool pm_row = rr >= 10 && rr < TS - 10;
        if (pm_row) {
            pm = cur.pm;
            if (cc >= 10 && cc < TS - 10) {
                float alt = (pm_prev[(cc - 1) >> 1] + pm_prev[(cc + 1) >> 1] + cur.pm_dl +
                             cur.pm_dr) /
                            4.0f;
                if (fabsf(0.5f - pm) < fabsf(0.5f - alt)) {
                    pm = alt;
                    pmwt[h] = pm;
                }
                rbint[h] = 0.5f * (cur.cfa + cur.rbm * (1.0f - pm) + cur.rbp * pm);
            }
        }
        __syncthreads();
        hv_prev[p] = hv;
        if (pm_row) pm_prev[p] = pm;
        __syncthreads();
    }
}

// ---------------------------------------------------------------- refinements
// nyquist.rs:204-236 (refine_green), diagonal.rs:193-247
// (refine_green_with_rb) and the G-B split of chroma.rs:26-32, in that order,
// each on the thread's own red/blue site.
__device__ __forceinline__ float green_from_rb(float sample, float centre, float other) {
    float ratio = sample * 2.0f / (EPS + centre + other);
    return fabsf(1.0f - ratio) < ARTHRESH ? centre * ratio : sample + 0.5f * (centre - other);
}
__device__ __forceinline__ void red_offset(int pattern, int& ey, int& ex) {
    int c00 = color_at(pattern, 0, 0), c10 = color_at(pattern, 1, 0);
    if (c00 == 1 && c10 == 0) { ey = 0; ex = 1; }
    else if (c00 == 1) { ey = 1; ex = 0; }
    else if (c00 == 0) { ey = 0; ex = 0; }
    else { ey = 1; ex = 1; }
}

__global__ __launch_bounds__(256) void amz_refine(AmazeArgs a, float clip_pt) {
    const uint3 blk = xcd_block();
    const Tile t = tile_of(a, blk.z);
    const int p = blk.x * 64 + (threadIdx.x & 63);
    const int rr = blk.y * 4 + (threadIdx.x >> 6);
    const int cc = 2 * p + gae(a.pattern, rr);
    const int i = rr * TS + cc;
    const int h = i >> 1;
    const float* cfa = t.full(CFA);
    float* dgrb0 = t.half(DGRB0);
    float* grb = t.half(GRB);
    Area ar;
    if (tile_area(a, blk.z, ar) && in_area(ar, a.pattern, rr, cc) && t.nyq2()[h] != 0) {
        auto average = [&](const float* pl) {
            return EPSSQ + (GQUINC[0] * pl[h] +
                            GQUINC[1] * (pl[(i - TS - 1) >> 1] + pl[(i - TS + 1) >> 1] +
                                         pl[(i + TS - 1) >> 1] + pl[(i + TS + 1) >> 1]) +
                            GQUINC[2] * (pl[(i - 2 * TS) >> 1] + pl[(i - 2) >> 1] +
                                         pl[(i + 2) >> 1] + pl[(i + 2 * TS) >> 1]) +
                            GQUINC[3] * (pl[(i - 2 * TS - 2) >> 1] + pl[(i - 2 * TS + 2) >> 1] +
                                         pl[(i + 2 * TS - 2) >> 1] + pl[(i + 2 * TS + 2) >> 1]));
        };
        float gvarh = average(t.half(D2H));
        float gvarv = average(t.half(D2V));
        float dg = (t.full(HCD)[i] * gvarv + t.full(VCD)[i] * gvarh) / (gvarv + gvarh);
        dgrb0[h] = dg;
        grb[h] = cfa[i] + dg;
    }
    if (rr >= 12 && rr < TS - 12 && cc >= 12 && cc < TS - 12) {
        const float* pm = t.half(PMWT);
        const float* hv = t.half(HVWT);
        if (!(fabsf(0.5f - pm[h]) < fabsf(0.5f - hv[h]))) {
            const float* rb = t.half(RBINT);
            const float* w0 = t.full(DW0);
            const float* w1 = t.full(DW1);
            float centre = rb[h];
            float gu = green_from_rb(cfa[i - TS], centre, rb[h - 2 * TSH]);
            float gd = green_from_rb(cfa[i + TS], centre, rb[h + 2 * TSH]);
            float gl = green_from_rb(cfa[i - 1], centre, rb[h - 1]);
            float gr = green_from_rb(cfa[i + 1], centre, rb[h + 1]);
            float gintv = (w0[i - TS] * gd + w0[i + TS] * gu) / (w0[i + TS] + w0[i - TS]);
            float ginth = (w1[i - 1] * gr + w1[i + 1] * gl) / (w1[i - 1] + w1[i + 1]);
            gintv = bound_towards(gintv, centre, cfa[i - TS], cfa[i + TS], clip_pt);
            ginth = bound_towards(ginth, centre, cfa[i - 1], cfa[i + 1], clip_pt);
            float green = ginth * (1.0f - hv[h]) + gintv * hv[h];
            grb[h] = green;
            dgrb0[h] = green - cfa[i];
        }
    }
    // Split G - B out of the combined colour difference on blue rows.
    int ey, ex;
    red_offset(a.pattern, ey, ex);
    if (rr >= 13 - ey && rr < TS - 12 && ((rr - (13 - ey)) & 1) == 0 && p >= ((13 - ex) >> 1) &&
        p < ((TS - 12) >> 1)) {
        t.half(DGRB1)[h] = dgrb0[h];
        dgrb0[h] = 0.0f;
    }
}

// ---------------------------------------------------------------- chroma
// chroma.rs:34-96
__global__ __launch_bounds__(256) void amz_chroma(AmazeArgs a) {
    const uint3 blk = xcd_block();
    const Tile t = tile_of(a, blk.z);
    const int p = blk.x * 64 + (threadIdx.x & 63);
    const int rr = blk.y * 4 + (threadIdx.x >> 6);
    if (rr < 14 || rr >= TS - 14) return;
    const int start = 14 + gae(a.pattern, rr);
    const int cc = 2 * p + gae(a.pattern, rr);
    if (cc < start || cc >= TS - 14) return;
    const int plane = 1 - fc(a.pattern, rr, start) / 2;
    float* pl = t.half(plane ? DGRB1 : DGRB0);
    auto d = [&](int dr, int dc) { return pl[(rr + dr) * TSH + ((cc + dc) >> 1)]; };
    float wtnw = 1.0f / (EPS + fabsf(d(-1, -1) - d(1, 1)) + fabsf(d(-1, -1) - d(-3, -3)) +
                         fabsf(d(1, 1) - d(-3, -3)));
    float wtne = 1.0f / (EPS + fabsf(d(-1, 1) - d(1, -1)) + fabsf(d(-1, 1) - d(-3, 3)) +
                         fabsf(d(1, -1) - d(-3, 3)));
    float wtsw = 1.0f / (EPS + fabsf(d(1, -1) - d(-1, 1)) + fabsf(d(1, -1) - d(3, 3)) +
                         fabsf(d(-1, 1) - d(3, -3)));
    float wtse = 1.0f / (EPS + fabsf(d(1, 1) - d(-1, -1)) + fabsf(d(1, 1) - d(3, -3)) +
                         fabsf(d(-1, -1) - d(3, 3)));
    float value = (wtnw * (1.325f * d(-1, -1) - 0.175f * d(-3, -3) - 0.075f * d(-1, -3) -
                           0.075f * d(-3, -1)) +
                   wtne * (1.325f * d(-1, 1) - 0.175f * d(-3, 3) - 0.075f * d(-1, 3) -
                           0.075f * d(1, 1)) +
                   wtsw * (1.325f * d(1, -1) - 0.175f * d(3, -3) - 0.075f * d(1, -3) -
                           0.075f * d(-1, -1)) +
                   wtse * (1.325f * d(1, 1) - 0.175f * d(3, 3) - 0.075f * d(1, 3) -
                           0.075f * d(3, 1))) /
                  (wtnw + wtne + wtsw + wtse);
    pl[rr * TSH + (cc >> 1)] = value;
}

// ---------------------------------------------------------------- finish + write
// chroma.rs:103-147 (finish_planes) fused with mod.rs:196-273 (write_tile).
__global__ __launch_bounds__(256) void amz_finish(AmazeArgs a) {
    const uint3 blk = xcd_block();
    const Tile t = tile_of(a, blk.z);
    const int cc = BORDER + blk.x * 64 + (threadIdx.x & 63);
    const int rr = BORDER + blk.y * 4 + (threadIdx.x >> 6);
    if (cc >= TS - BORDER || rr >= TS - BORDER) return;
    const int tx = t.index % a.tiles_x, ty = t.index / a.tiles_x;
    const int gx = tx * VALID + cc - BORDER, gy = ty * VALID + rr - BORDER;
    if (gx >= a.width || gy >= a.height) return;
    const float* cfa = t.full(CFA);
    const float* hv = t.half(HVWT);
    const float* g_r = t.half(DGRB0);
    const float* g_b = t.half(DGRB1);
    const int i = rr * TS + cc;
    const int ch = fc(a.pattern, rr, cc);
    const float sample = cfa[i];
    float green, red, blue;
    if (ch == 1) {
        green = sample;
        const int r0 = rr * TSH;
        float up = hv[r0 - TSH + (cc >> 1)], down = hv[r0 + TSH + (cc >> 1)];
        float left = hv[r0 + ((cc - 1) >> 1)], right = hv[r0 + ((cc + 1) >> 1)];
        float temp = 1.0f / (up + 2.0f - right - left + down);
        float br = (up * g_r[r0 - TSH + (cc >> 1)] + (1.0f - right) * g_r[r0 + ((cc + 1) >> 1)] +
                    (1.0f - left) * g_r[r0 + ((cc - 1) >> 1)] + down * g_r[r0 + TSH + (cc >> 1)]) *
                   temp;
        float bb = (up * g_b[r0 - TSH + (cc >> 1)] + (1.0f - right) * g_b[r0 + ((cc + 1) >> 1)] +
                    (1.0f - left) * g_b[r0 + ((cc - 1) >> 1)] + down * g_b[r0 + TSH + (cc >> 1)]) *
                   temp;
        red = sel_max(green - br, 0.0f);
        blue = sel_max(green - bb, 0.0f);
    } else {
        green = t.half(GRB)[i >> 1];
        red = sel_max(green - g_r[i >> 1], 0.0f);
        blue = sel_max(green - g_b[i >> 1], 0.0f);
    }
    float rgb[3] = {ch == 0 ? sample : red, ch == 1 ? sample : sel_max(green, 0.0f),
                    ch == 2 ? sample : blue};
    rgb[0] *= a.inverse[0];
    rgb[1] *= a.inverse[1];
    rgb[2] *= a.inverse[2];
    const size_t n = (size_t)a.width * a.height, o = (size_t)gy * a.width + gx;
    a.out[o] = rgb[0];
    a.out[n + o] = rgb[1];
    a.out[2 * n + o] = rgb[2];
    if (!(is_finite(rgb[0]) && is_finite(rgb[1]) && is_finite(rgb[2])))
        atomicOr(&a.stats->nonfinite, 2u);
}

}  // namespace

size_t amaze_scratch_bytes_per_tile() { return TILE_FLOATS * sizeof(float) + 4 * sizeof(int); }

// mod.rs:136-161: the stage sequence for `tiles` tiles starting at
// a.tile_base; a.scratch and a.areas hold one slot per tile of this launch.
hipError_t launch_amaze(const AmazeArgs& a, int tiles, hipStream_t s, Profiler* prof) {
    const float clip_pt8 = 0.8f * a.clip_point;
    const dim3 px_grid(TS / 64, TS / 4, tiles), pair_grid(TSH / 64, TS / 4, tiles), blk(256);
    // Algorithmic bytes: the planes each stage must read and write once, per
    // tile (f = full-plane floats per tile, h = half-plane floats per tile).
    const uint64_t f = 4ull * FULL * tiles, h = 4ull * HALF * tiles;
    RHD_TIMED(prof, s, "amz_load", f + f + h, hipLaunchKernelGGL(amz_load, px_grid, blk, 0, s, a));
    RHD_TIMED(prof, s, "amz_gradients", f + 3 * f + 4 * h,
              hipLaunchKernelGGL(amz_gradients, px_grid, blk, 0, s, a));
    RHD_TIMED(prof, s, "amz_colour_differences", 3 * f + 6 * f,
              hipLaunchKernelGGL(amz_colour_differences, px_grid, blk, 0, s, a, clip_pt8));
    RHD_TIMED(prof, s, "amz_bound_v", 3 * f + f,
              hipLaunchKernelGGL(amz_bound_v, dim3(TS / 64, 1, tiles), dim3(64), 0, s, a,
                                 a.clip_point));
    RHD_TIMED(prof, s, "amz_bound_h", 4 * f + f + f / 2,
              hipLaunchKernelGGL(amz_bound_h, dim3((TS - 8 + BH_ROWS - 1) / BH_ROWS, 1, tiles),
                                 dim3(BH_ROWS), 0, s, a, a.clip_point));
    RHD_TIMED(prof, s, "amz_weights", 9 * f / 2 + 7 * h + h / 2 + h / 8,
              hipLaunchKernelGGL(amz_weights, pair_grid, blk, 0, s, a, a.clip_point));
    RHD_TIMED(prof, s, "amz_box", 16ull * BOX_SLOTS * tiles,
              hipLaunchKernelGGL(amz_box, dim3(tiles), blk, 0, s, a));
    RHD_TIMED(prof, s, "amz_vote", h / 2, hipLaunchKernelGGL(amz_vote, pair_grid, blk, 0, s, a));
    RHD_TIMED(prof, s, "amz_area", h / 4, hipLaunchKernelGGL(amz_area, pair_grid, blk, 0, s, a));
    RHD_TIMED(prof, s, "amz_chains", 3 * f / 2 + 6 * h + 6 * h,
              hipLaunchKernelGGL(amz_chains, dim3(tiles), dim3(TSH), 0, s, a));
    RHD_TIMED(prof, s, "amz_refine", f / 2 + 2 * f / 2 + 5 * h + 3 * h,
              hipLaunchKernelGGL(amz_refine, pair_grid, blk, 0, s, a, a.clip_point));
    RHD_TIMED(prof, s, "amz_chroma", 2 * h + h, hipLaunchKernelGGL(amz_chroma, pair_grid, blk, 0, s, a));
    RHD_TIMED(prof, s, "amz_finish", 12ull * a.width * a.height * tiles / (a.tiles_x * a.tiles_y) + f / 2 + 4 * h,
              hipLaunchKernelGGL(amz_finish, dim3(VALID / 64 + 1, VALID / 4, tiles), blk, 0, s, a));
    return hipGetLastError();
}

}  // namespace rhd
