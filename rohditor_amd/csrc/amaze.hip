// AMaZE demosaic on gfx950.
//
// Reference: crates/demosaic/src/amaze/{mod,math,green,nyquist,diagonal,
// chroma}.rs with the tile geometry of crates/demosaic/src/amaze_geometry.rs.
// The result is bit-identical to the reference: every stage keeps its loop
// bounds, plane addressing and operation order, and the tile grid (224-pixel
// cores, 16-pixel mirrored halo, 256-pixel tile planes) is kept because the
// Nyquist bounding box and the three in-place chains are per tile.
//
// Layout: every tile owns one scratch block in HBM (5 full planes of 256x256
// f32, 10 half planes of 256x128 f32 indexed by pixel pair, 2 byte planes).
// Green is stored only at red/blue sites (the GRB half plane); at green
// sites it equals the CFA sample. Stencil quantities that are cheap to form
// from the CFA samples (the directional and diagonal gradients, delhvsqsum,
// cddiffsq, the alternative colour differences) are never stored: the kernels that read them form them in LDS
// or registers. Stages run as grid kernels over (tile, block of rows and
// columns or pairs); the three sequential chains run as
//   - bound_v: one thread per tile column walking down the rows (vcd),
//   - bound_h: one thread per tile row walking right in register chunks (hcd),
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

enum FullPlane { CFA, VCD, HCD, DGV, DGH, NUM_FULL };
enum HalfPlane { HVWT, PMWT, RBM, RBP, NUM_HALF };
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

// diagonal.rs:98-144 (one diagonal estimate) and diagonal.rs:79-94 and
// :184-199 (bound_diagonal, bound_green).
__device__ __forceinline__ float diagonal_estimate(float centre, float near_, float far_) {
    float ratio = 2.0f * near_ / (EPS + centre + far_);
    return fabsf(1.0f - ratio) < ARTHRESH ? centre * ratio : near_ + 0.5f * (centre - far_);
}
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

// ---------------------------------------------------------------- front
// mod.rs:171-194 (populate), green.rs:13-31 and diagonal.rs:15-42
// (gradients), green.rs:46-137 (colour differences) and diagonal.rs:98-144
// (diagonal estimates) for one FW x FH block of a tile. The block's CFA
// samples, 4 around, are staged in LDS from the mosaic; the directional
// gradients (2 around) and the diagonal gradients and squares (2 rows and
// one pair around) are formed in LDS from them. Only what later stages read
// leaves for HBM: the CFA, the two colour differences and their gradients,
// the diagonal estimates and weights, and the zeroed state of the planes
// that later stages write only in part (scratch.rs:113-145).
constexpr int FW = 64, FH = 32;
constexpr int FC_W = FW + 8, FC_H = FH + 8;      // CFA: columns c0 - 4 .., rows r0 - 4 ..
constexpr int FG_W = FW + 4, FG_H = FH + 4;      // gradients: columns c0 - 2 .., rows r0 - 2 ..
constexpr int FD_W = FW / 2 + 2, FD_H = FH + 4;  // diagonal planes: pairs p0 - 1 .., rows r0 - 2 ..
// Eight wavefronts per block: four blocks (the LDS of five) fill a CU's 32
// wavefront slots.
constexpr int F_WAVES = 8, F_THREADS = 64 * F_WAVES;
static_assert(TS % FW == 0 && TS % FH == 0 && FH % F_WAVES == 0, "whole blocks");

// The gradients are read only by the colour differences and the diagonal
// planes only by the diagonal estimates, which run after them: the two sets
// share one array, so five blocks fit a CU's LDS instead of three.
constexpr int FG_N = FG_H * FG_W, FD_N = FD_H * FD_W;
static_assert(4 * FD_N <= 2 * FG_N, "the diagonal planes fit the gradients' space");
struct alignas(16) FrontLds {
    float cfa[FC_H * FC_W];
    float work[2 * FG_N];  // w0, w1, then dp, dm, qp, qm
};

__device__ __forceinline__ float adaptive_or(float ratio, float centre, float fallback) {
    return fabsf(1.0f - ratio) < ARTHRESH ? centre * ratio : fallback;
}
// green.rs:13-31 at (r, c) of the CFA accessor f: the vertical (dirwts0)
// and horizontal (dirwts1) gradients and delhvsqsum.
template <class F>
__device__ __forceinline__ float vgrad(const F& f, int r, int c) {
    const float cur = f(r, c);
    const float delv = fabsf(f(r + 1, c) - f(r - 1, c));
    return EPS + fabsf(f(r + 2, c) - cur) + fabsf(cur - f(r - 2, c)) + delv;
}
template <class F>
__device__ __forceinline__ float hgrad(const F& f, int r, int c) {
    const float cur = f(r, c);
    const float delh = fabsf(f(r, c + 1) - f(r, c - 1));
    return EPS + fabsf(f(r, c + 2) - cur) + fabsf(cur - f(r, c - 2)) + delh;
}
template <class F>
__device__ __forceinline__ float delhv(const F& f, int r, int c) {
    const float delh = fabsf(f(r, c + 1) - f(r, c - 1));
    const float delv = fabsf(f(r + 1, c) - f(r - 1, c));
    return sqr(delh) + sqr(delv);
}

__global__ __launch_bounds__(F_THREADS) void amz_front(AmazeArgs a, float clip_pt8, float clip_pt) {
    const uint3 blk = xcd_block();
    const Tile t = tile_of(a, blk.z);
    const int c0 = blk.x * FW, r0 = blk.y * FH, p0 = c0 / 2;
    __shared__ FrontLds s;
    float* const s_w0 = s.work;
    float* const s_w1 = s.work + FG_N;
    float* const s_dp = s.work;
    float* const s_dm = s.work + FD_N;
    float* const s_qp = s.work + 2 * FD_N;
    float* const s_qm = s.work + 3 * FD_N;
    auto C = [&](int rr, int cc) { return s.cfa[(rr - r0 + 4) * FC_W + (cc - c0 + 4)]; };
    auto W0 = [&](int rr, int cc) { return s_w0[(rr - r0 + 2) * FG_W + (cc - c0 + 2)]; };
    auto W1 = [&](int rr, int cc) { return s_w1[(rr - r0 + 2) * FG_W + (cc - c0 + 2)]; };
    auto D = [&](const float* pl, int rr, int p) { return pl[(rr - r0 + 2) * FD_W + (p - p0 + 1)]; };

    // Lanes take columns, wavefronts rows: a lane's source column, mirror
    // and channel gains are formed once, a row's source row once per wave.
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));

    // CFA samples of the block and its halo (zero outside the tile, where
    // no stencil of a written site reaches). Columns c0 - 4 + lane and, for
    // the first FC_W - 64 lanes, c0 + 60 + lane.
    {
        static_assert(FC_W > 64 && FC_W <= 128, "two column slots per lane");
        const int tx = t.index % a.tiles_x, ty = t.index / a.tiles_x;
        const int ox = tx * VALID - BORDER, oy = ty * VALID - BORDER;
        int sx[2];
        bool in[2];
        float gain[2][2];  // [column slot][row parity]
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const int cc = c0 - 4 + lane + 64 * m;
            in[m] = (m == 0 || lane < FC_W - 64) && cc >= 0 && cc < TS;
            sx[m] = in[m] ? mirror_whole_near(ox + cc, a.width) : 0;
#pragma unroll
            for (int rp = 0; rp < 2; ++rp) {
                const int ch = fc(a.pattern, rp, cc & 1);
                gain[m][rp] = ch == 0 ? a.forward[0] : (ch == 1 ? a.forward[1] : a.forward[2]);
            }
        }
        for (int ly = wave; ly < FC_H; ly += F_WAVES) {
            const int rr = r0 - 4 + ly;
            const bool row_in = rr >= 0 && rr < TS;
            const float* src = a.mosaic + (size_t)(row_in ? mirror_whole_near(oy + rr, a.height) : 0) * a.width;
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                if (m == 1 && lane >= FC_W - 64) break;
                s.cfa[ly * FC_W + lane + 64 * m] = row_in && in[m] ? src[sx[m]] * gain[m][rr & 1] : 0.0f;
            }
        }
    }
    __syncthreads();
    // Gradients over [2, TS - 2); zero elsewhere, unread.
    static_assert(FG_W > 64 && FG_W <= 128 && FD_W <= 64, "column slots per lane");
    for (int ly = wave; ly < FG_H; ly += F_WAVES) {
        const int rr = r0 - 2 + ly;
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            if (m == 1 && lane >= FG_W - 64) break;
            const int lx = lane + 64 * m, cc = c0 - 2 + lx;
            float w0 = 0.0f, w1 = 0.0f;
            if (rr >= 2 && rr < TS - 2 && cc >= 2 && cc < TS - 2) {
                w0 = vgrad(C, rr, cc);
                w1 = hgrad(C, rr, cc);
            }
            s_w0[ly * FG_W + lx] = w0;
            s_w1[ly * FG_W + lx] = w1;
        }
    }
    __syncthreads();

    // Per site: the CFA plane and the colour differences (green.rs:46-137);
    // vcd and hcd are zero in the 4-pixel band, where later stages read them
    // but nothing writes them.
    const int cc = c0 + lane;
    for (int j = 0; j < FH / F_WAVES; ++j) {
        const int rr = r0 + wave * (FH / F_WAVES) + j;
        const int i = rr * TS + cc;
        const float cur = C(rr, cc);
        t.full(CFA)[i] = cur;
        if (!(rr >= 4 && rr < TS - 4 && cc >= 4 && cc < TS - 4)) {
            t.full(VCD)[i] = 0.0f;
            t.full(HCD)[i] = 0.0f;
            continue;
        }
        const bool fcswitch = (fc(a.pattern, rr, cc) & 1) == 1;  // green site
        const float up = C(rr - 1, cc), up2 = C(rr - 2, cc), down = C(rr + 1, cc),
                    down2 = C(rr + 2, cc);
        const float left = C(rr, cc - 1), left2 = C(rr, cc - 2), right = C(rr, cc + 1),
                    right2 = C(rr, cc + 2);
        const float w0u2 = W0(rr - 2, cc), w0u1 = W0(rr - 1, cc), w0c = W0(rr, cc),
                    w0d1 = W0(rr + 1, cc), w0d2 = W0(rr + 2, cc);
        const float w1l2 = W1(rr, cc - 2), w1l1 = W1(rr, cc - 1), w1c = W1(rr, cc),
                    w1r1 = W1(rr, cc + 1), w1r2 = W1(rr, cc + 2);
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
    }
    // Every read of the gradients is done before the diagonal planes
    // overwrite them: the planes over [6, TS - 6) (even columns, one entry
    // per pair); zero elsewhere, unread.
    __syncthreads();
    for (int ly = wave; ly < FD_H; ly += F_WAVES) {
        const int rr = r0 - 2 + ly;
        if (lane >= FD_W) continue;
        const int cc = c0 - 2 + 2 * lane;
        float delp = 0.0f, delm = 0.0f, sqp = 0.0f, sqm = 0.0f;
        if (rr >= 6 && rr < TS - 6 && cc >= 6 && cc < TS - 6) {
            if (gae(a.pattern, rr) == 0) {
                delp = fabsf(C(rr - 1, cc + 1) - C(rr + 1, cc - 1));
                delm = fabsf(C(rr + 1, cc + 1) - C(rr - 1, cc - 1));
                const float g = C(rr, cc + 1);
                sqp = sqr(g - C(rr + 1, cc)) + sqr(g - C(rr - 1, cc + 2));
                sqm = sqr(g - C(rr - 1, cc)) + sqr(g - C(rr + 1, cc + 2));
            } else {
                const float g = C(rr, cc);
                sqp = sqr(g - C(rr + 1, cc - 1)) + sqr(g - C(rr - 1, cc + 1));
                sqm = sqr(g - C(rr - 1, cc - 1)) + sqr(g - C(rr + 1, cc + 1));
                delp = fabsf(C(rr - 1, cc + 2) - C(rr + 1, cc));
                delm = fabsf(C(rr + 1, cc + 2) - C(rr - 1, cc));
            }
        }
        const int k = ly * FD_W + lane;
        s_dp[k] = delp;
        s_dm[k] = delm;
        s_qp[k] = sqp;
        s_qm[k] = sqm;
    }
    __syncthreads();

    // Per red/blue site (one per pair): the zeroed Nyquist flags and the
    // diagonal estimates over [8, TS - 8) (diagonal.rs:98-144).
    for (int k = threadIdx.x; k < FH * (FW / 2); k += F_THREADS) {
        const int rr = r0 + k / (FW / 2), p = p0 + k % (FW / 2);
        const int par = gae(a.pattern, rr);
        const int cc = 2 * p + par, hp = rr * TSH + p;
        t.nyq()[hp] = 0;
        t.nyq2()[hp] = 0;
        if (rr < 8 || rr >= TS - 8 || cc < 8 || cc >= TS - 8) continue;
        const float centre = C(rr, cc);
        const float se1 = C(rr + 1, cc + 1), se2 = C(rr + 2, cc + 2);
        const float nw1 = C(rr - 1, cc - 1), nw2 = C(rr - 2, cc - 2);
        const float ne1 = C(rr - 1, cc + 1), ne2 = C(rr - 2, cc + 2);
        const float sw1 = C(rr + 1, cc - 1), sw2 = C(rr + 2, cc - 2);
        const float dm_c = D(s_dm, rr, p), dp_c = D(s_dp, rr, p);
        const float dm_se1 = D(s_dm, rr + 1, p + par), dm_se2 = D(s_dm, rr + 2, p + 1);
        const float dm_nw1 = D(s_dm, rr - 1, p + par - 1), dm_nw2 = D(s_dm, rr - 2, p - 1);
        const float dp_ne1 = D(s_dp, rr - 1, p + par), dp_ne2 = D(s_dp, rr - 2, p + 1);
        const float dp_sw1 = D(s_dp, rr + 1, p + par - 1), dp_sw2 = D(s_dp, rr + 2, p - 1);
        // even_kernel (diagonal.rs:46-61) over rows rr - 2 .. rr + 2.
        auto even = [&](const float* q) {
            return GAUSSEVEN[0] * (D(q, rr - 1, p) + D(q, rr, p + par - 1) + D(q, rr, p + par) +
                                   D(q, rr + 1, p)) +
                   GAUSSEVEN[1] * (D(q, rr - 2, p + par - 1) + D(q, rr - 2, p + par) +
                                   D(q, rr - 1, p - 1) + D(q, rr - 1, p + 1) +
                                   D(q, rr + 1, p - 1) + D(q, rr + 1, p + 1) +
                                   D(q, rr + 2, p + par - 1) + D(q, rr + 2, p + par));
        };
        const float rbvarm_even = even(s_qm), rbvarp_even = even(s_qp);
        float rbse = diagonal_estimate(centre, se1, se2);
        float rbnw = diagonal_estimate(centre, nw1, nw2);
        float rbne = diagonal_estimate(centre, ne1, ne2);
        float rbsw = diagonal_estimate(centre, sw1, sw2);
        float wtse = EPS + dm_c + dm_se1 + dm_se2;
        float wtnw = EPS + dm_c + dm_nw1 + dm_nw2;
        float wtne = EPS + dp_c + dp_ne1 + dp_ne2;
        float wtsw = EPS + dp_c + dp_sw1 + dp_sw2;
        float rbm_est = (wtse * rbnw + wtnw * rbse) / (wtse + wtnw);
        float rbp_est = (wtne * rbsw + wtsw * rbne) / (wtne + wtsw);
        float rbvarm = EPSSQ + rbvarm_even;
        t.half(PMWT)[hp] = rbvarm / ((EPSSQ + rbvarp_even) + rbvarm);
        t.half(RBP)[hp] = bound_towards(rbp_est, centre, sw1, ne1, clip_pt);
        t.half(RBM)[hp] = bound_towards(rbm_est, centre, nw1, se1, clip_pt);
    }
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

// The alternative colour differences (vcdalt, hcdalt of green.rs:46-137):
// the Hamilton-Adams estimates along one axis, weighted by that axis's
// gradients. Each depends on the CFA samples of its own column (row) only,
// three either side, so the bound kernels form them from the CFA samples
// they walk instead of reading a stored plane. f(d) is the sample d away
// along the axis.
template <class F>
__device__ __forceinline__ float alt_difference(const F& f, bool green) {
    const float cur = f(0);
    const float w_m1 = EPS + fabsf(f(1) - f(-1)) + fabsf(f(-1) - f(-3)) + fabsf(f(0) - f(-2));
    const float w_p1 = EPS + fabsf(f(3) - f(1)) + fabsf(f(1) - f(-1)) + fabsf(f(2) - f(0));
    const float wt = w_m1 / (w_p1 + w_m1);
    const float ha_m = f(-1) + 0.5f * (cur - f(-2));
    const float ha_p = f(1) + 0.5f * (cur - f(2));
    const float gint = wt * ha_p + (1.0f - wt) * ha_m;
    return green ? cur - gint : gint - cur;
}

// green.rs:184-226, vertical half: vcd is bounded in place going down each
// column; the value two rows up is already bounded when read. Rows are
// processed four at a time with the next four rows' inputs (vcd, and the CFA
// samples that give vcdalt) loaded up front, so the loop waits for memory
// once per four rows.
__global__ __launch_bounds__(64) void amz_bound_v(AmazeArgs a, float clip_pt) {
    const uint3 blk = xcd_block();
    const Tile t = tile_of(a, blk.z);
    const int cc = blk.x * 64 + threadIdx.x;
    if (cc < 4 || cc >= TS - 4) return;
    float* __restrict__ vcd = t.full(VCD);
    const float* __restrict__ cfa = t.full(CFA);
    // vcdalt at row r (zero outside [4, TS - 4)) from the CFA samples of rows
    // r - 3 .. r + 3 held in w[0 .. 6].
    auto alt_at = [&](const float* w, int r, bool green) {
        return r < TS - 4 ? alt_difference([&](int d) { return w[3 + d]; }, green) : 0.0f;
    };
    float b_m2 = 0.0f, b_m1 = 0.0f;  // bounded vcd two and one rows up (rows 2, 3 are zero)
    float a_m2 = 0.0f, a_m1 = 0.0f;  // vcdalt two and one rows up
    bool green = (fc(a.pattern, 4, cc) & 1) == 1;
    float u0 = vcd[4 * TS + cc], u1 = vcd[5 * TS + cc];  // unbounded vcd at rr, rr + 1
    float a0, a1;
    {
        float w[9];  // CFA rows 1 .. 9
#pragma unroll
        for (int j = 0; j < 9; ++j) w[j] = cfa[(1 + j) * TS + cc];
        a0 = alt_at(w, 4, green);
        a1 = alt_at(w + 1, 5, !green);
    }
    float c_m1 = cfa[3 * TS + cc], c0 = cfa[4 * TS + cc];
    for (int rr = 4; rr < TS - 4; rr += 4) {
        // Rows rr + 2 .. rr + 5: unbounded vcd, and vcdalt from CFA rows
        // rr - 1 .. rr + 8 (w[j] is row rr - 1 + j; rows past the tile read
        // as zero and only feed vcdalt rows outside [4, TS - 4)).
        float U[4], w[10];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int r2 = rr + 2 + k;
            U[k] = r2 < TS - 4 ? vcd[r2 * TS + cc] : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < 10; ++j) w[j] = rr - 1 + j < TS ? cfa[(rr - 1 + j) * TS + cc] : 0.0f;
        float A[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) A[k] = alt_at(w + k, rr + 2 + k, green != ((k & 1) != 0));
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float u2 = U[k], a2 = A[k], c1 = w[k + 2];
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

// green.rs:184-226, horizontal half: hcd is bounded in place along each row.
// One thread per tile row walks it in 32-column chunks: each chunk's hcd
// (with the 4-column halo on the left, 8 on the right) and CFA samples (8
// and 8) arrive as aligned float4 loads straight into registers, hcdalt is
// formed from the CFA samples, the serial recurrence runs fully unrolled
// over them, and hcd leaves as float4 stores. cddiffsq (green.rs:219-223)
// is formed where it is read, by the Nyquist test, from the bounded vcd and
// hcd.
constexpr int BH_ROWS = 64, BH_CHUNK = 32, BH_IN = BH_CHUNK + 8, BH_CF = BH_CHUNK + 16;

__global__ __launch_bounds__(64) void amz_bound_h(AmazeArgs a, float clip_pt) {
    const uint3 blk = xcd_block();
    const Tile t = tile_of(a, blk.z);
    const int rr = 4 + blk.x * BH_ROWS + (int)threadIdx.x;
    if (rr >= TS - 4) return;
    float* __restrict__ hcd_row = t.full(HCD) + rr * TS;
    const float* __restrict__ cfa_row = t.full(CFA) + rr * TS;
    const bool green4 = (fc(a.pattern, rr, 4) & 1) == 1;  // colour at column 4 (and c0)
    float b_m2 = 0.0f, b_m1 = 0.0f;  // bounded hcd two and one columns left (cols 2, 3 are zero)
    for (int c0 = 4; c0 < TS - 4; c0 += BH_CHUNK) {
        // h[k] and al[k] hold column c0 - 4 + k, cf[k] column c0 - 8 + k.
        // hcd and hcdalt read as zero outside [4, TS - 4), the CFA as zero
        // outside the tile.
        float h[BH_IN], al[BH_IN], cf[BH_CF];
#pragma unroll
        for (int q = 0; q < BH_IN / 4; ++q) {
            const int col = c0 - 4 + 4 * q;
            float4 hv = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (col < TS) hv = *reinterpret_cast<const float4*>(hcd_row + col);
            const float hs[4] = {hv.x, hv.y, hv.z, hv.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) h[4 * q + e] = col + e >= 4 && col + e < TS - 4 ? hs[e] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < BH_CF / 4; ++q) {
            const int col = c0 - 8 + 4 * q;
            float4 cv = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (col >= 0 && col < TS) cv = *reinterpret_cast<const float4*>(cfa_row + col);
            cf[4 * q] = cv.x;
            cf[4 * q + 1] = cv.y;
            cf[4 * q + 2] = cv.z;
            cf[4 * q + 3] = cv.w;
        }
        // hcdalt at columns c0 - 2 .. c0 + 33, the ones the chunk reads.
#pragma unroll
        for (int j = 2; j < BH_CHUNK + 6; ++j) {
            const int col = c0 - 4 + j;
            al[j] = col >= 4 && col < TS - 4
                        ? alt_difference([&](int d) { return cf[j + 4 + d]; }, green4 != ((j & 1) != 0))
                        : 0.0f;
        }
        float out[BH_CHUNK];
#pragma unroll
        for (int k = 0; k < BH_CHUNK; ++k) {
            const bool green = green4 != ((k & 1) != 0);
            const float hcdvar = spread(b_m2, h[k + 4], h[k + 6]);
            const float hcdaltvar = spread(al[k + 2], al[k + 4], al[k + 6]);
            const float hv = hcdaltvar < hcdvar ? al[k + 4] : h[k + 4];
            const float bounded = green ? bound_at_green(hv, cf[k + 8], cf[k + 7], cf[k + 9], clip_pt)
                                        : bound_at_red_blue(hv, cf[k + 8], cf[k + 7], cf[k + 9], clip_pt);
            out[k] = bounded;
            b_m2 = b_m1;
            b_m1 = bounded;
        }
        // The last chunk holds 24 columns (TS - 8 = 248 = 7 * 32 + 24).
#pragma unroll
        for (int q = 0; q < BH_CHUNK / 4; ++q) {
            if (c0 + 4 * q < TS - 4) {
                *reinterpret_cast<float4*>(hcd_row + c0 + 4 * q) =
                    make_float4(out[4 * q], out[4 * q + 1], out[4 * q + 2], out[4 * q + 3]);
            }
        }
    }
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

// ---------------------------------------------------------------- weights, Nyquist test
// green.rs:229-292 (direction_weights) and nyquist.rs:25-89 (test and
// bounding box) for a WH x WW block of a tile, one thread per red/blue site
// (cc = 2 * pair + green_at_even(row); a wavefront holds 64 consecutive pairs
// of a row). The CFA, the bounded colour differences and their gradients
// are staged in LDS with the halo the stencils need (float4 loads of whole
// plane rows); the directional gradients and delhvsqsum (dhv) are formed
// from the CFA samples there, cddiffsq from the bounded colour differences.
// LDS rows are stored split by column parity (even columns, then odd), so
// the lanes of a wavefront, whose sites are two columns apart, read
// consecutive words. The arithmetic of each site is the reference's, term
// for term.
constexpr int WW = 128, WH = 8;
constexpr int WS_W = WW + 8;                 // staged columns c0 - 4 .. c0 + WW + 4
constexpr int WC_H = WH + 6;                 // CFA and vcd rows r0 - 3 .. r0 + WH + 3
constexpr int WK_H = WH + 4;                 // hcd and dgv rows r0 - 2 .. r0 + WH + 2
constexpr int WD_W = WW + 4, WD_H = WH + 4;  // dhv: columns c0 - 2 .., rows r0 - 2 ..
constexpr int WQ_W = WD_W / 2;               // cddiffsq: the red/blue sites of those columns
// One wavefront per block row: eight wavefronts, so the three blocks whose
// staging fits a CU's LDS keep six wavefronts per SIMD.
constexpr int W_WAVES = WH, W_THREADS = 64 * W_WAVES;
static_assert(TS % WW == 0 && TS % WH == 0 && WW == 128, "whole blocks");

struct alignas(16) WeightsLds {
    float cfa[WC_H * WS_W];
    float vcd[WC_H * WS_W];
    float hcd[WK_H * WS_W];
    float dgv[WK_H * WW];
    float dgh[WH * WS_W];
    float dhv[WD_H * WD_W];
    float cds[WD_H * WQ_W];
};

// Word of column lc (counted from the row's first staged column, which is
// even) in an LDS row of `width` words split by parity.
__device__ __forceinline__ int split_at(int width, int lc) { return (lc & 1) * (width / 2) + (lc >> 1); }

// Rows [row0, row0 + rows) of a tile plane, columns [col0, col0 + 4 * quads)
// (col0 a multiple of 4), into parity-split LDS rows of 4 * quads words;
// zero outside the tile.
__device__ __forceinline__ void stage_rows(float* dst, const float* plane, int row0, int rows,
                                           int col0, int quads) {
    for (int k = threadIdx.x; k < rows * quads; k += W_THREADS) {
        const int ly = k / quads, q = k - ly * quads;
        const int rr = row0 + ly, cc = col0 + 4 * q;
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (rr >= 0 && rr < TS && cc >= 0 && cc < TS)
            v = *reinterpret_cast<const float4*>(plane + rr * TS + cc);
        float* row = dst + ly * 4 * quads;
        *reinterpret_cast<float2*>(row + 2 * q) = make_float2(v.x, v.z);
        *reinterpret_cast<float2*>(row + 2 * quads + 2 * q) = make_float2(v.y, v.w);
    }
}

// One red/blue site of a weights block: row ly of the block, column
// c0 + 2 * lane + PAR. Every LDS word it reads sits at a fixed offset from
// the lane's own, so the offsets are immediates. Writes hvwt and returns
// the Nyquist flag.
template <int PAR>
__device__ __forceinline__ bool weights_site(const WeightsLds& s, float* hvwt, int ly, int lane) {
    constexpr int LS = 4 + PAR;  // staged column (from c0 - 4) of the site, less 2 * lane
    constexpr int LD = 2 + PAR;  // the same from c0 - 2
    auto at = [&](int dx, int half) { return ((LS + dx) & 1) * half + ((LS + dx) >> 1) + lane; };
    auto C = [&](int dy, int dx) { return s.cfa[(ly + 3 + dy) * WS_W + at(dx, WS_W / 2)]; };
    auto V = [&](int dy) { return s.vcd[(ly + 3 + dy) * WS_W + at(0, WS_W / 2)]; };
    auto H = [&](int dx) { return s.hcd[(ly + 2) * WS_W + at(dx, WS_W / 2)]; };
    auto G = [&](int dx) { return s.dgh[ly * WS_W + at(dx, WS_W / 2)]; };
    auto DV = [&](int dy) { return s.dgv[(ly + 2 + dy) * WW + PAR * (WW / 2) + lane]; };
    auto d = [&](int dy, int dx) {
        return s.dhv[(ly + 2 + dy) * WD_W + ((LD + dx) & 1) * (WD_W / 2) + ((LD + dx) >> 1) + lane];
    };
    auto c = [&](int dy, int dx) { return s.cds[(ly + 2 + dy) * WQ_W + ((LD + dx) >> 1) + lane]; };
    // HV weights (green.rs:229-292).
    float v[7], dv[5], hcd[7], dgh[5];
#pragma unroll
    for (int k = 0; k < 7; ++k) v[k] = V(k - 3);
#pragma unroll
    for (int k = 0; k < 5; ++k) dv[k] = DV(k - 2);
#pragma unroll
    for (int k = -3; k <= 3; ++k) hcd[k + 3] = H(k);
#pragma unroll
    for (int k = -2; k <= 2; ++k) dgh[k + 2] = G(k);
    const float w0u = vgrad(C, -1, 0), w0d = vgrad(C, 1, 0);
    const float w1l = hgrad(C, 0, -1), w1r = hgrad(C, 0, 1);
    float uave = v[3] + v[2] + v[1] + v[0];
    float dave = v[3] + v[4] + v[5] + v[6];
    float lave = hcd[3] + hcd[2] + hcd[1] + hcd[0];
    float rave = hcd[3] + hcd[4] + hcd[5] + hcd[6];
    float varu = sqr(v[3] - uave) + sqr(v[2] - uave) + sqr(v[1] - uave) + sqr(v[0] - uave);
    float vard = sqr(v[3] - dave) + sqr(v[4] - dave) + sqr(v[5] - dave) + sqr(v[6] - dave);
    float varl = sqr(hcd[3] - lave) + sqr(hcd[2] - lave) + sqr(hcd[1] - lave) + sqr(hcd[0] - lave);
    float varr = sqr(hcd[3] - rave) + sqr(hcd[4] - rave) + sqr(hcd[5] - rave) + sqr(hcd[6] - rave);
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
    *hvwt = ((0.5f - varwt) * (0.5f - diffwt) > 0.0f && fabsf(0.5f - diffwt) < fabsf(0.5f - varwt))
                ? varwt
                : diffwt;
    // Nyquist test statistic (nyquist.rs:28-57); cddiffsq at the red/blue
    // sites the stencil reads is (vcd - hcd)^2 of the bounded differences
    // (green.rs:219-223).
    float test =
        (GAUSSODD[0] * c(0, 0) + GAUSSODD[1] * (c(-1, -1) + c(-1, 1) + c(1, -1) + c(1, 1)) +
         GAUSSODD[2] * (c(-2, 0) + c(0, -2) + c(0, 2) + c(2, 0)) +
         GAUSSODD[3] * (c(-2, -2) + c(-2, 2) + c(2, -2) + c(2, 2))) -
        (GAUSSGRAD[0] * d(0, 0) + GAUSSGRAD[1] * (d(-1, 0) + d(0, 1) + d(0, -1) + d(1, 0)) +
         GAUSSGRAD[2] * (d(-1, -1) + d(-1, 1) + d(1, -1) + d(1, 1)) +
         GAUSSGRAD[3] * (d(-2, 0) + d(0, -2) + d(0, 2) + d(2, 0)) +
         GAUSSGRAD[4] * (d(-2, -1) + d(-2, 1) + d(-1, -2) + d(-1, 2) + d(1, -2) + d(1, 2) +
                         d(2, -1) + d(2, 1)) +
         GAUSSGRAD[5] * (d(-2, -2) + d(-2, 2) + d(2, -2) + d(2, 2)));
    return test > 0.0f;
}

__global__ __launch_bounds__(W_THREADS) void amz_weights(AmazeArgs a) {
    const uint3 blk = xcd_block();
    const Tile t = tile_of(a, blk.z);
    const int c0 = blk.x * WW, r0 = blk.y * WH;
    __shared__ WeightsLds s;
    stage_rows(s.cfa, t.full(CFA), r0 - 3, WC_H, c0 - 4, WS_W / 4);
    stage_rows(s.vcd, t.full(VCD), r0 - 3, WC_H, c0 - 4, WS_W / 4);
    stage_rows(s.hcd, t.full(HCD), r0 - 2, WK_H, c0 - 4, WS_W / 4);
    stage_rows(s.dgv, t.full(DGV), r0 - 2, WK_H, c0, WW / 4);
    stage_rows(s.dgh, t.full(DGH), r0, WH, c0 - 4, WS_W / 4);
    __syncthreads();
    auto C = [&](int rr, int cc) { return s.cfa[(rr - r0 + 3) * WS_W + split_at(WS_W, cc - c0 + 4)]; };
    auto V = [&](int rr, int cc) { return s.vcd[(rr - r0 + 3) * WS_W + split_at(WS_W, cc - c0 + 4)]; };
    auto H = [&](int rr, int cc) { return s.hcd[(rr - r0 + 2) * WS_W + split_at(WS_W, cc - c0 + 4)]; };
    // dhv over [2, TS - 2) and cddiffsq at the red/blue sites; zero
    // elsewhere, where no valid site's stencil reaches. Lanes take columns
    // c0 - 2 + lane + 64 m, wavefronts rows.
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    static_assert(WD_W > 128 && WD_W <= 192, "three column slots per lane");
    for (int ly = wave; ly < WD_H; ly += W_WAVES) {
        const int rr = r0 - 2 + ly;
        const int par = gae(a.pattern, rr);
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            if (m == 2 && lane >= WD_W - 128) break;
            const int lc = lane + 64 * m, cc = c0 - 2 + lc;
            const bool inside = rr >= 2 && rr < TS - 2 && cc >= 2 && cc < TS - 2;
            s.dhv[ly * WD_W + split_at(WD_W, lc)] = inside ? delhv(C, rr, cc) : 0.0f;
            if ((cc & 1) == par)
                s.cds[ly * WQ_W + (lc >> 1)] = inside ? sqr(V(rr, cc) - H(rr, cc)) : 0.0f;
        }
    }
    __syncthreads();
    const int p = blk.x * 64 + lane;
    {
        const int ly = wave, rr = r0 + ly;
        const int par = gae(a.pattern, rr);
        const int cc = 2 * p + par;
        bool flagged = false;
        if (rr >= 6 && rr < TS - 6 && cc >= 6 && cc < TS - 6) {
            float* hv = t.half(HVWT) + rr * TSH + p;
            flagged = par ? weights_site<1>(s, hv, ly, lane) : weights_site<0>(s, hv, ly, lane);
            if (flagged) t.nyq()[rr * TSH + p] = 1;
        }
        // The row's bounding box of flagged sites (nyquist.rs:59-89): cc
        // increases with the lane, so the flagged sites span [cc of the first
        // flagged lane, cc of the last]. Every wavefront row writes its slot
        // (empty: INT_MAX, -1, INT_MAX, -1); amz_box reduces the slots, so no
        // two rows touch the same address.
        const unsigned long long m = __ballot(flagged);
        int4 box = make_int4(INT_MAX, -1, INT_MAX, -1);
        if (m) {
            const int first = __ffsll((long long)m) - 1, last = 63 - __clzll((long long)m);
            box = make_int4(rr, rr, __shfl(cc, first, 64), __shfl(cc, last, 64));
        }
        if (lane == 0) reinterpret_cast<int4*>(t.boxes())[rr * 2 + blk.x] = box;
    }
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

// The Nyquist vote and area blocks: AR_BLOCK_ROWS rows x 64 pairs of a
// tile, so that a launch is a few workgroups per CU rather than dozens of
// short ones.
constexpr int AR_BLOCK_ROWS = 16;
static_assert(TS % AR_BLOCK_ROWS == 0 && AR_BLOCK_ROWS % 4 == 0, "whole blocks");

// nyquist.rs:100-118: neighbourhood vote.
__global__ __launch_bounds__(256) void amz_vote(AmazeArgs a) {
    const uint3 blk = xcd_block();
    Area ar;
    if (!tile_area(a, blk.z, ar)) return;
    const Tile t = tile_of(a, blk.z);
    const int p = blk.x * 64 + (threadIdx.x & 63);
    const uint8_t* n = t.nyq();
#pragma unroll
    for (int k = 0; k < AR_BLOCK_ROWS / 4; ++k) {
        const int rr = blk.y * AR_BLOCK_ROWS + 4 * k + (threadIdx.x >> 6);
        const int cc = 2 * p + gae(a.pattern, rr);
        if (!in_area(ar, a.pattern, rr, cc)) continue;
        const int i = rr * TS + cc;
        unsigned votes = n[(i - 2 * TS) >> 1] + n[(i - TS - 1) >> 1] + n[(i - TS + 1) >> 1] +
                         n[(i - 2) >> 1] + n[(i + 2) >> 1] + n[(i + TS - 1) >> 1] +
                         n[(i + TS + 1) >> 1] + n[(i + 2 * TS) >> 1];
        t.nyq2()[i >> 1] = votes >= 5 ? 1 : (votes == 4 ? n[i >> 1] : 0);
    }
}

// nyquist.rs:120-152: area interpolation of hvwt at the sites the vote
// kept (about a tenth of them). A workgroup (AR_BLOCK_ROWS rows x 64 pairs)
// first lists its kept sites in LDS; when it has any, it stages the CFA
// samples and the vote flags its 7x7 stencils reach, then its lanes take the
// listed sites densely, so the gathers run from LDS in full wavefronts; each
// site's result is independent of the order.
//   CFA: rows r0 - 7 .. r0 + AR_BLOCK_ROWS + 6, columns c_base .. c_base + 143
//        (c_base = 128 blk.x - 8): the taps and their four neighbours.
//   flags: rows r0 - 6 .. r0 + AR_BLOCK_ROWS + 5, pairs p_base .. p_base + 71
//        (p_base = 64 blk.x - 4).
// Both read as zero outside the tile, where no valid site's stencil reaches.
constexpr int AR_ROWS = AR_BLOCK_ROWS + 14, AR_COLS = 144;
constexpr int AR_FROWS = AR_BLOCK_ROWS + 12, AR_FPAIRS = 72;
constexpr int AR_SITES = AR_BLOCK_ROWS * 64;

__global__ __launch_bounds__(256) void amz_area(AmazeArgs a) {
    const uint3 blk = xcd_block();
    Area ar;
    if (!tile_area(a, blk.z, ar)) return;
    const Tile t = tile_of(a, blk.z);
    __shared__ int s_sites[AR_SITES];
    __shared__ int s_count;
    __shared__ __align__(16) float s_cfa[AR_ROWS * AR_COLS];
    __shared__ __align__(4) uint8_t s_flag[AR_FROWS * AR_FPAIRS];
    if (threadIdx.x == 0) s_count = 0;
    __syncthreads();
    const int r0 = blk.y * AR_BLOCK_ROWS;
    {
        const int p = blk.x * 64 + (threadIdx.x & 63);
        bool kept[AR_BLOCK_ROWS / 4];
#pragma unroll
        for (int k = 0; k < AR_BLOCK_ROWS / 4; ++k) {
            const int rr = r0 + 4 * k + (threadIdx.x >> 6);
            const int cc = 2 * p + gae(a.pattern, rr);
            kept[k] = in_area(ar, a.pattern, rr, cc) && t.nyq2()[(rr * TS + cc) >> 1] != 0;
        }
#pragma unroll
        for (int k = 0; k < AR_BLOCK_ROWS / 4; ++k) {
            const int rr = r0 + 4 * k + (threadIdx.x >> 6);
            if (kept[k]) s_sites[atomicAdd(&s_count, 1)] = rr * TS + 2 * p + gae(a.pattern, rr);
        }
    }
    __syncthreads();
    const int n = s_count;
    if (n == 0) return;
    const int c_base = blk.x * 128 - 8, p_base = blk.x * 64 - 4;
    const float* cfa = t.full(CFA);
    for (int k = threadIdx.x; k < AR_ROWS * (AR_COLS / 4); k += 256) {
        const int ly = k / (AR_COLS / 4), q = k - ly * (AR_COLS / 4);
        const int rr = r0 - 7 + ly, cc = c_base + 4 * q;
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (rr >= 0 && rr < TS && cc >= 0 && cc < TS)
            v = *reinterpret_cast<const float4*>(cfa + rr * TS + cc);
        *reinterpret_cast<float4*>(s_cfa + ly * AR_COLS + 4 * q) = v;
    }
    const uint8_t* n2 = t.nyq2();
    for (int k = threadIdx.x; k < AR_FROWS * (AR_FPAIRS / 4); k += 256) {
        const int ly = k / (AR_FPAIRS / 4), q = k - ly * (AR_FPAIRS / 4);
        const int rr = r0 - 6 + ly, pp = p_base + 4 * q;
        unsigned v = 0;
        if (rr >= 0 && rr < TS && pp >= 0 && pp < TSH)
            v = *reinterpret_cast<const unsigned*>(n2 + rr * TSH + pp);
        *reinterpret_cast<unsigned*>(s_flag + ly * AR_FPAIRS + 4 * q) = v;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += 256) {
        const int i = s_sites[k];
        const int rr = i / TS, cc = i % TS;
        // Row by row of taps: the row's flags, then the flagged taps' samples,
        // then the sums in the reference's order.
        float sumcfa = 0.0f, sumh = 0.0f, sumv = 0.0f, sumsqh = 0.0f, sumsqv = 0.0f, areawt = 0.0f;
        for (int di = -6; di < 7; di += 2) {
            const uint8_t* frow = s_flag + (rr + di - (r0 - 6)) * AR_FPAIRS - p_base;
            const float* crow = s_cfa + (rr + di - (r0 - 7)) * AR_COLS - c_base;
            bool f[7];
            float c[7], l[7], r[7], u[7], d[7];
#pragma unroll
            for (int j = 0; j < 7; ++j) f[j] = frow[(cc - 6 + 2 * j) >> 1] != 0;
#pragma unroll
            for (int j = 0; j < 7; ++j) {
                const int col = cc - 6 + 2 * j;
                c[j] = l[j] = r[j] = u[j] = d[j] = 0.0f;
                if (f[j]) {
                    c[j] = crow[col];
                    l[j] = crow[col - 1];
                    r[j] = crow[col + 1];
                    u[j] = crow[col - AR_COLS];
                    d[j] = crow[col + AR_COLS];
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

// ---------------------------------------------------------------- back end
// nyquist.rs:159-201 and diagonal.rs:149-168 (the row chains: hvwt in
// populate_green, pmwt in combine_diagonals), nyquist.rs:204-236 and
// diagonal.rs:193-247 (refine_green, refine_green_with_rb), chroma.rs:26-96
// (the G-B split and the diagonal chroma interpolation) and chroma.rs:103-147
// with mod.rs:196-273 (finish_planes and write_tile), for one tile in one
// workgroup as a row pipeline. Four groups of two wavefronts (thread p of a
// group owns pair p) work on different rows at step s:
//   phase 1: group 0 runs the row chains on row s, group 1 refines row
//            s - 3, group 2 interpolates chroma on row s - 7, group 3
//            finishes and writes row s - 9,
//   barrier,
//   phase 2: group 0 passes row s's weights on; the first two thread halves
//            store their share of the inputs of row s + 1 (loaded eight
//            steps earlier),
//   barrier.
// Each stage reads only rows that the stages before it completed in earlier
// steps, so every value equals the one the stage sequence produces. The
// rows still in use live in LDS rings; nothing but the RGB output goes back
// to HBM.
constexpr int BACK_THREADS = 512, BACK_AHEAD = 8;
constexpr int LAG_REFINE = 3, LAG_CHROMA = 7, LAG_FINISH = 9;
// Ring depths: the span of rows alive at once, from the lags above.
constexpr int R_LONG = 12;  // cfa, final hvwt, dgrb0, dgrb1, green
constexpr int R_MID = 6;    // d2h, d2v, rbint
constexpr int R_SHORT = 4;  // original hvwt / pmwt, final pmwt, vcd, hcd, Nyquist flags
constexpr int R_RB = 2;     // rbm, rbp
static_assert((TS - 16) % BACK_AHEAD == 0, "whole groups of steps");

struct BackLds {
    float cfa[R_LONG][TS];
    float hv[R_LONG][TSH], g0[R_LONG][TSH], g1[R_LONG][TSH], grb[R_LONG][TSH];
    float d2h[R_MID][TSH], d2v[R_MID][TSH], rbint[R_MID][TSH];
    float hvo[R_SHORT][TSH], pmo[R_SHORT][TSH], pm[R_SHORT][TSH];
    float vcd[R_SHORT][TSH], hcd[R_SHORT][TSH];
    float rbm[R_RB][TSH], rbp[R_RB][TSH];
    uint32_t nyq[R_SHORT][TSH / 4];
    float hv_prev[TSH], pm_prev[TSH];
};

// The inputs of step X that a thread loads ahead: its share of CFA row X + 1,
// of original hvwt / pmwt row X + 1, of rbm / rbp row X, of vcd / hcd at the
// red/blue sites of row X and of the Nyquist flags of row X.
struct BackIn {
    float cfa, w, rb, d;
    uint32_t nyq;
};

__device__ __forceinline__ int ring(int row, int depth) { return row % depth; }

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

__device__ __forceinline__ void back_load(const Tile& t, int pattern, int x, BackIn& in) {
    const int tid = threadIdx.x, p = tid & (TSH - 1);
    const bool lo = tid < TSH;
    in.cfa = t.full(CFA)[(x + 1) * TS + tid];
    in.w = (lo ? t.half(HVWT) : t.half(PMWT))[(x + 1) * TSH + p];
    in.rb = (lo ? t.half(RBM) : t.half(RBP))[x * TSH + p];
    in.d = t.full(lo ? VCD : HCD)[x * TS + 2 * p + gae(pattern, x)];
    in.nyq = tid < TSH / 4 ? reinterpret_cast<const uint32_t*>(t.nyq2() + x * TSH)[tid] : 0u;
}
__device__ __forceinline__ void back_store(BackLds& s, int x, const BackIn& in) {
    const int tid = threadIdx.x, p = tid & (TSH - 1);
    const bool lo = tid < TSH;
    s.cfa[ring(x + 1, R_LONG)][tid] = in.cfa;
    (lo ? s.hvo : s.pmo)[ring(x + 1, R_SHORT)][p] = in.w;
    (lo ? s.rbm : s.rbp)[ring(x, R_RB)][p] = in.rb;
    (lo ? s.vcd : s.hcd)[ring(x, R_SHORT)][p] = in.d;
    if (tid < TSH / 4) s.nyq[ring(x, R_SHORT)][tid] = in.nyq;
}
__device__ __forceinline__ bool nyq_at(const BackLds& s, int row, int p) {
    return ((s.nyq[ring(row, R_SHORT)][p >> 2] >> (8 * (p & 3))) & 0xffu) != 0;
}

// nyquist.rs:159-201 and diagonal.rs:149-168 for pair p of row rr.
__device__ __forceinline__ void back_chains(BackLds& s, int pattern, int rr, int p, float& hv_out,
                                            float& pm_out) {
    const int cc = 2 * p + gae(pattern, rr);
    const int rc = ring(rr, R_LONG), rs = ring(rr, R_SHORT), rn = ring(rr + 1, R_SHORT);
    const int dl = max((cc - 1) >> 1, 0), dr = min((cc + 1) >> 1, TSH - 1);
    const int cl = min(max(cc, 1), TS - 2);
    float hv = s.hvo[rs][p];
    float d2h = 0.0f, d2v = 0.0f;
    if (cc >= 8 && cc < TS - 8) {
        float alt = (s.hv_prev[(cc - 1) >> 1] + s.hv_prev[(cc + 1) >> 1] + s.hvo[rn][dl] +
                     s.hvo[rn][dr]) /
                    4.0f;
        if (fabsf(0.5f - hv) < fabsf(0.5f - alt)) hv = alt;
        const float cfa = s.cfa[rc][cl];
        float dg = intp(hv, s.vcd[rs][p], s.hcd[rs][p]);
        float green = cfa + dg;
        s.g0[rc][p] = dg;
        s.grb[rc][p] = green;
        if (nyq_at(s, rr, p)) {
            d2h = sqr(green - 0.5f * (s.cfa[rc][cl - 1] + s.cfa[rc][cl + 1]));
            d2v = sqr(green - 0.5f * (s.cfa[ring(rr - 1, R_LONG)][cl] + s.cfa[ring(rr + 1, R_LONG)][cl]));
        }
    }
    const int rm = ring(rr, R_MID);
    s.d2h[rm][p] = d2h;
    s.d2v[rm][p] = d2v;
    s.hv[rc][p] = hv;
    hv_out = hv;
    float pm = 0.0f;
    if (rr >= 10 && rr < TS - 10) {
        pm = s.pmo[rs][p];
        if (cc >= 10 && cc < TS - 10) {
            float alt = (s.pm_prev[(cc - 1) >> 1] + s.pm_prev[(cc + 1) >> 1] + s.pmo[rn][dl] +
                         s.pmo[rn][dr]) /
                        4.0f;
            if (fabsf(0.5f - pm) < fabsf(0.5f - alt)) pm = alt;
            s.rbint[rm][p] = 0.5f * (s.cfa[rc][cc] + s.rbm[ring(rr, R_RB)][p] * (1.0f - pm) +
                                     s.rbp[ring(rr, R_RB)][p] * pm);
        }
        s.pm[rs][p] = pm;
    }
    pm_out = pm;
}

// nyquist.rs:204-236, diagonal.rs:193-247 and the G-B split of
// chroma.rs:26-32, in that order, on pair p of row rr.
__device__ __forceinline__ void back_refine(BackLds& s, int pattern, int rr, int p, bool has_area,
                                            const Area& ar, float clip_pt) {
    const int cc = 2 * p + gae(pattern, rr);
    const int rc = ring(rr, R_LONG);
    auto F = [&](int r, int c) { return s.cfa[ring(r, R_LONG)][c]; };
    if (has_area && in_area(ar, pattern, rr, cc) && nyq_at(s, rr, p)) {
        auto pl = [&](const float (*plane)[TSH], int dr, int dc) {
            return plane[ring(rr + dr, R_MID)][(cc + dc) >> 1];
        };
        auto average = [&](const float (*plane)[TSH]) {
            return EPSSQ + (GQUINC[0] * pl(plane, 0, 0) +
                            GQUINC[1] * (pl(plane, -1, -1) + pl(plane, -1, 1) + pl(plane, 1, -1) +
                                         pl(plane, 1, 1)) +
                            GQUINC[2] * (pl(plane, -2, 0) + pl(plane, 0, -2) + pl(plane, 0, 2) +
                                         pl(plane, 2, 0)) +
                            GQUINC[3] * (pl(plane, -2, -2) + pl(plane, -2, 2) + pl(plane, 2, -2) +
                                         pl(plane, 2, 2)));
        };
        float gvarh = average(s.d2h);
        float gvarv = average(s.d2v);
        const int rs = ring(rr, R_SHORT);
        float dg = (s.hcd[rs][p] * gvarv + s.vcd[rs][p] * gvarh) / (gvarv + gvarh);
        s.g0[rc][p] = dg;
        s.grb[rc][p] = F(rr, cc) + dg;
    }
    if (rr >= 12 && rr < TS - 12 && cc >= 12 && cc < TS - 12) {
        const float pm = s.pm[ring(rr, R_SHORT)][p], hv = s.hv[rc][p];
        if (!(fabsf(0.5f - pm) < fabsf(0.5f - hv))) {
            const int rm = ring(rr, R_MID);
            const float centre = s.rbint[rm][p];
            float gu = green_from_rb(F(rr - 1, cc), centre, s.rbint[ring(rr - 2, R_MID)][p]);
            float gd = green_from_rb(F(rr + 1, cc), centre, s.rbint[ring(rr + 2, R_MID)][p]);
            float gl = green_from_rb(F(rr, cc - 1), centre, s.rbint[rm][p - 1]);
            float gr = green_from_rb(F(rr, cc + 1), centre, s.rbint[rm][p + 1]);
            const float w0u = vgrad(F, rr - 1, cc), w0d = vgrad(F, rr + 1, cc);
            const float w1l = hgrad(F, rr, cc - 1), w1r = hgrad(F, rr, cc + 1);
            float gintv = (w0u * gd + w0d * gu) / (w0d + w0u);
            float ginth = (w1l * gr + w1r * gl) / (w1l + w1r);
            gintv = bound_towards(gintv, centre, F(rr - 1, cc), F(rr + 1, cc), clip_pt);
            ginth = bound_towards(ginth, centre, F(rr, cc - 1), F(rr, cc + 1), clip_pt);
            float green = ginth * (1.0f - hv) + gintv * hv;
            s.grb[rc][p] = green;
            s.g0[rc][p] = green - F(rr, cc);
        }
    }
    // Split G - B out of the combined colour difference on blue rows.
    int ey, ex;
    red_offset(pattern, ey, ex);
    if (rr >= 13 - ey && rr < TS - 12 && ((rr - (13 - ey)) & 1) == 0 && p >= ((13 - ex) >> 1) &&
        p < ((TS - 12) >> 1)) {
        s.g1[rc][p] = s.g0[rc][p];
        s.g0[rc][p] = 0.0f;
    }
}

// chroma.rs:34-96 on pair p of row rr.
__device__ __forceinline__ void back_chroma(BackLds& s, int pattern, int rr, int p) {
    if (rr < 14 || rr >= TS - 14) return;
    const int start = 14 + gae(pattern, rr);
    const int cc = 2 * p + gae(pattern, rr);
    if (cc < start || cc >= TS - 14) return;
    const int plane = 1 - fc(pattern, rr, start) / 2;
    float (*pl)[TSH] = plane ? s.g1 : s.g0;
    auto d = [&](int dr, int dc) { return pl[ring(rr + dr, R_LONG)][(cc + dc) >> 1]; };
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
    pl[ring(rr, R_LONG)][cc >> 1] = value;
}

// chroma.rs:103-147 (finish_planes) with mod.rs:196-273 (write_tile) on the
// two pixels of pair p of row rr, inside the tile core.
__device__ __forceinline__ void back_finish(const BackLds& s, const AmazeArgs& a, const Tile& t,
                                            int rr, int p) {
    if (rr < BORDER || rr >= TS - BORDER) return;
    const int tx = t.index % a.tiles_x, ty = t.index / a.tiles_x;
    const int gy = ty * VALID + rr - BORDER;
    if (gy >= a.height) return;
    const int ru = ring(rr - 1, R_LONG), rc = ring(rr, R_LONG), rd = ring(rr + 1, R_LONG);
    const size_t n = (size_t)a.width * a.height;
    bool finite = true;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const int cc = 2 * p + e;
        const int gx = tx * VALID + cc - BORDER;
        if (cc < BORDER || cc >= TS - BORDER || gx >= a.width) continue;
        const int ch = fc(a.pattern, rr, cc);
        const float sample = s.cfa[rc][cc];
        float green, red, blue;
        if (ch == 1) {
            green = sample;
            const int h = cc >> 1, hl = (cc - 1) >> 1, hr = (cc + 1) >> 1;
            float up = s.hv[ru][h], down = s.hv[rd][h];
            float left = s.hv[rc][hl], right = s.hv[rc][hr];
            float temp = 1.0f / (up + 2.0f - right - left + down);
            float br = (up * s.g0[ru][h] + (1.0f - right) * s.g0[rc][hr] +
                        (1.0f - left) * s.g0[rc][hl] + down * s.g0[rd][h]) *
                       temp;
            float bb = (up * s.g1[ru][h] + (1.0f - right) * s.g1[rc][hr] +
                        (1.0f - left) * s.g1[rc][hl] + down * s.g1[rd][h]) *
                       temp;
            red = sel_max(green - br, 0.0f);
            blue = sel_max(green - bb, 0.0f);
        } else {
            green = s.grb[rc][cc >> 1];
            red = sel_max(green - s.g0[rc][cc >> 1], 0.0f);
            blue = sel_max(green - s.g1[rc][cc >> 1], 0.0f);
        }
        float rgb[3] = {ch == 0 ? sample : red, ch == 1 ? sample : sel_max(green, 0.0f),
                        ch == 2 ? sample : blue};
        rgb[0] *= a.inverse[0];
        rgb[1] *= a.inverse[1];
        rgb[2] *= a.inverse[2];
        const size_t o = (size_t)gy * a.width + gx;
        a.out[o] = rgb[0];
        a.out[n + o] = rgb[1];
        a.out[2 * n + o] = rgb[2];
        if (a.guide) {
            const CaptureInputs in =
                capture_inputs_fast(rgb[0], rgb[1], rgb[2], a.ceilings, a.ceilings_inv);
            a.guide[o] = in.guide;
            a.mask[o] = in.mask;
        }
        finite = finite && is_finite(rgb[0]) && is_finite(rgb[1]) && is_finite(rgb[2]);
    }
    // Bit 2 flags the capture input as well when its guide is made here.
    if (!finite) atomicOr(&a.stats->nonfinite, a.guide ? 6u : 2u);
}

__global__ __launch_bounds__(BACK_THREADS) void amz_back(AmazeArgs a, float clip_pt) {
    static_assert(BACK_THREADS == 4 * TSH && TS == 2 * TSH, "four groups of a thread per pair");
    const uint3 blk = xcd_block();
    const Tile t = tile_of(a, blk.x);
    __shared__ BackLds s;
    const int tid = threadIdx.x, p = tid & (TSH - 1);
    // Wavefront w issues on SIMD w % 4, so the thread halves 0 and 2 share two
    // SIMDs and halves 1 and 3 the other two. The halves take the stages in
    // the order chains, chroma, refinements, finish: the light chains share
    // their SIMDs with the heavy refinements, chroma with the finish.
    const int half = __builtin_amdgcn_readfirstlane(tid / TSH);
    constexpr int STAGE_OF_HALF[4] = {0, 2, 1, 3};
    const int group = STAGE_OF_HALF[half];
    const bool loader = tid < TS;  // halves 0 and 1 bring the inputs in
    Area ar;
    const bool has_area = tile_area(a, blk.x, ar);
    // Rows before the pipeline: CFA rows 0 .. 8, original hvwt / pmwt row 8,
    // step 8's inputs; d2h and d2v read as zero until the chains write them.
    if (loader) {
        for (int r = 0; r <= 8; ++r) s.cfa[ring(r, R_LONG)][tid] = t.full(CFA)[r * TS + tid];
        (half == 0 ? s.hvo : s.pmo)[ring(8, R_SHORT)][p] =
            (half == 0 ? t.half(HVWT) : t.half(PMWT))[8 * TSH + p];
        for (int r = 0; r < R_MID; ++r) (half == 0 ? s.d2h : s.d2v)[r][p] = 0.0f;
        BackIn in;
        back_load(t, a.pattern, 8, in);
        back_store(s, 8, in);
    }
    if (group == 0) {
        s.hv_prev[p] = t.half(HVWT)[7 * TSH + p];
        s.pm_prev[p] = t.half(PMWT)[9 * TSH + p];
    }
    BackIn buf[BACK_AHEAD];
    if (loader) {
#pragma unroll
        for (int k = 1; k < BACK_AHEAD; ++k) back_load(t, a.pattern, 8 + k, buf[k]);
    }
    __syncthreads();
    // Phase 1 of step st (the chains only while st is a chain row).
    auto phase1 = [&](int st, float& hv, float& pm) {
        if (group == 0) {
            if (st < TS - 8) back_chains(s, a.pattern, st, p, hv, pm);
        } else if (group == 1) {
            back_refine(s, a.pattern, st - LAG_REFINE, p, has_area, ar, clip_pt);
        } else if (group == 2) {
            back_chroma(s, a.pattern, st - LAG_CHROMA, p);
        } else {
            back_finish(s, a, t, st - LAG_FINISH, p);
        }
    };
    for (int s0 = 8; s0 < TS - 8; s0 += BACK_AHEAD) {
#pragma unroll
        for (int k = 0; k < BACK_AHEAD; ++k) {
            const int st = s0 + k;
            // Inputs of step st + BACK_AHEAD (the last step's, re-read, past
            // the end) into the register slot step st's inputs left.
            if (loader) back_load(t, a.pattern, min(st + BACK_AHEAD, TS - 9), buf[k]);
            float hv = 0.0f, pm = 0.0f;
            phase1(st, hv, pm);
            __syncthreads();
            if (group == 0) {
                s.hv_prev[p] = hv;
                if (st >= 10 && st < TS - 10) s.pm_prev[p] = pm;
            }
            if (loader && st + 1 < TS - 8) back_store(s, st + 1, buf[(k + 1) % BACK_AHEAD]);
            __syncthreads();
        }
    }
    // One more step drains the later stages (the last chroma and finish rows).
    float hv = 0.0f, pm = 0.0f;
    phase1(TS - 8, hv, pm);
}

}  // namespace

size_t amaze_scratch_bytes_per_tile() { return TILE_FLOATS * sizeof(float) + 4 * sizeof(int); }

// mod.rs:136-161: the stage sequence for `tiles` tiles starting at
// a.tile_base; a.scratch and a.areas hold one slot per tile of this launch.
hipError_t launch_amaze(const AmazeArgs& a, int tiles, hipStream_t s, Profiler* prof) {
    const float clip_pt8 = 0.8f * a.clip_point;
    const dim3 blk(256);
    // Algorithmic bytes: the planes each stage must read and write once, per
    // tile (f = full-plane floats per tile, h = half-plane floats per tile).
    const uint64_t f = 4ull * FULL * tiles, h = 4ull * HALF * tiles;
    RHD_TIMED(prof, s, "amz_front", f + 5 * f + 3 * h + h / 2,
              hipLaunchKernelGGL(amz_front, dim3(TS / FW, TS / FH, tiles), dim3(F_THREADS), 0, s,
                                 a, clip_pt8, a.clip_point));
    RHD_TIMED(prof, s, "amz_bound_v", 2 * f + f,
              hipLaunchKernelGGL(amz_bound_v, dim3(TS / 64, 1, tiles), dim3(64), 0, s, a,
                                 a.clip_point));
    RHD_TIMED(prof, s, "amz_bound_h", 2 * f + f,
              hipLaunchKernelGGL(amz_bound_h, dim3((TS - 8 + BH_ROWS - 1) / BH_ROWS, 1, tiles),
                                 dim3(BH_ROWS), 0, s, a, a.clip_point));
    RHD_TIMED(prof, s, "amz_weights", 5 * f + h,
              hipLaunchKernelGGL(amz_weights, dim3(TS / WW, TS / WH, tiles), dim3(W_THREADS), 0,
                                 s, a));
    RHD_TIMED(prof, s, "amz_box", 16ull * BOX_SLOTS * tiles,
              hipLaunchKernelGGL(amz_box, dim3(tiles), blk, 0, s, a));
    const dim3 area_grid(TSH / 64, TS / AR_BLOCK_ROWS, tiles);
    RHD_TIMED(prof, s, "amz_vote", h / 2, hipLaunchKernelGGL(amz_vote, area_grid, blk, 0, s, a));
    RHD_TIMED(prof, s, "amz_area", h / 4, hipLaunchKernelGGL(amz_area, area_grid, blk, 0, s, a));
    // The back end reads CFA, vcd, hcd (full planes), hvwt, pmwt, rbm, rbp
    // and the Nyquist flags, and writes the RGB core of the tile.
    RHD_TIMED(prof, s, "amz_back",
              3 * f + 4 * h + h / 4 + 12ull * a.width * a.height * tiles / (a.tiles_x * a.tiles_y),
              hipLaunchKernelGGL(amz_back, dim3(tiles), dim3(BACK_THREADS), 0, s, a, a.clip_point));
    return hipGetLastError();
}

}  // namespace rhd
