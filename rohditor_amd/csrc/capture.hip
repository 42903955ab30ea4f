// Masked Richardson-Lucy capture sharpening on planar camera RGB.
//
// Reference: crates/core/src/sharpening.rs:87-127 (apply), :180-310 (solve)
// and :312-384 (smoothstep, mirror, blur); Gaussian weights from
// crates/detail/src/gaussian.rs:9-17 (computed on the host with the platform
// expf, as Rust's f32::exp does on Linux). Results are bit-identical to the
// reference: every blurred value is computed at its image coordinate with the
// reference's tap order (0.0 start, taps in ascending order, -0.0 start in
// the mirrored horizontal border) and half-sample mirroring.
//
// Kernels: init (guide, highlight mask; the AMaZE finish writes both instead
// when capture directly follows it), prep (the mask's two blurs and
// epilogues, and the first iteration's ratio from the same blur of the
// guide), one fused kernel per RL iteration (blur estimate -> ratio -> blur
// ratio -> clamped update, all in LDS for a 64x32 output tile with a 2r
// halo; the first starts from prep's ratio on 64x16 tiles), the last of
// which applies the
// gain to RGB instead of storing its estimate, and a plain gain apply for
// zero iterations. Iterations ping-pong between two estimate planes. Per
// iteration the estimate (with halo) and the guide are read and the estimate
// written; the iterations are bound by vector-ALU issue (the blur taps and
// the ratio's division), not by HBM.
// Tiles at the image edge run the same strip code as interior ones on
// windows staged half-sample mirrored, so no tap computes a mirrored index.
#include <cstdint>
#include <type_traits>

#include "rhd_internal.hpp"

namespace rhd {
namespace {

constexpr int TW = 64, TH = 16, MAXR = 4, THREADS = 256;
// The RL iteration's tiles are taller, with a wider workgroup: the vertical
// halos are paid per tile.
constexpr int ITH = 32, ITHREADS = 512;
constexpr float FLOOR = CAPTURE_FLOOR;

struct Weights {
    float w[2 * MAXR + 1];
};

// Guide values below this bound let the RL ratio take div_in_range.
constexpr float GUIDE_LIMIT = 0x1p64f;

// a / b, correctly rounded, for a in [2^-20, 2^64] and b in [2^-20, 2^66]:
// the reciprocal refined by one Newton-Raphson step, then two residual
// corrections of the quotient (Markstein). This is the division sequence the
// compiler emits for a / b without its operand scaling (which only acts when
// an operand, the quotient or the reciprocal is denormal, or the exponents
// differ by 96 or more) and its special-case fixup (zeros, infinities,
// NaNs, overflow), none of which these operands reach; the results are
// identical.
__device__ __forceinline__ float div_in_range(float a, float b) {
    const float y0 = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, y0, 1.0f);
    const float y1 = __builtin_fmaf(e, y0, y0);
    const float q0 = a * y1;
    const float r0 = __builtin_fmaf(-b, q0, a);
    const float q1 = __builtin_fmaf(r0, y1, q0);
    const float r1 = __builtin_fmaf(-b, q1, a);
    return __builtin_fmaf(r1, y1, q1);
}

// sharpening.rs:194-218
__global__ __launch_bounds__(THREADS) void cap_init(const float* __restrict__ rgb, int n,
                                                    float l0, float l1, float l2,
                                                    float* __restrict__ guide,
                                                    float* __restrict__ mask,
                                                    DeviceStats* stats) {
    int i = blockIdx.x * THREADS + threadIdx.x;
    if (i >= n) return;
    float r = rgb[i], g = rgb[n + i], b = rgb[2 * (size_t)n + i];
    if (!(is_finite(r) && is_finite(g) && is_finite(b))) atomicOr(&stats->nonfinite, 4u);
    const CaptureInputs in = capture_inputs(r, g, b, l0, l1, l2);
    guide[i] = in.guide;
    mask[i] = in.mask;
}

// The reference's sums start from 0.0 (interior, vertical) or -0.0 (the
// mirrored horizontal border). Every plane blurred here is >= +0 (the mask
// in [0, 1], the guide >= its floor, the estimate and the ratio positive),
// so the first product p is >= +0 and 0.0 + p == -0.0 + p == p exactly: the
// sums start from the first product instead, one add fewer per output.

// A (rows x cols) window of a global plane held in registers, one value per
// lane per slot. The trip counts are compile-time bounds (MAXROWS x MAXCOLS),
// so every load of the window is issued before the first LDS store and the
// memory latency is paid once. A wavefront stages whole rows, so each row's
// address is wave-uniform and formed in scalar registers.
template <int MAXROWS, int MAXCOLS, int NT = THREADS>
struct Staged {
    static constexpr int WAVES = NT / 64;
    static constexpr int RI = (MAXROWS + WAVES - 1) / WAVES, CI = (MAXCOLS + 63) / 64;
    float v[RI * CI];

    // The window whose first element is src (inside the image).
    __device__ __forceinline__ void load(const float* __restrict__ src, size_t src_pitch, int rows,
                                         int cols) {
        const int ty = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), tx = threadIdx.x & 63;
#pragma unroll
        for (int i = 0; i < RI; ++i) {
            const int r = ty + WAVES * i;
            const float* row = src + (size_t)r * src_pitch;
#pragma unroll
            for (int j = 0; j < CI; ++j) {
                const int c = tx + 64 * j;
                v[i * CI + j] = (r < rows && c < cols) ? row[c] : 0.0f;
            }
        }
    }
    // The window whose first element is image position (top, left), every
    // position half-sample mirrored into the image (sharpening.rs:318-330).
    __device__ __forceinline__ void load_mirrored(const float* __restrict__ plane, int width,
                                                  int height, int top, int left, int rows,
                                                  int cols) {
        const int ty = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), tx = threadIdx.x & 63;
        int col[CI];
#pragma unroll
        for (int j = 0; j < CI; ++j) col[j] = mirror_near(left + tx + 64 * j, width);
#pragma unroll
        for (int i = 0; i < RI; ++i) {
            const int r = ty + WAVES * i;
            const float* row = plane + (size_t)mirror_near(top + r, height) * width;
#pragma unroll
            for (int j = 0; j < CI; ++j) {
                const int c = tx + 64 * j;
                v[i * CI + j] = (r < rows && c < cols) ? row[col[j]] : 0.0f;
            }
        }
    }
    __device__ __forceinline__ void store(float* lds, int pitch, int rows, int cols) const {
        const int ty = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), tx = threadIdx.x & 63;
#pragma unroll
        for (int i = 0; i < RI; ++i)
#pragma unroll
            for (int j = 0; j < CI; ++j) {
                const int r = ty + WAVES * i, c = tx + 64 * j;
                if (r < rows && c < cols) lds[r * pitch + c] = v[i * CI + j];
            }
    }
};

// Copies the (ROWS x COLS) window at image position (top, left) into LDS,
// mirrored at the image edges when BORDER.
template <int ROWS, int COLS, bool BORDER, int NT = THREADS>
__device__ __forceinline__ void stage(float* lds, int pitch, const float* __restrict__ plane,
                                      int width, int height, int top, int left) {
    Staged<ROWS, COLS, NT> w;
    if (BORDER)
        w.load_mirrored(plane, width, height, top, left, ROWS, COLS);
    else
        w.load(plane + (size_t)top * width + left, width, ROWS, COLS);
    w.store(lds, pitch, ROWS, COLS);
}

// Window geometry for radius R. Blurs and the RL iteration compute strips of
// four outputs per thread (4 columns of a row in the horizontal passes, read
// as float4 chunks; 4 rows of a column in the vertical ones), so loads,
// addressing and loop control are shared by 4 results; every output keeps
// the reference's tap order. Strips that reach past a window's edge compute
// throw-away values from the padding and are not stored.
//
// The LDS arrays of the iteration begin at 16-byte aligned image columns, so
// interior windows move between HBM and LDS as float4 and every horizontal
// strip reads aligned chunks. For a tile at (x0, y0), LDS column 0 of
//   the estimate window (EH rows from y0 - 2r)   is image column x0 - EL,
//   its horizontal blur                          is x0 - EL + r,
//   the guide / ratio window (GH rows from y0 - r) is x0 - r,
//   the ratio's horizontal blur                  is x0.
// Interior guide and ratio windows load as float4 from the aligned column
// x0 - 4 and land GS columns left of their LDS origin: the columns left of
// the window fall in the previous row's padding (in front of the array for
// the first row), those right of it in the row's own padding.
template <int R, int TILE_H = TH>
struct Windows {
    static constexpr int EW = TW + 4 * R, EH = TILE_H + 4 * R;   // estimate (2r halo)
    static constexpr int GW = TW + 2 * R, GH = TILE_H + 2 * R;   // guide / ratio (r halo)
    static constexpr int CHUNKS = (4 + 2 * R + 3) / 4;      // float4 reads per strip
    static constexpr int EL = (2 * R + 3) / 4 * 4;          // estimate halo rounded up to 4
    static constexpr int EP = TW + 2 * EL;                  // estimate pitch
    static constexpr int ED = EL - 2 * R;                   // LDS column of image column x0 - 2r
    static constexpr int GS = 4 - R;                        // aligned load start, left of x0 - r
    static constexpr int P = 72;                            // guide / ratio and blur pitch
    static constexpr int NS1 = (TW + EL) / 4;               // estimate strips per row
    static constexpr int NS3 = TW / 4;                      // ratio strips per row
    static constexpr int E_FLOATS = EH * EP;
    static constexpr int H_FLOATS = (EH + 4) * P;
    static constexpr int G_FLOATS = GH * P + 4;
    // The estimate window, then the guide / ratio window, share one array.
    static constexpr int A_FLOATS = E_FLOATS > G_FLOATS ? E_FLOATS : G_FLOATS;
    static_assert(GW + GS <= P && ED + EW <= EP, "windows and their load padding fit the pitch");
    static_assert(4 * NS1 <= P && 4 * (NS1 - 1) + 4 * CHUNKS <= EP,
                  "estimate strips cover the blur the ratio needs and stay in a row");
    static_assert((GH - 1) * P + 4 * (NS3 - 1) + 4 * CHUNKS <= G_FLOATS,
                  "the last ratio row's strips stay in the array");
};

// A window of ROWS rows of 4 * F4 floats from a global plane, its first
// element 16-byte aligned, held in registers as float4 while earlier LDS work
// runs. Lane l of a wavefront takes chunk l % F4 of row l / F4 of each group of
// 64 / F4 rows the wavefront owns, so row addresses stay wave-uniform and a
// lane's offset within the group is formed once.
template <int ROWS, int F4, int NT = THREADS>
struct Staged4 {
    using vf4 = float __attribute__((ext_vector_type(4)));
    static constexpr int RPG = 64 / F4, GROUPS = (ROWS + RPG - 1) / RPG;
    static constexpr int WAVES = NT / 64, SLOTS = (GROUPS + WAVES - 1) / WAVES;
    vf4 v[SLOTS];

    __device__ __forceinline__ void load(const float* __restrict__ src, size_t pitch) {
        const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
        const int dr = l / F4, dc = l - dr * F4;
        const unsigned lane = (unsigned)(dr * pitch + 4 * dc);
#pragma unroll
        for (int k = 0; k < SLOTS; ++k) {
            const int g = w + WAVES * k;
            const float* group = src + (size_t)(g * RPG) * pitch;
            if (g < GROUPS && dr < RPG && g * RPG + dr < ROWS)
                v[k] = *reinterpret_cast<const vf4*>(group + lane);
        }
    }
    // Row r's chunk c goes to lds + r * pitch + 4 c - SHIFT (float4 stores
    // when SHIFT keeps them aligned, single words otherwise).
    template <int SHIFT = 0>
    __device__ __forceinline__ void store(float* lds, int pitch) const {
        const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
        const int dr = l / F4, dc = l - dr * F4;
#pragma unroll
        for (int k = 0; k < SLOTS; ++k) {
            const int g = w + WAVES * k, r = g * RPG + dr;
            if (!(g < GROUPS && dr < RPG && r < ROWS)) continue;
            float* dst = lds + r * pitch + 4 * dc - SHIFT;
            if constexpr (SHIFT % 4 == 0) {
                *reinterpret_cast<vf4*>(dst) = v[k];
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) dst[q] = v[k][q];
            }
        }
    }
};

// Calls f(row, col) for every item of a ROWS x COLS grid that this thread
// owns (items threadIdx.x + THREADS * t in row-major order). The row and
// column advance by THREADS per trip instead of being divided out of the item
// each time; only the last trip tests the item against the grid, and a
// wavefront whose first item is past the grid skips that trip as a whole.
template <int ROWS, int COLS, int NT = THREADS, typename F>
__device__ __forceinline__ void for_items(F f) {
    constexpr int N = ROWS * COLS, TRIPS = (N + NT - 1) / NT;
    constexpr int DR = NT / COLS, DC = NT % COLS;
    const int wave_first = __builtin_amdgcn_readfirstlane((int)threadIdx.x) & ~63;
    int r = (int)threadIdx.x / COLS, c = (int)threadIdx.x % COLS;
#pragma unroll
    for (int t = 0; t < TRIPS; ++t) {
        if (t + 1 < TRIPS || (wave_first + NT * t < N && (int)threadIdx.x + NT * t < N))
            f(r, c);
        r += DR;
        c += DC;
        if (c >= COLS) {
            c -= COLS;
            ++r;
        }
    }
}

// 4 horizontal outputs from the row segment starting at `src` (16-byte aligned).
template <int R>
__device__ __forceinline__ float4 hstrip(const float* src, const Weights& k) {
    constexpr int N = Windows<R>::CHUNKS * 4;
    float x[N];
#pragma unroll
    for (int q = 0; q < N / 4; ++q) {
        const float4 v = *reinterpret_cast<const float4*>(src + 4 * q);
        x[4 * q] = v.x;
        x[4 * q + 1] = v.y;
        x[4 * q + 2] = v.z;
        x[4 * q + 3] = v.w;
    }
    float o[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float v = x[q] * k.w[0];
#pragma unroll
        for (int t = 1; t <= 2 * R; ++t) v += x[q + t] * k.w[t];
        o[q] = v;
    }
    return make_float4(o[0], o[1], o[2], o[3]);
}

// 4 vertical outputs down the column starting at `src`, rows `pitch` apart.
template <int R, int PITCH>
__device__ __forceinline__ void vstrip(const float* src, const Weights& k, float o[4]) {
    float x[4 + 2 * R];
#pragma unroll
    for (int i = 0; i < 4 + 2 * R; ++i) x[i] = src[i * PITCH];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float v = x[q] * k.w[0];
#pragma unroll
        for (int t = 1; t <= 2 * R; ++t) v += x[q + t] * k.w[t];
        o[q] = v;
    }
}

// Border tiles stage their windows half-sample mirrored, so every window row
// holds an image row and every horizontal or vertical result at an image
// position is the reference's. A result at a window position outside the
// image (within r of it) is not the reference's value there; the windows
// that later passes read across the edge are repaired by copying each such
// position from its mirrored image position: whole rows first, then columns,
// so corners take the doubly mirrored value. Sources lie inside the image and
// destinations outside, so no copy reads another copy's output.
template <int R, int ROWS, int COLS, int NT = THREADS>
__device__ void mirror_window_rows(float* s, int pitch, int top, int height) {
    for (int item = threadIdx.x; item < 2 * R * COLS; item += NT) {
        const int k = item / COLS, c = item - k * COLS;
        const int y = k < R ? k - R : height + (k - R);
        const int r = y - top, src = mirror_near(y, height) - top;
        if (r >= 0 && r < ROWS && src >= 0 && src < ROWS) s[r * pitch + c] = s[src * pitch + c];
    }
}
template <int R, int ROWS, int COLS, int NT = THREADS>
__device__ void mirror_window_cols(float* s, int pitch, int left, int width) {
    for (int item = threadIdx.x; item < 2 * R * ROWS; item += NT) {
        const int r = item / (2 * R), k = item - r * (2 * R);
        const int x = k < R ? k - R : width + (k - R);
        const int c = x - left, src = mirror_near(x, width) - left;
        if (c >= 0 && c < COLS && src >= 0 && src < COLS) s[r * pitch + c] = s[r * pitch + src];
    }
}

// The highlight mask's two blurs (sharpening.rs:227-262) in one pass over a
// 64x16 tile, and the first RL ratio. The estimate starts as the guide
// (sharpening.rs:219), so the first iteration's blur of the estimate is this
// same blur of the guide; its ratio guide / max(blur, floor) is formed here
// once, and the first iteration starts from it.
//   mask_out  = min(mask, blur(mask)^4) * detail gate(guide, blur(guide))
//   ratio_out = guide / max(blur(guide), floor)
// Only the tile's columns are blurred horizontally, over every window row, so
// a border tile needs no repair: its mirrored rows are image rows.
template <int R, bool BORDER>
__device__ void prep_strips(const float* __restrict__ mask, const float* __restrict__ guide,
                            float* __restrict__ mask_out, float* __restrict__ ratio_out,
                            int width, int height, const Weights& k, float noise_protection,
                            unsigned* range_flag, float* s_in, float* s_h) {
    // Window column 0 is image column x0 - r. Interior windows load as
    // float4 rows from the aligned column x0 - 4 and land GS columns left of
    // the window, in the previous row's padding (in front of s_in for the
    // first row) and the row's own.
    constexpr int W = TW + 2 * R, H = TH + 2 * R, P = 72, GS = 4 - R;
    static_assert(W + GS <= P, "the aligned rows' padding stays in the pitch");
    static_assert(4 * (TW / 4 - 1) + Windows<R>::CHUNKS * 4 <= P, "strips stay in a row");
    static_assert(H * P <= (TH + 2 * MAXR) * (TW + 2 * MAXR), "the window fits s_in");
    static_assert(TH * TW == 4 * THREADS, "one vertical strip per thread");
    const uint3 blk = xcd_block();
    const int x0 = blk.x * TW, y0 = blk.y * TH;
    const int c = threadIdx.x & 63;
    const int r0 = 4 * __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    // Both windows are requested up front; the guide's waits in registers.
    std::conditional_t<BORDER, Staged<H, W>, Staged4<H, P / 4>> mw, gw;
    if constexpr (BORDER) {
        mw.load_mirrored(mask, width, height, y0 - R, x0 - R, H, W);
        gw.load_mirrored(guide, width, height, y0 - R, x0 - R, H, W);
    } else {
        mw.load(mask + (size_t)(y0 - R) * width + x0 - 4, width);
        gw.load(guide + (size_t)(y0 - R) * width + x0 - 4, width);
    }
    auto put = [&](const auto& win) {
        if constexpr (BORDER)
            win.store(s_in, P, H, W);
        else
            win.template store<GS>(s_in, P);
    };
    float m[4], bm[4], g[4], bg[4];
    // One blur of the window in s_in into b (this thread's 4 outputs); the
    // window's values at the outputs go to centre.
    auto blur = [&](float b[4], float centre[4]) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) centre[q] = s_in[(r0 + q + R) * P + c + R];
        for (int item = threadIdx.x; item < H * (TW / 4); item += THREADS) {
            const int r = item / (TW / 4), j = item % (TW / 4);
            *reinterpret_cast<float4*>(s_h + r * TW + 4 * j) = hstrip<R>(s_in + r * P + 4 * j, k);
        }
        __syncthreads();
        vstrip<R, TW>(s_h + r0 * TW + c, k, b);
    };
    put(mw);
    blur(bm, m);
    // Every read of s_in precedes blur's second barrier, and the guide's
    // horizontal pass writes s_h only after its first one.
    put(gw);
    blur(bg, g);
    bool wide = false;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int y = y0 + r0 + q, x = x0 + c;
        if (BORDER && (y >= height || x >= width)) continue;
        wide |= !(g[q] < GUIDE_LIMIT);
        const size_t i = (size_t)y * width + x;
        const float expanded = num_min(m[q], (bm[q] * bm[q]) * (bm[q] * bm[q]));
        const float threshold =
            0.0005f + noise_protection * 0.008f + bg[q] * noise_protection * 0.02f;
        const float detail = fabsf(g[q] - bg[q]);
        mask_out[i] = expanded * (smoothstep(threshold, threshold * 3.0f, detail) *
                                  smoothstep(0.001f, 0.01f, g[q]));
        ratio_out[i] = g[q] / num_max(bg[q], FLOOR);
    }
    // Every pixel's guide passes through one tile's outputs here, so the
    // flag is set before the iterations if any guide value is out of range.
    if (wide) atomicOr(range_flag, 1u);
}

template <int R>
__global__ __launch_bounds__(THREADS) void cap_prep(const float* __restrict__ mask,
                                                    const float* __restrict__ guide,
                                                    float* __restrict__ mask_out,
                                                    float* __restrict__ ratio_out, int width,
                                                    int height, Weights k, float noise_protection,
                                                    unsigned* range_flag, bool aligned) {
    // Four floats in front of s_in take the first window row's load padding.
    __shared__ __align__(16) float s_in_mem[(TH + 2 * MAXR) * (TW + 2 * MAXR) + 4];
    float* const s_in = s_in_mem + 4;
    __shared__ __align__(16) float s_h[(TH + 2 * MAXR) * TW];
    const uint3 blk = xcd_block();
    const int x0 = blk.x * TW, y0 = blk.y * TH;
    // Interior: the aligned windows (image columns x0 - 4 to x0 + 67) lie
    // inside the image and the planes' rows start 16-byte aligned.
    const bool interior = aligned && x0 - 4 >= 0 && x0 + TW + 4 <= width && y0 - R >= 0 &&
                          y0 + TH + R <= height;
    if (interior)
        prep_strips<R, false>(mask, guide, mask_out, ratio_out, width, height, k, noise_protection,
                              range_flag, s_in, s_h);
    else
        prep_strips<R, true>(mask, guide, mask_out, ratio_out, width, height, k, noise_protection,
                             range_flag, s_in, s_h);
}

// The gain apply (sharpening.rs:105-126) at pixel i of planar RGB with n
// pixels per plane: gain = 1 + amount * mask * (estimate / guide - 1), and
// the pixel is scaled only when the gain differs from 1 and every scaled
// channel stays finite.
struct Apply {
    float* rgb;
    const float* mask;
    float amount;
    size_t n;
};
// `mask` is the mask value at i, read early by callers that can.
__device__ __forceinline__ void apply_gain(const Apply& ap, size_t i, float est, float guide,
                                           float mask) {
    const float gain = 1.0f + ap.amount * mask * (est / guide - 1.0f);
    if (gain == 1.0f) return;
    const float r = ap.rgb[i], g = ap.rgb[ap.n + i], b = ap.rgb[2 * ap.n + i];
    if (is_finite(r * gain) && is_finite(g * gain) && is_finite(b * gain)) {
        ap.rgb[i] = r * gain;
        ap.rgb[ap.n + i] = g * gain;
        ap.rgb[2 * ap.n + i] = b * gain;
    }
}

// One RL iteration (sharpening.rs:268-304) for a 64x16 tile: blur(estimate)
// over the tile widened by r, ratio = guide / max(b, floor), blur(ratio) over
// the tile, estimate = clamp(estimate * b, guide / 2, 2 guide). Two LDS
// arrays: the first holds the estimate with a 2r halo and, once its
// horizontal blur is done, the guide with an r halo, which the ratio
// overwrites in place; the second holds the horizontal blurs. Sharing the
// first array keeps eight workgroups (eight waves per SIMD) resident. The
// guide values of the tile's outputs are kept in registers before the ratio
// overwrites the guide window. APPLY: the last iteration applies the gain
// with its estimate instead of storing the estimate.
template <int R, bool APPLY, bool BORDER, bool IN_RANGE>
__device__ void iterate_strips(const float* __restrict__ est_in, const float* __restrict__ guide,
                               float* __restrict__ est_out, int width, int height,
                               const Weights& k, const Apply& ap, float* s_a, float* s_b) {
    using Win = Windows<R, ITH>;
    constexpr int EW = Win::EW, EH = Win::EH, GW = Win::GW, GH = Win::GH, P = Win::P;
    constexpr int EP = Win::EP, EL = Win::EL, NS1 = Win::NS1, NS3 = Win::NS3;
    constexpr int NR2 = (GH + 3) / 4;   // row strips of the blurred estimate
    constexpr int NR4 = ITH / 4;        // row strips of the output
    constexpr int OUT_STRIPS = NR4 * TW / ITHREADS;
    static_assert(TW == 64, "output strips map one wavefront to one tile row");
    static_assert(OUT_STRIPS >= 1, "every thread owns at least one output strip");
    const uint3 blk = xcd_block();
    const int x0 = blk.x * TW, y0 = blk.y * ITH;
    // The guide window waits in registers while the estimate window, which
    // shares its LDS, is blurred. Border tiles stage both windows element by
    // element, half-sample mirrored; interior ones as aligned float4 rows
    // that include the columns between the aligned start and the window.
    std::conditional_t<BORDER, Staged<GH, GW, ITHREADS>, Staged4<GH, P / 4, ITHREADS>> gw;
    // The applying iteration requests the mask at its outputs up front, so
    // its latency hides behind the blurs.
    float m_out[OUT_STRIPS][4];
    if constexpr (APPLY) {
#pragma unroll
        for (int q = 0; q < OUT_STRIPS; ++q) {
            const int c = threadIdx.x & 63, r0 = 4 * ((int)(threadIdx.x >> 6) + q * (ITHREADS / 64));
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int y = y0 + r0 + i, x = x0 + c;
                m_out[q][i] = !BORDER || (y < height && x < width)
                                  ? ap.mask[(size_t)y * width + x] : 0.0f;
            }
        }
    }
    if constexpr (BORDER) {
        gw.load_mirrored(guide, width, height, y0 - R, x0 - R, GH, GW);
        stage<EH, EW, true, ITHREADS>(s_a + Win::ED, EP, est_in, width, height, y0 - 2 * R,
                                      x0 - 2 * R);
    } else {
        gw.load(guide + (size_t)(y0 - R) * width + x0 - 4, width);
        Staged4<EH, EP / 4, ITHREADS> ew;
        ew.load(est_in + (size_t)(y0 - 2 * R) * width + x0 - EL, width);
        ew.store(s_a, EP);
    }
    __syncthreads();
    float e_out[OUT_STRIPS][4], g_out[OUT_STRIPS][4];
#pragma unroll
    for (int q = 0; q < OUT_STRIPS; ++q) {
        const int c = threadIdx.x & 63, r0 = 4 * ((int)(threadIdx.x >> 6) + q * (ITHREADS / 64));
#pragma unroll
        for (int i = 0; i < 4; ++i) e_out[q][i] = s_a[(r0 + i + 2 * R) * EP + c + EL];
    }
    for_items<EH, NS1, ITHREADS>([&](int r, int j) {
        *reinterpret_cast<float4*>(s_b + r * P + 4 * j) = hstrip<R>(s_a + r * EP + 4 * j, k);
    });
    __syncthreads();
    if constexpr (BORDER)
        gw.store(s_a, P, GH, GW);
    else
        gw.template store<Win::GS>(s_a, P);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < OUT_STRIPS; ++q) {
        const int c = threadIdx.x & 63, r0 = 4 * ((int)(threadIdx.x >> 6) + q * (ITHREADS / 64));
#pragma unroll
        for (int i = 0; i < 4; ++i) g_out[q][i] = s_a[(r0 + i + R) * P + c + R];
    }
    __syncthreads();
    // The ratio at guide column c takes the estimate's
    // blur at LDS column c + ED of s_b, the same image column.
    for_items<NR2, GW, ITHREADS>([&](int strip, int c) {
        const int r0 = 4 * strip;
        float b[4];
        vstrip<R, P>(s_b + r0 * P + c + Win::ED, k, b);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (r0 + i >= GH) continue;
            float* ratio = s_a + (r0 + i) * P + c;
            const float den = num_max(b[i], FLOOR);
            *ratio = IN_RANGE ? div_in_range(*ratio, den) : *ratio / den;
        }
    });
    __syncthreads();
    if constexpr (BORDER) {
        // The ratio's blur reads up to r positions across the image edge.
        mirror_window_rows<R, GH, GW, ITHREADS>(s_a, P, y0 - R, height);
        __syncthreads();
        mirror_window_cols<R, GH, GW, ITHREADS>(s_a, P, x0 - R, width);
        __syncthreads();
    }
    for_items<GH, NS3, ITHREADS>([&](int r, int j) {
        *reinterpret_cast<float4*>(s_b + r * P + 4 * j) = hstrip<R>(s_a + r * P + 4 * j, k);
    });
    __syncthreads();
#pragma unroll
    for (int q = 0; q < OUT_STRIPS; ++q) {
        // TW == 64: a wavefront's strips share their rows.
        const int c = threadIdx.x & 63;
        const int r0 = 4 * __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6) + q * (ITHREADS / 64));
        float b[4];
        vstrip<R, P>(s_b + r0 * P + c, k, b);
        const size_t o = (size_t)(y0 + r0) * width + x0 + c;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (BORDER && (y0 + r0 + i >= height || x0 + c >= width)) continue;
            const float g = g_out[q][i];
            const float v = e_out[q][i] * b[i];
            // clamp(v, g / 2, 2 g): g >= the floor, so the bounds are ordered,
            // and v is a product of positive values, never NaN for finite
            // input; the median of three is then the reference's clamp.
            const float est = __builtin_amdgcn_fmed3f(v, g * 0.5f, g * 2.0f);
            if (APPLY)
                apply_gain(ap, o + (size_t)i * width, est, g, m_out[q][i]);
            else
                est_out[o + (size_t)i * width] = est;
        }
    }
}

template <int R, bool APPLY>
__global__ __launch_bounds__(ITHREADS) void cap_iterate(const float* __restrict__ est_in,
                                                       const float* __restrict__ guide,
                                                       float* __restrict__ est_out, int width,
                                                       int height, Weights k, Apply ap,
                                                       bool aligned,
                                                       const unsigned* __restrict__ range_flag) {
    using Win = Windows<R, ITH>;
    // Four floats in front of s_a take the first guide row's load padding.
    __shared__ __align__(16) float s_a_mem[Win::A_FLOATS + 4];
    float* const s_a = s_a_mem + 4;
    __shared__ __align__(16) float s_b[Win::H_FLOATS];
    const uint3 blk = xcd_block();
    const int x0 = blk.x * TW, y0 = blk.y * ITH;
    // Interior: the aligned windows lie inside the image and the planes' rows
    // start 16-byte aligned (aligned: the width is a multiple of 4 and the
    // planes are 16-byte aligned).
    const bool interior = aligned && x0 - Win::EL >= 0 && x0 + TW + Win::EL <= width &&
                          y0 - 2 * R >= 0 && y0 + ITH + 2 * R <= height;
    // With every guide value in [floor, 2^64) (cap_prep's flag), every
    // estimate lies in [floor / 2, 2^65] (the clamp) and so does its blur,
    // whose weights sum to 1 within a few ulps: the ratio's operands are in
    // div_in_range's domain.
    if (interior && *range_flag == 0)
        iterate_strips<R, APPLY, false, true>(est_in, guide, est_out, width, height, k, ap, s_a,
                                              s_b);
    else if (interior)
        iterate_strips<R, APPLY, false, false>(est_in, guide, est_out, width, height, k, ap, s_a,
                                               s_b);
    else
        iterate_strips<R, APPLY, true, false>(est_in, guide, est_out, width, height, k, ap, s_a,
                                              s_b);
}

// The first RL iteration from cap_prep's ratio: blur(ratio) over the tile,
// estimate = clamp(guide * b, guide / 2, 2 guide) (the estimate it updates is
// the guide). The ratio window is staged half-sample mirrored for a border
// tile, which gives the positions across the edge their mirrored values.
template <int R, bool APPLY, bool BORDER>
__device__ void iterate_first_strips(const float* __restrict__ ratio,
                                     const float* __restrict__ guide, float* __restrict__ est_out,
                                     int width, int height, const Weights& k, const Apply& ap,
                                     float* s_a, float* s_b) {
    using Win = Windows<R>;
    constexpr int GW = Win::GW, GH = Win::GH, P = Win::P, NS3 = Win::NS3;
    constexpr int OUT_STRIPS = (TH / 4) * TW / THREADS;
    static_assert(TW == 64 && OUT_STRIPS >= 1, "output strips map one wavefront to one tile row");
    const uint3 blk = xcd_block();
    const int x0 = blk.x * TW, y0 = blk.y * TH;
    const int c = threadIdx.x & 63;
    float g_out[OUT_STRIPS][4];
#pragma unroll
    for (int q = 0; q < OUT_STRIPS; ++q) {
        const int r0 = 4 * ((int)(threadIdx.x >> 6) + q * (THREADS / 64));
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int y = y0 + r0 + i, x = x0 + c;
            g_out[q][i] = !BORDER || (y < height && x < width) ? guide[(size_t)y * width + x] : 1.0f;
        }
    }
    if constexpr (BORDER) {
        stage<GH, GW, true>(s_a, P, ratio, width, height, y0 - R, x0 - R);
    } else {
        Staged4<GH, P / 4> rw;
        rw.load(ratio + (size_t)(y0 - R) * width + x0 - 4, width);
        rw.template store<Win::GS>(s_a, P);
    }
    __syncthreads();
    for_items<GH, NS3>([&](int r, int j) {
        *reinterpret_cast<float4*>(s_b + r * P + 4 * j) = hstrip<R>(s_a + r * P + 4 * j, k);
    });
    __syncthreads();
#pragma unroll
    for (int q = 0; q < OUT_STRIPS; ++q) {
        const int r0 = 4 * __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6) + q * (THREADS / 64));
        float b[4];
        vstrip<R, P>(s_b + r0 * P + c, k, b);
        const size_t o = (size_t)(y0 + r0) * width + x0 + c;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (BORDER && (y0 + r0 + i >= height || x0 + c >= width)) continue;
            const float g = g_out[q][i];
            const float est = __builtin_amdgcn_fmed3f(g * b[i], g * 0.5f, g * 2.0f);
            if (APPLY)
                apply_gain(ap, o + (size_t)i * width, est, g, ap.mask[o + (size_t)i * width]);
            else
                est_out[o + (size_t)i * width] = est;
        }
    }
}

template <int R, bool APPLY>
__global__ __launch_bounds__(THREADS) void cap_iterate_first(const float* __restrict__ ratio,
                                                             const float* __restrict__ guide,
                                                             float* __restrict__ est_out,
                                                             int width, int height, Weights k,
                                                             Apply ap, bool aligned) {
    using Win = Windows<R>;
    // Four floats in front of s_a take the first ratio row's load padding.
    __shared__ __align__(16) float s_a_mem[Win::G_FLOATS + 4];
    float* const s_a = s_a_mem + 4;
    __shared__ __align__(16) float s_b[Win::G_FLOATS];
    const uint3 blk = xcd_block();
    const int x0 = blk.x * TW, y0 = blk.y * TH;
    // Interior: the aligned ratio window (image columns x0 - 4 to x0 + 67)
    // lies inside the image and the planes' rows start 16-byte aligned.
    const bool interior = aligned && x0 - 4 >= 0 && x0 + TW + 4 <= width && y0 - R >= 0 &&
                          y0 + TH + R <= height;
    if (interior)
        iterate_first_strips<R, APPLY, false>(ratio, guide, est_out, width, height, k, ap, s_a, s_b);
    else
        iterate_first_strips<R, APPLY, true>(ratio, guide, est_out, width, height, k, ap, s_a, s_b);
}

// sharpening.rs:105-126
__global__ __launch_bounds__(THREADS) void cap_apply(float* __restrict__ rgb, int n, float amount,
                                                     const float* __restrict__ guide,
                                                     const float* __restrict__ mask,
                                                     const float* __restrict__ est) {
    int i = blockIdx.x * THREADS + threadIdx.x;
    if (i >= n) return;
    apply_gain(Apply{rgb, mask, amount, (size_t)n}, i, est[i], guide[i], mask[i]);
}

}  // namespace

size_t capture_scratch_bytes(int width, int height) {
    return (size_t)width * height * sizeof(float) * 5;
}

hipError_t launch_capture(const CaptureArgs& a, hipStream_t s, Profiler* prof) {
    if (a.radius < 1 || a.radius > MAXR || !a.stats) return hipErrorInvalidValue;
    const int n = a.width * a.height;
    const uint64_t px = (uint64_t)n;
    Weights k{};
    for (int t = 0; t <= 2 * a.radius; ++t) k.w[t] = a.weights[t];
    const int blocks = (n + THREADS - 1) / THREADS;
    if (!a.inputs_ready)
        RHD_TIMED(prof, s, "cap_init", 20 * px,
                  hipLaunchKernelGGL(cap_init, dim3(blocks), dim3(THREADS), 0, s, a.rgb, n,
                                     a.ceilings[0], a.ceilings[1], a.ceilings[2], a.guide,
                                     a.mask, a.stats));
    const dim3 tiles((a.width + TW - 1) / TW, (a.height + TH - 1) / TH);
    const dim3 iter_tiles((a.width + TW - 1) / TW, (a.height + ITH - 1) / ITH);
    // Interior tiles move their windows as float4 when every row of the
    // planes they read starts 16-byte aligned.
    auto aligned16 = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
    const bool aligned = a.width % 4 == 0 && aligned16(a.guide) && aligned16(a.mask) &&
                         aligned16(a.est_a) && aligned16(a.est_b);
    // The final mask goes to tmp and the first iteration's ratio to est_b,
    // which the first iteration reads before the second one overwrites it.
    float* mask = a.tmp;
    float* ratio = a.est_b;
    const Apply ap{a.rgb, mask, a.amount, (size_t)n};
    // The radius is a template parameter so the tap loops unroll.
    auto prep = [&] {
        switch (a.radius) {
            case 1: hipLaunchKernelGGL(cap_prep<1>, tiles, dim3(THREADS), 0, s, a.mask, a.guide, mask, ratio, a.width, a.height, k, a.noise_protection, &a.stats->capture_range, aligned); break;
            case 2: hipLaunchKernelGGL(cap_prep<2>, tiles, dim3(THREADS), 0, s, a.mask, a.guide, mask, ratio, a.width, a.height, k, a.noise_protection, &a.stats->capture_range, aligned); break;
            case 3: hipLaunchKernelGGL(cap_prep<3>, tiles, dim3(THREADS), 0, s, a.mask, a.guide, mask, ratio, a.width, a.height, k, a.noise_protection, &a.stats->capture_range, aligned); break;
            default: hipLaunchKernelGGL(cap_prep<4>, tiles, dim3(THREADS), 0, s, a.mask, a.guide, mask, ratio, a.width, a.height, k, a.noise_protection, &a.stats->capture_range, aligned); break;
        }
    };
    auto first = [&](float* out, auto apply) {
        constexpr bool APPLY = decltype(apply)::value;
        switch (a.radius) {
            case 1: hipLaunchKernelGGL((cap_iterate_first<1, APPLY>), tiles, dim3(THREADS), 0, s, ratio, a.guide, out, a.width, a.height, k, ap, aligned); break;
            case 2: hipLaunchKernelGGL((cap_iterate_first<2, APPLY>), tiles, dim3(THREADS), 0, s, ratio, a.guide, out, a.width, a.height, k, ap, aligned); break;
            case 3: hipLaunchKernelGGL((cap_iterate_first<3, APPLY>), tiles, dim3(THREADS), 0, s, ratio, a.guide, out, a.width, a.height, k, ap, aligned); break;
            default: hipLaunchKernelGGL((cap_iterate_first<4, APPLY>), tiles, dim3(THREADS), 0, s, ratio, a.guide, out, a.width, a.height, k, ap, aligned); break;
        }
    };
    auto iterate = [&](const float* in, float* out, auto apply) {
        constexpr bool APPLY = decltype(apply)::value;
        switch (a.radius) {
            case 1: hipLaunchKernelGGL((cap_iterate<1, APPLY>), iter_tiles, dim3(ITHREADS), 0, s, in, a.guide, out, a.width, a.height, k, ap, aligned, &a.stats->capture_range); break;
            case 2: hipLaunchKernelGGL((cap_iterate<2, APPLY>), iter_tiles, dim3(ITHREADS), 0, s, in, a.guide, out, a.width, a.height, k, ap, aligned, &a.stats->capture_range); break;
            case 3: hipLaunchKernelGGL((cap_iterate<3, APPLY>), iter_tiles, dim3(ITHREADS), 0, s, in, a.guide, out, a.width, a.height, k, ap, aligned, &a.stats->capture_range); break;
            default: hipLaunchKernelGGL((cap_iterate<4, APPLY>), iter_tiles, dim3(ITHREADS), 0, s, in, a.guide, out, a.width, a.height, k, ap, aligned, &a.stats->capture_range); break;
        }
    };
    // Reads the mask and the guide, writes the final mask and the ratio.
    RHD_TIMED(prof, s, "cap_prep", 16 * px, prep());
    // The last iteration applies the gain itself; with no iterations the
    // estimate is the guide and cap_apply does it. The others alternate
    // between the two estimate planes.
    if (a.iterations == 0) {
        RHD_TIMED(prof, s, "cap_apply", 36 * px,
                  hipLaunchKernelGGL(cap_apply, dim3(blocks), dim3(THREADS), 0, s, a.rgb, n,
                                     a.amount, a.guide, mask, a.guide));
    } else if (a.iterations == 1) {
        // Reads the ratio, the guide, the mask and RGB, writes RGB.
        RHD_TIMED(prof, s, "cap_iterate_first", 36 * px, first(nullptr, std::true_type{}));
    } else {
        // Reads the ratio and the guide, writes the estimate: 12 B per pixel.
        RHD_TIMED(prof, s, "cap_iterate_first", 12 * px, first(a.est_a, std::false_type{}));
        const float* cur = a.est_a;
        for (int it = 1; it + 1 < a.iterations; ++it) {
            float* out = (it & 1) ? a.est_b : a.est_a;
            // Reads the estimate and the guide, writes the estimate: 12 B per pixel.
            RHD_TIMED(prof, s, "cap_iterate", 12 * px, iterate(cur, out, std::false_type{}));
            cur = out;
        }
        // Reads the estimate, the guide, the mask and RGB, writes RGB.
        RHD_TIMED(prof, s, "cap_iterate_apply", 36 * px, iterate(cur, nullptr, std::true_type{}));
    }
    return hipGetLastError();
}

}  // namespace rhd
