// Region-aware Opposed highlight reconstruction on gfx950.
//
// Reference: crates/highlight/src/opposed.rs:63-364 (driver and per-site
// estimate), cells.rs:41-113 (2x2 cell summaries), regions.rs:34-271
// (Chebyshev-1 dilation and 8-connected labelling with compact raster-order
// labels), region_correction.rs:53-300 (two-pass trimmed means in f64,
// raster order) and support.rs:26-125 (support sets, cube-root reference,
// log2 chromaticity gate).
//
// Pipeline (all on the context stream, no host round trip):
//   cells      one thread per 2x2 cell: means and flags (float4 per cell)
//   dilate     fg = a clipped cell within Chebyshev distance 1
//   union      concurrent union-find; every set's root is its smallest index,
//              i.e. the component's first cell in raster order
//   flatten    parent -> root
//   scan       exclusive scan of the root flags -> compact label = rank + 1
//              (the reference's numbering by first raster appearance)
//   regions    per-region clipped channels and wanted slots (atomic OR),
//              and the ring flags (label != 0, no clipped channel)
//   scan       ring cells compacted in raster order, plus each region's
//              first/last ring index
//   contrib    every ring cell's nine correction samples, in parallel
//   sums       one wavefront per (group, slot), group 0 = image-wide pool,
//              group r = region r: both passes over its ring entries in
//              raster order, accumulated one f64 add at a time, so the sums
//              are bit-identical to the reference's
//   sites      one thread per photosite: the per-site estimate, the bound,
//              the neutral ceiling, and the statistics
// cbrt and log2 are evaluated in double and rounded to f32 (correctly
// rounded in practice); the reference calls the platform libm's cbrtf and
// log2f, which are within 1 ulp of that. Results therefore agree with the
// reference within the stated f32 tolerance rather than bit for bit.
#include "rhd_internal.hpp"

namespace rhd {
namespace {

constexpr int THREADS = 256;
constexpr int SCAN_PER_THREAD = 16;
constexpr int SCAN_BLOCK = THREADS * SCAN_PER_THREAD;
constexpr unsigned NONE = 0xffffffffu;
constexpr int SLOTS = 9;
constexpr unsigned MIN_BORDER_CELLS = 3;

struct Layout {
    int w, h, cw, ch, cells, cap;  // cap: region-count bound (+1)
    float4* cell;                  // means (x, y, z), flags as bits in w
    unsigned* label;               // parent during labelling, then compact label
    unsigned* rank;                // root rank, then scratch
    unsigned* block_sums;
    unsigned* ring_cell;
    unsigned* ring_label;          // label of each ring entry
    unsigned* contrib_flags;       // per ring entry: valid slots, then valid ratios << 9
    float* contrib_value;          // per ring entry: nine correction samples
    float* contrib_ratio;          // per ring entry: three two-channel support ratios
    unsigned* region_clipped;      // per label
    unsigned* region_wanted;
    unsigned* region_first;
    unsigned* region_last;
    float4* corrections;           // per (group, slot): offset, ratio, valid, has_ratio
    unsigned* scalars;             // [0] region count, [1] ring count, [2] any wanted
};

__device__ __forceinline__ unsigned flags_of(const float4& c) { return __float_as_uint(c.w); }
__device__ __forceinline__ bool usable(unsigned f, int c) { return f & (1u << c); }
__device__ __forceinline__ bool clean(unsigned f, int c) { return !(f & (1u << (c + 3))); }

__device__ __forceinline__ float mean_of(const float4& c, int ch) {
    return ch == 0 ? c.x : (ch == 1 ? c.y : c.z);
}

// support.rs:26-33
__device__ __forceinline__ bool bright_clean(const float4& cell, int c, const float* lv) {
    unsigned f = flags_of(cell);
    return usable(f, c) && clean(f, c) && mean_of(cell, c) > 0.2f * lv[c];
}
// support.rs:37-46
__device__ __forceinline__ unsigned support_mask(const float4& cell, int target, const float* lv) {
    unsigned m = 0;
    for (int c = 0; c < 3; ++c)
        if (c != target && bright_clean(cell, c, lv)) m |= 1u << c;
    return m;
}
__device__ __forceinline__ float cbrt_f32(float v) { return (float)cbrt((double)v); }
__device__ __forceinline__ float log2_f32(float v) { return (float)log2((double)v); }

// support.rs:51-72
__device__ __forceinline__ bool reference(const float4& cell, unsigned support, float& out) {
    if (__popc(support) == 1) {
        float v = mean_of(cell, __ffs(support) - 1);
        if (is_finite(v) && v > 0.0f) {
            out = v;
            return true;
        }
        return false;
    }
    float sum = 0.0f;
    unsigned count = 0;
    for (int c = 0; c < 3; ++c) {
        if (!(support & (1u << c))) continue;
        float v = mean_of(cell, c);
        if (!is_finite(v) || v <= 0.0f) return false;
        sum += cbrt_f32(v);
        ++count;
    }
    if (count == 0) return false;
    float root = sum / (float)count;
    out = root * root * root;
    return true;
}
// support.rs:78-95
__device__ __forceinline__ bool correction(const float4& cell, int target, unsigned support,
                                           const float* lv, float& out) {
    if (!bright_clean(cell, target, lv)) return false;
    for (int c = 0; c < 3; ++c)
        if ((support & (1u << c)) && !bright_clean(cell, c, lv)) return false;
    float ref;
    if (!reference(cell, support, ref)) return false;
    float v = mean_of(cell, target) - ref;
    if (!is_finite(v)) return false;
    out = v;
    return true;
}
// support.rs:101-110
__device__ __forceinline__ bool support_log_ratio(const float4& cell, unsigned support, float& out) {
    if (__popc(support) != 2) return false;
    int first = __ffs(support) - 1;
    int second = __ffs(support & ~(1u << first)) - 1;
    float ratio = mean_of(cell, first) / mean_of(cell, second);
    if (!(is_finite(ratio) && ratio > 0.0f)) return false;
    out = log2_f32(ratio);
    return true;
}
// support.rs:116-125
__device__ __forceinline__ bool chromaticity_agrees(unsigned support, bool ht, float t, bool he,
                                                    float e) {
    if (__popc(support) < 2) return true;
    if (ht && he) return fabsf(t - e) <= 0.75f;
    return false;
}
// region_correction.rs:288-300
__device__ __forceinline__ unsigned support_pattern(int channel, int k) {
    int first = channel == 0 ? 1 : 0;
    int second = channel == 2 ? 1 : 2;
    return k == 0 ? (1u << first) : (k == 1 ? (1u << second) : ((1u << first) | (1u << second)));
}
__device__ __forceinline__ int pattern_position(int channel, unsigned support) {
    for (int k = 0; k < 3; ++k)
        if (support_pattern(channel, k) == support) return k;
    return -1;
}

// ---------------------------------------------------------------- cells
// cells.rs:41-113
__global__ __launch_bounds__(THREADS) void opp_cells(const float* __restrict__ m, Layout L,
                                                     int pattern, float l0, float l1, float l2) {
    int i = blockIdx.x * THREADS + threadIdx.x;
    if (i >= L.cells) return;
    const float lv[3] = {l0, l1, l2};
    int cx = i % L.cw, cy = i / L.cw;
    float sums[3] = {0.0f, 0.0f, 0.0f};
    unsigned counts[3] = {0, 0, 0};
    unsigned flags = 0;
    for (int y = 2 * cy; y < min(2 * cy + 2, L.h); ++y)
        for (int x = 2 * cx; x < min(2 * cx + 2, L.w); ++x) {
            int c = color_at(pattern, x, y);
            float s = m[(size_t)y * L.w + x];
            if (s >= lv[c]) {
                flags |= 1u << (c + 3);
            } else if (s > 0.0f) {
                sums[c] += s;
                counts[c] += 1;
            }
        }
    float means[3] = {0.0f, 0.0f, 0.0f};
    for (int c = 0; c < 3; ++c)
        if (counts[c] != 0) {
            means[c] = sums[c] / (float)counts[c];
            flags |= 1u << c;
        }
    L.cell[i] = make_float4(means[0], means[1], means[2], __uint_as_float(flags));
}

// regions.rs:126-183: Chebyshev radius-1 dilation of the clipped cells.
__global__ __launch_bounds__(THREADS) void opp_dilate(Layout L) {
    int i = blockIdx.x * THREADS + threadIdx.x;
    if (i >= L.cells) return;
    int cx = i % L.cw, cy = i / L.cw;
    bool any = false;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            int x = cx + dx, y = cy + dy;
            if (x < 0 || y < 0 || x >= L.cw || y >= L.ch) continue;
            any |= (flags_of(L.cell[y * L.cw + x]) >> 3) != 0;
        }
    L.label[i] = any ? (unsigned)i : NONE;
}

__device__ __forceinline__ unsigned load_parent(unsigned* p, unsigned x) {
    return __hip_atomic_load(p + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned find_root(unsigned* p, unsigned x) {
    unsigned y = load_parent(p, x);
    while (y != x) {
        x = y;
        y = load_parent(p, x);
    }
    return x;
}
// Union by linking the larger root under the smaller one, so every set's root
// is its smallest member.
__device__ void merge(unsigned* p, unsigned a, unsigned b) {
    while (true) {
        a = find_root(p, a);
        b = find_root(p, b);
        if (a == b) return;
        if (a > b) {
            unsigned t = a;
            a = b;
            b = t;
        }
        unsigned old = atomicMin(p + b, a);
        if (old == b) return;
        b = old;
    }
}

// 8-connected unions with the already-visited neighbours W, NW, N, NE
// (regions.rs:200-214).
__global__ __launch_bounds__(THREADS) void opp_union(Layout L) {
    int i = blockIdx.x * THREADS + threadIdx.x;
    if (i >= L.cells || L.label[i] == NONE) return;
    int cx = i % L.cw, cy = i / L.cw;
    const int nx[4] = {-1, -1, 0, 1}, ny[4] = {0, -1, -1, -1};
    for (int k = 0; k < 4; ++k) {
        int x = cx + nx[k], y = cy + ny[k];
        if (x < 0 || y < 0 || x >= L.cw) continue;
        unsigned j = (unsigned)(y * L.cw + x);
        if (load_parent(L.label, j) != NONE) merge(L.label, (unsigned)i, j);
    }
}

__global__ __launch_bounds__(THREADS) void opp_flatten(Layout L) {
    int i = blockIdx.x * THREADS + threadIdx.x;
    if (i >= L.cells) return;
    unsigned p = L.label[i];
    if (p == NONE) return;
    unsigned r = find_root(L.label, (unsigned)i);
    if (r != p) L.label[i] = r;
}

// ---------------------------------------------------------------- scan
// Exclusive scan of a predicate over the cells, in blocks of SCAN_BLOCK.
enum ScanMode { ROOTS = 0, RING = 1 };

__device__ __forceinline__ unsigned predicate(const Layout& L, int mode, int i) {
    if (i >= L.cells) return 0;
    if (mode == ROOTS) return L.label[i] == (unsigned)i ? 1u : 0u;
    unsigned lab = L.label[i];
    return lab != 0 && (flags_of(L.cell[i]) >> 3) == 0 ? 1u : 0u;
}

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ unsigned block_exclusive_scan(unsigned v, unsigned& exclusive) {
    __shared__ unsigned wave_totals[THREADS / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned inclusive = v;
    for (int off = 1; off < 64; off <<= 1) {
        unsigned n = __shfl_up(inclusive, off, 64);
        if (lane >= off) inclusive += n;
    }
    if (lane == 63) wave_totals[wave] = inclusive;
    __syncthreads();
    unsigned before = 0, total = 0;
    for (int k = 0; k < THREADS / 64; ++k) {
        if (k < wave) before += wave_totals[k];
        total += wave_totals[k];
    }
    __syncthreads();
    exclusive = before + inclusive - v;
    return total;
}

__global__ __launch_bounds__(THREADS) void opp_scan_count(Layout L, int mode) {
    const int base = blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_PER_THREAD;
    unsigned n = 0;
    for (int k = 0; k < SCAN_PER_THREAD; ++k) n += predicate(L, mode, base + k);
    unsigned excl;
    unsigned total = block_exclusive_scan(n, excl);
    if (threadIdx.x == 0) L.block_sums[blockIdx.x] = total;
}

// Single block: exclusive scan of the block sums; the grand total goes to
// scalars[mode == ROOTS ? 0 : 1].
__global__ __launch_bounds__(THREADS) void opp_scan_blocks(Layout L, int blocks, int mode) {
    unsigned carry = 0;
    for (int start = 0; start < blocks; start += THREADS) {
        int b = start + threadIdx.x;
        unsigned v = b < blocks ? L.block_sums[b] : 0;
        unsigned excl;
        unsigned total = block_exclusive_scan(v, excl);
        if (b < blocks) L.block_sums[b] = carry + excl;
        carry += total;
    }
    if (threadIdx.x == 0) L.scalars[mode == ROOTS ? 0 : 1] = carry;
}

__global__ __launch_bounds__(THREADS) void opp_scan_emit(Layout L, int mode) {
    const int base = blockIdx.x * SCAN_BLOCK + threadIdx.x * SCAN_PER_THREAD;
    unsigned flags = 0, n = 0;
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        unsigned p = predicate(L, mode, base + k);
        flags |= p << k;
        n += p;
    }
    unsigned excl;
    block_exclusive_scan(n, excl);
    unsigned pos = L.block_sums[blockIdx.x] + excl;
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        if (!(flags & (1u << k))) continue;
        int i = base + k;
        if (mode == ROOTS) {
            L.rank[i] = pos;
        } else {
            L.ring_cell[pos] = (unsigned)i;
            unsigned lab = L.label[i];
            if (lab < (unsigned)L.cap) {
                atomicMin(L.region_first + lab, pos);
                atomicMax(L.region_last + lab, pos);
            }
        }
        ++pos;
    }
}

// Compact labels (regions.rs:254-268): a root's rank + 1.
__global__ __launch_bounds__(THREADS) void opp_relabel(Layout L) {
    int i = blockIdx.x * THREADS + threadIdx.x;
    if (i >= L.cells) return;
    unsigned p = L.label[i];
    L.label[i] = p == NONE ? 0u : L.rank[p] + 1u;
}

// Zero the per-region arrays for labels 0..count.
__global__ __launch_bounds__(THREADS) void opp_region_init(Layout L, DeviceStats* stats) {
    const unsigned count = min(L.scalars[0], (unsigned)L.cap - 1);
    for (unsigned r = blockIdx.x * THREADS + threadIdx.x; r <= count; r += gridDim.x * THREADS) {
        L.region_clipped[r] = 0;
        L.region_wanted[r] = 0;
        L.region_first[r] = NONE;
        L.region_last[r] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        L.scalars[2] = 0;
        stats->regions = L.scalars[0];
    }
}

// regions.rs:44-55 (clipped channels per region) and region_correction.rs:
// 125-187 (WantedSlots::scan).
__global__ __launch_bounds__(THREADS) void opp_regions(Layout L, float l0, float l1, float l2) {
    int i = blockIdx.x * THREADS + threadIdx.x;
    const float lv[3] = {l0, l1, l2};
    const int lane = threadIdx.x & 63;
    unsigned lab = 0, clipped = 0, bits = 0;
    bool has = false;
    if (i < L.cells) {
        lab = L.label[i];
        float4 cell = L.cell[i];
        unsigned f = flags_of(cell);
        clipped = f >> 3;
        if (lab != 0 && lab < (unsigned)L.cap && clipped != 0) {
            has = true;
            for (int c = 0; c < 3; ++c) {
                if (!(clipped & (1u << c))) continue;
                bool has_clean = false;
                for (int o = 0; o < 3; ++o)
                    if (o != c && usable(f, o) && clean(f, o)) has_clean = true;
                if (!has_clean) continue;
                int p = pattern_position(c, support_mask(cell, c, lv));
                if (p >= 0) bits |= 1u << (c * 3 + p);
            }
        }
    }
    // Neighbouring cells share their region, so OR the bits per label across
    // the wavefront and let one lane per distinct label touch memory; skip the
    // atomic when every bit is already set.
    unsigned long long active = __ballot(has);
    unsigned any = 0;
    while (active) {
        const int leader = __ffsll((long long)active) - 1;
        const unsigned target = __shfl(lab, leader, 64);
        const bool mine = has && lab == target;
        unsigned vc = mine ? clipped : 0u, vw = mine ? bits : 0u;
        for (int off = 32; off > 0; off >>= 1) {
            vc |= __shfl_xor(vc, off, 64);
            vw |= __shfl_xor(vw, off, 64);
        }
        if (lane == leader) {
            if ((L.region_clipped[target] | vc) != L.region_clipped[target])
                atomicOr(L.region_clipped + target, vc);
            if (vw && (L.region_wanted[target] | vw) != L.region_wanted[target])
                atomicOr(L.region_wanted + target, vw);
        }
        any |= vw;
        active &= ~__ballot(mine);
    }
    if (lane == 0 && any && (L.scalars[2] | any) != L.scalars[2]) atomicOr(L.scalars + 2, any);
}

// ---------------------------------------------------------------- corrections
// region_correction.rs:53-110 (Slot) and :196-261 (two raster-order passes).
//
// The per-sample correction of a ring cell does not depend on which
// accumulator receives it, so opp_contrib computes every ring cell's nine
// (channel, support set) samples once, in parallel. opp_sums then gives each
// (group, slot) pair one wavefront: lanes load 64 consecutive ring entries,
// and the accumulation walks the valid ones in raster order, one f64 add at a
// time, exactly as the reference's sequential loop does.
__global__ __launch_bounds__(THREADS) void opp_contrib(Layout L, float l0, float l1, float l2) {
    const float lv[3] = {l0, l1, l2};
    const unsigned ring = L.scalars[1];
    const unsigned any = L.scalars[2];
    for (unsigned k = blockIdx.x * THREADS + threadIdx.x; k < ring; k += gridDim.x * THREADS) {
        const unsigned i = L.ring_cell[k];
        const unsigned lab = L.label[i];
        const float4 cell = L.cell[i];
        const unsigned rc = lab < (unsigned)L.cap ? L.region_clipped[lab] : 0u;
        unsigned flags = 0;
        for (int c = 0; c < 3; ++c) {
            if (!(rc & (1u << c))) continue;
            if (mean_of(cell, c) < 0.5f * lv[c]) continue;
            for (int p = 0; p < 3; ++p) {
                const int off = c * 3 + p;
                if (!(any & (1u << off))) continue;
                float value;
                if (!correction(cell, c, support_pattern(c, p), lv, value)) continue;
                L.contrib_value[(size_t)k * SLOTS + off] = value;
                flags |= 1u << off;
                if (p == 2) {
                    float ratio;
                    if (support_log_ratio(cell, support_pattern(c, 2), ratio)) {
                        L.contrib_ratio[(size_t)k * 3 + c] = ratio;
                        flags |= 1u << (SLOTS + c);
                    }
                }
            }
        }
        L.ring_label[k] = lab;
        L.contrib_flags[k] = flags;
    }
}

__global__ __launch_bounds__(THREADS) void opp_sums(Layout L) {
    const int lane = threadIdx.x & 63;
    const unsigned count = min(L.scalars[0], (unsigned)L.cap - 1);
    const unsigned ring = L.scalars[1];
    const unsigned any = L.scalars[2];
    const unsigned waves = gridDim.x * (THREADS / 64);
    for (unsigned w = blockIdx.x * (THREADS / 64) + (threadIdx.x >> 6); w < (count + 1) * SLOTS;
         w += waves) {
        const unsigned g = w / SLOTS;
        const int s = (int)(w % SLOTS), c = s / 3, p = s % 3;
        unsigned begin = 0, end = ring, use = any;
        if (g != 0) {
            begin = L.region_first[g];
            end = begin == NONE ? 0u : L.region_last[g] + 1u;
            if (begin == NONE) begin = 0;
            use = L.region_wanted[g];
        }
        unsigned n = 0, kept = 0;
        double accum = 0.0, second = 0.0;
        float mean = 0.0f, tolerance = 0.0f;
        if (use & (1u << s)) {
            for (int pass = 0; pass < 2; ++pass) {
                for (unsigned base = begin; base < end; base += 64) {
                    const unsigned k = base + lane;
                    bool valid = false;
                    float v = 0.0f, r = 0.0f;
                    if (k < end && (g == 0 || L.ring_label[k] == g)) {
                        const unsigned f = L.contrib_flags[k];
                        if (f & (1u << s)) {
                            valid = true;
                            v = L.contrib_value[(size_t)k * SLOTS + s];
                            if (p == 2 && (f & (1u << (SLOTS + c))))
                                r = L.contrib_ratio[(size_t)k * 3 + c];
                        }
                    }
                    unsigned long long mask = __ballot(valid);
                    while (mask) {
                        const int j = __ffsll((long long)mask) - 1;
                        mask &= mask - 1;
                        const float vj = __shfl(v, j, 64);
                        const float rj = __shfl(r, j, 64);
                        const double wide = (double)vj;
                        if (pass == 0) {
                            n += 1;
                            accum += wide;
                            second += wide * wide;
                        } else if (fabsf(vj - mean) <= tolerance) {
                            kept += 1;
                            accum += wide;
                            second += (double)rj;
                        }
                    }
                }
                if (pass == 0) {
                    if (n != 0) {
                        const double cnt = (double)n;
                        const double m = accum / cnt;
                        const double variance = fmax(second / cnt - m * m, 0.0);
                        mean = (float)m;
                        tolerance = (float)(2.0 * sqrt(variance)) + 1e-5f;
                    }
                    accum = 0.0;
                    second = 0.0;
                }
            }
        }
        if (lane == 0) {
            // Slot::value: fewer than the minimum samples or nothing kept
            // leaves the slot empty.
            float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (n >= MIN_BORDER_CELLS && kept != 0) {
                const double kd = (double)kept;
                out = make_float4((float)(accum / kd), (float)(second / kd), 1.0f, 0.0f);
            }
            L.corrections[(size_t)g * SLOTS + s] = out;
        }
    }
}

// ---------------------------------------------------------------- sites
__device__ __forceinline__ int total_key(float v) {
    int i = __float_as_int(v);
    return i ^ (int)(((unsigned)(i >> 31)) >> 1);
}

struct Site {
    const Layout* L;
    const float* lv;

    __device__ bool lookup(unsigned label, int target, unsigned support, float& offset,
                           bool& has_ratio, float& ratio) const {
        int p = pattern_position(target, support);
        if (p < 0) return false;
        int off = target * 3 + p;
        bool want_ratio = __popc(support) == 2;
        float4 v = label < (unsigned)L->cap ? L->corrections[(size_t)label * SLOTS + off]
                                            : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (v.z == 0.0f) v = L->corrections[off];
        if (v.z == 0.0f) return false;
        offset = v.x;
        has_ratio = want_ratio;
        ratio = v.y;
        return true;
    }

    // opposed.rs:282-327
    __device__ bool local_correction(int target, unsigned support, bool ht, float tr, int tx,
                                     int ty, float& out, int& out_radius) const {
        float corr[48];
        int count = 0;
        for (int radius = 1; radius <= 3; ++radius) {
            int y0 = max(ty - radius, 0), y1 = min(ty + radius, L->ch - 1);
            int x0 = max(tx - radius, 0), x1 = min(tx + radius, L->cw - 1);
            for (int cy = y0; cy <= y1; ++cy)
                for (int cx = x0; cx <= x1; ++cx) {
                    if (max(abs(cx - tx), abs(cy - ty)) != radius) continue;
                    float4 cand = L->cell[cy * L->cw + cx];
                    unsigned f = flags_of(cand);
                    if (!(usable(f, target) && clean(f, target))) continue;
                    float value;
                    if (!correction(cand, target, support, lv, value)) continue;
                    float cr = 0.0f;
                    bool hc = support_log_ratio(cand, support, cr);
                    if (!chromaticity_agrees(support, ht, tr, hc, cr)) continue;
                    if (count < 48) corr[count++] = value;
                }
            if (count >= 3) {
                // Median in Rust's total order: the element(s) of rank
                // count / 2 (and count / 2 - 1 when even), ties by position.
                int middle = count / 2;
                float lo = 0.0f, hi = 0.0f;
                for (int a = 0; a < count; ++a) {
                    int ka = total_key(corr[a]);
                    int rank = 0;
                    for (int b = 0; b < count; ++b) {
                        int kb = total_key(corr[b]);
                        rank += (kb < ka) || (kb == ka && b < a);
                    }
                    if (rank == middle) hi = corr[a];
                    if (rank == middle - 1) lo = corr[a];
                }
                out = count % 2 == 0 ? (lo + hi) * 0.5f : hi;
                out_radius = radius;
                return true;
            }
        }
        return false;
    }

    // opposed.rs:243-278
    __device__ bool estimate(const float4& target, int channel, int tx, int ty, float& out) const {
        unsigned support = support_mask(target, channel, lv);
        float target_reference;
        if (!reference(target, support, target_reference)) return false;
        float tr = 0.0f;
        bool ht = support_log_ratio(target, support, tr);
        float local = 0.0f;
        int radius = 0;
        bool has_local = local_correction(channel, support, ht, tr, tx, ty, local, radius);
        unsigned label = L->label[ty * L->cw + tx];
        float region = 0.0f, br = 0.0f;
        bool bhr = false;
        bool has_region = lookup(label, channel, support, region, bhr, br) &&
                          chromaticity_agrees(support, ht, tr, bhr, br);
        float corr;
        if (has_local && has_region) {
            float weight = (float)(3 + 1 - radius) / 3.0f;
            corr = weight * local + (1.0f - weight) * region;
        } else if (has_local) {
            corr = local;
        } else if (has_region) {
            corr = region;
        } else {
            return false;
        }
        float e = target_reference + corr;
        if (!(is_finite(e) && e > 0.0f)) return false;
        out = e;
        return true;
    }
};

// opposed.rs:97-162
__global__ __launch_bounds__(THREADS) void opp_sites(float* __restrict__ m, Layout L, int pattern,
                                                     float l0, float l1, float l2, int has_ceiling,
                                                     float c0, float c1, float c2,
                                                     DeviceStats* stats) {
    const float lv[3] = {l0, l1, l2};
    const float ceiling[3] = {c0, c1, c2};
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    unsigned reconstructed = 0, changed = 0, fallback = 0, unsupported = 0, neutralised = 0;
    if (x < L.w && y < L.h) {
        float* sp = m + (size_t)y * L.w + x;
        float sample = *sp;
        int c = color_at(pattern, x, y);
        float level = lv[c];
        if (sample >= level) {
            int tx = x / 2, ty = y / 2;
            float4 target = L.cell[ty * L.cw + tx];
            unsigned f = flags_of(target);
            int support_count = 0;
            for (int o = 0; o < 3; ++o)
                if (o != c && usable(f, o) && clean(f, o)) ++support_count;
            Site site{&L, lv};
            float est = 0.0f;
            bool has = support_count != 0 && site.estimate(target, c, tx, ty, est);
            bool neutralise = false;
            if (!has) {
                fallback = 1;
                unsupported = support_count == 0;
                neutralise = true;
            } else {
                reconstructed = 1;
                float bounded = fminf(est, 4.0f * level);
                if (bounded > sample) {
                    *sp = bounded;
                    changed = 1;
                } else {
                    neutralise = true;
                }
            }
            if (neutralise && has_ceiling && sample > ceiling[c]) {
                *sp = ceiling[c];
                neutralised = 1;
            }
        }
    }
    unsigned v[5] = {reconstructed, changed, fallback, unsupported, neutralised};
    for (int k = 0; k < 5; ++k) {
        unsigned t = wave_sum(v[k]);
        if ((threadIdx.x & 63) == 0 && t) atomicAdd(stats->opposed + k, (unsigned long long)t);
    }
}

Layout layout_of(const OpposedArgs& a) {
    Layout L{};
    L.w = a.width;
    L.h = a.height;
    L.cw = a.width / 2 + a.width % 2;
    L.ch = a.height / 2 + a.height % 2;
    L.cells = L.cw * L.ch;
    // regions.rs:108-114: separate regions are at least 4 cells apart.
    L.cap = ((L.cw + 3) / 4) * ((L.ch + 3) / 4) + 1;
    uintptr_t p = (uintptr_t)a.scratch;
    auto take = [&](size_t bytes) {
        uintptr_t q = p;
        p += (bytes + 255) & ~(size_t)255;
        return (void*)q;
    };
    const size_t cells = (size_t)L.cells, cap = (size_t)L.cap;
    const size_t scan_blocks = (cells + SCAN_BLOCK - 1) / SCAN_BLOCK + 1;
    L.cell = (float4*)take(cells * sizeof(float4));
    L.label = (unsigned*)take(cells * 4);
    L.rank = (unsigned*)take(cells * 4);
    L.block_sums = (unsigned*)take(scan_blocks * 4);
    L.ring_cell = (unsigned*)take(cells * 4);
    L.ring_label = (unsigned*)take(cells * 4);
    L.contrib_flags = (unsigned*)take(cells * 4);
    L.contrib_value = (float*)take(cells * SLOTS * 4);
    L.contrib_ratio = (float*)take(cells * 3 * 4);
    L.region_clipped = (unsigned*)take(cap * 4);
    L.region_wanted = (unsigned*)take(cap * 4);
    L.region_first = (unsigned*)take(cap * 4);
    L.region_last = (unsigned*)take(cap * 4);
    L.corrections = (float4*)take(cap * SLOTS * sizeof(float4));
    L.scalars = (unsigned*)take(64);
    return L;
}

}  // namespace

size_t opposed_scratch_bytes(int width, int height) {
    OpposedArgs a{};
    a.width = width;
    a.height = height;
    a.scratch = nullptr;
    Layout L = layout_of(a);
    return (size_t)(uintptr_t)L.scalars + 256;
}

hipError_t launch_opposed(const OpposedArgs& a, hipStream_t s, Profiler* prof) {
    Layout L = layout_of(a);
    if ((size_t)((uintptr_t)L.scalars - (uintptr_t)a.scratch) + 64 > a.scratch_bytes)
        return hipErrorInvalidValue;
    const float l0 = a.levels[0], l1 = a.levels[1], l2 = a.levels[2];
    const int cell_blocks = (L.cells + THREADS - 1) / THREADS;
    const int scan_blocks = (L.cells + SCAN_BLOCK - 1) / SCAN_BLOCK;
    const uint64_t px = (uint64_t)L.w * L.h, cells = (uint64_t)L.cells;
    RHD_TIMED(prof, s, "opp_cells", 4 * px + 16 * cells,
              hipLaunchKernelGGL(opp_cells, dim3(cell_blocks), dim3(THREADS), 0, s, a.mosaic, L,
                                 a.pattern, l0, l1, l2));
    RHD_TIMED(prof, s, "opp_label", 16 * cells + 4 * cells + 8 * cells + 8 * cells,
              hipLaunchKernelGGL(opp_dilate, dim3(cell_blocks), dim3(THREADS), 0, s, L);
              hipLaunchKernelGGL(opp_union, dim3(cell_blocks), dim3(THREADS), 0, s, L);
              hipLaunchKernelGGL(opp_flatten, dim3(cell_blocks), dim3(THREADS), 0, s, L);
              hipLaunchKernelGGL(opp_scan_count, dim3(scan_blocks), dim3(THREADS), 0, s, L,
                                 (int)ROOTS);
              hipLaunchKernelGGL(opp_scan_blocks, dim3(1), dim3(THREADS), 0, s, L, scan_blocks,
                                 (int)ROOTS);
              hipLaunchKernelGGL(opp_scan_emit, dim3(scan_blocks), dim3(THREADS), 0, s, L,
                                 (int)ROOTS);
              hipLaunchKernelGGL(opp_relabel, dim3(cell_blocks), dim3(THREADS), 0, s, L));
    RHD_TIMED(prof, s, "opp_regions", 24 * cells,
              hipLaunchKernelGGL(opp_region_init, dim3(256), dim3(THREADS), 0, s, L, a.stats);
              hipLaunchKernelGGL(opp_regions, dim3(cell_blocks), dim3(THREADS), 0, s, L, l0, l1,
                                 l2);
              hipLaunchKernelGGL(opp_scan_count, dim3(scan_blocks), dim3(THREADS), 0, s, L,
                                 (int)RING);
              hipLaunchKernelGGL(opp_scan_blocks, dim3(1), dim3(THREADS), 0, s, L, scan_blocks,
                                 (int)RING);
              hipLaunchKernelGGL(opp_scan_emit, dim3(scan_blocks), dim3(THREADS), 0, s, L,
                                 (int)RING));
    RHD_TIMED(prof, s, "opp_corrections", 0,
              hipLaunchKernelGGL(opp_contrib, dim3(2048), dim3(THREADS), 0, s, L, l0, l1, l2);
              hipLaunchKernelGGL(opp_sums, dim3(1024), dim3(THREADS), 0, s, L));
    RHD_TIMED(prof, s, "opp_sites", 8 * px,
              hipLaunchKernelGGL(opp_sites, dim3((L.w + 63) / 64, (L.h + 3) / 4), dim3(THREADS), 0,
                                 s, a.mosaic, L, a.pattern, l0, l1, l2, a.has_ceiling,
                                 a.ceiling[0], a.ceiling[1], a.ceiling[2], a.stats));
    return hipGetLastError();
}

}  // namespace rhd
