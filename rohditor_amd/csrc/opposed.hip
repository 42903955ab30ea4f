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
//   cells      one thread per 2x2 cell: means and flags (float4 per cell), and
//              per excluded channel the two-channel cube-root reference and
//              log2 ratio of the other two channels
//   label      per 64x16-cell tile in LDS: the Chebyshev-1 dilation and
//              union-find of the dilated cells; then the unions across tile
//              borders in global memory, and a flatten. Every set's root is
//              its smallest index, i.e. the component's first raster cell
//   scan       exclusive scan of the root flags -> compact label = rank + 1
//              (the reference's numbering by first raster appearance)
//   regions    per-region clipped channels and wanted slots (atomic OR),
//              and the ring flags (label != 0, no clipped channel)
//   scan       ring cells compacted in raster order, plus each region's
//              first/last ring index
//   contrib    every ring cell's nine correction samples, in parallel
//   sums       one workgroup per group (0 = image-wide pool, r = region r),
//              four wavefronts staging the ring entries and lane s of a fifth
//              owning slot s: both passes over the group's ring
//              entries in raster order, accumulated one f64 add at a time, so
//              the sums are bit-identical to the reference's
//   sites      one thread per (cell, channel) with a suspected site: the
//              estimate, then per site the bound, the neutral ceiling and the
//              statistics
// cbrt and log2 are the platform libm's cbrtf and log2f bit for bit
// (libm_f32.hpp), as the reference's f32::cbrt and f32::log2 are, so every
// output is bit-identical to the reference's.
#include "libm_f32.hpp"
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
    uint8_t* flag8;                // the same flags, one byte per cell, for the
                                   // kernels that read nothing else of a cell
    float* pair_ref;               // [3][cells]: cube-root reference of the two channels
                                   // other than the plane's index, NaN when undefined
    float* pair_lr;                // [3][cells]: log2 ratio of those two channels, or NaN
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
    unsigned* site_counts;         // per opp_sites workgroup: its five statistics
    unsigned* site_list;           // (sites block, channel) pairs with a suspected cell
    float2* site_value;            // per (channel, cell): target reference, local median
    uint8_t* site_code;            // per (channel, cell): LOCAL_REFERENCE | ring radius
    unsigned* scalars;             // [0] region count, [1] ring count, [2] any wanted,
                                   // [3] site_list length, [4] next site_list item
    const uint32_t* log2_bitmap;   // libm_f32.hpp exception bitmap
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
__device__ __forceinline__ float cbrt_f32(float v) { return libm::cbrt_libm(v); }
__device__ __forceinline__ float log2_f32(float v, const uint32_t* bitmap) {
    return libm::log2_libm(v, bitmap);
}

// The two-channel parts of support.rs:51-72 (reference) and :101-110
// (support_log_ratio) depend only on a cell's means, and a two-channel support
// set of a target channel is always the other two channels. opp_cells
// therefore evaluates them once per cell and excluded channel; NaN marks the
// cases where the reference returns None.
__device__ __forceinline__ bool positive_finite(float v) { return is_finite(v) && v > 0.0f; }
// `roots[c]` is cbrt_f32 of channel c's mean wherever that mean is positive
// and finite (each channel's root serves two of the pairs).
__device__ __forceinline__ void pair_values(const float4& cell, const float* roots, int excluded,
                                            const uint32_t* bitmap, float& ref, float& lr) {
    const int a = excluded == 0 ? 1 : 0, b = excluded == 2 ? 1 : 2;
    const float ma = mean_of(cell, a), mb = mean_of(cell, b);
    const float nan = __int_as_float(0x7fc00000);
    ref = nan;
    if (positive_finite(ma) && positive_finite(mb)) {
        float sum = 0.0f;
        sum += roots[a];
        sum += roots[b];
        const float root = sum / 2.0f;
        ref = root * root * root;
    }
    const float ratio = ma / mb;
    lr = is_finite(ratio) && ratio > 0.0f ? log2_f32(ratio, bitmap) : nan;
}

// support.rs:51-72, with the two-channel value taken from pair_values.
__device__ __forceinline__ bool reference(const float4& cell, unsigned support, float pair_ref,
                                          float& out) {
    const int n = __popc(support);
    if (n == 1) {
        float v = mean_of(cell, __ffs(support) - 1);
        if (is_finite(v) && v > 0.0f) {
            out = v;
            return true;
        }
        return false;
    }
    if (n != 2 || pair_ref != pair_ref) return false;
    out = pair_ref;
    return true;
}
// support.rs:78-95
__device__ __forceinline__ bool correction(const float4& cell, int target, unsigned support,
                                           float pair_ref, const float* lv, float& out) {
    if (!bright_clean(cell, target, lv)) return false;
    for (int c = 0; c < 3; ++c)
        if ((support & (1u << c)) && !bright_clean(cell, c, lv)) return false;
    float ref;
    if (!reference(cell, support, pair_ref, ref)) return false;
    float v = mean_of(cell, target) - ref;
    if (!is_finite(v)) return false;
    out = v;
    return true;
}
// support.rs:101-110, with the ratio's log2 taken from pair_values.
__device__ __forceinline__ bool support_log_ratio(unsigned support, float pair_lr, float& out) {
    if (__popc(support) != 2 || pair_lr != pair_lr) return false;
    out = pair_lr;
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
    // The site list's length and next item (opp_site_list, opp_sites_local).
    if (i == 0) {
        L.scalars[3] = 0;
        L.scalars[4] = 0;
    }
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
    const float4 cell = make_float4(means[0], means[1], means[2], __uint_as_float(flags));
    L.cell[i] = cell;
    L.flag8[i] = (uint8_t)flags;
}

// The pair reference and log ratio of every cell for each excluded channel
// (pair_values); only the site searches and the ring contributions read
// them, so this runs on the auxiliary stream beside the labelling.
__global__ __launch_bounds__(THREADS) void opp_pairs(Layout L) {
    const int i = blockIdx.x * THREADS + threadIdx.x;
    if (i >= L.cells) return;
    const float4 cell = L.cell[i];
    const float means[3] = {cell.x, cell.y, cell.z};
    float roots[3];
    for (int c = 0; c < 3; ++c) roots[c] = positive_finite(means[c]) ? cbrt_f32(means[c]) : 0.0f;
    for (int e = 0; e < 3; ++e) {
        float ref, lr;
        pair_values(cell, roots, e, L.log2_bitmap, ref, lr);
        L.pair_ref[(size_t)e * L.cells + i] = ref;
        L.pair_lr[(size_t)e * L.cells + i] = lr;
    }
}

// Union-find with union by linking the larger root under the smaller one, so
// every set's root is its smallest member: the component's first cell in
// raster order, which is the order the reference numbers its labels in.
// Parents are read with relaxed atomic loads at the scope the array lives at.
template <int SCOPE>
__device__ __forceinline__ unsigned load_parent(unsigned* p, unsigned x) {
    return __hip_atomic_load(p + x, __ATOMIC_RELAXED, SCOPE);
}
// The root of x, pointing every node passed on the way at its grandparent
// (path splitting), so that chains of tile roots built by the border unions
// are walked once rather than by every cell. Only a node seen with a parent
// other than itself is written, and roots never become non-roots again, so a
// root's link is never touched. A parent only moves to a smaller member of
// the same set. A split store may land after a concurrent merge lowered the
// same node's parent with atomicMin and undo that link; that merge saw the
// old parent returned and goes on to unite it with its other side, so no
// union is lost (the pointer jumping of ECL-CC, Jaiganesh and Burtscher,
// SC 2018, with atomicMin hooking).
template <int SCOPE>
__device__ __forceinline__ unsigned find_root(unsigned* p, unsigned x) {
    unsigned y = load_parent<SCOPE>(p, x);
    while (y != x) {
        const unsigned z = load_parent<SCOPE>(p, y);
        if (z != y) __hip_atomic_store(p + x, z, __ATOMIC_RELAXED, SCOPE);
        x = y;
        y = z;
    }
    return x;
}
template <int SCOPE>
__device__ void merge(unsigned* p, unsigned a, unsigned b) {
    while (true) {
        a = find_root<SCOPE>(p, a);
        b = find_root<SCOPE>(p, b);
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

// Labelling tiles: LX x LY cells per workgroup.
constexpr int LX = 64, LY = 16;
constexpr int LHX = LX + 2, LHY = LY + 2;

// regions.rs:126-183 (Chebyshev radius-1 dilation of the clipped cells) and
// regions.rs:188-252 (8-connected unions with the already-visited neighbours
// W, NW, N, NE), inside one tile in LDS. Each cell's parent leaves as the
// global index of its tile-local root, which is the smallest index of its
// tile-local component; cells outside the mask get NONE.
__global__ __launch_bounds__(THREADS) void opp_label_tiles(Layout L) {
    __shared__ unsigned char s_clip[LHY * LHX];
    __shared__ unsigned s_par[LY * LX];
    const uint3 blk = xcd_block();
    const int x0 = blk.x * LX, y0 = blk.y * LY;
    bool any = false;
    for (int k = threadIdx.x; k < LHY * LHX; k += THREADS) {
        const int x = x0 + k % LHX - 1, y = y0 + k / LHX - 1;
        bool clipped = false;
        if (x >= 0 && y >= 0 && x < L.cw && y < L.ch)
            clipped = (L.flag8[(size_t)y * L.cw + x] >> 3) != 0;
        s_clip[k] = clipped;
        any |= clipped;
    }
    // A tile with no clipped cell in its window has no dilated cell either.
    if (!__syncthreads_or(any)) {
        for (int k = threadIdx.x; k < LY * LX; k += THREADS) {
            const int x = x0 + k % LX, y = y0 + k / LX;
            if (x < L.cw && y < L.ch) L.label[(size_t)y * L.cw + x] = NONE;
        }
        return;
    }
    for (int k = threadIdx.x; k < LY * LX; k += THREADS) {
        const int lx = k % LX, ly = k / LX;
        bool fg = false;
        if (x0 + lx < L.cw && y0 + ly < L.ch)
            for (int dy = 0; dy < 3; ++dy)
                for (int dx = 0; dx < 3; ++dx) fg |= s_clip[(ly + dy) * LHX + lx + dx] != 0;
        s_par[k] = fg ? (unsigned)k : NONE;
    }
    __syncthreads();
    // Unions with the visited in-tile neighbours, leaving out those that
    // reach the cell through another union: with N in the mask, NW and NE
    // are N's own W and E neighbours; otherwise W, when in the mask, has NW
    // as its N. By induction over the rows every 8-adjacent pair still ends
    // in one set, so the components (and their smallest-index roots) are
    // those of all four unions. Neighbours in other tiles are left to
    // opp_label_borders, which takes all of them.
    for (int k = threadIdx.x; k < LY * LX; k += THREADS) {
        if (s_par[k] == NONE) continue;
        const int lx = k % LX, ly = k / LX;
        auto in_mask = [&](int x, int y) {
            return x >= 0 && y >= 0 && x < LX &&
                   load_parent<__HIP_MEMORY_SCOPE_WORKGROUP>(s_par, (unsigned)(y * LX + x)) != NONE;
        };
        auto unite = [&](int x, int y) {
            merge<__HIP_MEMORY_SCOPE_WORKGROUP>(s_par, (unsigned)k, (unsigned)(y * LX + x));
        };
        if (in_mask(lx, ly - 1)) {
            unite(lx, ly - 1);
            continue;
        }
        if (in_mask(lx - 1, ly))
            unite(lx - 1, ly);
        else if (in_mask(lx - 1, ly - 1))
            unite(lx - 1, ly - 1);
        if (in_mask(lx + 1, ly - 1)) unite(lx + 1, ly - 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < LY * LX; k += THREADS) {
        const int lx = k % LX, ly = k / LX;
        const int x = x0 + lx, y = y0 + ly;
        if (x >= L.cw || y >= L.ch) continue;
        unsigned out = NONE;
        if (s_par[k] != NONE) {
            const unsigned r = find_root<__HIP_MEMORY_SCOPE_WORKGROUP>(s_par, (unsigned)k);
            out = (unsigned)((y0 + (int)(r / LX)) * L.cw + x0 + (int)(r % LX));
        }
        L.label[(size_t)y * L.cw + x] = out;
    }
}

// The unions that cross tile borders: a cell in a tile's first row, first
// column or last column with a W, NW, N or NE neighbour in another tile.
__global__ __launch_bounds__(THREADS) void opp_label_borders(Layout L, int tiles_x) {
    const int per_tile = LX + 2 * (LY - 1);
    const int t = blockIdx.x * THREADS + threadIdx.x;
    const int tile = t / per_tile, e = t % per_tile;
    if (tile >= tiles_x * ((L.ch + LY - 1) / LY)) return;
    const int x0 = (tile % tiles_x) * LX, y0 = (tile / tiles_x) * LY;
    int lx, ly;
    if (e < LX) {
        lx = e;
        ly = 0;
    } else if (e < LX + LY - 1) {
        lx = 0;
        ly = e - LX + 1;
    } else {
        lx = LX - 1;
        ly = e - (LX + LY - 1) + 1;
    }
    const int cx = x0 + lx, cy = y0 + ly;
    if (cx >= L.cw || cy >= L.ch) return;
    const unsigned i = (unsigned)(cy * L.cw + cx);
    if (load_parent<__HIP_MEMORY_SCOPE_AGENT>(L.label, i) == NONE) return;
    const int nx[4] = {-1, -1, 0, 1}, ny[4] = {0, -1, -1, -1};
    for (int q = 0; q < 4; ++q) {
        const int x = cx + nx[q], y = cy + ny[q];
        if (x < 0 || y < 0 || x >= L.cw) continue;
        if (x >= x0 && x < x0 + LX && y >= y0) continue;  // same tile: already united
        const unsigned j = (unsigned)(y * L.cw + x);
        if (load_parent<__HIP_MEMORY_SCOPE_AGENT>(L.label, j) != NONE)
            merge<__HIP_MEMORY_SCOPE_AGENT>(L.label, i, j);
    }
}

__global__ __launch_bounds__(THREADS) void opp_flatten(Layout L) {
    int i = blockIdx.x * THREADS + threadIdx.x;
    if (i >= L.cells) return;
    unsigned p = L.label[i];
    if (p == NONE) return;
    unsigned r = find_root<__HIP_MEMORY_SCOPE_AGENT>(L.label, (unsigned)i);
    if (r != p) L.label[i] = r;
}

// ---------------------------------------------------------------- scan
// Exclusive scan of a predicate over the cells, in blocks of SCAN_BLOCK.
enum ScanMode { ROOTS = 0, RING = 1 };

__device__ __forceinline__ unsigned predicate(const Layout& L, int mode, int i) {
    if (i >= L.cells) return 0;
    if (mode == ROOTS) return L.label[i] == (unsigned)i ? 1u : 0u;
    unsigned lab = L.label[i];
    return lab != 0 && (L.flag8[i] >> 3) == 0 ? 1u : 0u;
}

// The predicate bits of cells base .. base + SCAN_PER_THREAD - 1 (bit k for
// cell base + k). A thread's labels and flag bytes are contiguous and 16-byte
// aligned (the layout's arrays start on 256-byte boundaries), so a whole run
// arrives in four uint4 label loads and one uint4 flag load.
// The cells' labels are left in lab.
__device__ __forceinline__ unsigned predicate_bits(const Layout& L, int mode, int base,
                                                   unsigned (&lab)[SCAN_PER_THREAD]) {
    static_assert(SCAN_PER_THREAD == 16, "four uint4 of labels, one of flag bytes");
    unsigned bits = 0;
    if (base + SCAN_PER_THREAD <= L.cells) {
        const uint4* lp = reinterpret_cast<const uint4*>(L.label + base);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint4 v = lp[q];
            lab[4 * q] = v.x;
            lab[4 * q + 1] = v.y;
            lab[4 * q + 2] = v.z;
            lab[4 * q + 3] = v.w;
        }
        if (mode == ROOTS) {
#pragma unroll
            for (int k = 0; k < SCAN_PER_THREAD; ++k)
                bits |= (lab[k] == (unsigned)(base + k) ? 1u : 0u) << k;
        } else {
            const uint4 f = *reinterpret_cast<const uint4*>(L.flag8 + base);
            const unsigned fw[4] = {f.x, f.y, f.z, f.w};
#pragma unroll
            for (int k = 0; k < SCAN_PER_THREAD; ++k) {
                const unsigned flag = (fw[k >> 2] >> (8 * (k & 3))) & 0xffu;
                bits |= (lab[k] != 0 && (flag >> 3) == 0 ? 1u : 0u) << k;
            }
        }
        return bits;
    }
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        lab[k] = base + k < L.cells ? L.label[base + k] : 0u;
        bits |= predicate(L, mode, base + k) << k;
    }
    return bits;
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
    unsigned lab[SCAN_PER_THREAD];
    const unsigned n = __popc(predicate_bits(L, mode, base, lab));
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
    unsigned lab[SCAN_PER_THREAD];
    const unsigned flags = predicate_bits(L, mode, base, lab), n = __popc(flags);
    unsigned excl;
    block_exclusive_scan(n, excl);
    unsigned pos = L.block_sums[blockIdx.x] + excl;
    // Ring entries of one label come in runs; each run updates its region's
    // first/last index once, and only when it would change them (the values
    // only move one way, so an observed value that is already better holds;
    // unconditional atomics contend on the large regions and are slower).
    unsigned run_lab = NONE, run_first = 0, run_last = 0;
    auto flush = [&]() {
        if (run_lab >= (unsigned)L.cap) return;
        if (run_first < L.region_first[run_lab]) atomicMin(L.region_first + run_lab, run_first);
        if (run_last > L.region_last[run_lab]) atomicMax(L.region_last + run_lab, run_last);
    };
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        if (!(flags & (1u << k))) continue;
        int i = base + k;
        if (mode == ROOTS) {
            L.rank[i] = pos;
        } else {
            L.ring_cell[pos] = (unsigned)i;
            if (lab[k] != run_lab) {
                flush();
                run_lab = lab[k];
                run_first = pos;
            }
            run_last = pos;
        }
        ++pos;
    }
    if (mode != ROOTS) flush();
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
// Each workgroup walks REGION_STRIDE consecutive blocks of cells; a
// wavefront's successive slices lie along the same image rows, so they mostly
// belong to the region it saw last. The wavefront remembers the bits it has
// already set for that region and issues an atomic only for new ones, without
// reading the region's words back; the next slice's cells load while the
// current one is processed.
constexpr int REGION_STRIDE = 32;
__global__ __launch_bounds__(THREADS) void opp_regions(Layout L, float l0, float l1, float l2) {
    const float lv[3] = {l0, l1, l2};
    const int lane = threadIdx.x & 63;
    unsigned any = 0;
    unsigned seen_label = 0, seen_clipped = 0, seen_wanted = 0;  // wavefront-uniform
    const int begin = blockIdx.x * REGION_STRIDE * THREADS;
    const int end = min(begin + REGION_STRIDE * THREADS, L.cells);
    int i = begin + (int)threadIdx.x;
    unsigned next_lab = i < end ? L.label[i] : 0u;
    float4 next_cell = i < end ? L.cell[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (; i - (int)threadIdx.x < end; i += THREADS) {
        const unsigned lab = next_lab;
        const float4 cell = next_cell;
        const int j = i + THREADS;
        if (j < end) {
            next_lab = L.label[j];
            next_cell = L.cell[j];
        }
        unsigned clipped = 0, bits = 0;
        bool has = false;
        if (i < end) {
            const unsigned f = flags_of(cell);
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
        // OR the bits per label across the wavefront; one lane per distinct
        // label sets the bits this wavefront has not set for it yet.
        unsigned long long active = __ballot(has);
        while (active) {
            const int leader = __ffsll((long long)active) - 1;
            const unsigned target = __shfl(lab, leader, 64);
            const bool mine = has && lab == target;
            unsigned vc = mine ? clipped : 0u, vw = mine ? bits : 0u;
            for (int off = 32; off > 0; off >>= 1) {
                vc |= __shfl_xor(vc, off, 64);
                vw |= __shfl_xor(vw, off, 64);
            }
            if (target != seen_label) {
                seen_label = target;
                seen_clipped = 0;
                seen_wanted = 0;
            }
            const unsigned new_c = vc & ~seen_clipped, new_w = vw & ~seen_wanted;
            if (lane == leader) {
                if (new_c) atomicOr(L.region_clipped + target, new_c);
                if (new_w) atomicOr(L.region_wanted + target, new_w);
            }
            seen_clipped |= vc;
            seen_wanted |= vw;
            any |= vw;
            active &= ~__ballot(mine);
        }
    }
    if (lane == 0 && any && (L.scalars[2] | any) != L.scalars[2]) atomicOr(L.scalars + 2, any);
}

// ---------------------------------------------------------------- corrections
// region_correction.rs:53-110 (Slot) and :196-261 (two raster-order passes).
//
// The per-sample correction of a ring cell does not depend on which
// accumulator receives it, so opp_contrib computes every ring cell's nine
// (channel, support set) samples once, in parallel. opp_sums then gives each
// group (0 = the image-wide pool, r = region r) one workgroup. Producer
// wavefronts stage SUM_BATCH consecutive ring entries at a time in LDS, with
// the next batch's loads in flight, and lane s of the consumer wavefront walks
// the staged entries in raster order for slot s, one f64 add at a time,
// exactly as the reference's sequential loop does. Entries that do not count
// add +0.0, which leaves an accumulator that starts at +0.0 unchanged, so the
// adds run without branches.
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
        // The pair values of the ring cell, formed here as opp_pairs forms
        // them (the pair planes are written beside the region sums).
        float roots[3];
        bool rooted = false;
        for (int c = 0; c < 3; ++c) {
            if (!(rc & (1u << c))) continue;
            if (mean_of(cell, c) < 0.5f * lv[c]) continue;
            if (!rooted) {
                for (int q = 0; q < 3; ++q)
                    roots[q] = positive_finite(mean_of(cell, q)) ? cbrt_f32(mean_of(cell, q)) : 0.0f;
                rooted = true;
            }
            float pref, plr;
            pair_values(cell, roots, c, L.log2_bitmap, pref, plr);
            for (int p = 0; p < 3; ++p) {
                const int off = c * 3 + p;
                if (!(any & (1u << off))) continue;
                float value;
                if (!correction(cell, c, support_pattern(c, p), pref, lv, value)) continue;
                L.contrib_value[(size_t)k * SLOTS + off] = value;
                flags |= 1u << off;
                if (p == 2) {
                    float ratio;
                    if (support_log_ratio(support_pattern(c, 2), plr, ratio)) {
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

constexpr int SUM_BATCH = 256;
constexpr int SUM_PRODUCERS = SUM_BATCH / 64;
// Two consumer wavefronts: wave 0 adds the first accumulator of every slot,
// wave 1 the second. They run on different SIMDs, so each carries one
// dependent f64 chain instead of two interleaved ones.
constexpr int SUM_CONSUMERS = 2;
constexpr int SUM_THREADS = 64 * (SUM_CONSUMERS + SUM_PRODUCERS);
constexpr int SUM_UNROLL = 4;
static_assert((SUM_BATCH / 2) % (4 * SUM_UNROLL) == 0, "whole register groups per batch");

// One ring entry in a producer lane's registers: its slot flags (zero when it
// belongs to another region), nine samples and three ratios.
struct SumEntry {
    unsigned fl;
    float val[SLOTS];
    float rat[3];
};

__device__ __forceinline__ void sum_load(const Layout& L, unsigned g, unsigned k, unsigned end,
                                         SumEntry& b) {
    const bool in = k < end;
    const unsigned lab = in ? L.ring_label[k] : 0u;
    const unsigned f = in ? L.contrib_flags[k] : 0u;
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) b.val[s] = in ? L.contrib_value[(size_t)k * SLOTS + s] : 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) b.rat[c] = in ? L.contrib_ratio[(size_t)k * 3 + c] : 0.0f;
    b.fl = in && (g == 0 || lab == g) ? f : 0u;
}

// Staging buffers of one batch: per slot the exact f64 terms it adds, and per
// producer wavefront the number of contributing entries of each slot and
// whether any of its entries belongs to the group.
// Slot rows are SUM_PITCH doubles apart (2 more than a batch) so that the
// nine lanes reading them start in different LDS banks.
constexpr int SUM_PITCH = SUM_BATCH + 2;
struct SumStage {
    double a[SLOTS * SUM_PITCH];
    double b[SLOTS * SUM_PITCH];
    unsigned count[SUM_PRODUCERS][SLOTS];
    unsigned staged[SUM_PRODUCERS];
};

// One pass of every slot of one group over ring entries [begin, end). The
// SUM_PRODUCERS producer waves each load a quarter of batch i (their share of
// the next batch's loads in flight behind it) and turn it into the terms each
// slot adds: +0.0 for an entry that does not count, which leaves an
// accumulator that starts at +0.0 unchanged, and ballot counts. Meanwhile
// lane s of consumer wave 0 (first accumulator) and of consumer wave 1
// (second accumulator) add slot s's terms of batch i - 1 in raster order, one
// f64 add at a time, as the reference's sequential loop does. The two stage
// buffers alternate.
template <int PASS>
__device__ __forceinline__ void sum_pass(const Layout& L, unsigned g, unsigned begin,
                                         unsigned end, const float* means, const float* tols,
                                         SumStage* st, unsigned& count, double& accum,
                                         double& second) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int q = wave - SUM_CONSUMERS, e = q * 64 + lane;
    const unsigned batches = begin < end ? (end - begin + SUM_BATCH - 1) / SUM_BATCH : 0u;
    SumEntry b;
    if (wave >= SUM_CONSUMERS && batches != 0) sum_load(L, g, begin + e, end, b);
    for (unsigned i = 0; i <= batches; ++i) {
        if (wave >= SUM_CONSUMERS && i < batches) {
            SumStage& out = st[i & 1];
            const unsigned base = begin + i * SUM_BATCH;
            // Every producer writes its terms, so a batch that another
            // producer stages never holds stale ones.
            unsigned counts[SLOTS];
#pragma unroll
            for (int t = 0; t < SLOTS; ++t) {
                const float v = b.val[t];
                bool take = (b.fl >> t) & 1u;
                double ta, tb;
                if (PASS == 0) {
                    ta = take ? (double)v : 0.0;
                    tb = ta * ta;
                } else {
                    take = take && fabsf(v - means[t]) <= tols[t];
                    const bool ratio = t % 3 == 2 && ((b.fl >> (SLOTS + t / 3)) & 1u);
                    ta = take ? (double)v : 0.0;
                    tb = take && ratio ? (double)b.rat[t / 3] : 0.0;
                }
                out.a[t * SUM_PITCH + e] = ta;
                out.b[t * SUM_PITCH + e] = tb;
                counts[t] = __popcll(__ballot(take));
            }
            if (lane < SLOTS) {
                unsigned c = 0;
#pragma unroll
                for (int t = 0; t < SLOTS; ++t) c = lane == t ? counts[t] : c;
                out.count[q][lane] = c;
            }
            const bool staged = __ballot(b.fl != 0) != 0;
            if (lane == 0) out.staged[q] = staged;
            if (base + SUM_BATCH < end) sum_load(L, g, base + SUM_BATCH + e, end, b);
        }
        if (wave < SUM_CONSUMERS && i >= 1 && lane < SLOTS) {
            const SumStage& in = st[(i - 1) & 1];
            unsigned staged = 0;
#pragma unroll
            for (int p = 0; p < SUM_PRODUCERS; ++p) staged |= in.staged[p];
            if (staged) {
                if (wave == 0) {
#pragma unroll
                    for (int p = 0; p < SUM_PRODUCERS; ++p) count += in.count[p][lane];
                }
                // Wave 0 adds slot `lane`'s first terms, wave 1 its second.
                const double2* pa = reinterpret_cast<const double2*>(
                    (wave == 0 ? in.a : in.b) + lane * SUM_PITCH);
                double acc = wave == 0 ? accum : second;
                // Two register groups alternate: the next group's LDS reads
                // are in flight while the current group's terms are added.
                double2 xa[2 * SUM_UNROLL], ya[2 * SUM_UNROLL];
#pragma unroll
                for (int u = 0; u < 2 * SUM_UNROLL; ++u) xa[u] = pa[u];
                for (int j = 0; j < SUM_BATCH / 2; j += 4 * SUM_UNROLL) {
#pragma unroll
                    for (int u = 0; u < 2 * SUM_UNROLL; ++u) ya[u] = pa[j + 2 * SUM_UNROLL + u];
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int u = 0; u < 2 * SUM_UNROLL; ++u) {
                        acc += xa[u].x;
                        acc += xa[u].y;
                    }
                    if (j + 4 * SUM_UNROLL < SUM_BATCH / 2) {
#pragma unroll
                        for (int u = 0; u < 2 * SUM_UNROLL; ++u) xa[u] = pa[j + 4 * SUM_UNROLL + u];
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int u = 0; u < 2 * SUM_UNROLL; ++u) {
                        acc += ya[u].x;
                        acc += ya[u].y;
                    }
                }
                if (wave == 0)
                    accum = acc;
                else
                    second = acc;
            }
        }
        __syncthreads();
    }
}

// One workgroup (two consumer wavefronts and SUM_PRODUCERS producers) per group
// (0 = the image-wide pool, r = region r); lane s of wave 0 owns slot s
// (region_correction.rs:196-261).
__global__ __launch_bounds__(SUM_THREADS) void opp_sums(Layout L) {
    __shared__ __align__(16) SumStage s_stage[2];
    __shared__ float s_mean[SLOTS], s_tol[SLOTS];
    __shared__ double s_second[SLOTS];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const unsigned count = min(L.scalars[0], (unsigned)L.cap - 1);
    const unsigned ring = L.scalars[1];
    const unsigned any = L.scalars[2];
    for (unsigned g = blockIdx.x; g <= count; g += gridDim.x) {
        unsigned begin = 0, end = ring, use = any;
        if (g != 0) {
            begin = L.region_first[g];
            end = begin == NONE ? 0u : L.region_last[g] + 1u;
            if (begin == NONE) begin = 0;
            use = L.region_wanted[g];
        }
        if (use == 0) end = begin;
        unsigned n = 0, kept = 0;
        double accum = 0.0, second = 0.0;
        float means[SLOTS], tols[SLOTS];
#pragma unroll
        for (int t = 0; t < SLOTS; ++t) means[t] = tols[t] = 0.0f;
        sum_pass<0>(L, g, begin, end, means, tols, s_stage, n, accum, second);
        if (wave == 1 && lane < SLOTS) s_second[lane] = second;
        __syncthreads();
        if (wave == 0 && lane < SLOTS) {
            second = s_second[lane];
            float mean = 0.0f, tolerance = 0.0f;
            if (n != 0) {
                const double cnt = (double)n;
                const double m = accum / cnt;
                const double variance = fmax(second / cnt - m * m, 0.0);
                mean = (float)m;
                tolerance = (float)(2.0 * sqrt(variance)) + 1e-5f;
            }
            s_mean[lane] = mean;
            s_tol[lane] = tolerance;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < SLOTS; ++t) {
            means[t] = s_mean[t];
            tols[t] = s_tol[t];
        }
        accum = 0.0;
        second = 0.0;
        sum_pass<1>(L, g, begin, end, means, tols, s_stage, kept, accum, second);
        if (wave == 1 && lane < SLOTS) s_second[lane] = second;
        __syncthreads();
        if (wave == 0 && lane < SLOTS) {
            second = s_second[lane];
            // Slot::value: fewer than the minimum samples or nothing kept
            // leaves the slot empty.
            float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if ((use & (1u << lane)) && n >= MIN_BORDER_CELLS && kept != 0) {
                const double kd = (double)kept;
                out = make_float4((float)(accum / kd), (float)(second / kd), 1.0f, 0.0f);
            }
            L.corrections[(size_t)g * SLOTS + lane] = out;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- sites
__device__ __forceinline__ int total_key(float v) {
    int i = __float_as_int(v);
    return i ^ (int)(((unsigned)(i >> 31)) >> 1);
}
__device__ __forceinline__ float from_total_key(int k) {
    return __int_as_float(k ^ (int)(((unsigned)(k >> 31)) >> 1));
}

// At most two candidates carry over from radius 2 (three or more end the
// search), and the radius-3 ring holds 24 cells.
constexpr int MAX_LOCAL = 26;

// Sites blocks: SX x SY cells of one channel; the candidate window adds the
// three-cell search radius on every side and is staged in LDS.
constexpr int SX = 64, SY = 8, SR = 3;
constexpr int WX = SX + 2 * SR, WY = SY + 2 * SR;

// The estimate for one (cell, channel): every suspected site of that channel
// in the cell gets the same one (opposed.rs:243-278 reads only cell data).
// Its parts that need no region data (the target reference and the local
// ring median, opposed.rs:282-327) are formed by opp_sites_local, which
// depends only on the cells and so runs on a second stream beside the
// labelling and the region sums; opp_sites combines them with the region
// correction once the sums are done.
// Cell data comes from the block's LDS window, where cells outside the image
// carry no flags, so they fail every candidate test as the reference's
// image-clamped rings skip them.
struct SiteLocal {
    const float* lv;
    const float4* cells;  // LDS window
    const float* pref;    // LDS window: pair reference for the target channel
    const float* plr;     // LDS window: pair log ratio for the target channel
    int target;
    unsigned support;
    bool two;  // two-channel support: reference and ratio come from the pair planes
    bool ht;
    float tr;
    int* keys;  // this thread's LDS column, stride THREADS

    // One candidate of the local search (opposed.rs:295-312), at window index j.
    __device__ void candidate(int j, int& count) const {
        const float4 cand = cells[j];
        const unsigned f = flags_of(cand);
        if (!(usable(f, target) && clean(f, target))) return;
        float value;
        if (!correction(cand, target, support, pref[j], lv, value)) return;
        float cr = 0.0f;
        const bool hc = two && support_log_ratio(support, plr[j], cr);
        if (!chromaticity_agrees(support, ht, tr, hc, cr)) return;
        if (count < MAX_LOCAL) keys[count * THREADS] = total_key(value);
        ++count;
    }

    // opposed.rs:282-327: rings of radius 1..3 cells around the target cell
    // (window index `centre`); the first radius with at least three
    // candidates gives their median.
    __device__ bool local_correction(int centre, float& out, int& out_radius) const {
        int count = 0;
        for (int radius = 1; radius <= SR; ++radius) {
            for (int dy = -radius; dy <= radius; ++dy) {
                const int row = centre + dy * WX;
                if (dy == -radius || dy == radius) {
                    for (int dx = -radius; dx <= radius; ++dx) candidate(row + dx, count);
                } else {
                    candidate(row - radius, count);
                    candidate(row + radius, count);
                }
            }
            if (count >= 3) {
                // Median in Rust's total order: the element(s) of rank
                // count / 2 (and count / 2 - 1 when even). Equal keys are
                // equal bit patterns, so ties need no position order.
                const int n = min(count, MAX_LOCAL);
                const int middle = n / 2;
                int lo = 0, hi = 0;
                for (int a = 0; a < n; ++a) {
                    const int ka = keys[a * THREADS];
                    int below = 0, equal_before = 0;
                    for (int b = 0; b < n; ++b) {
                        const int kb = keys[b * THREADS];
                        below += kb < ka;
                        equal_before += (kb == ka) & (b < a);
                    }
                    const int rank = below + equal_before;
                    if (rank == middle) hi = ka;
                    if (rank == middle - 1) lo = ka;
                }
                out = n % 2 == 0 ? (from_total_key(lo) + from_total_key(hi)) * 0.5f
                                 : from_total_key(hi);
                out_radius = radius;
                return true;
            }
        }
        return false;
    }
};

// Per (channel, cell) result of the local part: the target reference and the
// local median, and a byte with bit 7 set when the reference exists and the
// median's ring radius (0: no local median) in bits 0-2.
constexpr uint8_t LOCAL_REFERENCE = 0x80;

// The (sites block, channel) pairs with a suspected cell of the channel,
// appended to site_list in any order. A wavefront looks at LIST_PER_WAVE
// sites blocks in turn (SX == 64: a lane per cell column); the workgroup
// gathers its pairs in LDS and reserves their places in the list with one
// atomic. The local parts are independent of one another, so
// opp_sites_local may take the pairs in any order.
constexpr int LIST_PER_WAVE = 4, LIST_PER_GROUP = LIST_PER_WAVE * (THREADS / 64);
__global__ __launch_bounds__(THREADS) void opp_site_list(Layout L) {
    static_assert(SX == 64, "a lane per cell column");
    __shared__ unsigned s_ids[3 * LIST_PER_GROUP];
    __shared__ unsigned s_n, s_base;
    const int gx = (L.cw + SX - 1) / SX, gy = (L.ch + SY - 1) / SY;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    for (int i = 0; i < LIST_PER_WAVE; ++i) {
        const int b = blockIdx.x * LIST_PER_GROUP + wave * LIST_PER_WAVE + i;
        if (b >= gx * gy) break;
        const int cx = (b % gx) * SX + lane, by0 = (b / gx) * SY;
        unsigned suspected = 0;
#pragma unroll
        for (int r = 0; r < SY; ++r)
            if (cx < L.cw && by0 + r < L.ch)
                suspected |= (unsigned)L.flag8[(size_t)(by0 + r) * L.cw + cx] >> 3;
        for (int c = 0; c < 3; ++c)
            if (__ballot((suspected >> c) & 1u) && lane == 0)
                s_ids[atomicAdd(&s_n, 1u)] = (unsigned)(b + c * gx * gy);
    }
    __syncthreads();
    const unsigned n = s_n;
    if (threadIdx.x == 0 && n) s_base = atomicAdd(L.scalars + 3, n);
    __syncthreads();
    for (unsigned k = threadIdx.x; k < n; k += THREADS) L.site_list[s_base + k] = s_ids[k];
}

// opposed.rs:243-327, the parts that need no region data, for every
// suspected (cell, channel) with a supporting channel. A resident grid of
// workgroups takes the listed (sites block, channel) pairs from a shared
// counter until none is left, so the long searches of large clipped areas
// spread over every workgroup. The block's cells are staged in LDS with the
// three-cell search radius around them; the cells that need a search are
// listed in LDS and taken densely, so the ring searches run in full
// wavefronts.
__global__ __launch_bounds__(THREADS) void opp_sites_local(Layout L, float l0, float l1, float l2) {
    __shared__ float4 s_cell[WY * WX];
    __shared__ float s_pref[WY * WX];
    __shared__ float s_plr[WY * WX];
    __shared__ int s_keys[MAX_LOCAL * THREADS];
    __shared__ unsigned short s_task[SX * SY];
    __shared__ int s_ntask;
    __shared__ unsigned s_item;
    const float lv[3] = {l0, l1, l2};
    const int gx = (L.cw + SX - 1) / SX, gy = (L.ch + SY - 1) / SY;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    constexpr int ROWS_PER_THREAD = SY / (THREADS / 64);
    const unsigned listed = L.scalars[3];
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(L.scalars + 4, 1u);
        __syncthreads();
        const unsigned work = s_item;
        // Every thread reads the same item, so the whole workgroup leaves
        // together once the list is exhausted.
        if (work >= listed) break;
        const unsigned id = L.site_list[work];
        const int channel = (int)(id / (unsigned)(gx * gy));
        const int rest = (int)(id % (unsigned)(gx * gy));
        const int bx0 = (rest % gx) * SX, by0 = (rest / gx) * SY;
        const size_t plane = (size_t)channel * L.cells;
        for (int k = threadIdx.x; k < WY * WX; k += THREADS) {
            const int x = bx0 - SR + k % WX, y = by0 - SR + k / WX;
            float4 cell = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            float pr = 0.0f, pl = 0.0f;
            if (x >= 0 && y >= 0 && x < L.cw && y < L.ch) {
                const size_t j = (size_t)y * L.cw + x;
                cell = L.cell[j];
                pr = L.pair_ref[plane + j];
                pl = L.pair_lr[plane + j];
            }
            s_cell[k] = cell;
            s_pref[k] = pr;
            s_plr[k] = pl;
        }
        if (threadIdx.x == 0) s_ntask = 0;
        __syncthreads();
        // The suspected cells of the channel with a supporting channel.
#pragma unroll
        for (int q = 0; q < ROWS_PER_THREAD; ++q) {
            const int r = ty + 4 * q, c = tx;
            if (bx0 + c >= L.cw || by0 + r >= L.ch) continue;
            const unsigned f = flags_of(s_cell[(r + SR) * WX + c + SR]);
            if (!(f & (1u << (channel + 3)))) continue;
            bool supported = false;
            for (int o = 0; o < 3; ++o)
                if (o != channel && usable(f, o) && clean(f, o)) supported = true;
            if (supported) s_task[atomicAdd(&s_ntask, 1)] = (unsigned short)(r * SX + c);
        }
        __syncthreads();
        for (int k = threadIdx.x; k < s_ntask; k += THREADS) {
            const int code = s_task[k], r = code / SX, c = code % SX;
            const int centre = (r + SR) * WX + c + SR;
            const float4 target = s_cell[centre];
            SiteLocal site{lv, s_cell, s_pref, s_plr, channel, support_mask(target, channel, lv),
                           false, false, 0.0f, s_keys + threadIdx.x};
            site.two = __popc(site.support) == 2;
            site.ht = site.two && support_log_ratio(site.support, s_plr[centre], site.tr);
            float reference_value = 0.0f, local = 0.0f;
            int radius = 0;
            uint8_t code_out = 0;
            if (reference(target, site.support, s_pref[centre], reference_value)) {
                code_out = LOCAL_REFERENCE;
                if (site.local_correction(centre, local, radius)) code_out |= (uint8_t)radius;
            }
            const size_t j = plane + (size_t)(by0 + r) * L.cw + bx0 + c;
            L.site_value[j] = make_float2(reference_value, local);
            L.site_code[j] = code_out;
        }
        __syncthreads();
    }
}

// opposed.rs:97-162 and :243-278: for every suspected (cell, channel), the
// estimate from the local part and the region correction, applied to each
// suspected site of that channel in the cell. One thread per cell, in a
// grid of SITE_GROUPS workgroups striding over the cells.
constexpr int SITE_GROUPS = 1024;
__global__ __launch_bounds__(THREADS) void opp_sites(float* __restrict__ m, Layout L, int pattern,
                                                     float l0, float l1, float l2, int has_ceiling,
                                                     float c0, float c1, float c2) {
    const float lv[3] = {l0, l1, l2};
    const float ceiling[3] = {c0, c1, c2};
    unsigned reconstructed = 0, changed = 0, fallback = 0, unsupported = 0, neutralised = 0;
    for (int i = blockIdx.x * THREADS + threadIdx.x; i < L.cells; i += gridDim.x * THREADS) {
    const unsigned suspected = (unsigned)L.flag8[i] >> 3;
    if (suspected) {
        const int cx = i % L.cw, cy = i / L.cw;
        const float4 cell = L.cell[i];
        const unsigned f = flags_of(cell);
        for (int channel = 0; channel < 3; ++channel) {
            if (!((suspected >> channel) & 1u)) continue;
            int support_count = 0;
            for (int o = 0; o < 3; ++o)
                if (o != channel && usable(f, o) && clean(f, o)) ++support_count;
            float est = 0.0f;
            bool has = false;
            if (support_count != 0) {
                const size_t j = (size_t)channel * L.cells + i;
                const uint8_t code = L.site_code[j];
                if (code & LOCAL_REFERENCE) {
                    const float2 v = L.site_value[j];
                    const unsigned support = support_mask(cell, channel, lv);
                    const bool two = __popc(support) == 2;
                    float tr = 0.0f;
                    const bool ht = two && support_log_ratio(support, L.pair_lr[j], tr);
                    const int radius = code & 7;
                    const bool has_local = radius != 0;
                    // region_correction.rs:278-300 (lookup) with the pool fallback.
                    float region = 0.0f, br = 0.0f;
                    bool has_region = false;
                    const int p = pattern_position(channel, support);
                    if (p >= 0) {
                        const unsigned label = L.label[i];
                        const int off = channel * 3 + p;
                        float4 s = label < (unsigned)L.cap
                                       ? L.corrections[(size_t)label * SLOTS + off]
                                       : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        if (s.z == 0.0f) s = L.corrections[off];
                        if (s.z != 0.0f) {
                            region = s.x;
                            br = s.y;
                            has_region = chromaticity_agrees(support, ht, tr, two, br);
                        }
                    }
                    float corr = 0.0f;
                    bool any = true;
                    if (has_local && has_region) {
                        float weight = (float)(3 + 1 - radius) / 3.0f;
                        corr = weight * v.y + (1.0f - weight) * region;
                    } else if (has_local) {
                        corr = v.y;
                    } else if (has_region) {
                        corr = region;
                    } else {
                        any = false;
                    }
                    if (any) {
                        const float e = v.x + corr;
                        if (is_finite(e) && e > 0.0f) {
                            est = e;
                            has = true;
                        }
                    }
                }
            }
            const float level = lv[channel];
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    const int x = 2 * cx + dx, y = 2 * cy + dy;
                    if (x >= L.w || y >= L.h || color_at(pattern, x, y) != channel) continue;
                    float* sp = m + (size_t)y * L.w + x;
                    const float sample = *sp;
                    if (!(sample >= level)) continue;
                    bool neutralise = false;
                    if (!has) {
                        fallback += 1;
                        unsupported += support_count == 0;
                        neutralise = true;
                    } else {
                        reconstructed += 1;
                        const float bounded = fminf(est, 4.0f * level);
                        if (bounded > sample) {
                            *sp = bounded;
                            changed += 1;
                        } else {
                            neutralise = true;
                        }
                    }
                    if (neutralise && has_ceiling && sample > ceiling[channel]) {
                        *sp = ceiling[channel];
                        neutralised += 1;
                    }
                }
        }
    }
    }
    // Per-workgroup totals go to the workgroup's own slot (no shared
    // counters); opp_site_stats adds them up.
    __shared__ unsigned s_part[THREADS / 64][5];
    unsigned v[5] = {reconstructed, changed, fallback, unsupported, neutralised};
    for (int k = 0; k < 5; ++k) {
        const unsigned t = wave_sum(v[k]);
        if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6][k] = t;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        unsigned t = 0;
        for (int w = 0; w < THREADS / 64; ++w) t += s_part[w][threadIdx.x];
        L.site_counts[5 * (size_t)blockIdx.x + threadIdx.x] = t;
    }
}

// Sum of the opp_sites block totals into the develop statistics.
__global__ __launch_bounds__(THREADS) void opp_site_stats(Layout L, unsigned blocks,
                                                          DeviceStats* stats) {
    unsigned long long v[5] = {0, 0, 0, 0, 0};
    for (unsigned b = blockIdx.x * THREADS + threadIdx.x; b < blocks; b += gridDim.x * THREADS)
        for (int k = 0; k < 5; ++k) v[k] += L.site_counts[5 * (size_t)b + k];
    for (int k = 0; k < 5; ++k) {
        const unsigned long long t = wave_sum(v[k]);
        if ((threadIdx.x & 63) == 0 && t) atomicAdd(stats->opposed + k, t);
    }
}

size_t site_blocks(const Layout& L) {
    return (size_t)((L.cw + SX - 1) / SX) * ((L.ch + SY - 1) / SY) * 3;
}
// Resident opp_sites_local workgroups: three per CU of the MI355X (LDS-bound),
// never more than there are (sites block, channel) pairs.
constexpr int SITE_WORKERS = 3 * 256;
int site_workers(const Layout& L) { return (int)std::min<size_t>(SITE_WORKERS, site_blocks(L)); }

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
    L.flag8 = (uint8_t*)take(cells);
    L.pair_ref = (float*)take(cells * 3 * 4);
    L.pair_lr = (float*)take(cells * 3 * 4);
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
    L.site_counts = (unsigned*)take(SITE_GROUPS * 5 * sizeof(unsigned));
    L.site_list = (unsigned*)take(site_blocks(L) * sizeof(unsigned));
    L.site_value = (float2*)take(cells * 3 * sizeof(float2));
    L.site_code = (uint8_t*)take(cells * 3);
    L.scalars = (unsigned*)take(64);
    L.log2_bitmap = a.log2_bitmap;
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
    const int tiles_x = (L.cw + LX - 1) / LX, tiles_y = (L.ch + LY - 1) / LY;
    const int border_cells = tiles_x * tiles_y * (LX + 2 * (LY - 1));
    const uint64_t px = (uint64_t)L.w * L.h, cells = (uint64_t)L.cells;
    // Algorithmic bytes per launch. The ring-entry kernels (contrib, sums)
    // depend on the ring size, which only the device knows, and report 0.
    RHD_TIMED(prof, s, "opp_cells", 4 * px + 17 * cells,
              hipLaunchKernelGGL(opp_cells, dim3(cell_blocks), dim3(THREADS), 0, s, a.mosaic, L,
                                 a.pattern, l0, l1, l2));
    // The pair planes and the local part of the site estimates need only the
    // cells: they run on the auxiliary stream beside the region sums, and the
    // main stream waits for them before the final site pass. Without an
    // auxiliary stream everything runs in order on s.
    hipStream_t side = a.aux ? a.aux : s;
    RHD_TIMED(prof, s, "opp_label_tiles", cells + 4 * cells,
              hipLaunchKernelGGL(opp_label_tiles, dim3(tiles_x, tiles_y), dim3(THREADS), 0, s, L));
    RHD_TIMED(prof, s, "opp_label_borders", 20ull * border_cells,
              hipLaunchKernelGGL(opp_label_borders, dim3((border_cells + THREADS - 1) / THREADS),
                                 dim3(THREADS), 0, s, L, tiles_x));
    RHD_TIMED(prof, s, "opp_flatten", 8 * cells,
              hipLaunchKernelGGL(opp_flatten, dim3(cell_blocks), dim3(THREADS), 0, s, L));
    RHD_TIMED(prof, s, "opp_scan_roots", 8 * cells,
              hipLaunchKernelGGL(opp_scan_count, dim3(scan_blocks), dim3(THREADS), 0, s, L,
                                 (int)ROOTS);
              hipLaunchKernelGGL(opp_scan_blocks, dim3(1), dim3(THREADS), 0, s, L, scan_blocks,
                                 (int)ROOTS);
              hipLaunchKernelGGL(opp_scan_emit, dim3(scan_blocks), dim3(THREADS), 0, s, L,
                                 (int)ROOTS));
    RHD_TIMED(prof, s, "opp_relabel", 12 * cells,
              hipLaunchKernelGGL(opp_relabel, dim3(cell_blocks), dim3(THREADS), 0, s, L));
    RHD_TIMED(prof, s, "opp_regions", 20 * cells,
              hipLaunchKernelGGL(opp_region_init, dim3(256), dim3(THREADS), 0, s, L, a.stats);
              hipLaunchKernelGGL(opp_regions, dim3((cell_blocks + REGION_STRIDE - 1) / REGION_STRIDE),
                                 dim3(THREADS), 0, s, L, l0, l1,
                                 l2));
    RHD_TIMED(prof, s, "opp_scan_ring", 40 * cells,
              hipLaunchKernelGGL(opp_scan_count, dim3(scan_blocks), dim3(THREADS), 0, s, L,
                                 (int)RING);
              hipLaunchKernelGGL(opp_scan_blocks, dim3(1), dim3(THREADS), 0, s, L, scan_blocks,
                                 (int)RING);
              hipLaunchKernelGGL(opp_scan_emit, dim3(scan_blocks), dim3(THREADS), 0, s, L,
                                 (int)RING));
    // The pair planes and the local site searches wait until the labelling is
    // done, so that they run beside the region sums, whose sequential f64
    // chains leave most of the device idle, rather than beside the labelling
    // kernels. The ring contributions form their own pair values.
    if (a.aux) {
        hipError_t e = hipEventRecord(a.fork, s);
        if (e == hipSuccess) e = hipStreamWaitEvent(a.aux, a.fork, 0);
        if (e != hipSuccess) return e;
    }
    RHD_TIMED(prof, side, "opp_pairs", 16 * cells + 24 * cells,
              hipLaunchKernelGGL(opp_pairs, dim3(cell_blocks), dim3(THREADS), 0, side, L));
    const int sites_blocks = (L.cw + SX - 1) / SX * ((L.ch + SY - 1) / SY);
    RHD_TIMED(prof, side, "opp_site_list", cells,
              hipLaunchKernelGGL(opp_site_list, dim3((sites_blocks + LIST_PER_GROUP - 1) / LIST_PER_GROUP),
                                 dim3(THREADS), 0, side, L));
    RHD_TIMED(prof, side, "opp_sites_local", 16 * cells + 8 * cells,
              hipLaunchKernelGGL(opp_sites_local, dim3(site_workers(L)), dim3(THREADS), 0, side, L,
                                 l0, l1, l2));
    if (a.aux) {
        hipError_t e = hipEventRecord(a.join, a.aux);
        if (e != hipSuccess) return e;
    }
    RHD_TIMED(prof, s, "opp_contrib", 0,
              hipLaunchKernelGGL(opp_contrib, dim3(2048), dim3(THREADS), 0, s, L, l0, l1, l2));
    RHD_TIMED(prof, s, "opp_sums", 0,
              hipLaunchKernelGGL(opp_sums, dim3(1024), dim3(SUM_THREADS), 0, s, L));
    if (a.aux) {
        hipError_t e = hipStreamWaitEvent(s, a.join, 0);
        if (e != hipSuccess) return e;
    }
    const int site_groups = std::min(cell_blocks, SITE_GROUPS);
    RHD_TIMED(prof, s, "opp_sites", 8 * px + cells,
              hipLaunchKernelGGL(opp_sites, dim3(site_groups), dim3(THREADS), 0, s, a.mosaic, L,
                                 a.pattern, l0, l1, l2, a.has_ceiling, a.ceiling[0], a.ceiling[1],
                                 a.ceiling[2]));
    RHD_TIMED(prof, s, "opp_site_stats", 20 * site_groups,
              hipLaunchKernelGGL(opp_site_stats, dim3(64), dim3(THREADS), 0, s, L,
                                 (unsigned)site_groups, a.stats));
    return hipGetLastError();
}

}  // namespace rhd
