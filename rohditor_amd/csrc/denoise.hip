// Noise reduction of the detail stage on planar RGB in HBM.
//
// Noise profile (crates/detail/src/noise.rs:48-170 over the rows of
// crates/detail/src/denoise/mod.rs:112-150). The whole estimate stays on the
// device so the develop never waits on the host:
//   nr_est_split   balanced Y, C1, C2 of the sampled rows (non-finite -> 0)
//   nr_b3_h/_v     finest B3 detail band of each plane
//   nr_est_tiles   per 16x16 tile: mean band^2 and mean luma, f64 in raster
//                  order, narrowed to f32
//   nr_keys*/sort  the reference's two stable sorts (by brightness, then by
//                  energy inside each brightness bin) as bitonic sorts of
//                  unique 64-bit keys: total-order key above the position
//                  that breaks ties, which is what stability means
//   nr_select      per (plane, bin): median and MAD of the flattest quarter's
//                  band samples by 4 x 8-bit radix select in total order,
//                  and the mean brightness of those tiles in their order
//   nr_fit         the least-squares line of noise.rs:142-170 in f64
// Chroma (crates/detail/src/denoise/chroma.rs:88-425): five luma-guided 3x3
// binomial a-trous levels of C1/C2 with soft thresholds, rebuilt into RGB in
// the last level's kernel. Luma (crates/detail/src/denoise/luma.rs:128-366):
// five B3 a-trous levels of the balanced luma, each band soft-thresholded
// against its 7x7 energy (BayesShrink), applied back as a common RGB gain in
// the last level's kernel. All f32 in the reference's operation order.
#include "rhd_internal.hpp"

namespace rhd {
namespace {

constexpr float B3_0 = 1.0f / 16.0f, B3_1 = 4.0f / 16.0f, B3_2 = 6.0f / 16.0f;

__device__ __forceinline__ float b3(float t0, float t1, float t2, float t3, float t4) {
    return B3_0 * t0 + B3_1 * t1 + B3_2 * t2 + B3_1 * t3 + B3_0 * t4;
}

// f32::total_cmp as an unsigned key, and back.
__device__ __forceinline__ uint32_t order_key(float v) {
    const uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_value(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__device__ __forceinline__ float finite_or_zero(float v) { return is_finite(v) ? v : 0.0f; }

struct Bal {
    float b[3];
};

// Y, C1, C2 of split_opponent (planes.rs:125-152), non-finite -> 0.
__device__ __forceinline__ void split(float r0, float g0, float b0, const Bal& p, float& y,
                                      float& c1, float& c2) {
    const float r = r0 * p.b[0], g = g0 * p.b[1], b = b0 * p.b[2];
    y = finite_or_zero((r + 2.0f * g + b) * 0.25f);
    c1 = finite_or_zero(r - g);
    c2 = finite_or_zero(b - g);
}

// luma_plane (planes.rs:108-121).
__device__ __forceinline__ float luma_of(float r0, float g0, float b0, const Bal& p) {
    return finite_or_zero((r0 * p.b[0] + 2.0f * g0 * p.b[1] + b0 * p.b[2]) * 0.25f);
}

// chroma.rs:376-378 / luma.rs:274-277: signum(v) * max(|v| - t, 0).
__device__ __forceinline__ float soft_threshold(float v, float t) {
    const float sign = __builtin_isnan(v) ? v : (__builtin_signbit(v) ? -1.0f : 1.0f);
    return sign * num_max(fabsf(v) - t, 0.0f);
}

// noise.rs:36-46.
__device__ __forceinline__ float variance(const float* prof, int c, float x) {
    return prof[2 * c] * num_max(x, 0.0f) + prof[2 * c + 1];
}

// ---------------------------------------------------------------- profile

// Stacked sampled row -> image row (every fourth 64-row band when h >= 1024).
__device__ __forceinline__ int sampled_row(int r, int h) {
    return h < 1024 ? r : (r >> 6) * 256 + (r & 63);
}

__global__ __launch_bounds__(256) void nr_est_split(const float* __restrict__ rgb, size_t n,
                                                    int w, int h, Bal p,
                                                    float* __restrict__ planes, size_t pitch) {
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t s = (size_t)sampled_row(y, h) * w + x, o = (size_t)y * w + x;
    float Y, c1, c2;
    split(rgb[s], rgb[n + s], rgb[2 * n + s], p, Y, c1, c2);
    planes[o] = Y;
    planes[pitch + o] = c1;
    planes[2 * pitch + o] = c2;
}

// Horizontal B3 of plane z (atrous.rs:99-120), taps `step` apart, mirrored.
__global__ __launch_bounds__(256) void nr_b3_h(const float* __restrict__ src,
                                               float* __restrict__ dst, int w, size_t pitch,
                                               int step) {
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const float* line = src + blockIdx.z * pitch + (size_t)y * w;
    float t[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) t[k] = line[mirror_near(x + (k - 2) * step, w)];
    dst[blockIdx.z * pitch + (size_t)y * w + x] = b3(t[0], t[1], t[2], t[3], t[4]);
}

// Vertical B3 (atrous.rs:79-96) and the detail band plane - smoothing of
// plane z. The band of plane 0 goes to band0, the others overwrite their
// plane (each pixel reads only its own plane value).
__global__ __launch_bounds__(256) void nr_est_band(float* planes, const float* __restrict__ tmp,
                                                   float* __restrict__ band0, int w, int rows,
                                                   size_t pitch) {
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    if (x >= w) return;
    const float* t = tmp + z * pitch + x;
    const float s = b3(t[(size_t)mirror_near(y - 2, rows) * w], t[(size_t)mirror_near(y - 1, rows) * w],
                       t[(size_t)y * w], t[(size_t)mirror_near(y + 1, rows) * w],
                       t[(size_t)mirror_near(y + 2, rows) * w]);
    const size_t o = (size_t)y * w + x;
    const float d = planes[z * pitch + o] - s;
    if (z == 0)
        band0[o] = d;
    else
        planes[z * pitch + o] = d;
}

// noise.rs:66-84: per tile, mean band^2 (f64, raster order) of plane y and,
// for plane 0, the mean luma.
__global__ __launch_bounds__(256) void nr_est_tiles(const float* __restrict__ band0,
                                                    const float* __restrict__ planes,
                                                    size_t pitch, int w, int tiles_x, int nt,
                                                    float* __restrict__ energy,
                                                    float* __restrict__ bright) {
    const int t = blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
    if (t >= nt) return;
    const size_t origin = (size_t)(t / tiles_x) * 16 * w + (size_t)(t % tiles_x) * 16;
    const float* b = (c == 0 ? band0 : planes + c * pitch) + origin;
    const float* luma = planes + origin;
    double e = 0.0, m = 0.0;
    for (int y = 0; y < 16; ++y)
        for (int x = 0; x < 16; ++x) {
            const double v = b[(size_t)y * w + x];
            e += v * v;
            if (c == 0) m += (double)luma[(size_t)y * w + x];
        }
    energy[(size_t)c * nt + t] = (float)(e / 256.0);
    if (c == 0) bright[t] = (float)(m / 256.0);
}

constexpr int POS_BITS = 28;
constexpr uint64_t POS_MASK = (1ull << POS_BITS) - 1;

// Brightness order: key above tile index (a stable sort by brightness).
__global__ __launch_bounds__(256) void nr_keys_bright(const float* __restrict__ bright, int nt,
                                                      int N, uint64_t* __restrict__ keys) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    keys[i] = i < nt ? ((uint64_t)order_key(bright[i]) << 32) | (uint32_t)i : ~0ull;
}

// Inside each brightness bin, energy order: bin, key, brightness position.
__global__ __launch_bounds__(256) void nr_keys_energy(const uint64_t* __restrict__ by_bright,
                                                      const float* __restrict__ energy, int nt,
                                                      int N, int bins,
                                                      uint64_t* __restrict__ keys) {
    const int p = blockIdx.x * 256 + threadIdx.x, c = blockIdx.y;
    if (p >= N) return;
    uint64_t* out = keys + (size_t)c * N;
    if (p >= nt) {
        out[p] = ~0ull;
        return;
    }
    const uint32_t tile = (uint32_t)by_bright[p];
    int b = (int)((int64_t)p * bins / nt);
    while (b + 1 < bins && (int64_t)(b + 1) * nt / bins <= p) ++b;
    while (b > 0 && (int64_t)b * nt / bins > p) --b;
    out[p] = ((uint64_t)b << 60) | ((uint64_t)order_key(energy[(size_t)c * nt + tile]) << POS_BITS) |
             (uint64_t)p;
}

// Bitonic sort of N (a power of two) unique keys per array (blockIdx.y):
// chunks of SORT_CHUNK sort in LDS, larger strides step through HBM.
constexpr int SORT_CHUNK = 2048;

__device__ __forceinline__ void order_pair(uint64_t& a, uint64_t& b, bool up) {
    if ((a > b) == up) {
        const uint64_t t = a;
        a = b;
        b = t;
    }
}

// All stages k <= kmax (first pass, kmax = chunk) or only the strides below
// the chunk of stage k (merge pass).
__global__ __launch_bounds__(1024) void nr_sort_chunk(uint64_t* keys, int N, int k_only) {
    __shared__ uint64_t s[SORT_CHUNK];
    const int L = N < SORT_CHUNK ? N : SORT_CHUNK;
    uint64_t* a = keys + (size_t)blockIdx.y * N + (size_t)blockIdx.x * L;
    const unsigned base = blockIdx.x * L;
    for (int i = threadIdx.x; i < L; i += 1024) s[i] = a[i];
    __syncthreads();
    const int k_first = k_only ? k_only : 2, k_last = k_only ? k_only : L;
    for (int k = k_first; k <= k_last; k <<= 1)
        for (int j = (k_only ? L : k) >> 1; j > 0; j >>= 1) {
            for (int t = threadIdx.x; t < L / 2; t += 1024) {
                const int i = (t / j) * 2 * j + (t % j);
                order_pair(s[i], s[i + j], ((base + i) & k) == 0);
            }
            __syncthreads();
        }
    for (int i = threadIdx.x; i < L; i += 1024) a[i] = s[i];
}

__global__ __launch_bounds__(256) void nr_sort_step(uint64_t* keys, int N, int k, int j) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= N / 2) return;
    uint64_t* a = keys + (size_t)blockIdx.y * N;
    const int i = (t / j) * 2 * j + (t % j);
    uint64_t x = a[i], y = a[i + j];
    const uint64_t lo = x < y ? x : y, hi = x < y ? y : x;
    const bool up = (i & k) == 0;
    a[i] = up ? lo : hi;
    a[i + j] = up ? hi : lo;
}

// Sorts each brightness bin of each plane's energy keys in LDS (the bins
// are contiguous and already hold their keys; L >= the longest bin).
template <int L>
__global__ __launch_bounds__(1024) void nr_sort_bins(uint64_t* keys, int N, int nt, int bins) {
    __shared__ uint64_t s[L];
    const int b = blockIdx.x;
    const int64_t start = (int64_t)b * nt / bins, len = (int64_t)(b + 1) * nt / bins - start;
    uint64_t* a = keys + (size_t)blockIdx.y * N + start;
    for (int i = threadIdx.x; i < L; i += 1024) s[i] = i < len ? a[i] : ~0ull;
    __syncthreads();
    for (int k = 2; k <= L; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = threadIdx.x; t < L / 2; t += 1024) {
                const int i = (t / j) * 2 * j + (t % j);
                order_pair(s[i], s[i + j], (i & k) == 0);
            }
            __syncthreads();
        }
    for (int i = threadIdx.x; i < len; i += 1024) a[i] = s[i];
}

void sort_keys(uint64_t* keys, int N, int arrays, hipStream_t s, Profiler* prof) {
    const int L = N < SORT_CHUNK ? N : SORT_CHUNK;
    const dim3 chunks(N / L, arrays);
    RHD_TIMED(prof, s, "nr_sort", 16 * (uint64_t)N * arrays,
              nr_sort_chunk<<<chunks, 1024, 0, s>>>(keys, N, 0));
    for (int k = 2 * L; k <= N; k <<= 1) {
        for (int j = k >> 1; j >= L; j >>= 1)
            RHD_TIMED(prof, s, "nr_sort", 32 * (uint64_t)N * arrays,
                      nr_sort_step<<<dim3((N / 2 + 255) / 256, arrays), 256, 0, s>>>(keys, N, k, j));
        RHD_TIMED(prof, s, "nr_sort", 16 * (uint64_t)N * arrays,
                  nr_sort_chunk<<<chunks, 1024, 0, s>>>(keys, N, k));
    }
}

constexpr float DETAIL_NOISE_GAIN = 0.8908f;
constexpr float MAD_TO_SIGMA = 1.4826f;
constexpr int MAX_BINS = 16;
constexpr int HIST_CHUNK = 4096;  // samples per histogram workgroup
constexpr int BRIGHT_CHUNK = 4096;

// Brightness bin b of nt tiles in `bins` bins: its first sorted position and
// the size of its flattest quarter (noise.rs:96-109, 119-121).
__host__ __device__ __forceinline__ void segment(int nt, int bins, int b, int64_t& start,
                                                 int64_t& flat) {
    start = (int64_t)b * nt / bins;
    const int64_t len = (int64_t)(b + 1) * nt / bins - start;
    flat = len / 4 > 1 ? len / 4 : 1;
}

// The bin that holds item `i` of a per-plane list where bin b owns `per(b)`
// consecutive items; returns the bin and the index of its first item.
template <class Per>
__device__ __forceinline__ int locate(int64_t i, int bins, Per per, int64_t& first) {
    first = 0;
    for (int b = 0; b < bins; ++b) {
        const int64_t n = per(b);
        if (i < first + n) return b;
        first += n;
    }
    return -1;
}

struct Selection {
    int nt, bins, N, tiles_x, w;
    size_t pitch;
    int64_t plane_samples;  // flat samples of one plane, all bins
};

// Copies the flattest tiles' band samples of every (plane, bin) into one
// contiguous run per bin: one workgroup per flat tile, 16x16 threads.
__global__ __launch_bounds__(256) void nr_gather(Selection q, const float* __restrict__ band0,
                                                 const float* __restrict__ planes,
                                                 const uint64_t* __restrict__ by_bright,
                                                 const uint64_t* __restrict__ by_energy,
                                                 float* __restrict__ samples) {
    const int c = blockIdx.y;
    int64_t first;
    const int b = locate(blockIdx.x, q.bins, [&](int k) {
        int64_t st, fl;
        segment(q.nt, q.bins, k, st, fl);
        return fl;
    }, first);
    if (b < 0) return;
    int64_t start, flat;
    segment(q.nt, q.bins, b, start, flat);
    const int64_t i = blockIdx.x - first;
    const uint32_t tile =
        (uint32_t)by_bright[by_energy[(size_t)c * q.N + start + i] & POS_MASK];
    const float* band = c == 0 ? band0 : planes + c * q.pitch;
    const size_t y = (size_t)(tile / q.tiles_x) * 16 + (threadIdx.x >> 4);
    const size_t x = (size_t)(tile % q.tiles_x) * 16 + (threadIdx.x & 15);
    // Samples of bin b start after the flat samples of the bins before it.
    samples[(size_t)c * q.plane_samples + (size_t)first * 256 + (size_t)i * 256 + threadIdx.x] =
        band[y * q.w + x];
}

// Radix-select state of one (plane, bin): the key prefix found so far, the
// rank still sought below it, and the two results.
struct SelectState {
    uint32_t prefix, rank;
    float median, mad;
};

__global__ void nr_select_init(Selection q, SelectState* __restrict__ st,
                               uint32_t* __restrict__ hist) {
    const int seg = blockIdx.x, t = threadIdx.x;
    hist[seg * 256 + t] = 0;
    if (t == 0) {
        int64_t start, flat;
        segment(q.nt, q.bins, seg % q.bins, start, flat);
        st[seg].prefix = 0;
        st[seg].rank = (uint32_t)(flat * 256 / 2);
        st[seg].median = 0.0f;
        st[seg].mad = 0.0f;
    }
}

// One 8-bit radix pass over HIST_CHUNK samples of one (plane, bin): the
// histogram of the digit below the prefix, of the samples (stage 0, the
// median) or of their distance to the median (stage 1, the MAD), in total
// order.
__global__ __launch_bounds__(256) void nr_select_hist(Selection q, int stage, int pass,
                                                      const float* __restrict__ samples,
                                                      const SelectState* __restrict__ st,
                                                      uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[256];
    const int c = blockIdx.y, t = threadIdx.x;
    int64_t first;
    const int b = locate(blockIdx.x, q.bins, [&](int k) {
        int64_t s0, fl;
        segment(q.nt, q.bins, k, s0, fl);
        return (fl * 256 + HIST_CHUNK - 1) / HIST_CHUNK;
    }, first);
    if (b < 0) return;
    int64_t sample0 = 0;
    for (int k = 0; k < b; ++k) {
        int64_t s0, fl;
        segment(q.nt, q.bins, k, s0, fl);
        sample0 += fl * 256;
    }
    int64_t start, flat;
    segment(q.nt, q.bins, b, start, flat);
    const int seg = c * q.bins + b;
    const int64_t lo = (blockIdx.x - first) * (int64_t)HIST_CHUNK;
    const int64_t hi = flat * 256 < lo + HIST_CHUNK ? flat * 256 : lo + HIST_CHUNK;
    const float* v = samples + (size_t)c * q.plane_samples + sample0;
    const uint32_t prefix = st[seg].prefix;
    const float median = st[seg].median;
    const int shift = 24 - 8 * pass;
    const uint32_t mask = pass == 0 ? 0u : 0xffffffffu << (32 - 8 * pass);
    h[t] = 0;
    __syncthreads();
    for (int64_t i = lo + t; i < hi; i += 256) {
        const float x = stage ? fabsf(v[i] - median) : v[i];
        const uint32_t key = order_key(x);
        if ((key & mask) == prefix) atomicAdd(&h[(key >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (h[t]) atomicAdd(&hist[seg * 256 + t], h[t]);
}

// Picks the digit that holds the sought rank, clears the histogram for the
// next pass and, after the last pass, stores the selected value.
__global__ __launch_bounds__(256) void nr_select_digit(Selection q, int stage, int pass,
                                                       SelectState* __restrict__ st,
                                                       uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[256];
    const int seg = blockIdx.x, t = threadIdx.x;
    // Read before the barriers: one thread rewrites the state afterwards.
    const uint32_t rank = st[seg].rank, old_prefix = st[seg].prefix;
    const uint32_t mine = hist[seg * 256 + t];
    hist[seg * 256 + t] = 0;
    h[t] = mine;
    __syncthreads();
    // Inclusive prefix sums of the digit counts.
    for (int d = 1; d < 256; d <<= 1) {
        const uint32_t add = t >= d ? h[t - d] : 0u;
        __syncthreads();
        h[t] += add;
        __syncthreads();
    }
    const uint32_t above = h[t], below = above - mine;
    // The one digit whose counts span the sought rank.
    if (!(rank >= below && rank < above)) return;
    const int shift = 24 - 8 * pass;
    const uint32_t prefix = old_prefix | ((uint32_t)t << shift);
    if (pass < 3) {
        st[seg].prefix = prefix;
        st[seg].rank = rank - below;
        return;
    }
    if (stage == 0) {
        // The second selection seeks the same index, len / 2, among the
        // distances to this median.
        int64_t start, flat;
        segment(q.nt, q.bins, seg % q.bins, start, flat);
        st[seg].median = key_value(prefix);
        st[seg].prefix = 0;
        st[seg].rank = (uint32_t)(flat * 256 / 2);
    } else {
        st[seg].mad = key_value(prefix);
    }
}

// noise.rs:104-108: sigma = 1.4826 MAD / 0.8908 and x = the mean brightness
// of the flat tiles, summed in their energy order (f32 Sum folds from -0.0).
__global__ __launch_bounds__(1024) void nr_points(Selection q,
                                                  const uint64_t* __restrict__ by_bright,
                                                  const uint64_t* __restrict__ by_energy,
                                                  const float* __restrict__ bright,
                                                  const SelectState* __restrict__ st,
                                                  float* __restrict__ points) {
    __shared__ float s_bright[BRIGHT_CHUNK];
    const int b = blockIdx.x, c = blockIdx.y, tid = threadIdx.x;
    int64_t start, flat;
    segment(q.nt, q.bins, b, start, flat);
    const uint64_t* order = by_energy + (size_t)c * q.N + start;
    float sum = -0.0f;
    for (int64_t base = 0; base < flat; base += BRIGHT_CHUNK) {
        const int64_t m = flat - base < BRIGHT_CHUNK ? flat - base : BRIGHT_CHUNK;
        for (int64_t j = tid; j < m; j += 1024)
            s_bright[j] = bright[(uint32_t)by_bright[order[base + j] & POS_MASK]];
        __syncthreads();
        if (tid == 0)
            for (int64_t j = 0; j < m; ++j) sum += s_bright[j];
        __syncthreads();
    }
    if (tid == 0) {
        const float sigma = MAD_TO_SIGMA * st[c * q.bins + b].mad / DETAIL_NOISE_GAIN;
        points[(c * MAX_BINS + b) * 2] = sum / (float)flat;
        points[(c * MAX_BINS + b) * 2 + 1] = sigma * sigma;
    }
}

// noise.rs:113-124, 142-170: the flat fallback (bins < 8) or the clamped
// least-squares line. prof = {a, b} x {luma, C1, C2}, then the bin counts.
__global__ void nr_fit(const float* __restrict__ points, int bins, float* __restrict__ prof) {
    const int c = threadIdx.x;
    if (c >= 3) return;
    const float* pt = points + c * MAX_BINS * 2;
    if (bins < 8) {
        prof[2 * c] = 0.0f;
        prof[2 * c + 1] = pt[1];
        prof[6 + c] = 0.0f;
        return;
    }
    const double n = (double)bins;
    double sx = 0.0, sy = 0.0, sxx = 0.0, sxy = 0.0;
    for (int i = 0; i < bins; ++i) {
        const double x = pt[2 * i], y = pt[2 * i + 1];
        sx += x;
        sy += y;
        sxx += x * x;
        sxy += x * y;
    }
    const double den = n * sxx - sx * sx;
    const double slope = fabs(den) < 1e-18 ? 0.0 : (n * sxy - sx * sy) / den;
    const double a = fmax(slope, 0.0);
    const double b = fmax((sy - a * sx) / n, 0.0);
    prof[2 * c] = (float)a;
    prof[2 * c + 1] = (float)b;
    prof[6 + c] = (float)bins;
}

// ---------------------------------------------------------------- chroma

constexpr float KERNEL[3] = {0.25f, 0.5f, 0.25f};

// chroma.rs:184-192.
__device__ __forceinline__ float fast_exp_neg(float t) {
    float x = num_max(1.0f - t * 0.125f, 0.0f);
    x = x * x;
    x = x * x;
    return x * x;
}

struct Chroma {
    Bal bal;
    float ceilings[3];
    float strength;
};

enum { FIRST = 0, MIDDLE = 1, LAST = 2 };

// One a-trous level of both colour planes (chroma.rs:198-374). FIRST splits
// the taps from RGB and stores the guide Y; LAST adds the residual and
// rebuilds RGB (chroma.rs:382-425).
template <int MODE>
__global__ __launch_bounds__(256) void nr_chroma_level(
    float* rgb, size_t n, int w, int h, int step, float level_weight, Chroma p,
    const float* __restrict__ prof, float* guide, const float* __restrict__ in1,
    const float* __restrict__ in2, float* __restrict__ out1, float* __restrict__ out2,
    float* __restrict__ acc1, float* __restrict__ acc2) {
    const uint3 blk = xcd_block();
    const int x = blk.x * 64 + (threadIdx.x & 63);
    const int y = blk.y * 4 + (threadIdx.x >> 6);
    if (x >= w || y >= h) return;
    int rows[3], cols[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        rows[k] = mirror_near(y + (k - 1) * step, h);
        cols[k] = mirror_near(x + (k - 1) * step, w);
    }
    float g[3][3], s1[3][3], s2[3][3];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const size_t q = (size_t)rows[j] * w + cols[k];
            if (MODE == FIRST) {
                split(rgb[q], rgb[n + q], rgb[2 * n + q], p.bal, g[j][k], s1[j][k], s2[j][k]);
            } else {
                g[j][k] = guide[q];
                s1[j][k] = in1[q];
                s2[j][k] = in2[q];
            }
        }
    const size_t i = (size_t)y * w + x;
    const float centre = g[1][1];
    // Shrinkage::inv_range_variance (chroma.rs:176-181).
    const float sg = num_max(sqrtf(variance(prof, 0, centre)), 1e-6f) * 4.0f;
    const float inv = 1.0f / (sg * sg);
    float sum1 = 0.0f, sum2 = 0.0f, wsum = 0.0f;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            // The centre tap is the guide itself (finite, as zeroed), and inv
            // is finite, so its weight is exactly 0.25 * 0.25 * 1.
            const float d = g[j][k] - centre;
            const float wt = (j == 1 && k == 1)
                                 ? KERNEL[1] * KERNEL[1]
                                 : KERNEL[j] * KERNEL[k] * fast_exp_neg(d * d * inv);
            sum1 += wt * s1[j][k];
            sum2 += wt * s2[j][k];
            wsum += wt;
        }
    const float o1 = sum1 / wsum, o2 = sum2 / wsum;
    const float scale = level_weight * p.strength * 5.0f;
    const float t1 = scale * sqrtf(variance(prof, 1, centre));
    const float t2 = scale * sqrtf(variance(prof, 2, centre));
    const float d1 = soft_threshold(s1[1][1] - o1, t1);
    const float d2 = soft_threshold(s2[1][1] - o2, t2);
    if (MODE == FIRST) {
        guide[i] = centre;
        acc1[i] = 0.0f + d1;
        acc2[i] = 0.0f + d2;
    } else if (MODE == MIDDLE) {
        acc1[i] = acc1[i] + d1;
        acc2[i] = acc2[i] + d2;
    }
    if (MODE != LAST) {
        out1[i] = o1;
        out2[i] = o2;
        return;
    }
    const float c1 = (acc1[i] + d1) + o1, c2 = (acc2[i] + d2) + o2;
    const float r0 = rgb[i], g0 = rgb[n + i], b0 = rgb[2 * n + i];
    const bool finite = is_finite(r0) && is_finite(g0) && is_finite(b0);
    const bool blown = r0 >= p.ceilings[0] || g0 >= p.ceilings[1] || b0 >= p.ceilings[2];
    if (!finite || blown) return;
    const float gg = centre - (c1 + c2) * 0.25f;
    const float r = (gg + c1) / p.bal.b[0], gn = gg / p.bal.b[1], b = (gg + c2) / p.bal.b[2];
    if (is_finite(r) && is_finite(gn) && is_finite(b)) {
        rgb[i] = r;
        rgb[n + i] = gn;
        rgb[2 * n + i] = b;
    }
}

// ---------------------------------------------------------------- luma

constexpr float BAND_NOISE_GAIN[5] = {0.8908f, 0.2006f, 0.0856f, 0.0413f, 0.0203f};
constexpr float SIGNAL_FLOOR = 0.4f;
constexpr float DETAIL_RELIEF = 0.8f;
constexpr float MAX_GAIN = 4.0f;
constexpr float LUMA_FLOOR = 1.0e-6f;

struct Luma {
    Bal bal;
    float ceilings[3];
    float strength, detail;
};

// Horizontal B3 of the level input; level 0 reads the luma of RGB.
template <bool FROM_RGB>
__global__ __launch_bounds__(256) void nr_luma_h(const float* __restrict__ src, size_t n, int w,
                                                 int step, Bal bal, float* __restrict__ tmp) {
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t row = (size_t)y * w;
    float t[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const size_t q = row + mirror_near(x + (k - 2) * step, w);
        t[k] = FROM_RGB ? luma_of(src[q], src[n + q], src[2 * n + q], bal) : src[q];
    }
    tmp[row + x] = b3(t[0], t[1], t[2], t[3], t[4]);
}

// Vertical B3 into `smooth`; the band (input - smoothing) replaces the input
// (luma.rs:153-166). Level 0 reads its input from RGB and writes the band.
template <bool FROM_RGB>
__global__ __launch_bounds__(256) void nr_luma_v(const float* src, size_t n, int w, int h,
                                                 int step, Bal bal, const float* __restrict__ tmp,
                                                 float* __restrict__ smooth, float* band) {
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const float* t = tmp + x;
    const float s = b3(t[(size_t)mirror_near(y - 2 * step, h) * w],
                       t[(size_t)mirror_near(y - step, h) * w], t[(size_t)y * w],
                       t[(size_t)mirror_near(y + step, h) * w],
                       t[(size_t)mirror_near(y + 2 * step, h) * w]);
    const size_t i = (size_t)y * w + x;
    const float cur = FROM_RGB ? luma_of(src[i], src[n + i], src[2 * n + i], bal) : src[i];
    smooth[i] = s;
    band[i] = cur - s;
}

// Horizontal 7-tap sums of band^2 (luma.rs:291-305).
__global__ __launch_bounds__(256) void nr_luma_energy_h(const float* __restrict__ band, int w,
                                                        float* __restrict__ out) {
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const float* line = band + (size_t)y * w;
    float sum = 0.0f;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        const float v = line[mirror_near(x + k - 3, w)];
        sum += v * v;
    }
    out[(size_t)y * w + x] = sum;
}

// Vertical 7-tap sums into the 7x7 mean square (luma.rs:306-320), the
// BayesShrink threshold (luma.rs:237-250) and the shrunk band into the
// accumulator; LAST adds the residual and applies the gain (luma.rs:323-366).
template <int MODE>
__global__ __launch_bounds__(256) void nr_luma_shrink(const float* __restrict__ hsum,
                                                      const float* __restrict__ band,
                                                      const float* __restrict__ smooth,
                                                      float* __restrict__ acc, float* rgb,
                                                      size_t n, int w, int h, int level, Luma p,
                                                      const float* __restrict__ prof) {
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    float energy = 0.0f;
#pragma unroll
    for (int k = 0; k < 7; ++k) energy += hsum[(size_t)mirror_near(y + k - 3, h) * w + x];
    energy *= 1.0f / 49.0f;
    const size_t i = (size_t)y * w + x;
    const float coarse = smooth[i];
    const float relief = level < 2 ? 1.0f - DETAIL_RELIEF * p.detail : 1.0f;
    const float scale = p.strength * 1.0f * relief;
    const float gain_sq = BAND_NOISE_GAIN[level] * BAND_NOISE_GAIN[level];
    const float nv = variance(prof, 0, coarse) * gain_sq;
    const float floor = SIGNAL_FLOOR * SIGNAL_FLOOR * nv;
    const float sigma_signal = sqrtf(num_max(energy - nv, floor));
    const float threshold = sigma_signal > 0.0f ? scale * nv / sigma_signal : 0.0f;
    const float d = soft_threshold(band[i], threshold);
    if (MODE == FIRST) {
        acc[i] = 0.0f + d;
        return;
    }
    if (MODE == MIDDLE) {
        acc[i] = acc[i] + d;
        return;
    }
    const float filtered = (acc[i] + d) + coarse;
    const float r0 = rgb[i], g0 = rgb[n + i], b0 = rgb[2 * n + i];
    const bool finite = is_finite(r0) && is_finite(g0) && is_finite(b0);
    const bool blown = r0 >= p.ceilings[0] || g0 >= p.ceilings[1] || b0 >= p.ceilings[2];
    if (!finite || blown) return;
    const float old = (r0 * p.bal.b[0] + 2.0f * g0 * p.bal.b[1] + b0 * p.bal.b[2]) * 0.25f;
    if (old <= LUMA_FLOOR) return;
    float gain = filtered / old;
    if (gain < 0.0f) gain = 0.0f;
    if (gain > MAX_GAIN) gain = MAX_GAIN;
    const float r = r0 * gain, g = g0 * gain, b = b0 * gain;
    if (is_finite(r) && is_finite(g) && is_finite(b)) {
        rgb[i] = r;
        rgb[n + i] = g;
        rgb[2 * n + i] = b;
    }
}

size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

}  // namespace

NoiseGeometry noise_geometry(int width, int height) {
    NoiseGeometry g{};
    g.rows = height < 1024 ? height : 0;
    if (height >= 1024)
        for (int start = 0; start < height; start += 256)
            g.rows += (height - start < 64 ? height - start : 64);
    g.tiles_x = width / 16;
    g.tiles = (int64_t)g.tiles_x * (g.rows / 16);
    g.bins = (int)(g.tiles / 4 / 8 < 16 ? g.tiles / 4 / 8 : 16);
    int64_t N = 2;
    while (N < g.tiles) N <<= 1;
    g.sort_n = N;
    const size_t plane = (size_t)width * g.rows * sizeof(float);
    g.band0 = align_up(3 * plane);
    g.tmp = align_up(g.band0 + plane);
    g.energy = align_up(g.tmp + 3 * plane);
    g.bright = align_up(g.energy + 3 * (size_t)g.tiles * sizeof(float));
    g.keys = align_up(g.bright + (size_t)g.tiles * sizeof(float));
    g.keys2 = align_up(g.keys + (size_t)N * sizeof(uint64_t));
    g.points = align_up(g.keys2 + 3 * (size_t)N * sizeof(uint64_t));
    const int bins = g.bins >= 8 ? g.bins : 1;
    g.flat_tiles = 0;
    g.hist_blocks = 0;
    for (int b = 0; b < bins && g.tiles > 0; ++b) {
        int64_t start, flat;
        segment((int)g.tiles, bins, b, start, flat);
        g.flat_tiles += flat;
        g.hist_blocks += (flat * 256 + HIST_CHUNK - 1) / HIST_CHUNK;
    }
    g.state = align_up(g.points + 3 * MAX_BINS * 2 * sizeof(float));
    g.hist = align_up(g.state + 3 * MAX_BINS * sizeof(SelectState));
    g.samples = align_up(g.hist + 3 * MAX_BINS * 256 * sizeof(uint32_t));
    g.bytes = align_up(g.samples + 3 * (size_t)g.flat_tiles * 256 * sizeof(float));
    return g;
}

size_t noise_filter_bytes(int width, int height) {
    return (size_t)width * height * sizeof(float) * 7;
}

hipError_t launch_noise_profile(const float* rgb, int width, int height, const float balance[3],
                                void* scratch, float* prof, hipStream_t s, Profiler* pr) {
    const NoiseGeometry g = noise_geometry(width, height);
    if (g.tiles <= 0 || g.tiles > MAX_NOISE_TILES) return hipErrorInvalidValue;
    const size_t n = (size_t)width * height, pitch = (size_t)width * g.rows;
    char* base = (char*)scratch;
    float* planes = (float*)base;
    float* band0 = (float*)(base + g.band0);
    float* tmp = (float*)(base + g.tmp);
    float* energy = (float*)(base + g.energy);
    float* bright = (float*)(base + g.bright);
    uint64_t* keys = (uint64_t*)(base + g.keys);
    uint64_t* keys2 = (uint64_t*)(base + g.keys2);
    float* points = (float*)(base + g.points);
    Bal bal{{balance[0], balance[1], balance[2]}};
    const unsigned bx = (unsigned)((width + 255) / 256);
    const uint64_t px = (uint64_t)width * g.rows;
    RHD_TIMED(pr, s, "nr_est_split", 24 * px,
              nr_est_split<<<dim3(bx, g.rows), 256, 0, s>>>(rgb, n, width, height, bal, planes,
                                                            pitch));
    RHD_TIMED(pr, s, "nr_b3_h", 24 * px,
              nr_b3_h<<<dim3(bx, g.rows, 3), 256, 0, s>>>(planes, tmp, width, pitch, 1));
    RHD_TIMED(pr, s, "nr_est_band", 36 * px,
              nr_est_band<<<dim3(bx, g.rows, 3), 256, 0, s>>>(planes, tmp, band0, width, g.rows,
                                                              pitch));
    const int nt = (int)g.tiles, N = (int)g.sort_n;
    RHD_TIMED(pr, s, "nr_est_tiles", 16 * px,
              nr_est_tiles<<<dim3((nt + 255) / 256, 3), 256, 0, s>>>(
                  band0, planes, pitch, width, g.tiles_x, nt, energy, bright));
    RHD_TIMED(pr, s, "nr_keys", 12 * (uint64_t)N,
              nr_keys_bright<<<(N + 255) / 256, 256, 0, s>>>(bright, nt, N, keys));
    sort_keys(keys, N, 1, s, pr);
    const int bins = g.bins >= 8 ? g.bins : 1;
    RHD_TIMED(pr, s, "nr_keys", 48 * (uint64_t)N,
              nr_keys_energy<<<dim3((N + 255) / 256, 3), 256, 0, s>>>(keys, energy, nt, N, bins,
                                                                      keys2));
    // Inside each bin only: whole bins in LDS when they fit, else the
    // full-array network (the bin field keeps the bins apart).
    const int64_t longest = ((int64_t)nt + bins - 1) / bins;
    const dim3 per_bin(bins, 3);
    if (longest <= 2048)
        RHD_TIMED(pr, s, "nr_sort", 16 * 3 * (uint64_t)nt,
                  nr_sort_bins<2048><<<per_bin, 1024, 0, s>>>(keys2, N, nt, bins));
    else if (longest <= 4096)
        RHD_TIMED(pr, s, "nr_sort", 16 * 3 * (uint64_t)nt,
                  nr_sort_bins<4096><<<per_bin, 1024, 0, s>>>(keys2, N, nt, bins));
    else if (longest <= 8192)
        RHD_TIMED(pr, s, "nr_sort", 16 * 3 * (uint64_t)nt,
                  nr_sort_bins<8192><<<per_bin, 1024, 0, s>>>(keys2, N, nt, bins));
    else
        sort_keys(keys2, N, 3, s, pr);
    Selection q{nt, bins, N, g.tiles_x, width, pitch, g.flat_tiles * 256};
    SelectState* state = (SelectState*)(base + g.state);
    uint32_t* hist = (uint32_t*)(base + g.hist);
    float* samples = (float*)(base + g.samples);
    const uint64_t nsamples = 3 * (uint64_t)g.flat_tiles * 256;
    RHD_TIMED(pr, s, "nr_select", 0,
              nr_select_init<<<3 * bins, 256, 0, s>>>(q, state, hist));
    RHD_TIMED(pr, s, "nr_gather", 8 * nsamples,
              nr_gather<<<dim3((unsigned)g.flat_tiles, 3), 256, 0, s>>>(q, band0, planes, keys,
                                                                       keys2, samples));
    for (int stage = 0; stage < 2; ++stage)
        for (int pass = 0; pass < 4; ++pass) {
            RHD_TIMED(pr, s, "nr_select_hist", 4 * nsamples,
                      nr_select_hist<<<dim3((unsigned)g.hist_blocks, 3), 256, 0, s>>>(
                          q, stage, pass, samples, state, hist));
            RHD_TIMED(pr, s, "nr_select", 0,
                      nr_select_digit<<<3 * bins, 256, 0, s>>>(q, stage, pass, state, hist));
        }
    RHD_TIMED(pr, s, "nr_select", 0,
              nr_points<<<dim3(bins, 3), 1024, 0, s>>>(q, keys, keys2, bright, state, points));
    RHD_TIMED(pr, s, "nr_fit", 0, nr_fit<<<1, 64, 0, s>>>(points, g.bins, prof));
    return hipGetLastError();
}

hipError_t launch_chroma_nr(float* rgb, int width, int height, float strength,
                            const float balance[3], const float ceilings[3], const float* prof,
                            float* scratch, hipStream_t s, Profiler* pr) {
    const size_t n = (size_t)width * height;
    Chroma p{};
    for (int c = 0; c < 3; ++c) {
        p.bal.b[c] = balance[c];
        p.ceilings[c] = ceilings[c];
    }
    p.strength = strength;
    float* guide = scratch;
    float* a1 = scratch + n;
    float* a2 = scratch + 2 * n;
    float* x[2][2] = {{scratch + 3 * n, scratch + 4 * n}, {scratch + 5 * n, scratch + 6 * n}};
    const float weights[5] = {1.0f, 0.8f, 0.6f, 0.4f, 0.3f};
    const dim3 grid((width + 63) / 64, (height + 3) / 4);
    RHD_TIMED(pr, s, "nr_chroma_level", 32 * (uint64_t)n,
              nr_chroma_level<FIRST><<<grid, 256, 0, s>>>(rgb, n, width, height, 1, weights[0], p,
                                                          prof, guide, nullptr, nullptr, x[0][0],
                                                          x[0][1], a1, a2));
    for (int level = 1; level < 4; ++level) {
        float** in = x[(level - 1) & 1];
        float** out = x[level & 1];
        RHD_TIMED(pr, s, "nr_chroma_level", 36 * (uint64_t)n,
                  nr_chroma_level<MIDDLE><<<grid, 256, 0, s>>>(rgb, n, width, height, 1 << level,
                                                               weights[level], p, prof, guide,
                                                               in[0], in[1], out[0], out[1], a1,
                                                               a2));
    }
    RHD_TIMED(pr, s, "nr_chroma_last", 44 * (uint64_t)n,
              nr_chroma_level<LAST><<<grid, 256, 0, s>>>(rgb, n, width, height, 16, weights[4], p,
                                                         prof, guide, x[1][0], x[1][1], nullptr,
                                                         nullptr, a1, a2));
    return hipGetLastError();
}

hipError_t launch_luma_nr(float* rgb, int width, int height, float strength, float detail,
                          const float balance[3], const float ceilings[3], const float* prof,
                          float* scratch, hipStream_t s, Profiler* pr) {
    const size_t n = (size_t)width * height;
    Luma p{};
    for (int c = 0; c < 3; ++c) {
        p.bal.b[c] = balance[c];
        p.ceilings[c] = ceilings[c];
    }
    p.strength = strength;
    p.detail = detail;
    float* acc = scratch;
    float* tmp = scratch + n;
    float* buf[2] = {scratch + 2 * n, scratch + 3 * n};
    const dim3 grid((width + 255) / 256, height);
    for (int level = 0; level < 5; ++level) {
        const int step = 1 << level;
        float* band = buf[level & 1];     // this level's input, then its band
        float* smooth = buf[~level & 1];  // the next level's input
        if (level == 0) {
            RHD_TIMED(pr, s, "nr_luma_h", 16 * (uint64_t)n,
                      nr_luma_h<true><<<grid, 256, 0, s>>>(rgb, n, width, step, p.bal, tmp));
            RHD_TIMED(pr, s, "nr_luma_v", 24 * (uint64_t)n,
                      nr_luma_v<true><<<grid, 256, 0, s>>>(rgb, n, width, height, step, p.bal,
                                                           tmp, smooth, band));
        } else {
            RHD_TIMED(pr, s, "nr_luma_h", 8 * (uint64_t)n,
                      nr_luma_h<false><<<grid, 256, 0, s>>>(band, n, width, step, p.bal, tmp));
            RHD_TIMED(pr, s, "nr_luma_v", 16 * (uint64_t)n,
                      nr_luma_v<false><<<grid, 256, 0, s>>>(band, n, width, height, step, p.bal,
                                                            tmp, smooth, band));
        }
        RHD_TIMED(pr, s, "nr_luma_energy_h", 8 * (uint64_t)n,
                  nr_luma_energy_h<<<grid, 256, 0, s>>>(band, width, tmp));
        if (level == 0)
            RHD_TIMED(pr, s, "nr_luma_shrink", 16 * (uint64_t)n,
                      nr_luma_shrink<FIRST><<<grid, 256, 0, s>>>(tmp, band, smooth, acc, rgb, n,
                                                                 width, height, level, p, prof));
        else if (level < 4)
            RHD_TIMED(pr, s, "nr_luma_shrink", 20 * (uint64_t)n,
                      nr_luma_shrink<MIDDLE><<<grid, 256, 0, s>>>(tmp, band, smooth, acc, rgb, n,
                                                                  width, height, level, p, prof));
        else
            RHD_TIMED(pr, s, "nr_luma_last", 40 * (uint64_t)n,
                      nr_luma_shrink<LAST><<<grid, 256, 0, s>>>(tmp, band, smooth, acc, rgb, n,
                                                                width, height, level, p, prof));
    }
    return hipGetLastError();
}

}  // namespace rhd
