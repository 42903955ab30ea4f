// Oracle (test infrastructure only): demosaic drivers restated from
// crates/demosaic/src/{lib,balance,bilinear,malvar_he_cutler}.rs and the tile
// loop of crates/demosaic/src/amaze/mod.rs. Output is interleaved RGB with a
// row stride of 3 * width, like LinearRgbImage.
#pragma once

#include <atomic>
#include <thread>
#include <vector>

#include "amaze.hpp"
#include "common.hpp"

namespace rhdo {

enum Algorithm : int { BILINEAR = 0, MHC = 1, RCD = 2, AMAZE = 3 };

// Runs `fn(job)` for job in [0, jobs) on `threads` workers pulling from a shared
// counter, like rayon's work stealing for independent tile and row jobs.
template <class F> void parallel_for(size_t jobs, int threads, F fn) {
    if (threads <= 1 || jobs <= 1) {
        for (size_t j = 0; j < jobs; ++j) fn(j);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            for (size_t j = next.fetch_add(1); j < jobs; j = next.fetch_add(1)) fn(j);
        });
    for (auto& th : pool) th.join();
}

template <class T> struct Mosaic {
    const T* data;
    size_t width, height, stride;
    int pattern;
    T at(size_t x, size_t y) const { return data[y * stride + x]; }
};

// bilinear.rs:82-113 (average_offsets)
template <class T>
T average_offsets(const Mosaic<T>& m, size_t x, size_t y, const int (*offsets)[2], int n) {
    T sum = T(0);
    unsigned count = 0;
    for (int k = 0; k < n; ++k) {
        ptrdiff_t nx = (ptrdiff_t)x + offsets[k][0], ny = (ptrdiff_t)y + offsets[k][1];
        if (nx < 0 || ny < 0) continue;
        if ((size_t)nx < m.width && (size_t)ny < m.height) {
            sum += m.at((size_t)nx, (size_t)ny);
            ++count;
        }
    }
    return count == 0 ? m.at(x, y) : sum / (T)count;
}

// bilinear.rs:57-80 (reconstruct_pixel)
template <class T> void bilinear_pixel(const Mosaic<T>& m, size_t x, size_t y, T rgb[3]) {
    static const int cross[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
    static const int diag[4][2] = {{-1, -1}, {1, -1}, {-1, 1}, {1, 1}};
    static const int horiz[2][2] = {{-1, 0}, {1, 0}};
    static const int vert[2][2] = {{0, -1}, {0, 1}};
    int c = color_at(m.pattern, x, y);
    if (c == 0) {
        rgb[0] = m.at(x, y);
        rgb[1] = average_offsets(m, x, y, cross, 4);
        rgb[2] = average_offsets(m, x, y, diag, 4);
    } else if (c == 2) {
        rgb[0] = average_offsets(m, x, y, diag, 4);
        rgb[1] = average_offsets(m, x, y, cross, 4);
        rgb[2] = m.at(x, y);
    } else {
        bool red_horizontal = color_at(m.pattern, x + 1, y) == 0;
        rgb[0] = average_offsets(m, x, y, red_horizontal ? horiz : vert, 2);
        rgb[1] = m.at(x, y);
        rgb[2] = average_offsets(m, x, y, red_horizontal ? vert : horiz, 2);
    }
}

template <class T> inline void apply_gains(T rgb[3], const T gains[3]) {
    rgb[0] *= gains[0];
    rgb[1] *= gains[1];
    rgb[2] *= gains[2];
}

template <class T> inline bool all_finite(const T rgb[3]) {
    return finite(rgb[0]) && finite(rgb[1]) && finite(rgb[2]);
}

// Status codes shared with the product's C ABI.
constexpr int OK = 0;
constexpr int ERR_NONFINITE = 8;  // InvalidInput (non-finite image data)
constexpr int ERR_DIMENSIONS = 4;

template <class T>
int bilinear_reconstruct(const Mosaic<T>& m, const T gains[3], T* out, int threads) {
    std::atomic<int> status{OK};
    parallel_for(m.height, threads, [&](size_t y) {
        for (size_t x = 0; x < m.width; ++x) {
            T rgb[3];
            bilinear_pixel(m, x, y, rgb);
            apply_gains(rgb, gains);
            if (!all_finite(rgb)) status = ERR_NONFINITE;
            T* p = out + (y * m.width + x) * 3;
            p[0] = rgb[0]; p[1] = rgb[1]; p[2] = rgb[2];
        }
    });
    return status;
}

// malvar_he_cutler.rs:16-38
static const int MHC_G_AT_RB[5][5] = {
    {0, 0, -2, 0, 0}, {0, 0, 4, 0, 0}, {-2, 4, 8, 4, -2}, {0, 0, 4, 0, 0}, {0, 0, -2, 0, 0}};
static const int MHC_AT_G_ROW[5][5] = {
    {0, 0, 1, 0, 0}, {0, -2, 0, -2, 0}, {-2, 8, 10, 8, -2}, {0, -2, 0, -2, 0}, {0, 0, 1, 0, 0}};
static const int MHC_RB_AT_BR[5][5] = {
    {0, 0, -3, 0, 0}, {0, 4, 0, 4, 0}, {-3, 0, 12, 0, -3}, {0, 4, 0, 4, 0}, {0, 0, -3, 0, 0}};

// malvar_he_cutler.rs:97-119: row-major sum of the non-zero taps, times 1/16.
template <class T>
T mhc_kernel(const Mosaic<T>& m, size_t x, size_t y, const int k[5][5], bool transpose) {
    T sum = T(0);
    for (int ky = 0; ky < 5; ++ky)
        for (int kx = 0; kx < 5; ++kx) {
            int c = transpose ? k[kx][ky] : k[ky][kx];
            if (c != 0) sum += (T)c * m.at(x + kx - 2, y + ky - 2);
        }
    return sum * (T(1) / T(16));
}

template <class T> int mhc_reconstruct(const Mosaic<T>& m, const T gains[3], T* out, int threads) {
    std::atomic<int> status{OK};
    parallel_for(m.height, threads, [&](size_t y) {
        for (size_t x = 0; x < m.width; ++x) {
            T rgb[3];
            bool border = x < 2 || y < 2 || x + 2 >= m.width || y + 2 >= m.height;
            if (m.width < 2 || m.height < 2) border = true;
            if (border) {
                bilinear_pixel(m, x, y, rgb);
            } else {
                T measured = m.at(x, y);
                int c = color_at(m.pattern, x, y);
                if (c == 0) {
                    rgb[0] = measured;
                    rgb[1] = mhc_kernel(m, x, y, MHC_G_AT_RB, false);
                    rgb[2] = mhc_kernel(m, x, y, MHC_RB_AT_BR, false);
                } else if (c == 2) {
                    rgb[0] = mhc_kernel(m, x, y, MHC_RB_AT_BR, false);
                    rgb[1] = mhc_kernel(m, x, y, MHC_G_AT_RB, false);
                    rgb[2] = measured;
                } else {
                    bool red_horizontal = color_at(m.pattern, x + 1, y) == 0;
                    rgb[0] = mhc_kernel(m, x, y, MHC_AT_G_ROW, !red_horizontal);
                    rgb[1] = measured;
                    rgb[2] = mhc_kernel(m, x, y, MHC_AT_G_ROW, red_horizontal);
                }
            }
            apply_gains(rgb, gains);
            if (!all_finite(rgb)) status = ERR_NONFINITE;
            T* p = out + (y * m.width + x) * 3;
            p[0] = rgb[0]; p[1] = rgb[1]; p[2] = rgb[2];
        }
    });
    return status;
}

// amaze/mod.rs:171-194 (populate)
template <class T>
void amaze_populate(amaze::Scratch<T>& s, const Mosaic<T>& m, size_t origin_x, size_t origin_y) {
    using namespace amaze;
    for (int row = 0; row < TS; ++row) {
        size_t sy = mirror_coordinate((ptrdiff_t)origin_y + row - BORDER, m.height);
        for (int col = 0; col < TS; ++col) {
            size_t sx = mirror_coordinate((ptrdiff_t)origin_x + col - BORDER, m.width);
            T v = m.at(sx, sy);
            s.cfa[row * TS + col] = v;
            s.green[row * TS + col] = v;
        }
    }
}

// amaze/mod.rs:196-273 (write_tile)
template <class T>
bool amaze_write_tile(const amaze::Scratch<T>& s, const Mosaic<T>& m, const T gains[3],
                      size_t origin_x, size_t origin_y, T* out) {
    using namespace amaze;
    size_t x_end = origin_x + VALID < m.width ? origin_x + VALID : m.width;
    size_t y_end = origin_y + VALID < m.height ? origin_y + VALID : m.height;
    bool ok = true;
    for (size_t gy = origin_y; gy < y_end; ++gy) {
        size_t ly = gy - origin_y + BORDER;
        for (size_t gx = origin_x; gx < x_end; ++gx) {
            size_t at = ly * TS + (gx - origin_x + BORDER);
            T sample = m.at(gx, gy);
            int ch = color_at(m.pattern, gx, gy);
            T rgb[3] = {ch == 0 ? sample : s.red[at], ch == 1 ? sample : sel_max(s.green[at], T(0)),
                        ch == 2 ? sample : s.blue[at]};
            apply_gains(rgb, gains);
            ok &= all_finite(rgb);
            T* p = out + (gy * m.width + gx) * 3;
            p[0] = rgb[0]; p[1] = rgb[1]; p[2] = rgb[2];
        }
    }
    return ok;
}

template <class T>
int amaze_reconstruct(const Mosaic<T>& m, const T gains[3], T clip_point, T* out, int threads,
                      bool poison) {
    using namespace amaze;
    if (m.width <= 2 * BORDER || m.height <= 2 * BORDER) return bilinear_reconstruct(m, gains, out, threads);
    size_t cols = (m.width + VALID - 1) / VALID, rows = (m.height + VALID - 1) / VALID;
    std::atomic<int> status{OK};
    int workers = threads < 1 ? 1 : threads;
    std::vector<amaze::Scratch<T>*> pool(workers, nullptr);
    std::atomic<size_t> next{0};
    auto work = [&](int w) {
        if (!pool[w]) pool[w] = new amaze::Scratch<T>();
        amaze::Scratch<T>& s = *pool[w];
        for (size_t j = next.fetch_add(1); j < cols * rows; j = next.fetch_add(1)) {
            size_t ox = (j % cols) * VALID, oy = (j / cols) * VALID;
            s.clear(poison);
            amaze_populate(s, m, ox, oy);
            reconstruct_tile(s, shifted(m.pattern, 0, oy), clip_point);
            if (!amaze_write_tile(s, m, gains, ox, oy, out)) status = ERR_NONFINITE;
        }
    };
    if (workers == 1) {
        work(0);
    } else {
        std::vector<std::thread> th;
        for (int w = 0; w < workers; ++w) th.emplace_back(work, w);
        for (auto& t : th) t.join();
    }
    for (auto* p : pool) delete p;
    return status;
}

// balance.rs:33-44 (ChannelBalance::new), :60-66 (clip_point).
template <class T> struct Balance {
    T forward[3];
    explicit Balance(const T gains[3]) {
        T largest = num_max(num_max(gains[0], gains[1]), gains[2]);
        for (int c = 0; c < 3; ++c) forward[c] = gains[c] / largest;
    }
    bool identity() const { return forward[0] == T(1) && forward[1] == T(1) && forward[2] == T(1); }
    T clip_point() const { return num_min(num_min(forward[0], forward[1]), forward[2]); }
};

template <class T>
int rcd_reconstruct(const Mosaic<T>& m, const T gains[3], T* out, int threads);

// lib.rs:302-375 (demosaic_cancellable) with the balance logic of
// balance.rs: RCD and AMaZE reconstruct on a scaled copy and undo it through
// the output gains.
template <class T>
int demosaic(const T* data, size_t width, size_t height, size_t stride, int pattern,
             const T balance_gains[3], int algorithm, T* out, int threads, bool poison) {
    if (width < 2 || height < 2) return ERR_DIMENSIONS;
    for (size_t y = 0; y < height; ++y)
        for (size_t x = 0; x < width; ++x)
            if (!finite(data[y * stride + x])) return ERR_NONFINITE;
    Balance<T> balance(balance_gains);
    Mosaic<T> m{data, width, height, stride, pattern};
    std::vector<T> scaled;
    T gains[3] = {T(1), T(1), T(1)};
    T clip_point = T(1);
    if ((algorithm == RCD || algorithm == AMAZE) && !balance.identity()) {
        scaled.resize(width * height);
        for (size_t y = 0; y < height; ++y)
            for (size_t x = 0; x < width; ++x)
                scaled[y * width + x] = data[y * stride + x] * balance.forward[color_at(pattern, x, y)];
        m = Mosaic<T>{scaled.data(), width, height, width, pattern};
        for (int c = 0; c < 3; ++c) gains[c] = T(1) / balance.forward[c];
        clip_point = balance.clip_point();
    }
    switch (algorithm) {
        case BILINEAR: return bilinear_reconstruct(m, gains, out, threads);
        case MHC: return mhc_reconstruct(m, gains, out, threads);
        case RCD: return rcd_reconstruct(m, gains, out, threads);
        case AMAZE: return amaze_reconstruct(m, gains, clip_point, out, threads, poison);
        default: return ERR_DIMENSIONS;
    }
}

}  // namespace rhdo
