// Oracle (test infrastructure only): tiled RCD demosaic restated from
// crates/demosaic/src/rcd.rs:36-299 (tiles, populate, write), rcd_stages.rs
// (the stages) and rcd_geometry.rs (194-pixel tiles, 10-pixel mirrored halo).
// `interpolate` is a fused multiply-add in the reference (rcd_stages.rs:
// 455-456), so it is std::fma here.
#pragma once

#include <cmath>
#include <vector>

#include "common.hpp"
#include "demosaic.hpp"

namespace rhdo {
namespace rcd {

constexpr int TS = 194, BORDER = 10, STEP = TS - 2 * BORDER, MIN_DIMENSION = 2 * BORDER + 4;

template <class T> struct Scratch {
    std::vector<T> cfa, rgb0, rgb1, rgb2, vh, pq, p, q;
    Scratch() : cfa(TS * TS), rgb0(TS * TS), rgb1(TS * TS), rgb2(TS * TS), vh(TS * TS),
                pq(TS * TS / 2), p(TS * TS / 2), q(TS * TS / 2) {}
    std::vector<T>& rgb(int c) { return c == 0 ? rgb0 : (c == 1 ? rgb1 : rgb2); }
    void clear() {
        for (auto* v : {&cfa, &rgb0, &rgb1, &rgb2, &vh, &pq, &p, &q})
            std::fill(v->begin(), v->end(), T(0));
    }
};

template <class T> inline T eps() { return RHDO_LIT(T, 1.0e-5); }
template <class T> inline T eps_sq() { return RHDO_LIT(T, 1.0e-10); }
template <class T> inline T square(T v) { return v * v; }

// rcd_stages.rs:432-445
template <class T> inline T vhp(const T* v, size_t i) {
    return (v[i - 3 * TS] - v[i - TS] - v[i + TS] + v[i + 3 * TS]) -
           T(3) * (v[i - 2 * TS] + v[i + 2 * TS]) + T(6) * v[i];
}
template <class T> inline T hhp(const T* v, size_t i) {
    return (v[i - 3] - v[i - 1] - v[i + 1] + v[i + 3]) - T(3) * (v[i - 2] + v[i + 2]) + T(6) * v[i];
}
// rcd_stages.rs:425-430
inline size_t off(size_t i, int dx, int dy) { return (size_t)((ptrdiff_t)i + dy * TS + dx); }
// rcd_stages.rs:395-406
template <class T> inline T cardinal_gradient(const T* v, size_t i, int dx, int dy) {
    size_t one = off(i, dx, dy), opp = off(i, -dx, -dy), two = off(i, 2 * dx, 2 * dy),
           three = off(i, 3 * dx, 3 * dy), four = off(i, 4 * dx, 4 * dy);
    return eps<T>() + std::fabs(v[one] - v[opp]) + std::fabs(v[i] - v[two]) +
           std::fabs(v[one] - v[three]) + std::fabs(v[two] - v[four]);
}
// rcd_stages.rs:408-423
template <class T> inline T diagonal_gradient(const T* v, const T* g, size_t i, int dx, int dy) {
    size_t near_ = off(i, dx, dy), opp = off(i, -dx, -dy), far_ = off(i, 3 * dx, 3 * dy),
           g2 = off(i, 2 * dx, 2 * dy);
    return eps<T>() + std::fabs(v[near_] - v[opp]) + std::fabs(v[near_] - v[far_]) +
           std::fabs(g[i] - g[g2]);
}
// rcd_stages.rs:465-473
template <class T> inline T nonzero_denominator(T v) {
    if (std::fabs(v) >= eps<T>()) return v;
    return std::signbit(v) ? -eps<T>() : eps<T>();
}
template <class T> inline T ratio_estimate(T sample, T lp, T nlp) {
    T den = nonzero_denominator(eps<T>() + lp + nlp);
    return sample * (T(2) * lp) / den;
}
template <class T> inline T weighted_pair(T w1, T v1, T w2, T v2) {
    return (w1 * v1 + w2 * v2) / nonzero_denominator(w1 + w2);
}
// rcd_stages.rs:475-489
template <class T> inline T refined_direction(T c, T nw, T ne, T sw, T se) {
    T hood = RHDO_LIT(T, 0.25) * (nw + ne + sw + se);
    T half = RHDO_LIT(T, 0.5);
    T sel = std::fabs(half - c) < std::fabs(half - hood) ? hood : c;
    return sel < T(0) ? T(0) : (sel > T(1) ? T(1) : sel);
}
template <class T> inline T interpolate(T d, T h, T v) { return std::fma(d, h, (T(1) - d) * v); }

template <class T> void find_directions(Scratch<T>& s) {
    std::vector<T> vb[3] = {std::vector<T>(TS), std::vector<T>(TS), std::vector<T>(TS)};
    std::vector<T> hb(TS);
    const T* cfa = s.cfa.data();
    int initial_end = std::min(TS - 3, 5);
    for (int row = 3; row < initial_end; ++row)
        for (int col = 4; col < TS - 4; ++col) vb[row - 3][col - 4] = square(vhp(cfa, row * TS + col));
    int v0 = 0, v1 = 1, v2 = 2;
    for (int row = 4; row < TS - 4; ++row) {
        for (int col = 3; col < TS - 3; ++col) hb[col - 3] = square(hhp(cfa, row * TS + col));
        for (int col = 4; col < TS - 4; ++col) vb[v2][col - 4] = square(vhp(cfa, (row + 1) * TS + col));
        for (int col = 4; col < TS - 4; ++col) {
            T vs = num_max(vb[v0][col - 4] + vb[v1][col - 4] + vb[v2][col - 4], eps_sq<T>());
            T hs = num_max(hb[col - 4] + hb[col - 3] + hb[col - 2], eps_sq<T>());
            s.vh[row * TS + col] = vs / (vs + hs);
        }
        int t = v0; v0 = v1; v1 = v2; v2 = t;
    }
}

template <class T> void low_pass(Scratch<T>& s, int pattern) {
    const T* c = s.cfa.data();
    for (int row = 2; row < TS - 2; ++row) {
        int start = 2 + (color_at(pattern, 0, row) & 1);
        for (int col = start; col < TS - 2; col += 2) {
            size_t i = row * TS + col;
            s.pq[i / 2] = c[i] + RHDO_LIT(T, 0.5) * (c[i - TS] + c[i + TS] + c[i - 1] + c[i + 1]) +
                          RHDO_LIT(T, 0.25) * (c[i - TS - 1] + c[i - TS + 1] + c[i + TS - 1] + c[i + TS + 1]);
        }
    }
}

template <class T> void interpolate_green(Scratch<T>& s, int pattern) {
    const T* c = s.cfa.data();
    for (int row = 4; row < TS - 4; ++row) {
        int start = 4 + (color_at(pattern, 0, row) & 1);
        for (int col = start; col < TS - 4; col += 2) {
            size_t i = row * TS + col;
            T ng = cardinal_gradient(c, i, 0, -1), sg = cardinal_gradient(c, i, 0, 1);
            T wg = cardinal_gradient(c, i, -1, 0), eg = cardinal_gradient(c, i, 1, 0);
            size_t li = i / 2;
            T lp = s.pq[li];
            T n_est = ratio_estimate(c[i - TS], lp, s.pq[li - TS]);
            T s_est = ratio_estimate(c[i + TS], lp, s.pq[li + TS]);
            T w_est = ratio_estimate(c[i - 1], lp, s.pq[li - 1]);
            T e_est = ratio_estimate(c[i + 1], lp, s.pq[li + 1]);
            T vert = weighted_pair(sg, n_est, ng, s_est);
            T horiz = weighted_pair(wg, e_est, eg, w_est);
            T d = refined_direction(s.vh[i], s.vh[i - TS - 1], s.vh[i - TS + 1], s.vh[i + TS - 1],
                                    s.vh[i + TS + 1]);
            s.rgb1[i] = interpolate(d, horiz, vert);
        }
    }
}

template <class T> void diagonal_high_pass(Scratch<T>& s) {
    const T* c = s.cfa.data();
    for (int row = 3; row < TS - 3; ++row)
        for (int col = 3; col < TS - 3; col += 2) {
            size_t i = row * TS + col;
            s.p[i / 2] = square((c[i - 3 * TS - 3] - c[i - TS - 1] - c[i + TS + 1] + c[i + 3 * TS + 3]) -
                                T(3) * (c[i - 2 * TS - 2] + c[i + 2 * TS + 2]) + T(6) * c[i]);
            s.q[i / 2] = square((c[i - 3 * TS + 3] - c[i - TS + 1] - c[i + TS - 1] + c[i + 3 * TS - 3]) -
                                T(3) * (c[i - 2 * TS + 2] + c[i + 2 * TS - 2]) + T(6) * c[i]);
        }
}

template <class T> void diagonal_directions(Scratch<T>& s, int pattern) {
    for (int row = 4; row < TS - 4; ++row) {
        int start = 4 + (color_at(pattern, 0, row) & 1);
        for (int col = start; col < TS - 4; col += 2) {
            size_t i = row * TS + col, h = i / 2, prev = (i - TS - 1) / 2, next = (i + TS - 1) / 2;
            T ps = num_max(s.p[prev] + s.p[h] + s.p[next + 1], eps_sq<T>());
            T qs = num_max(s.q[prev + 1] + s.q[h] + s.q[next], eps_sq<T>());
            s.pq[h] = ps / (ps + qs);
        }
    }
}

template <class T> void opposite_color(Scratch<T>& s, int pattern) {
    for (int row = 4; row < TS - 4; ++row) {
        int start = 4 + (color_at(pattern, 0, row) & 1);
        for (int col = start; col < TS - 4; col += 2) {
            size_t i = row * TS + col;
            int ch = 2 - color_at(pattern, col, row);
            size_t h = i / 2, prev = (i - TS - 1) / 2, next = (i + TS - 1) / 2;
            T d = refined_direction(s.pq[h], s.pq[prev], s.pq[prev + 1], s.pq[next], s.pq[next + 1]);
            const T* v = s.rgb(ch).data();
            const T* g = s.rgb1.data();
            T nwg = diagonal_gradient(v, g, i, -1, -1), neg = diagonal_gradient(v, g, i, 1, -1);
            T swg = diagonal_gradient(v, g, i, -1, 1), seg = diagonal_gradient(v, g, i, 1, 1);
            T nwe = v[i - TS - 1] - g[i - TS - 1], nee = v[i - TS + 1] - g[i - TS + 1];
            T swe = v[i + TS - 1] - g[i + TS - 1], see = v[i + TS + 1] - g[i + TS + 1];
            T dp = weighted_pair(nwg, see, seg, nwe);
            T dq = weighted_pair(neg, swe, swg, nee);
            s.rgb(ch)[i] = g[i] + interpolate(d, dq, dp);
        }
    }
}

template <class T> void at_green(Scratch<T>& s, int pattern) {
    const T EPS = eps<T>();
    for (int row = 4; row < TS - 4; ++row) {
        int start = 4 + (color_at(pattern, 1, row) & 1);
        for (int col = start; col < TS - 4; col += 2) {
            size_t i = row * TS + col;
            const T* g = s.rgb1.data();
            T d = refined_direction(s.vh[i], s.vh[i - TS - 1], s.vh[i - TS + 1], s.vh[i + TS - 1],
                                    s.vh[i + TS + 1]);
            T green = g[i];
            T n1 = EPS + std::fabs(green - g[i - 2 * TS]);
            T s1 = EPS + std::fabs(green - g[i + 2 * TS]);
            T w1 = EPS + std::fabs(green - g[i - 2]);
            T e1 = EPS + std::fabs(green - g[i + 2]);
            T gn = g[i - TS], gs = g[i + TS], gw = g[i - 1], ge = g[i + 1];
            for (int ch = 0; ch <= 2; ch += 2) {
                T* v = s.rgb(ch).data();
                T ns = std::fabs(v[i - TS] - v[i + TS]);
                T ew = std::fabs(v[i - 1] - v[i + 1]);
                T ng = n1 + ns + std::fabs(v[i - TS] - v[i - 3 * TS]);
                T sg = s1 + ns + std::fabs(v[i + TS] - v[i + 3 * TS]);
                T wg = w1 + ew + std::fabs(v[i - 1] - v[i - 3]);
                T eg = e1 + ew + std::fabs(v[i + 1] - v[i + 3]);
                T ne = v[i - TS] - gn, se = v[i + TS] - gs, we = v[i - 1] - gw, ee = v[i + 1] - ge;
                T vert = weighted_pair(ng, se, sg, ne);
                T horiz = weighted_pair(eg, we, wg, ee);
                v[i] = green + interpolate(d, horiz, vert);
            }
        }
    }
}

// Stage probes of crates/demosaic/src/rcd.rs:355-466: values after each
// stage of the first tile at tile indices `red` and `green`.
template <class T> void probe_first_tile(const Mosaic<T>& m, size_t red, size_t green, T out[9]) {
    Scratch<T> s;
    s.clear();
    int pattern = m.pattern;
    for (int row = 0; row < TS; ++row) {
        size_t sy = mirror_coordinate((ptrdiff_t)row - BORDER, m.height);
        int c0 = color_at(pattern, 0, row), c1 = color_at(pattern, 1, row);
        for (int col = 0; col < TS; ++col) {
            T v = m.at(mirror_coordinate((ptrdiff_t)col - BORDER, m.width), sy);
            s.cfa[row * TS + col] = v;
            s.rgb(c0)[row * TS + col] = v;
            s.rgb(c1)[row * TS + col] = v;
        }
    }
    find_directions(s);
    out[0] = s.vh[red];
    low_pass(s, pattern);
    out[1] = s.pq[red / 2];
    interpolate_green(s, pattern);
    out[2] = s.rgb1[red];
    diagonal_high_pass(s);
    out[3] = s.p[green / 2];
    out[4] = s.q[green / 2];
    diagonal_directions(s, pattern);
    out[5] = s.pq[red / 2];
    opposite_color(s, pattern);
    out[6] = s.rgb2[red];
    at_green(s, pattern);
    out[7] = s.rgb0[green];
    out[8] = s.rgb2[green];
}

}  // namespace rcd

// rcd.rs:36-113 with populate (:210-257) and write_tile (:261-299).
template <class T>
int rcd_reconstruct(const Mosaic<T>& m, const T gains[3], T* out, int threads) {
    using namespace rcd;
    if (m.width <= (size_t)MIN_DIMENSION || m.height <= (size_t)MIN_DIMENSION)
        return bilinear_reconstruct(m, gains, out, threads);
    size_t cols = (m.width + STEP - 1) / STEP, rows = (m.height + STEP - 1) / STEP;
    std::atomic<int> status{OK};
    std::atomic<size_t> next{0};
    auto work = [&]() {
        rcd::Scratch<T> s;
        for (size_t j = next.fetch_add(1); j < cols * rows; j = next.fetch_add(1)) {
            size_t ox = (j % cols) * STEP, oy = (j / cols) * STEP;
            int pattern = shifted(m.pattern, ox, oy);
            s.clear();
            ptrdiff_t wx = (ptrdiff_t)ox - BORDER, wy = (ptrdiff_t)oy - BORDER;
            for (int row = 0; row < TS; ++row) {
                size_t sy = mirror_coordinate(wy + row, m.height);
                int c0 = color_at(pattern, 0, row), c1 = color_at(pattern, 1, row);
                for (int col = 0; col < TS; ++col) {
                    T v = m.at(mirror_coordinate(wx + col, m.width), sy);
                    s.cfa[row * TS + col] = v;
                    s.rgb(c0)[row * TS + col] = v;
                    s.rgb(c1)[row * TS + col] = v;
                }
            }
            find_directions(s);
            low_pass(s, pattern);
            interpolate_green(s, pattern);
            diagonal_high_pass(s);
            diagonal_directions(s, pattern);
            opposite_color(s, pattern);
            at_green(s, pattern);
            size_t x1 = std::min(ox + STEP, m.width), y1 = std::min(oy + STEP, m.height);
            for (size_t gy = oy; gy < y1; ++gy)
                for (size_t gx = ox; gx < x1; ++gx) {
                    size_t idx = (gy - oy + BORDER) * TS + (gx - ox + BORDER);
                    T rgb[3] = {s.rgb0[idx], s.rgb1[idx], s.rgb2[idx]};
                    rgb[color_at(m.pattern, gx, gy)] = m.at(gx, gy);
                    apply_gains(rgb, gains);
                    if (!all_finite(rgb)) status = ERR_NONFINITE;
                    T* p = out + (gy * m.width + gx) * 3;
                    p[0] = rgb[0]; p[1] = rgb[1]; p[2] = rgb[2];
                }
        }
    };
    int workers = threads < 1 ? 1 : threads;
    if (workers == 1) {
        work();
    } else {
        std::vector<std::thread> th;
        for (int w = 0; w < workers; ++w) th.emplace_back(work);
        for (auto& t : th) t.join();
    }
    return status;
}

}  // namespace rhdo
