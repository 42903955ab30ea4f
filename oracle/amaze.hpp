// Oracle (test infrastructure only): restatement of the reference's tiled
// AMaZE demosaic, crates/demosaic/src/amaze/{mod,math,scratch,green,nyquist,
// diagonal,chroma}.rs and crates/demosaic/src/amaze_geometry.rs. Every stage
// keeps the reference's loop bounds, plane addressing (`index >> 1` half
// planes) and operation order so the f32 instantiation is bit-identical; the
// f64 instantiation is the FP64 twin.
#pragma once

#include <cstring>
#include <vector>

#include "common.hpp"

namespace rhdo {
namespace amaze {

constexpr int TS = 256;          // amaze_geometry.rs:15
constexpr int BORDER = 16;       // amaze_geometry.rs:16
constexpr int VALID = TS - 2 * BORDER;
constexpr int TSH = TS / 2;
constexpr int FULL = TS * TS;
constexpr int HALF = TS * TSH;
constexpr ptrdiff_t V1 = TS, V2 = 2 * TS, P1 = -TS + 1, P2 = -2 * TS + 2, M1 = TS + 1,
                    M2 = 2 * TS + 2;  // math.rs:15-20

template <class T> struct K {
    static T eps() { return RHDO_LIT(T, 1.0e-5); }
    static T epssq() { return RHDO_LIT(T, 1.0e-10); }
    static T arthresh() { return RHDO_LIT(T, 0.75); }
    // math.rs:28-53. GAUSSGRAD is NYQTHRESH (0.5) times the literal; the
    // product is exact, so it is computed the same way here.
    static T gaussodd(int k) {
        static const T v[4] = {RHDO_LIT(T, 0.14659727707323927), RHDO_LIT(T, 0.103592713382435),
                               RHDO_LIT(T, 0.0732036125103057), RHDO_LIT(T, 0.0365543548389495)};
        return v[k];
    }
    static T gaussgrad(int k) {
        static const T v[6] = {
            RHDO_LIT(T, 0.5) * RHDO_LIT(T, 0.07384411893421103),
            RHDO_LIT(T, 0.5) * RHDO_LIT(T, 0.06207511968171489),
            RHDO_LIT(T, 0.5) * RHDO_LIT(T, 0.0521818194747806),
            RHDO_LIT(T, 0.5) * RHDO_LIT(T, 0.03687419286733595),
            RHDO_LIT(T, 0.5) * RHDO_LIT(T, 0.03099732204057846),
            RHDO_LIT(T, 0.5) * RHDO_LIT(T, 0.018413194161458882)};
        return v[k];
    }
    static T gausseven(int k) {
        static const T v[2] = {RHDO_LIT(T, 0.13719494435797422), RHDO_LIT(T, 0.05640252782101291)};
        return v[k];
    }
    static T gquinc(int k) {
        static const T v[4] = {RHDO_LIT(T, 0.169917), RHDO_LIT(T, 0.108947), RHDO_LIT(T, 0.069855),
                               RHDO_LIT(T, 0.0287182)};
        return v[k];
    }
};

template <class T> inline T sqr(T x) { return x * x; }
template <class T> inline T median3(T a, T b, T c) {
    return sel_max(sel_min(a, b), sel_min(sel_max(b, a), c));
}
template <class T> inline T intp(T a, T b, T c) { return a * (b - c) + c; }

inline int fc(int pattern, int row, int col) { return color_at(pattern, (size_t)col, (size_t)row); }
inline int green_at_even(int pattern, int row) { return fc(pattern, row, 2) & 1; }

// scratch.rs: one plane per quantity, zeroed at allocation.
template <class T> struct Scratch {
    std::vector<T> cfa, green, delhvsqsum, dirwts0, dirwts1, vcd, hcd, vcdalt, hcdalt, cddiffsq,
        dgintv, dginth;
    std::vector<T> hvwt, dgrb0, dgrb1, dgrb2_h, dgrb2_v, delp, delm, sq1m, sq1p, nyqutest, pmwt,
        rbm, rbp, rbint;
    std::vector<uint8_t> nyquist, nyquist2;
    std::vector<T> red, blue;
    Scratch()
        : cfa(FULL), green(FULL), delhvsqsum(FULL), dirwts0(FULL), dirwts1(FULL), vcd(FULL),
          hcd(FULL), vcdalt(FULL), hcdalt(FULL), cddiffsq(FULL), dgintv(FULL), dginth(FULL),
          hvwt(HALF), dgrb0(HALF), dgrb1(HALF), dgrb2_h(HALF), dgrb2_v(HALF), delp(HALF),
          delm(HALF), sq1m(HALF), sq1p(HALF), nyqutest(HALF), pmwt(HALF), rbm(HALF), rbp(HALF),
          rbint(HALF), nyquist(HALF), nyquist2(HALF), red(FULL), blue(FULL) {}
    // scratch.rs:113-145. `poison` fills every plane the stages must write
    // before reading with NaN, so a read of stale data shows in the output.
    void clear(bool poison) {
        std::fill(nyquist.begin(), nyquist.end(), 0);
        std::fill(nyquist2.begin(), nyquist2.end(), 0);
        std::fill(dgrb2_h.begin(), dgrb2_h.end(), T(0));
        std::fill(dgrb2_v.begin(), dgrb2_v.end(), T(0));
        if (poison) {
            const T nan = T(NAN);
            for (auto* p : {&dgrb0, &dgrb1, &delhvsqsum, &dirwts0, &dirwts1, &cddiffsq, &dgintv,
                            &dginth, &hvwt, &delp, &delm, &sq1m, &sq1p, &nyqutest, &pmwt, &rbm,
                            &rbp, &rbint, &red, &blue})
                std::fill(p->begin(), p->end(), nan);
        }
    }
};

// green.rs:13-31
template <class T> void gradients(Scratch<T>& s) {
    const T EPS = K<T>::eps();
    for (int rr = 2; rr < TS - 2; ++rr) {
        const T* c0 = &s.cfa[(rr - 2) * TS];
        const T* c1 = &s.cfa[(rr - 1) * TS];
        const T* c2 = &s.cfa[rr * TS];
        const T* c3 = &s.cfa[(rr + 1) * TS];
        const T* c4 = &s.cfa[(rr + 2) * TS];
        for (int cc = 2; cc < TS - 2; ++cc) {
            T delh = std::fabs(c2[cc + 1] - c2[cc - 1]);
            T delv = std::fabs(c3[cc] - c1[cc]);
            s.dirwts0[rr * TS + cc] =
                EPS + std::fabs(c4[cc] - c2[cc]) + std::fabs(c2[cc] - c0[cc]) + delv;
            s.dirwts1[rr * TS + cc] =
                EPS + std::fabs(c2[cc + 2] - c2[cc]) + std::fabs(c2[cc] - c2[cc - 2]) + delh;
            s.delhvsqsum[rr * TS + cc] = sqr(delh) + sqr(delv);
        }
    }
}

template <class T> inline T adaptive_or(T ratio, T centre, T fallback) {
    return std::fabs(T(1) - ratio) < K<T>::arthresh() ? centre * ratio : fallback;
}

// green.rs:46-137
template <class T> void colour_differences(Scratch<T>& s, int pattern, T clip_pt8) {
    const T EPS = K<T>::eps();
    const T half = RHDO_LIT(T, 0.5);
    for (int rr = 4; rr < TS - 4; ++rr) {
        const T* cfa[5];
        const T* w0[5];
        for (int k = 0; k < 5; ++k) {
            cfa[k] = &s.cfa[(rr + k - 2) * TS];
            w0[k] = &s.dirwts0[(rr + k - 2) * TS];
        }
        const T* w1 = &s.dirwts1[rr * TS];
        bool green_first = (fc(pattern, rr, 4) & 1) == 1;
        for (int cc = 4; cc < TS - 4; ++cc) {
            bool fcswitch = green_first == ((cc & 1) == 0);
            T cur = cfa[2][cc];
            T up = cfa[1][cc], up2 = cfa[0][cc], down = cfa[3][cc], down2 = cfa[4][cc];
            T left = cfa[2][cc - 1], left2 = cfa[2][cc - 2], right = cfa[2][cc + 1],
              right2 = cfa[2][cc + 2];
            T cru = up * (w0[0][cc] + w0[2][cc]) /
                    (w0[0][cc] * (EPS + cur) + w0[2][cc] * (EPS + up2));
            T crd = down * (w0[4][cc] + w0[2][cc]) /
                    (w0[4][cc] * (EPS + cur) + w0[2][cc] * (EPS + down2));
            T crl = left * (w1[cc - 2] + w1[cc]) / (w1[cc - 2] * (EPS + cur) + w1[cc] * (EPS + left2));
            T crr = right * (w1[cc + 2] + w1[cc]) /
                    (w1[cc + 2] * (EPS + cur) + w1[cc] * (EPS + right2));
            T guha = up + half * (cur - up2);
            T gdha = down + half * (cur - down2);
            T glha = left + half * (cur - left2);
            T grha = right + half * (cur - right2);
            T guar = adaptive_or(cru, cur, guha);
            T gdar = adaptive_or(crd, cur, gdha);
            T glar = adaptive_or(crl, cur, glha);
            T grar = adaptive_or(crr, cur, grha);
            T hwt = w1[cc - 1] / (w1[cc - 1] + w1[cc + 1]);
            T vwt = w0[1][cc] / (w0[3][cc] + w0[1][cc]);
            T gintvha = vwt * gdha + (T(1) - vwt) * guha;
            T ginthha = hwt * grha + (T(1) - hwt) * glha;
            T vcd, hcd, vcdalt, hcdalt;
            if (fcswitch) {
                vcd = cur - (vwt * gdar + (T(1) - vwt) * guar);
                hcd = cur - (hwt * grar + (T(1) - hwt) * glar);
                vcdalt = cur - gintvha;
                hcdalt = cur - ginthha;
            } else {
                vcd = (vwt * gdar + (T(1) - vwt) * guar) - cur;
                hcd = (hwt * grar + (T(1) - hwt) * glar) - cur;
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
            int i = rr * TS + cc;
            s.dgintv[i] = sel_min(sqr(guha - gdha), sqr(guar - gdar));
            s.dginth[i] = sel_min(sqr(glha - grha), sqr(glar - grar));
            s.vcd[i] = vcd;
            s.hcd[i] = hcd;
            s.vcdalt[i] = vcdalt;
            s.hcdalt[i] = hcdalt;
        }
    }
}

// green.rs:142-174
template <class T> inline T bound_at_green(T d, T cur, T a, T b, T clip_pt) {
    const T EPS = K<T>::eps();
    T g = -d + cur;
    T alt = -median3(g, a, b) + cur;
    T bounded = d;
    if (d > T(0)) {
        if (T(3) * d > g + cur) {
            bounded = alt;
        } else {
            T weight = T(1) - T(3) * d / (EPS + g + cur);
            bounded = weight * d + (T(1) - weight) * alt;
        }
    }
    return g > clip_pt ? alt : bounded;
}
template <class T> inline T bound_at_red_blue(T d, T cur, T a, T b, T clip_pt) {
    const T EPS = K<T>::eps();
    T g = d + cur;
    T alt = median3(g, a, b) - cur;
    T bounded = d;
    if (d < T(0)) {
        if (T(3) * d < -(g + cur)) {
            bounded = alt;
        } else {
            T weight = T(1) + T(3) * d / (EPS + g + cur);
            bounded = weight * d + (T(1) - weight) * alt;
        }
    }
    return g > clip_pt ? alt : bounded;
}
template <class T> inline T spread(T a, T b, T c) {
    return T(3) * (sqr(a) + sqr(b) + sqr(c)) - sqr(a + b + c);
}

// green.rs:184-226: in place; `vcd` two rows up and `hcd` two columns left
// are already bounded when read.
template <class T> void bound_differences(Scratch<T>& s, int pattern, T clip_pt) {
    for (int rr = 4; rr < TS - 4; ++rr) {
        T* vcd = &s.vcd[rr * TS];
        const T* vcd_up = &s.vcd[(rr - 2) * TS];
        const T* vcd_down = &s.vcd[(rr + 2) * TS];
        const T* alt_up = &s.vcdalt[(rr - 2) * TS];
        const T* alt_down = &s.vcdalt[(rr + 2) * TS];
        const T* vcdalt = &s.vcdalt[rr * TS];
        const T* hcdalt = &s.hcdalt[rr * TS];
        T* hcd = &s.hcd[rr * TS];
        const T* cfa = &s.cfa[rr * TS];
        const T* cfa_up = &s.cfa[(rr - 1) * TS];
        const T* cfa_down = &s.cfa[(rr + 1) * TS];
        T* cddiffsq = &s.cddiffsq[rr * TS];
        bool green = (fc(pattern, rr, 4) & 1) == 1;
        for (int cc = 4; cc < TS - 4; ++cc) {
            T hcdvar = spread(hcd[cc - 2], hcd[cc], hcd[cc + 2]);
            T hcdaltvar = spread(hcdalt[cc - 2], hcdalt[cc], hcdalt[cc + 2]);
            T vcdvar = spread(vcd_up[cc], vcd[cc], vcd_down[cc]);
            T vcdaltvar = spread(alt_up[cc], vcdalt[cc], alt_down[cc]);
            T h = hcdaltvar < hcdvar ? hcdalt[cc] : hcd[cc];
            T v = vcdaltvar < vcdvar ? vcdalt[cc] : vcd[cc];
            T cur = cfa[cc];
            if (green) {
                hcd[cc] = bound_at_green(h, cur, cfa[cc - 1], cfa[cc + 1], clip_pt);
                vcd[cc] = bound_at_green(v, cur, cfa_up[cc], cfa_down[cc], clip_pt);
            } else {
                hcd[cc] = bound_at_red_blue(h, cur, cfa[cc - 1], cfa[cc + 1], clip_pt);
                vcd[cc] = bound_at_red_blue(v, cur, cfa_up[cc], cfa_down[cc], clip_pt);
                cddiffsq[cc] = sqr(vcd[cc] - hcd[cc]);
            }
            green = !green;
        }
    }
}

// green.rs:229-292
template <class T> void direction_weights(Scratch<T>& s, int pattern) {
    const T EPSSQ = K<T>::epssq();
    const T half = RHDO_LIT(T, 0.5);
    for (int rr = 6; rr < TS - 6; ++rr) {
        const T* vcd[7];
        const T* dgintv[5];
        for (int k = 0; k < 7; ++k) vcd[k] = &s.vcd[(rr + k - 3) * TS];
        for (int k = 0; k < 5; ++k) dgintv[k] = &s.dgintv[(rr + k - 2) * TS];
        const T* hcd = &s.hcd[rr * TS];
        const T* dginth = &s.dginth[rr * TS];
        const T* w0_up = &s.dirwts0[(rr - 1) * TS];
        const T* w0_down = &s.dirwts0[(rr + 1) * TS];
        const T* w1 = &s.dirwts1[rr * TS];
        T* hvwt = &s.hvwt[rr * TSH];
        for (int cc = 6 + green_at_even(pattern, rr); cc < TS - 6; cc += 2) {
            T uave = vcd[3][cc] + vcd[2][cc] + vcd[1][cc] + vcd[0][cc];
            T dave = vcd[3][cc] + vcd[4][cc] + vcd[5][cc] + vcd[6][cc];
            T lave = hcd[cc] + hcd[cc - 1] + hcd[cc - 2] + hcd[cc - 3];
            T rave = hcd[cc] + hcd[cc + 1] + hcd[cc + 2] + hcd[cc + 3];
            T dgrbvvaru = sqr(vcd[3][cc] - uave) + sqr(vcd[2][cc] - uave) +
                          sqr(vcd[1][cc] - uave) + sqr(vcd[0][cc] - uave);
            T dgrbvvard = sqr(vcd[3][cc] - dave) + sqr(vcd[4][cc] - dave) +
                          sqr(vcd[5][cc] - dave) + sqr(vcd[6][cc] - dave);
            T dgrbhvarl = sqr(hcd[cc] - lave) + sqr(hcd[cc - 1] - lave) + sqr(hcd[cc - 2] - lave) +
                          sqr(hcd[cc - 3] - lave);
            T dgrbhvarr = sqr(hcd[cc] - rave) + sqr(hcd[cc + 1] - rave) + sqr(hcd[cc + 2] - rave) +
                          sqr(hcd[cc + 3] - rave);
            T hwt = w1[cc - 1] / (w1[cc - 1] + w1[cc + 1]);
            T vwt = w0_up[cc] / (w0_down[cc] + w0_up[cc]);
            T vcdvar = EPSSQ + vwt * dgrbvvard + (T(1) - vwt) * dgrbvvaru;
            T hcdvar = EPSSQ + hwt * dgrbhvarr + (T(1) - hwt) * dgrbhvarl;
            T fluct_u = dgintv[2][cc] + dgintv[1][cc] + dgintv[0][cc];
            T fluct_d = dgintv[2][cc] + dgintv[3][cc] + dgintv[4][cc];
            T fluct_l = dginth[cc] + dginth[cc - 1] + dginth[cc - 2];
            T fluct_r = dginth[cc] + dginth[cc + 1] + dginth[cc + 2];
            T vcdvar1 = EPSSQ + vwt * fluct_d + (T(1) - vwt) * fluct_u;
            T hcdvar1 = EPSSQ + hwt * fluct_r + (T(1) - hwt) * fluct_l;
            T varwt = hcdvar / (vcdvar + hcdvar);
            T diffwt = hcdvar1 / (vcdvar1 + hcdvar1);
            hvwt[cc >> 1] = ((half - varwt) * (half - diffwt) > T(0) &&
                             std::fabs(half - diffwt) < std::fabs(half - varwt))
                                ? varwt
                                : diffwt;
        }
    }
}

struct NyquistArea {
    int start_row, end_row, start_col, end_col;
};

// nyquist.rs:25-89. Returns false when the tile has no area to refine.
template <class T> bool nyquist_detect(Scratch<T>& s, int pattern, NyquistArea& area) {
    for (int rr = 6; rr < TS - 6; ++rr) {
        const T* c[5];
        const T* d[5];
        for (int k = 0; k < 5; ++k) {
            c[k] = &s.cddiffsq[(rr + k - 2) * TS];
            d[k] = &s.delhvsqsum[(rr + k - 2) * TS];
        }
        T* out = &s.nyqutest[rr * TSH];
        for (int cc = 6 + green_at_even(pattern, rr); cc < TS - 6; cc += 2) {
            out[cc >> 1] =
                (K<T>::gaussodd(0) * c[2][cc] +
                 K<T>::gaussodd(1) * (c[1][cc - 1] + c[1][cc + 1] + c[3][cc - 1] + c[3][cc + 1]) +
                 K<T>::gaussodd(2) * (c[0][cc] + c[2][cc - 2] + c[2][cc + 2] + c[4][cc]) +
                 K<T>::gaussodd(3) * (c[0][cc - 2] + c[0][cc + 2] + c[4][cc - 2] + c[4][cc + 2])) -
                (K<T>::gaussgrad(0) * d[2][cc] +
                 K<T>::gaussgrad(1) * (d[1][cc] + d[2][cc + 1] + d[2][cc - 1] + d[3][cc]) +
                 K<T>::gaussgrad(2) * (d[1][cc - 1] + d[1][cc + 1] + d[3][cc - 1] + d[3][cc + 1]) +
                 K<T>::gaussgrad(3) * (d[0][cc] + d[2][cc - 2] + d[2][cc + 2] + d[4][cc]) +
                 K<T>::gaussgrad(4) * (d[0][cc - 1] + d[0][cc + 1] + d[1][cc - 2] + d[1][cc + 2] +
                                       d[3][cc - 2] + d[3][cc + 2] + d[4][cc - 1] + d[4][cc + 1]) +
                 K<T>::gaussgrad(5) * (d[0][cc - 2] + d[0][cc + 2] + d[4][cc - 2] + d[4][cc + 2]));
        }
    }
    int start_row = 0, end_row = 0, start_col = TS + 1, end_col = 0;
    for (int rr = 6; rr < TS - 6; ++rr) {
        for (int cc = 6 + green_at_even(pattern, rr); cc < TS - 6; cc += 2) {
            int i = rr * TS + cc;
            if (s.nyqutest[i >> 1] > T(0)) {
                s.nyquist[i >> 1] = 1;
                if (start_row == 0) start_row = rr;
                end_row = rr;
                if (cc < start_col) start_col = cc;
                if (cc > end_col) end_col = cc;
            }
        }
    }
    if (start_row == end_row || start_col == end_col) return false;
    end_row += 1;
    end_col += 1;
    start_col -= start_col & 1;
    area.start_row = start_row > 8 ? start_row : 8;
    area.end_row = end_row < TS - 8 ? end_row : TS - 8;
    area.start_col = start_col > 8 ? start_col : 8;
    area.end_col = end_col < TS - 8 ? end_col : TS - 8;
    return true;
}

// nyquist.rs:100-153
template <class T> void area_interpolation(Scratch<T>& s, int pattern, const NyquistArea& a) {
    const T EPSSQ = K<T>::epssq();
    const T half = RHDO_LIT(T, 0.5);
    for (int rr = a.start_row; rr < a.end_row; ++rr) {
        for (int i = rr * TS + a.start_col + green_at_even(pattern, rr); i < rr * TS + a.end_col;
             i += 2) {
            const uint8_t* n = s.nyquist.data();
            unsigned votes = n[(i - V2) >> 1] + n[(i - M1) >> 1] + n[(i + P1) >> 1] +
                             n[(i - 2) >> 1] + n[(i + 2) >> 1] + n[(i - P1) >> 1] +
                             n[(i + M1) >> 1] + n[(i + V2) >> 1];
            s.nyquist2[i >> 1] = votes >= 5 ? 1 : (votes == 4 ? s.nyquist[i >> 1] : 0);
        }
    }
    for (int rr = a.start_row; rr < a.end_row; ++rr) {
        for (int i = rr * TS + a.start_col + green_at_even(pattern, rr); i < rr * TS + a.end_col;
             i += 2) {
            if (s.nyquist2[i >> 1] == 0) continue;
            T sumcfa = 0, sumh = 0, sumv = 0, sumsqh = 0, sumsqv = 0, areawt = 0;
            for (int di = -6; di < 7; di += 2) {
                ptrdiff_t i1 = i + di * V1 - 6;
                for (int dj = -6; dj < 7; dj += 2) {
                    if (s.nyquist2[i1 >> 1] != 0) {
                        T c = s.cfa[i1];
                        sumcfa += c;
                        sumh += s.cfa[i1 - 1] + s.cfa[i1 + 1];
                        sumv += s.cfa[i1 - V1] + s.cfa[i1 + V1];
                        sumsqh += sqr(c - s.cfa[i1 - 1]) + sqr(c - s.cfa[i1 + 1]);
                        sumsqv += sqr(c - s.cfa[i1 - V1]) + sqr(c - s.cfa[i1 + V1]);
                        areawt += T(1);
                    }
                    i1 += 2;
                }
            }
            sumh = sumcfa - half * sumh;
            sumv = sumcfa - half * sumv;
            areawt *= half;
            T hcdvar = EPSSQ + std::fabs(areawt * sumsqh - sumh * sumh);
            T vcdvar = EPSSQ + std::fabs(areawt * sumsqv - sumv * sumv);
            s.hvwt[i >> 1] = hcdvar / (vcdvar + hcdvar);
        }
    }
}

// nyquist.rs:159-201: rows in order, `hvwt` updated in place.
template <class T> void populate_green(Scratch<T>& s, int pattern) {
    const T half = RHDO_LIT(T, 0.5);
    for (int rr = 8; rr < TS - 8; ++rr) {
        T* hv = &s.hvwt[rr * TSH];
        const T* hv_up = &s.hvwt[(rr - 1) * TSH];
        const T* hv_down = &s.hvwt[(rr + 1) * TSH];
        T* green = &s.green[rr * TS];
        const T* green_up = &s.green[(rr - 1) * TS];
        const T* green_down = &s.green[(rr + 1) * TS];
        const T* cfa = &s.cfa[rr * TS];
        const T* vcd = &s.vcd[rr * TS];
        const T* hcd = &s.hcd[rr * TS];
        const uint8_t* nyq = &s.nyquist2[rr * TSH];
        T* dgrb = &s.dgrb0[rr * TSH];
        T* ch = &s.dgrb2_h[rr * TSH];
        T* cv = &s.dgrb2_v[rr * TSH];
        for (int cc = 8 + green_at_even(pattern, rr); cc < TS - 8; cc += 2) {
            int i1 = cc >> 1;
            T alt = (hv_up[(cc - 1) >> 1] + hv_up[(cc + 1) >> 1] + hv_down[(cc - 1) >> 1] +
                     hv_down[(cc + 1) >> 1]) /
                    T(4);
            if (std::fabs(half - hv[i1]) < std::fabs(half - alt)) hv[i1] = alt;
            dgrb[i1] = intp(hv[i1], vcd[cc], hcd[cc]);
            green[cc] = cfa[cc] + dgrb[i1];
            bool flagged = nyq[i1] != 0;
            ch[i1] = flagged ? sqr(green[cc] - half * (green[cc - 1] + green[cc + 1])) : T(0);
            cv[i1] = flagged ? sqr(green[cc] - half * (green_up[cc] + green_down[cc])) : T(0);
        }
    }
}

// nyquist.rs:204-236
template <class T> void refine_green(Scratch<T>& s, int pattern, const NyquistArea& a) {
    const T EPSSQ = K<T>::epssq();
    for (int rr = a.start_row; rr < a.end_row; ++rr) {
        for (int i = rr * TS + a.start_col + green_at_even(pattern, rr); i < rr * TS + a.end_col;
             i += 2) {
            if (s.nyquist2[i >> 1] == 0) continue;
            auto average = [&](const std::vector<T>& p) {
                return EPSSQ + (K<T>::gquinc(0) * p[i >> 1] +
                                K<T>::gquinc(1) * (p[(i - M1) >> 1] + p[(i + P1) >> 1] +
                                                   p[(i - P1) >> 1] + p[(i + M1) >> 1]) +
                                K<T>::gquinc(2) * (p[(i - V2) >> 1] + p[(i - 2) >> 1] +
                                                   p[(i + 2) >> 1] + p[(i + V2) >> 1]) +
                                K<T>::gquinc(3) * (p[(i - M2) >> 1] + p[(i + P2) >> 1] +
                                                   p[(i - P2) >> 1] + p[(i + M2) >> 1]));
            };
            T gvarh = average(s.dgrb2_h);
            T gvarv = average(s.dgrb2_v);
            s.dgrb0[i >> 1] = (s.hcd[i] * gvarv + s.vcd[i] * gvarh) / (gvarv + gvarh);
            s.green[i] = s.cfa[i] + s.dgrb0[i >> 1];
        }
    }
}

// diagonal.rs:15-42
template <class T> void diagonal_gradients(Scratch<T>& s, int pattern) {
    for (int rr = 6; rr < TS - 6; ++rr) {
        const T* up = &s.cfa[(rr - 1) * TS];
        const T* mid = &s.cfa[rr * TS];
        const T* down = &s.cfa[(rr + 1) * TS];
        T* delp = &s.delp[rr * TSH];
        T* delm = &s.delm[rr * TSH];
        T* sqp = &s.sq1p[rr * TSH];
        T* sqm = &s.sq1m[rr * TSH];
        if (green_at_even(pattern, rr) == 0) {
            for (int cc = 6; cc < TS - 6; cc += 2) {
                int i1 = cc >> 1;
                delp[i1] = std::fabs(up[cc + 1] - down[cc - 1]);
                delm[i1] = std::fabs(down[cc + 1] - up[cc - 1]);
                T g = mid[cc + 1];
                sqp[i1] = sqr(g - down[cc]) + sqr(g - up[cc + 2]);
                sqm[i1] = sqr(g - up[cc]) + sqr(g - down[cc + 2]);
            }
        } else {
            for (int cc = 6; cc < TS - 6; cc += 2) {
                int i1 = cc >> 1;
                T g = mid[cc];
                sqp[i1] = sqr(g - down[cc - 1]) + sqr(g - up[cc + 1]);
                sqm[i1] = sqr(g - up[cc - 1]) + sqr(g - down[cc + 1]);
                delp[i1] = std::fabs(up[cc + 2] - down[cc]);
                delm[i1] = std::fabs(down[cc + 2] - up[cc]);
            }
        }
    }
}

// diagonal.rs:46-60
template <class T> inline T even_kernel(const T* const rows[5], int cc) {
    return K<T>::gausseven(0) * (rows[1][cc >> 1] + rows[2][(cc - 1) >> 1] +
                                 rows[2][(cc + 1) >> 1] + rows[3][cc >> 1]) +
           K<T>::gausseven(1) * (rows[0][(cc - 1) >> 1] + rows[0][(cc + 1) >> 1] +
                                 rows[1][(cc - 2) >> 1] + rows[1][(cc + 2) >> 1] +
                                 rows[3][(cc - 2) >> 1] + rows[3][(cc + 2) >> 1] +
                                 rows[4][(cc - 1) >> 1] + rows[4][(cc + 1) >> 1]);
}

// diagonal.rs:66-75
template <class T> inline T diagonal_estimate(T centre, T near_, T far_) {
    T ratio = T(2) * near_ / (K<T>::eps() + centre + far_);
    return std::fabs(T(1) - ratio) < K<T>::arthresh() ? centre * ratio
                                                       : near_ + RHDO_LIT(T, 0.5) * (centre - far_);
}

// diagonal.rs:79-94 and :174-189 (bound_diagonal and bound_green are the same).
template <class T> inline T bound_towards(T estimate, T centre, T a, T b, T clip_pt) {
    T value = estimate;
    if (value < centre) {
        if (T(2) * value < centre) {
            value = median3(value, a, b);
        } else {
            T weight = T(2) * (centre - value) / (K<T>::eps() + value + centre);
            value = weight * value + (T(1) - weight) * median3(value, a, b);
        }
    }
    return value > clip_pt ? median3(value, a, b) : value;
}

// diagonal.rs:98-144
template <class T> void diagonal_estimates(Scratch<T>& s, int pattern, T clip_pt) {
    const T EPS = K<T>::eps();
    const T EPSSQ = K<T>::epssq();
    for (int rr = 8; rr < TS - 8; ++rr) {
        const T* cfa[5];
        const T *delm[5], *delp[5], *sqm[5], *sqp[5];
        for (int k = 0; k < 5; ++k) {
            cfa[k] = &s.cfa[(rr + k - 2) * TS];
            delm[k] = &s.delm[(rr + k - 2) * TSH];
            delp[k] = &s.delp[(rr + k - 2) * TSH];
            sqm[k] = &s.sq1m[(rr + k - 2) * TSH];
            sqp[k] = &s.sq1p[(rr + k - 2) * TSH];
        }
        T* pmwt = &s.pmwt[rr * TSH];
        T* rbp = &s.rbp[rr * TSH];
        T* rbm = &s.rbm[rr * TSH];
        for (int cc = 8 + green_at_even(pattern, rr); cc < TS - 8; cc += 2) {
            int i1 = cc >> 1;
            T centre = cfa[2][cc];
            T se1 = cfa[3][cc + 1], se2 = cfa[4][cc + 2];
            T nw1 = cfa[1][cc - 1], nw2 = cfa[0][cc - 2];
            T ne1 = cfa[1][cc + 1], ne2 = cfa[0][cc + 2];
            T sw1 = cfa[3][cc - 1], sw2 = cfa[4][cc - 2];
            T rbse = diagonal_estimate(centre, se1, se2);
            T rbnw = diagonal_estimate(centre, nw1, nw2);
            T rbne = diagonal_estimate(centre, ne1, ne2);
            T rbsw = diagonal_estimate(centre, sw1, sw2);
            T wtse = EPS + delm[2][i1] + delm[3][(cc + 1) >> 1] + delm[4][(cc + 2) >> 1];
            T wtnw = EPS + delm[2][i1] + delm[1][(cc - 1) >> 1] + delm[0][(cc - 2) >> 1];
            T wtne = EPS + delp[2][i1] + delp[1][(cc + 1) >> 1] + delp[0][(cc + 2) >> 1];
            T wtsw = EPS + delp[2][i1] + delp[3][(cc - 1) >> 1] + delp[4][(cc - 2) >> 1];
            T rbm_est = (wtse * rbnw + wtnw * rbse) / (wtse + wtnw);
            T rbp_est = (wtne * rbsw + wtsw * rbne) / (wtne + wtsw);
            T rbvarm = EPSSQ + even_kernel(sqm, cc);
            pmwt[i1] = rbvarm / ((EPSSQ + even_kernel(sqp, cc)) + rbvarm);
            rbp[i1] = bound_towards(rbp_est, centre, sw1, ne1, clip_pt);
            rbm[i1] = bound_towards(rbm_est, centre, nw1, se1, clip_pt);
        }
    }
}

// diagonal.rs:149-168: rows in order, `pmwt` updated in place.
template <class T> void combine_diagonals(Scratch<T>& s, int pattern) {
    const T half = RHDO_LIT(T, 0.5);
    for (int rr = 10; rr < TS - 10; ++rr) {
        T* pm = &s.pmwt[rr * TSH];
        const T* pm_up = &s.pmwt[(rr - 1) * TSH];
        const T* pm_down = &s.pmwt[(rr + 1) * TSH];
        const T* cfa = &s.cfa[rr * TS];
        const T* rbm = &s.rbm[rr * TSH];
        const T* rbp = &s.rbp[rr * TSH];
        T* rbint = &s.rbint[rr * TSH];
        for (int cc = 10 + green_at_even(pattern, rr); cc < TS - 10; cc += 2) {
            int i1 = cc >> 1;
            T alt = (pm_up[(cc - 1) >> 1] + pm_up[(cc + 1) >> 1] + pm_down[(cc - 1) >> 1] +
                     pm_down[(cc + 1) >> 1]) /
                    T(4);
            if (std::fabs(half - pm[i1]) < std::fabs(half - alt)) pm[i1] = alt;
            rbint[i1] = half * (cfa[cc] + rbm[i1] * (T(1) - pm[i1]) + rbp[i1] * pm[i1]);
        }
    }
}

// diagonal.rs:174-182
template <class T> inline T green_from_rb(T sample, T centre, T other) {
    T ratio = sample * T(2) / (K<T>::eps() + centre + other);
    return std::fabs(T(1) - ratio) < K<T>::arthresh() ? centre * ratio
                                                       : sample + RHDO_LIT(T, 0.5) * (centre - other);
}

// diagonal.rs:193-247
template <class T> void refine_green_with_rb(Scratch<T>& s, int pattern, T clip_pt) {
    const T half = RHDO_LIT(T, 0.5);
    for (int rr = 12; rr < TS - 12; ++rr) {
        const T* cfa_up = &s.cfa[(rr - 1) * TS];
        const T* cfa = &s.cfa[rr * TS];
        const T* cfa_down = &s.cfa[(rr + 1) * TS];
        const T* w0_up = &s.dirwts0[(rr - 1) * TS];
        const T* w0_down = &s.dirwts0[(rr + 1) * TS];
        const T* w1 = &s.dirwts1[rr * TS];
        const T* rb_up = &s.rbint[(rr - 2) * TSH];
        const T* rb = &s.rbint[rr * TSH];
        const T* rb_down = &s.rbint[(rr + 2) * TSH];
        const T* pm = &s.pmwt[rr * TSH];
        const T* hv = &s.hvwt[rr * TSH];
        T* green = &s.green[rr * TS];
        T* dgrb = &s.dgrb0[rr * TSH];
        for (int cc = 12 + green_at_even(pattern, rr); cc < TS - 12; cc += 2) {
            int i1 = cc >> 1;
            if (std::fabs(half - pm[i1]) < std::fabs(half - hv[i1])) continue;
            T centre = rb[i1];
            T gu = green_from_rb(cfa_up[cc], centre, rb_up[i1]);
            T gd = green_from_rb(cfa_down[cc], centre, rb_down[i1]);
            T gl = green_from_rb(cfa[cc - 1], centre, rb[i1 - 1]);
            T gr = green_from_rb(cfa[cc + 1], centre, rb[i1 + 1]);
            T gintv = (w0_up[cc] * gd + w0_down[cc] * gu) / (w0_down[cc] + w0_up[cc]);
            T ginth = (w1[cc - 1] * gr + w1[cc + 1] * gl) / (w1[cc - 1] + w1[cc + 1]);
            gintv = bound_towards(gintv, centre, cfa_up[cc], cfa_down[cc], clip_pt);
            ginth = bound_towards(ginth, centre, cfa[cc - 1], cfa[cc + 1], clip_pt);
            green[cc] = ginth * (T(1) - hv[i1]) + gintv * hv[i1];
            dgrb[i1] = green[cc] - cfa[cc];
        }
    }
}

// chroma.rs:11-19
inline void red_offset(int pattern, int& ey, int& ex) {
    int c00 = color_at(pattern, 0, 0), c10 = color_at(pattern, 1, 0);
    if (c00 == 1 && c10 == 0) { ey = 0; ex = 1; }
    else if (c00 == 1) { ey = 1; ex = 0; }
    else if (c00 == 0) { ey = 0; ex = 0; }
    else { ey = 1; ex = 1; }
}

// chroma.rs:23-97
template <class T> void interpolate_chroma(Scratch<T>& s, int pattern) {
    const T EPS = K<T>::eps();
    int ey, ex;
    red_offset(pattern, ey, ex);
    for (int rr = 13 - ey; rr < TS - 12; rr += 2) {
        for (int i1 = (rr * TS + 13 - ex) >> 1; i1 < (rr * TS + TS - 12) >> 1; ++i1) {
            s.dgrb1[i1] = s.dgrb0[i1];
            s.dgrb0[i1] = T(0);
        }
    }
    const T c1325 = RHDO_LIT(T, 1.325), c175 = RHDO_LIT(T, 0.175), c075 = RHDO_LIT(T, 0.075);
    for (int rr = 14; rr < TS - 14; ++rr) {
        int start = 14 + green_at_even(pattern, rr);
        int plane = 1 - fc(pattern, rr, start) / 2;
        std::vector<T>& p = plane ? s.dgrb1 : s.dgrb0;
        auto d = [&](int dr, int cc, int dc) { return p[(rr + dr) * TSH + ((cc + dc) >> 1)]; };
        for (int cc = start; cc < TS - 14; cc += 2) {
            T wtnw = T(1) / (EPS + std::fabs(d(-1, cc, -1) - d(1, cc, 1)) +
                             std::fabs(d(-1, cc, -1) - d(-3, cc, -3)) +
                             std::fabs(d(1, cc, 1) - d(-3, cc, -3)));
            T wtne = T(1) / (EPS + std::fabs(d(-1, cc, 1) - d(1, cc, -1)) +
                             std::fabs(d(-1, cc, 1) - d(-3, cc, 3)) +
                             std::fabs(d(1, cc, -1) - d(-3, cc, 3)));
            T wtsw = T(1) / (EPS + std::fabs(d(1, cc, -1) - d(-1, cc, 1)) +
                             std::fabs(d(1, cc, -1) - d(3, cc, 3)) +
                             std::fabs(d(-1, cc, 1) - d(3, cc, -3)));
            T wtse = T(1) / (EPS + std::fabs(d(1, cc, 1) - d(-1, cc, -1)) +
                             std::fabs(d(1, cc, 1) - d(3, cc, -3)) +
                             std::fabs(d(-1, cc, -1) - d(3, cc, 3)));
            T value = (wtnw * (c1325 * d(-1, cc, -1) - c175 * d(-3, cc, -3) -
                               c075 * d(-1, cc, -3) - c075 * d(-3, cc, -1)) +
                       wtne * (c1325 * d(-1, cc, 1) - c175 * d(-3, cc, 3) - c075 * d(-1, cc, 3) -
                               c075 * d(1, cc, 1)) +
                       wtsw * (c1325 * d(1, cc, -1) - c175 * d(3, cc, -3) - c075 * d(1, cc, -3) -
                               c075 * d(-1, cc, -1)) +
                       wtse * (c1325 * d(1, cc, 1) - c175 * d(3, cc, 3) - c075 * d(1, cc, 3) -
                               c075 * d(3, cc, 1))) /
                      (wtnw + wtne + wtsw + wtse);
            p[rr * TSH + (cc >> 1)] = value;
        }
    }
}

// chroma.rs:103-147
template <class T> void finish_planes(Scratch<T>& s, int pattern) {
    for (int rr = 16; rr < TS - 16; ++rr) {
        const T* green = &s.green[rr * TS];
        const T* hv_up = &s.hvwt[(rr - 1) * TSH];
        const T* hv = &s.hvwt[rr * TSH];
        const T* hv_down = &s.hvwt[(rr + 1) * TSH];
        const T* gr_up = &s.dgrb0[(rr - 1) * TSH];
        const T* gr = &s.dgrb0[rr * TSH];
        const T* gr_down = &s.dgrb0[(rr + 1) * TSH];
        const T* gb_up = &s.dgrb1[(rr - 1) * TSH];
        const T* gb = &s.dgrb1[rr * TSH];
        const T* gb_down = &s.dgrb1[(rr + 1) * TSH];
        T* red = &s.red[rr * TS];
        T* blue = &s.blue[rr * TS];
        auto at_green = [&](int cc, T& r, T& b) {
            T up = hv_up[cc >> 1], down = hv_down[cc >> 1];
            T left = hv[(cc - 1) >> 1], right = hv[(cc + 1) >> 1];
            T temp = T(1) / (up + T(2) - right - left + down);
            auto blend = [&](const T* du, const T* d, const T* dd) {
                return (up * du[cc >> 1] + (T(1) - right) * d[(cc + 1) >> 1] +
                        (T(1) - left) * d[(cc - 1) >> 1] + down * dd[cc >> 1]) *
                       temp;
            };
            r = green[cc] - blend(gr_up, gr, gr_down);
            b = green[cc] - blend(gb_up, gb, gb_down);
        };
        auto at_rb = [&](int cc, T& r, T& b) {
            r = green[cc] - gr[cc >> 1];
            b = green[cc] - gb[cc >> 1];
        };
        bool green_first = fc(pattern, rr, 16) == 1;
        for (int cc = 16; cc < TS - 16; cc += 2) {
            T r0, b0, r1, b1;
            if (green_first) {
                at_green(cc, r0, b0);
                at_rb(cc + 1, r1, b1);
            } else {
                at_rb(cc, r0, b0);
                at_green(cc + 1, r1, b1);
            }
            red[cc] = sel_max(r0, T(0));
            blue[cc] = sel_max(b0, T(0));
            red[cc + 1] = sel_max(r1, T(0));
            blue[cc + 1] = sel_max(b1, T(0));
        }
    }
}

// mod.rs:136-161: the stage sequence of one populated tile.
template <class T> void reconstruct_tile(Scratch<T>& s, int pattern, T clip_point) {
    T clip_pt8 = RHDO_LIT(T, 0.8) * clip_point;
    gradients(s);
    colour_differences(s, pattern, clip_pt8);
    bound_differences(s, pattern, clip_point);
    direction_weights(s, pattern);
    NyquistArea area;
    bool has_area = nyquist_detect(s, pattern, area);
    if (has_area) area_interpolation(s, pattern, area);
    populate_green(s, pattern);
    if (has_area) refine_green(s, pattern, area);
    diagonal_gradients(s, pattern);
    diagonal_estimates(s, pattern, clip_point);
    combine_diagonals(s, pattern);
    refine_green_with_rb(s, pattern, clip_point);
    interpolate_chroma(s, pattern);
    finish_planes(s, pattern);
}

}  // namespace amaze
}  // namespace rhdo
