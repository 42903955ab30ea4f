// Oracle (test infrastructure only): region-aware Opposed highlight
// reconstruction restated from crates/highlight/src/opposed.rs:63-364,
// cells.rs:41-113, regions.rs:34-271, region_correction.rs:53-300 and
// support.rs:26-125. f32 arithmetic with the reference's f64 accumulators;
// cbrt/log2 come from the platform libm, which is what Rust's f32 methods
// call on Linux. Templated on the value type: T = double is the FP64 twin
// (the same steps in double precision, with double libm calls).
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace rhdo {
namespace opposed {

constexpr int FIRST_RADIUS = 1, LAST_RADIUS = 3, MAX_CANDIDATES = 48, MIN_CANDIDATES = 3;
constexpr uint32_t MIN_BORDER_CELLS = 3;

struct Stats {
    uint64_t suspected, reconstructed, changed, fallback, fully_unsupported, neutralised,
        suspected_by_channel[3];
};

template <class T>
struct Cell {
    T means[3];
    uint8_t flags;
    bool usable(int c) const { return flags & (1 << c); }
    bool clean(int c) const { return !(flags & (1 << (c + 3))); }
};

inline int popcount(unsigned m) { return __builtin_popcount(m); }

// support.rs:26-33
template <class T>
inline bool bright_clean(const Cell<T>& cell, int c, const T levels[3]) {
    return cell.usable(c) && cell.clean(c) && cell.means[c] > RHDO_LIT(T, 0.2) * levels[c];
}

// support.rs:37-46
template <class T>
inline unsigned support_mask(const Cell<T>& cell, int target, const T levels[3]) {
    unsigned m = 0;
    for (int c = 0; c < 3; ++c)
        if (c != target && bright_clean(cell, c, levels)) m |= 1u << c;
    return m;
}

// support.rs:51-72
template <class T>
inline bool reference(const Cell<T>& cell, unsigned support, T& out) {
    if (popcount(support) == 1) {
        T v = cell.means[__builtin_ctz(support)];
        if (finite(v) && v > T(0)) {
            out = v;
            return true;
        }
        return false;
    }
    T sum = T(0);
    unsigned count = 0;
    for (int c = 0; c < 3; ++c) {
        if (!(support & (1u << c))) continue;
        T v = cell.means[c];
        if (!finite(v) || v <= T(0)) return false;
        sum += RHDO_MUTANT(9) ? v : std::cbrt(v);  // 9: arithmetic instead of cube-root mean
        ++count;
    }
    if (count == 0) return false;
    T root = sum / (T)count;
    out = RHDO_MUTANT(9) ? root : root * root * root;
    return true;
}

// support.rs:78-95
template <class T>
inline bool correction(const Cell<T>& cell, int target, unsigned support, const T levels[3],
                       T& out) {
    if (!bright_clean(cell, target, levels)) return false;
    for (int c = 0; c < 3; ++c)
        if ((support & (1u << c)) && !bright_clean(cell, c, levels)) return false;
    T ref;
    if (!reference(cell, support, ref)) return false;
    T v = cell.means[target] - ref;
    if (!finite(v)) return false;
    out = v;
    return true;
}

// support.rs:101-110
template <class T>
inline bool support_log_ratio(const Cell<T>& cell, unsigned support, T& out) {
    if (popcount(support) != 2) return false;
    int first = __builtin_ctz(support);
    int second = __builtin_ctz(support & ~(1u << first));
    T ratio = cell.means[first] / cell.means[second];
    if (!(finite(ratio) && ratio > T(0))) return false;
    out = std::log2(ratio);
    return true;
}

// support.rs:116-125
template <class T>
inline bool chromaticity_agrees(unsigned support, bool has_t, T t, bool has_e, T e) {
    if (popcount(support) < 2) return true;
    if (has_t && has_e) return std::fabs(t - e) <= (RHDO_MUTANT(3) ? RHDO_LIT(T, 1.5) : RHDO_LIT(T, 0.75));  // 3: gate
    return false;
}

// region_correction.rs:288-300
inline void support_patterns(int channel, unsigned out[3]) {
    int first = -1, second = -1;
    for (int c = 0; c < 3; ++c)
        if (c != channel) (first < 0 ? first : second) = c;
    out[0] = 1u << first;
    out[1] = 1u << second;
    out[2] = (1u << first) | (1u << second);
}
inline int pattern_position(int channel, unsigned support) {
    unsigned p[3];
    support_patterns(channel, p);
    for (int k = 0; k < 3; ++k)
        if (p[k] == support) return k;
    return -1;
}

// cells.rs:41-113
template <class T>
inline std::vector<Cell<T>> build_cells(const T* m, size_t w, size_t h, int pattern,
                                     const T levels[3], size_t& cw, size_t& ch) {
    cw = w / 2 + w % 2;
    ch = h / 2 + h % 2;
    std::vector<Cell<T>> cells(cw * ch);
    for (size_t cy = 0; cy < ch; ++cy)
        for (size_t cx = 0; cx < cw; ++cx) {
            T sums[3] = {0, 0, 0};
            uint32_t counts[3] = {0, 0, 0};
            uint8_t flags = 0;
            for (size_t y = cy * 2; y < std::min(cy * 2 + 2, h); ++y)
                for (size_t x = cx * 2; x < std::min(cx * 2 + 2, w); ++x) {
                    int c = color_at(pattern, x, y);
                    T s = m[y * w + x];
                    if (s >= levels[c]) {
                        flags |= 1 << (c + 3);
                    } else if (s > T(0)) {
                        sums[c] += s;
                        counts[c] += 1;
                    }
                }
            Cell<T>& cell = cells[cy * cw + cx];
            for (int c = 0; c < 3; ++c) {
                cell.means[c] = T(0);
                if (counts[c] != 0) {
                    cell.means[c] = sums[c] / (T)counts[c];
                    flags |= 1 << c;
                }
            }
            cell.flags = flags;
        }
    return cells;
}

// regions.rs:188-271 computes compact labels numbered by the first raster
// appearance of each 8-connected component. This restatement reaches the
// same numbering with a raster-order flood fill (an independent algorithm,
// so a labelling bug in either shows up as a mismatch).
inline uint32_t label_components(uint32_t* labels, size_t w, size_t h) {
    std::vector<uint32_t> out(w * h, 0);
    std::vector<size_t> stack;
    uint32_t next = 0;
    for (size_t i = 0; i < w * h; ++i) {
        if (labels[i] == 0 || out[i] != 0) continue;
        ++next;
        out[i] = next;
        stack.push_back(i);
        while (!stack.empty()) {
            size_t j = stack.back();
            stack.pop_back();
            ptrdiff_t x = (ptrdiff_t)(j % w), y = (ptrdiff_t)(j / w);
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    ptrdiff_t nx = x + dx, ny = y + dy;
                    if (nx < 0 || ny < 0 || nx >= (ptrdiff_t)w || ny >= (ptrdiff_t)h) continue;
                    size_t k = (size_t)ny * w + (size_t)nx;
                    if (labels[k] != 0 && out[k] == 0) {
                        out[k] = next;
                        stack.push_back(k);
                    }
                }
        }
    }
    std::memcpy(labels, out.data(), w * h * sizeof(uint32_t));
    return next;
}

// region_correction.rs:53-110
template <class T>
struct Slot {
    uint32_t count = 0, kept = 0;
    double accum = 0, second = 0;
    T mean = 0, tolerance = 0;
    void add(bool trim, T value, bool has_ratio, T ratio) {
        double wide = (double)value;
        if (!trim) {
            count += 1;
            accum += wide;
            second += wide * wide;
        } else if (std::fabs(value - mean) <= tolerance) {
            kept += 1;
            accum += wide;
            second += (double)(has_ratio ? ratio : T(0));
        }
    }
    void finish_moments() {
        if (count != 0) {
            double c = (double)count;
            double m = accum / c;
            double variance = num_max(second / c - m * m, 0.0);
            mean = (T)m;
            // 1: a 3-sigma trim window instead of 2 sigma
            tolerance = (T)((RHDO_MUTANT(1) ? 3.0 : 2.0) * std::sqrt(variance)) + RHDO_LIT(T, 1e-5);
        }
        accum = 0;
        second = 0;
    }
    bool value(bool has_ratio, T& offset, bool& ratio_present, T& ratio) const {
        if (count < MIN_BORDER_CELLS || kept == 0) return false;
        double k = (double)kept;
        offset = (T)(accum / k);
        ratio_present = has_ratio;
        ratio = has_ratio ? (T)(second / k) : T(0);
        return true;
    }
};

// Rust's f32::total_cmp ordering (f64::total_cmp for the f64 twin).
template <class T>
inline bool total_less(T a, T b) {
    if constexpr (sizeof(T) == 4) {
        int32_t ia, ib;
        std::memcpy(&ia, &a, 4);
        std::memcpy(&ib, &b, 4);
        ia ^= (int32_t)(((uint32_t)(ia >> 31)) >> 1);
        ib ^= (int32_t)(((uint32_t)(ib >> 31)) >> 1);
        return ia < ib;
    } else {
        int64_t ia, ib;
        std::memcpy(&ia, &a, 8);
        std::memcpy(&ib, &b, 8);
        ia ^= (int64_t)(((uint64_t)(ia >> 63)) >> 1);
        ib ^= (int64_t)(((uint64_t)(ib >> 63)) >> 1);
        return ia < ib;
    }
}

template <class T>
struct Context {
    const T* levels;
    const std::vector<Cell<T>>* cells;
    size_t cw, ch;
    const std::vector<uint32_t>* labels;
    const std::vector<Slot<T>>* slots;

    // opposed.rs:282-327
    bool local_correction(int target, unsigned support, bool has_tr, T tr, size_t tx,
                          size_t ty, T& out, int& out_radius) const {
        T corr[MAX_CANDIDATES];
        int count = 0;
        for (int radius = FIRST_RADIUS; radius <= LAST_RADIUS; ++radius) {
            size_t y0 = ty >= (size_t)radius ? ty - radius : 0;
            size_t y1 = std::min(ty + radius, ch - 1);
            size_t x0 = tx >= (size_t)radius ? tx - radius : 0;
            size_t x1 = std::min(tx + radius, cw - 1);
            for (size_t cy = y0; cy <= y1; ++cy)
                for (size_t cx = x0; cx <= x1; ++cx) {
                    size_t dx = cx > tx ? cx - tx : tx - cx, dy = cy > ty ? cy - ty : ty - cy;
                    if (std::max(dx, dy) != (size_t)radius) continue;
                    const Cell<T>& cand = (*cells)[cy * cw + cx];
                    if (!(cand.usable(target) && cand.clean(target))) continue;
                    T value;
                    if (!correction(cand, target, support, levels, value)) continue;
                    T cr = 0;
                    bool has_cr = support_log_ratio(cand, support, cr);
                    if (!chromaticity_agrees(support, has_tr, tr, has_cr, cr)) continue;
                    if (count < MAX_CANDIDATES) corr[count++] = value;
                }
            if (count >= MIN_CANDIDATES) {
                std::sort(corr, corr + count, total_less<T>);
                int middle = count / 2;
                // 4: the upper middle element for even counts, lower for odd
                out = RHDO_MUTANT(4) ? corr[count % 2 == 0 ? middle : middle - 1]
                      : count % 2 == 0 ? (corr[middle - 1] + corr[middle]) * RHDO_LIT(T, 0.5) : corr[middle];
                out_radius = radius;
                return true;
            }
        }
        return false;
    }

    // region_correction.rs:270-282 (lookup)
    bool lookup(uint32_t label, int target, unsigned support, T& offset, bool& has_ratio,
                T& ratio) const {
        int p = pattern_position(target, support);
        if (p < 0) return false;
        int offs = target * 3 + p;
        bool want_ratio = popcount(support) == 2;
        if ((*slots)[label * 9 + offs].value(want_ratio, offset, has_ratio, ratio)) return true;
        return (*slots)[offs].value(want_ratio, offset, has_ratio, ratio);
    }

    // opposed.rs:243-278
    bool estimate_site(const Cell<T>& target, int channel, size_t tx, size_t ty, T& out) const {
        unsigned support = support_mask(target, channel, levels);
        T target_reference;
        if (!reference(target, support, target_reference)) return false;
        T tr = 0;
        bool has_tr = support_log_ratio(target, support, tr);
        T local = 0;
        int radius = 0;
        bool has_local = local_correction(channel, support, has_tr, tr, tx, ty, local, radius);
        uint32_t label = (*labels)[ty * cw + tx];
        T region = 0, border_ratio = 0;
        bool border_has_ratio = false;
        bool has_region = lookup(label, channel, support, region, border_has_ratio, border_ratio) &&
                          chromaticity_agrees(support, has_tr, tr, border_has_ratio, border_ratio);
        T corr;
        if (has_local && has_region) {
            T weight = (T)(LAST_RADIUS + 1 - radius) / (T)LAST_RADIUS;
            corr = RHDO_MUTANT(2) ? (T(1) - weight) * local + weight * region  // 2: swapped blend
                                  : weight * local + (T(1) - weight) * region;
        } else if (has_local) {
            corr = local;
        } else if (has_region) {
            corr = region;
        } else {
            return false;
        }
        T estimate = target_reference + corr;
        if (!(finite(estimate) && estimate > T(0))) return false;
        out = estimate;
        return true;
    }
};

template <class T>
inline void neutralise(T& sample, const T* ceiling, int channel, uint64_t& neutralised) {
    if (!ceiling) return;
    T limit = ceiling[channel];
    if (sample > limit) {
        sample = limit;
        neutralised += 1;
    }
}

// opposed.rs:63-180. `labels_out` (cell grid, optional) receives the region
// labels for parity checks of the segmentation.
template <class T>
inline int reconstruct(T* m, size_t w, size_t h, int pattern, const T levels[3],
                       const T* neutral_ceiling, Stats* stats_out, uint32_t* labels_out,
                       uint32_t* region_count_out) {
    Stats st{};
    // lib.rs:108-119 (ChannelDetectionLevels::validate) and the same check on
    // the neutral ceiling: finite and positive, before any sample is read.
    for (int c = 0; c < 3; ++c) {
        if (!(finite(levels[c]) && levels[c] > T(0))) return 8;
        if (neutral_ceiling && !(finite(neutral_ceiling[c]) && neutral_ceiling[c] > T(0)))
            return 8;
    }
    for (size_t y = 0; y < h; ++y)
        for (size_t x = 0; x < w; ++x) {
            T s = m[y * w + x];
            if (!finite(s)) return 8;
            int c = color_at(pattern, x, y);
            if (s >= levels[c]) {
                st.suspected++;
                st.suspected_by_channel[c]++;
            }
        }
    size_t cw = w / 2 + w % 2, chh = h / 2 + h % 2;
    if (region_count_out) *region_count_out = 0;
    if (labels_out) std::memset(labels_out, 0, cw * chh * sizeof(uint32_t));
    if (st.suspected == 0) {
        if (stats_out) *stats_out = st;
        return 0;
    }
    std::vector<Cell<T>> cells = build_cells(m, w, h, pattern, levels, cw, chh);
    auto clipped_at = [&](size_t x, size_t y) { return (unsigned)(cells[y * cw + x].flags >> 3); };
    // regions.rs:126-183: Chebyshev radius-1 dilation of the clipped cells.
    std::vector<uint32_t> labels(cw * chh, 0);
    for (size_t y = 0; y < chh; ++y)
        for (size_t x = 0; x < cw; ++x) {
            bool any = false;
            for (ptrdiff_t dy = -1; dy <= 1 && !any; ++dy)
                for (ptrdiff_t dx = -1; dx <= 1 && !any; ++dx) {
                    ptrdiff_t nx = (ptrdiff_t)x + dx, ny = (ptrdiff_t)y + dy;
                    if (nx < 0 || ny < 0 || nx >= (ptrdiff_t)cw || ny >= (ptrdiff_t)chh) continue;
                    any = clipped_at((size_t)nx, (size_t)ny) != 0;
                }
            labels[y * cw + x] = any ? 1u : 0u;
        }
    uint32_t count = label_components(labels.data(), cw, chh);
    std::vector<uint8_t> region_clipped(count + 1, 0);
    for (size_t i = 0; i < cw * chh; ++i)
        if (labels[i]) region_clipped[labels[i]] |= (uint8_t)(cells[i].flags >> 3);
    // region_correction.rs:125-187 (WantedSlots::scan)
    std::vector<uint16_t> wanted(count + 1, 0);
    uint16_t any_wanted = 0;
    for (size_t y = 0; y < chh; ++y)
        for (size_t x = 0; x < cw; ++x) {
            size_t i = y * cw + x;
            uint32_t label = labels[i];
            unsigned clipped = clipped_at(x, y);
            if (label == 0 || clipped == 0) continue;
            const Cell<T>& cell = cells[i];
            auto has_clean = [&](int c) {
                for (int o = 0; o < 3; ++o)
                    if (o != c && cell.usable(o) && cell.clean(o)) return true;
                return false;
            };
            for (int c = 0; c < 3; ++c) {
                if (!((clipped & (1u << c)) && has_clean(c))) continue;
                unsigned support = support_mask(cell, c, levels);
                int p = pattern_position(c, support);
                if (p >= 0) {
                    uint16_t bit = (uint16_t)(1u << (c * 3 + p));
                    wanted[label] |= bit;
                    any_wanted |= bit;
                }
            }
        }
    // region_correction.rs:196-261: two raster-order passes over the rings.
    std::vector<Slot<T>> slots((size_t)(count + 1) * 9);
    for (int pass = 0; pass < 2; ++pass) {
        for (size_t y = 0; y < chh; ++y)
            for (size_t x = 0; x < cw; ++x) {
                size_t i = y * cw + x;
                uint32_t label = labels[i];
                if (label == 0 || clipped_at(x, y) != 0) continue;
                const Cell<T>& cell = cells[i];
                for (int c = 0; c < 3; ++c) {
                    if (!(region_clipped[label] & (1u << c))) continue;
                    // 10: ring evidence admitted from a quarter of the level
                    if (cell.means[c] < (RHDO_MUTANT(10) ? RHDO_LIT(T, 0.25) : RHDO_LIT(T, 0.5)) * levels[c]) continue;
                    unsigned pats[3];
                    support_patterns(c, pats);
                    for (int p = 0; p < 3; ++p) {
                        int offs = c * 3 + p;
                        bool for_region = wanted[label] & (1u << offs);
                        bool for_pool = any_wanted & (1u << offs);
                        if (!for_region && !for_pool) continue;
                        T value;
                        if (!correction(cell, c, pats[p], levels, value)) continue;
                        T ratio = 0;
                        bool has_ratio = support_log_ratio(cell, pats[p], ratio);
                        if (for_region) slots[label * 9 + offs].add(pass == 1, value, has_ratio, ratio);
                        if (for_pool) slots[offs].add(pass == 1, value, has_ratio, ratio);
                    }
                }
            }
        if (pass == 0)
            for (Slot<T>& s : slots) s.finish_moments();
    }
    Context<T> ctx{levels, &cells, cw, chh, &labels, &slots};
    for (size_t y = 0; y < h; ++y)
        for (size_t x = 0; x < w; ++x) {
            T& sample = m[y * w + x];
            int c = color_at(pattern, x, y);
            T level = levels[c];
            if (!(sample >= level)) continue;
            size_t tx = x / 2, ty = y / 2;
            const Cell<T>& target = cells[ty * cw + tx];
            int support_count = 0;
            for (int o = 0; o < 3; ++o)
                if (o != c && target.usable(o) && target.clean(o)) ++support_count;
            T estimate = 0;
            bool has = support_count != 0 && ctx.estimate_site(target, c, tx, ty, estimate);
            if (!has) {
                st.fallback++;
                if (support_count == 0) st.fully_unsupported++;
                neutralise(sample, neutral_ceiling, c, st.neutralised);
                continue;
            }
            st.reconstructed++;
            T bounded = num_min(estimate, (RHDO_MUTANT(11) ? RHDO_LIT(T, 2.0) : RHDO_LIT(T, 4.0)) * level);  // 11
            if (bounded > sample) {
                sample = bounded;
                st.changed++;
            } else {
                neutralise(sample, neutral_ceiling, c, st.neutralised);
            }
        }
    if (labels_out) std::memcpy(labels_out, labels.data(), cw * chh * sizeof(uint32_t));
    if (region_count_out) *region_count_out = count;
    if (stats_out) *stats_out = st;
    return 0;
}

}  // namespace opposed
}  // namespace rhdo
