// Oracle (test infrastructure only): region-aware Opposed highlight
// reconstruction restated from crates/highlight/src/opposed.rs:63-364,
// cells.rs:41-113, regions.rs:34-271, region_correction.rs:53-300 and
// support.rs:26-125. f32 arithmetic with the reference's f64 accumulators;
// cbrt/log2 come from the platform libm, which is what Rust's f32 methods
// call on Linux.
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

struct Cell {
    float means[3];
    uint8_t flags;
    bool usable(int c) const { return flags & (1 << c); }
    bool clean(int c) const { return !(flags & (1 << (c + 3))); }
};

inline int popcount(unsigned m) { return __builtin_popcount(m); }

// support.rs:26-33
inline bool bright_clean(const Cell& cell, int c, const float levels[3]) {
    return cell.usable(c) && cell.clean(c) && cell.means[c] > 0.2f * levels[c];
}

// support.rs:37-46
inline unsigned support_mask(const Cell& cell, int target, const float levels[3]) {
    unsigned m = 0;
    for (int c = 0; c < 3; ++c)
        if (c != target && bright_clean(cell, c, levels)) m |= 1u << c;
    return m;
}

// support.rs:51-72
inline bool reference(const Cell& cell, unsigned support, float& out) {
    if (popcount(support) == 1) {
        float v = cell.means[__builtin_ctz(support)];
        if (finite(v) && v > 0.0f) {
            out = v;
            return true;
        }
        return false;
    }
    float sum = 0.0f;
    unsigned count = 0;
    for (int c = 0; c < 3; ++c) {
        if (!(support & (1u << c))) continue;
        float v = cell.means[c];
        if (!finite(v) || v <= 0.0f) return false;
        sum += RHDO_MUTANT(9) ? v : std::cbrt(v);  // 9: arithmetic instead of cube-root mean
        ++count;
    }
    if (count == 0) return false;
    float root = sum / (float)count;
    out = RHDO_MUTANT(9) ? root : root * root * root;
    return true;
}

// support.rs:78-95
inline bool correction(const Cell& cell, int target, unsigned support, const float levels[3],
                       float& out) {
    if (!bright_clean(cell, target, levels)) return false;
    for (int c = 0; c < 3; ++c)
        if ((support & (1u << c)) && !bright_clean(cell, c, levels)) return false;
    float ref;
    if (!reference(cell, support, ref)) return false;
    float v = cell.means[target] - ref;
    if (!finite(v)) return false;
    out = v;
    return true;
}

// support.rs:101-110
inline bool support_log_ratio(const Cell& cell, unsigned support, float& out) {
    if (popcount(support) != 2) return false;
    int first = __builtin_ctz(support);
    int second = __builtin_ctz(support & ~(1u << first));
    float ratio = cell.means[first] / cell.means[second];
    if (!(finite(ratio) && ratio > 0.0f)) return false;
    out = std::log2(ratio);
    return true;
}

// support.rs:116-125
inline bool chromaticity_agrees(unsigned support, bool has_t, float t, bool has_e, float e) {
    if (popcount(support) < 2) return true;
    if (has_t && has_e) return std::fabs(t - e) <= (RHDO_MUTANT(3) ? 1.5f : 0.75f);  // 3: gate
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
inline std::vector<Cell> build_cells(const float* m, size_t w, size_t h, int pattern,
                                     const float levels[3], size_t& cw, size_t& ch) {
    cw = w / 2 + w % 2;
    ch = h / 2 + h % 2;
    std::vector<Cell> cells(cw * ch);
    for (size_t cy = 0; cy < ch; ++cy)
        for (size_t cx = 0; cx < cw; ++cx) {
            float sums[3] = {0, 0, 0};
            uint32_t counts[3] = {0, 0, 0};
            uint8_t flags = 0;
            for (size_t y = cy * 2; y < std::min(cy * 2 + 2, h); ++y)
                for (size_t x = cx * 2; x < std::min(cx * 2 + 2, w); ++x) {
                    int c = color_at(pattern, x, y);
                    float s = m[y * w + x];
                    if (s >= levels[c]) {
                        flags |= 1 << (c + 3);
                    } else if (s > 0.0f) {
                        sums[c] += s;
                        counts[c] += 1;
                    }
                }
            Cell& cell = cells[cy * cw + cx];
            for (int c = 0; c < 3; ++c) {
                cell.means[c] = 0.0f;
                if (counts[c] != 0) {
                    cell.means[c] = sums[c] / (float)counts[c];
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
struct Slot {
    uint32_t count = 0, kept = 0;
    double accum = 0, second = 0;
    float mean = 0, tolerance = 0;
    void add(bool trim, float value, bool has_ratio, float ratio) {
        double wide = (double)value;
        if (!trim) {
            count += 1;
            accum += wide;
            second += wide * wide;
        } else if (std::fabs(value - mean) <= tolerance) {
            kept += 1;
            accum += wide;
            second += (double)(has_ratio ? ratio : 0.0f);
        }
    }
    void finish_moments() {
        if (count != 0) {
            double c = (double)count;
            double m = accum / c;
            double variance = num_max(second / c - m * m, 0.0);
            mean = (float)m;
            // 1: a 3-sigma trim window instead of 2 sigma
            tolerance = (float)((RHDO_MUTANT(1) ? 3.0 : 2.0) * std::sqrt(variance)) + 1e-5f;
        }
        accum = 0;
        second = 0;
    }
    bool value(bool has_ratio, float& offset, bool& ratio_present, float& ratio) const {
        if (count < MIN_BORDER_CELLS || kept == 0) return false;
        double k = (double)kept;
        offset = (float)(accum / k);
        ratio_present = has_ratio;
        ratio = has_ratio ? (float)(second / k) : 0.0f;
        return true;
    }
};

// Rust's f32::total_cmp ordering.
inline bool total_less(float a, float b) {
    int32_t ia, ib;
    std::memcpy(&ia, &a, 4);
    std::memcpy(&ib, &b, 4);
    ia ^= (int32_t)(((uint32_t)(ia >> 31)) >> 1);
    ib ^= (int32_t)(((uint32_t)(ib >> 31)) >> 1);
    return ia < ib;
}

struct Context {
    const float* levels;
    const std::vector<Cell>* cells;
    size_t cw, ch;
    const std::vector<uint32_t>* labels;
    const std::vector<Slot>* slots;

    // opposed.rs:282-327
    bool local_correction(int target, unsigned support, bool has_tr, float tr, size_t tx,
                          size_t ty, float& out, int& out_radius) const {
        float corr[MAX_CANDIDATES];
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
                    const Cell& cand = (*cells)[cy * cw + cx];
                    if (!(cand.usable(target) && cand.clean(target))) continue;
                    float value;
                    if (!correction(cand, target, support, levels, value)) continue;
                    float cr = 0;
                    bool has_cr = support_log_ratio(cand, support, cr);
                    if (!chromaticity_agrees(support, has_tr, tr, has_cr, cr)) continue;
                    if (count < MAX_CANDIDATES) corr[count++] = value;
                }
            if (count >= MIN_CANDIDATES) {
                std::sort(corr, corr + count, total_less);
                int middle = count / 2;
                // 4: the upper middle element for even counts, lower for odd
                out = RHDO_MUTANT(4) ? corr[count % 2 == 0 ? middle : middle - 1]
                      : count % 2 == 0 ? (corr[middle - 1] + corr[middle]) * 0.5f : corr[middle];
                out_radius = radius;
                return true;
            }
        }
        return false;
    }

    // region_correction.rs:270-282 (lookup)
    bool lookup(uint32_t label, int target, unsigned support, float& offset, bool& has_ratio,
                float& ratio) const {
        int p = pattern_position(target, support);
        if (p < 0) return false;
        int offs = target * 3 + p;
        bool want_ratio = popcount(support) == 2;
        if ((*slots)[label * 9 + offs].value(want_ratio, offset, has_ratio, ratio)) return true;
        return (*slots)[offs].value(want_ratio, offset, has_ratio, ratio);
    }

    // opposed.rs:243-278
    bool estimate_site(const Cell& target, int channel, size_t tx, size_t ty, float& out) const {
        unsigned support = support_mask(target, channel, levels);
        float target_reference;
        if (!reference(target, support, target_reference)) return false;
        float tr = 0;
        bool has_tr = support_log_ratio(target, support, tr);
        float local = 0;
        int radius = 0;
        bool has_local = local_correction(channel, support, has_tr, tr, tx, ty, local, radius);
        uint32_t label = (*labels)[ty * cw + tx];
        float region = 0, border_ratio = 0;
        bool border_has_ratio = false;
        bool has_region = lookup(label, channel, support, region, border_has_ratio, border_ratio) &&
                          chromaticity_agrees(support, has_tr, tr, border_has_ratio, border_ratio);
        float corr;
        if (has_local && has_region) {
            float weight = (float)(LAST_RADIUS + 1 - radius) / (float)LAST_RADIUS;
            corr = RHDO_MUTANT(2) ? (1.0f - weight) * local + weight * region  // 2: swapped blend
                                  : weight * local + (1.0f - weight) * region;
        } else if (has_local) {
            corr = local;
        } else if (has_region) {
            corr = region;
        } else {
            return false;
        }
        float estimate = target_reference + corr;
        if (!(finite(estimate) && estimate > 0.0f)) return false;
        out = estimate;
        return true;
    }
};

inline void neutralise(float& sample, const float* ceiling, int channel, uint64_t& neutralised) {
    if (!ceiling) return;
    float limit = ceiling[channel];
    if (sample > limit) {
        sample = limit;
        neutralised += 1;
    }
}

// opposed.rs:63-180. `labels_out` (cell grid, optional) receives the region
// labels for parity checks of the segmentation.
inline int reconstruct(float* m, size_t w, size_t h, int pattern, const float levels[3],
                       const float* neutral_ceiling, Stats* stats_out, uint32_t* labels_out,
                       uint32_t* region_count_out) {
    Stats st{};
    // lib.rs:108-119 (ChannelDetectionLevels::validate) and the same check on
    // the neutral ceiling: finite and positive, before any sample is read.
    for (int c = 0; c < 3; ++c) {
        if (!(finite(levels[c]) && levels[c] > 0.0f)) return 8;
        if (neutral_ceiling && !(finite(neutral_ceiling[c]) && neutral_ceiling[c] > 0.0f))
            return 8;
    }
    for (size_t y = 0; y < h; ++y)
        for (size_t x = 0; x < w; ++x) {
            float s = m[y * w + x];
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
    std::vector<Cell> cells = build_cells(m, w, h, pattern, levels, cw, chh);
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
            const Cell& cell = cells[i];
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
    std::vector<Slot> slots((size_t)(count + 1) * 9);
    for (int pass = 0; pass < 2; ++pass) {
        for (size_t y = 0; y < chh; ++y)
            for (size_t x = 0; x < cw; ++x) {
                size_t i = y * cw + x;
                uint32_t label = labels[i];
                if (label == 0 || clipped_at(x, y) != 0) continue;
                const Cell& cell = cells[i];
                for (int c = 0; c < 3; ++c) {
                    if (!(region_clipped[label] & (1u << c))) continue;
                    // 10: ring evidence admitted from a quarter of the level
                    if (cell.means[c] < (RHDO_MUTANT(10) ? 0.25f : 0.5f) * levels[c]) continue;
                    unsigned pats[3];
                    support_patterns(c, pats);
                    for (int p = 0; p < 3; ++p) {
                        int offs = c * 3 + p;
                        bool for_region = wanted[label] & (1u << offs);
                        bool for_pool = any_wanted & (1u << offs);
                        if (!for_region && !for_pool) continue;
                        float value;
                        if (!correction(cell, c, pats[p], levels, value)) continue;
                        float ratio = 0;
                        bool has_ratio = support_log_ratio(cell, pats[p], ratio);
                        if (for_region) slots[label * 9 + offs].add(pass == 1, value, has_ratio, ratio);
                        if (for_pool) slots[offs].add(pass == 1, value, has_ratio, ratio);
                    }
                }
            }
        if (pass == 0)
            for (Slot& s : slots) s.finish_moments();
    }
    Context ctx{levels, &cells, cw, chh, &labels, &slots};
    for (size_t y = 0; y < h; ++y)
        for (size_t x = 0; x < w; ++x) {
            float& sample = m[y * w + x];
            int c = color_at(pattern, x, y);
            float level = levels[c];
            if (!(sample >= level)) continue;
            size_t tx = x / 2, ty = y / 2;
            const Cell& target = cells[ty * cw + tx];
            int support_count = 0;
            for (int o = 0; o < 3; ++o)
                if (o != c && target.usable(o) && target.clean(o)) ++support_count;
            float estimate = 0;
            bool has = support_count != 0 && ctx.estimate_site(target, c, tx, ty, estimate);
            if (!has) {
                st.fallback++;
                if (support_count == 0) st.fully_unsupported++;
                neutralise(sample, neutral_ceiling, c, st.neutralised);
                continue;
            }
            st.reconstructed++;
            float bounded = num_min(estimate, (RHDO_MUTANT(11) ? 2.0f : 4.0f) * level);  // 11
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
