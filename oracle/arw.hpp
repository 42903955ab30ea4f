// Oracle (test infrastructure only): the Sony ARW decode in front of the
// stage. The reference decodes with rawler 0.7.2 (crates/raw/src/
// rawler_adapter.rs:83-170: `decoder.raw_image(...)` then the u16 samples
// become RawFrame::mosaic), a dependency that is not vendored. Restated here
// from its published algorithm for Sony's compressed ARW (TIFF compression
// 32767, 8 bits per pixel, the format of crates/raw/tests/fixtures/
// sony_a6400_expectations.json):
//   * the tone curve: four SonyCurve (tag 0x7010) knots, each (v >> 2) &
//     0xfff, between 0 and 4095; segment i of the table climbs 2^i per step;
//   * the dithering lookup: per entry (base, delta) with
//     base = center - (upper - lower + 2) / 4 (0 for center 0) and
//     delta = upper - lower from the neighbouring entries, and
//     pixel = base + ((delta * (rand & 2047) + 1024) >> 12), after which
//     rand = 15700 * (rand & 65535) + (rand >> 16);
//   * each row holds `width` bytes read LSB first; rand starts as the row's
//     first 16 bits; every 32 pixels come from two 16-byte blocks, the first
//     for the even columns and the second for the odd ones; a block holds an
//     11-bit max, an 11-bit min, 4-bit positions of the max and the min, and
//     fourteen 7-bit deltas scaled by 2^shift with
//     shift = max(0, bit_length(max - min) - 7), clamped to 0x7ff;
//   * the sample written is dither(curve, value << 1).
// Trailing columns past the last whole 32-pixel group stay 0. u32 and u16
// arithmetic wraps as in a release build. A block whose max and min positions
// coincide reads fifteen deltas (135 bits), so the rest of its row is
// misaligned; the bit reader runs on past the row's bytes into the next rows,
// and reads zeros past the strip's end (there rawler's reader leaves the
// strip; what it then reads is not fixed here). Parity with rawler itself is
// unpinned (no decoded pixels ship with the reference); the tests pin the
// format by hand-built blocks and closed forms.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

namespace rhdo {
namespace arw {

struct Lookup {
    std::vector<uint16_t> base, delta;
};

// The 4096-entry tone curve from the four tag values.
inline std::vector<uint16_t> curve_table(const uint16_t tag[4]) {
    size_t knots[6] = {0, 0, 0, 0, 0, 4095};
    for (int i = 0; i < 4; ++i) knots[i + 1] = (size_t)((tag[i] >> 2) & 0xfff);
    std::vector<uint16_t> out(knots[5] + 1, 0);
    for (int i = 0; i < 5; ++i)
        for (size_t j = knots[i] + 1; j < knots[i + 1] + 1; ++j)
            out[j] = (uint16_t)(out[j - 1] + (1u << i));
    return out;
}

inline Lookup lookup(const std::vector<uint16_t>& table) {
    Lookup l;
    const size_t n = table.size();
    l.base.resize(n);
    l.delta.resize(n);
    for (size_t i = 0; i < n; ++i) {
        const uint16_t center = table[i];
        const uint16_t lower = i > 0 ? table[i - 1] : center;
        const uint16_t upper = i + 1 < n ? table[i + 1] : center;
        const uint16_t spread = (uint16_t)(upper - lower);
        l.base[i] = center == 0 ? 0 : (uint16_t)(center - (uint16_t)((uint16_t)(spread + 2) / 4));
        l.delta[i] = spread;
    }
    return l;
}

inline uint16_t dither(const Lookup& l, uint16_t value, uint32_t& rand) {
    const uint32_t pixel = (uint32_t)l.base[value] + (((uint32_t)l.delta[value] * (rand & 2047) + 1024) >> 12);
    rand = 15700u * (rand & 65535u) + (rand >> 16);
    return (uint16_t)pixel;
}

// LSB-first bits of a byte row.
struct Bits {
    const uint8_t* p;
    size_t n, pos = 0;
    uint32_t peek(int count) const {
        uint32_t v = 0;
        for (int k = 0; k < count; ++k) {
            const size_t b = pos + k;
            const uint32_t bit = (b >> 3) < n ? (p[b >> 3] >> (b & 7)) & 1u : 0u;
            v |= bit << k;
        }
        return v;
    }
    uint32_t get(int count) {
        const uint32_t v = peek(count);
        pos += count;
        return v;
    }
};

inline int bit_length(uint32_t v) {
    int n = 0;
    while (v) {
        ++n;
        v >>= 1;
    }
    return n;
}

// Decodes width x height pixels from width * height bytes. Returns 1 when
// the strip is shorter than that.
inline int decode_arw2(const uint8_t* strip, size_t bytes, size_t width, size_t height,
                       const uint16_t curve_tag[4], uint16_t* out) {
    if (bytes < width * height) return 1;
    const Lookup l = lookup(curve_table(curve_tag));
    std::memset(out, 0, width * height * sizeof(uint16_t));
    for (size_t row = 0; row < height; ++row) {
        // A row's stream runs on into the following rows (a block whose max
        // and min share a position carries a 15th delta); past the strip's
        // end it reads zeros.
        Bits bits{strip + row * width, bytes - row * width};
        uint32_t rand = bits.peek(16);
        uint16_t* line = out + row * width;
        for (size_t chunk = 0; chunk + 32 <= width; chunk += 32)
            for (int j = 0; j < 2; ++j) {
                const uint32_t max = bits.get(11), min = bits.get(11);
                const uint32_t delta = max - min;
                const int shift = bit_length(delta) - 7 > 0 ? bit_length(delta) - 7 : 0;
                const uint32_t imax = bits.get(4), imin = bits.get(4);
                for (uint32_t i = 0; i < 16; ++i) {
                    uint32_t val;
                    if (i == imax) {
                        val = max;
                    } else if (i == imin) {
                        val = min;
                    } else {
                        val = (bits.get(7) << shift) + min;
                        if (val > 0x7ff) val = 0x7ff;
                    }
                    line[chunk + j + 2 * i] = dither(l, (uint16_t)(val << 1), rand);
                }
            }
    }
    return 0;
}

// ---------------------------------------------------------------- TIFF

// What the decoder needs from the raw IFD.
struct Layout {
    uint32_t width = 0, height = 0, bits_per_sample = 0, compression = 0;
    uint64_t offset = 0, bytes = 0;
    uint16_t curve[4] = {0, 0, 0, 0};
    int has_curve = 0;
    int kind = 0;  // 1 compressed 8 bpp, 2 uncompressed 16-bit (little-endian files)
};

// Status codes: 0 ok, 1 not a TIFF container, 2 corrupt (offsets outside the
// file, loops), 3 a TIFF without a supported raw IFD.
struct Reader {
    const uint8_t* p;
    size_t n;
    bool le = true;
    bool ok(uint64_t at, uint64_t len) const { return at <= n && len <= n - at; }
    uint32_t u16(uint64_t at) const {
        return le ? (uint32_t)(p[at] | p[at + 1] << 8) : (uint32_t)(p[at] << 8 | p[at + 1]);
    }
    uint32_t u32(uint64_t at) const {
        return le ? (uint32_t)p[at] | (uint32_t)p[at + 1] << 8 | (uint32_t)p[at + 2] << 16 |
                        (uint32_t)p[at + 3] << 24
                  : (uint32_t)p[at] << 24 | (uint32_t)p[at + 1] << 16 | (uint32_t)p[at + 2] << 8 |
                        (uint32_t)p[at + 3];
    }
};

struct Entry {
    uint32_t tag, type, count;
    uint64_t value_at;  // where the value bytes start
};

inline uint32_t type_size(uint32_t type) {
    switch (type) {
        case 1: case 2: case 6: case 7: return 1;
        case 3: case 8: return 2;
        case 4: case 9: case 11: case 13: return 4;
        case 5: case 10: case 12: return 8;
        default: return 0;
    }
}

// Reads one IFD; returns 2 when it does not fit.
inline int read_ifd(const Reader& r, uint64_t at, std::vector<Entry>& out, uint64_t& next) {
    if (!r.ok(at, 2)) return 2;
    const uint32_t count = r.u16(at);
    if (!r.ok(at + 2, (uint64_t)count * 12 + 4)) return 2;
    for (uint32_t k = 0; k < count; ++k) {
        const uint64_t e = at + 2 + (uint64_t)k * 12;
        Entry en{r.u16(e), r.u16(e + 2), r.u32(e + 4), e + 8};
        const uint64_t size = (uint64_t)type_size(en.type) * en.count;
        if (size > 4) {
            en.value_at = r.u32(e + 8);
            if (!r.ok(en.value_at, size)) return 2;
        }
        out.push_back(en);
    }
    next = r.u32(at + 2 + (uint64_t)count * 12);
    return 0;
}

inline uint32_t value(const Reader& r, const Entry& e, uint32_t index) {
    if (index >= e.count) return 0;
    switch (type_size(e.type)) {
        case 1: return r.p[e.value_at + index];
        case 2: return r.u16(e.value_at + 2 * (uint64_t)index);
        case 4: return r.u32(e.value_at + 4 * (uint64_t)index);
        default: return 0;
    }
}

inline int probe(const uint8_t* file, size_t size, Layout& out) {
    Reader r{file, size};
    if (size < 8) return 1;
    if (file[0] == 'I' && file[1] == 'I')
        r.le = true;
    else if (file[0] == 'M' && file[1] == 'M')
        r.le = false;
    else
        return 1;
    if (r.u16(2) != 42) return 1;
    // Depth-first over the IFD chain and SubIFDs, bounded.
    std::vector<uint64_t> stack{r.u32(4)};
    std::vector<uint64_t> seen;
    int visited = 0;
    while (!stack.empty()) {
        const uint64_t at = stack.back();
        stack.pop_back();
        if (at == 0) continue;
        for (uint64_t s : seen)
            if (s == at) return 2;
        seen.push_back(at);
        if (++visited > 64) return 2;
        std::vector<Entry> entries;
        uint64_t next = 0;
        if (read_ifd(r, at, entries, next)) return 2;
        Layout l;
        bool strips = false, counts = false;
        std::vector<uint64_t> subs;
        for (const Entry& e : entries) {
            switch (e.tag) {
                case 0x0100: l.width = value(r, e, 0); break;
                case 0x0101: l.height = value(r, e, 0); break;
                case 0x0102: l.bits_per_sample = value(r, e, 0); break;
                case 0x0103: l.compression = value(r, e, 0); break;
                case 0x0111: l.offset = value(r, e, 0); strips = e.count == 1; break;
                case 0x0117: l.bytes = value(r, e, 0); counts = e.count == 1; break;
                case 0x7010:
                    if (e.count >= 4 && type_size(e.type) == 2) {
                        for (uint32_t k = 0; k < 4; ++k) l.curve[k] = (uint16_t)value(r, e, k);
                        l.has_curve = 1;
                    }
                    break;
                case 0x014a:
                    for (uint32_t k = 0; k < e.count && k < 16; ++k) subs.push_back(value(r, e, k));
                    break;
            }
        }
        const uint64_t px = (uint64_t)l.width * l.height;
        if (strips && counts && px > 0) {
            if (l.compression == 32767 && l.bytes == px && l.has_curve) {
                l.kind = 1;
            } else if (l.compression == 1 && l.bits_per_sample == 16 && l.bytes == 2 * px &&
                       r.le) {
                l.kind = 2;
            }
            if (l.kind) {
                if (!r.ok(l.offset, l.bytes)) return 2;
                out = l;
                return 0;
            }
        }
        stack.push_back(next);
        for (size_t k = subs.size(); k-- > 0;) stack.push_back(subs[k]);
    }
    return 3;
}

}  // namespace arw
}  // namespace rhdo
