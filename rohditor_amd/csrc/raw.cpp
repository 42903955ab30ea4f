// RAW ingest in front of the stage: the container walk that finds a Sony
// ARW's raw strip (host), and the strip's upload and decode into a resident
// mosaic (device). The reference opens files with rawler 0.7.2 behind
// crates/raw/src/rawler_adapter.rs (open :38-61, raw_image :131-170); its
// errors map to RawError::{Unsupported, Corrupt} (:304-320).
#include <cstring>
#include <vector>

#include "rhd_context.hpp"

namespace rhd {
namespace {

struct Tiff {
    const uint8_t* p;
    uint64_t n;
    bool le;
    bool fits(uint64_t at, uint64_t len) const { return at <= n && len <= n - at; }
    uint32_t u16(uint64_t at) const {
        return le ? (uint32_t)p[at] | (uint32_t)p[at + 1] << 8
                  : (uint32_t)p[at] << 8 | (uint32_t)p[at + 1];
    }
    uint32_t u32(uint64_t at) const {
        return le ? u16(at) | u16(at + 2) << 16 : u16(at) << 16 | u16(at + 2);
    }
};

// Bytes per value of a TIFF field type (0 for unknown types).
uint32_t field_size(uint32_t type) {
    static const uint8_t sizes[14] = {0, 1, 1, 2, 4, 8, 1, 1, 2, 4, 8, 4, 8, 4};
    return type < 14 ? sizes[type] : 0;
}

struct Field {
    uint32_t tag, type, count;
    uint64_t at;  // first value byte
    uint32_t get(const Tiff& t, uint32_t i) const {
        if (i >= count) return 0;
        switch (field_size(type)) {
            case 1: return t.p[at + i];
            case 2: return t.u16(at + 2ull * i);
            case 4: return t.u32(at + 4ull * i);
            default: return 0;
        }
    }
};

// RAW_CORRUPT when the IFD or one of its values leaves the file.
rhd_status read_fields(const Tiff& t, uint64_t at, std::vector<Field>& fields, uint64_t& next) {
    if (!t.fits(at, 2)) return RHD_RAW_CORRUPT;
    const uint32_t count = t.u16(at);
    if (!t.fits(at + 2, 12ull * count + 4)) return RHD_RAW_CORRUPT;
    fields.clear();
    for (uint32_t k = 0; k < count; ++k) {
        const uint64_t e = at + 2 + 12ull * k;
        Field f{t.u16(e), t.u16(e + 2), t.u32(e + 4), e + 8};
        const uint64_t bytes = (uint64_t)field_size(f.type) * f.count;
        if (bytes > 4) {
            f.at = t.u32(e + 8);
            if (!t.fits(f.at, bytes)) return RHD_RAW_CORRUPT;
        }
        fields.push_back(f);
    }
    next = t.u32(at + 2 + 12ull * count);
    return RHD_OK;
}

}  // namespace

hipError_t launch_arw2(const uint8_t* strip, uint64_t bytes, int width, int height,
                       const uint32_t* lut, const uint32_t* jump, uint32_t* first_bad,
                       uint16_t* out, hipStream_t s, Profiler* prof);

}  // namespace rhd

namespace rhd {

// Uploads the raw strip of `l` and decodes it on lane L into a new mosaic.
rhd_status decode_arw_on(rhd_context* ctx, Lane& L, const uint8_t* file, uint64_t size,
                         const rhd_arw_layout* l, rhd_mosaic** out, float* gpu_ms) {
    *out = nullptr;
    if (gpu_ms) *gpu_ms = 0.0f;
    if (l->width == 0 || l->height == 0 || l->width > (1u << 20) || l->height > 65535)
        return fail(ctx, RHD_INVALID_DIMENSIONS, "raw dimensions out of range");
    if (l->offset > size || l->bytes > size - l->offset)
        return fail(ctx, RHD_RAW_CORRUPT, "raw strip outside the file");
    const uint64_t px = (uint64_t)l->width * l->height;
    if (l->kind == RHD_ARW_UNCOMPRESSED16) {
        if (l->bytes != 2 * px || (l->offset & 1))
            return fail(ctx, RHD_RAW_CORRUPT, "uncompressed strip size does not match");
        (void)hipSetDevice(ctx->device);
        return mosaic_upload(ctx, L, (const uint16_t*)(file + l->offset), l->width, l->width,
                             l->height, out);
    }
    if (l->kind != RHD_ARW_COMPRESSED) return fail(ctx, RHD_RAW_UNSUPPORTED, "unknown raw kind");
    if (l->bytes != px) return fail(ctx, RHD_RAW_CORRUPT, "compressed strip size does not match");

    // Tone curve (4096 entries, segment i climbs 2^i) and its dithering
    // entries (base | delta << 16); the stream's jump per 16-pixel block.
    uint32_t knots[6] = {0, 0, 0, 0, 0, 4095};
    for (int i = 0; i < 4; ++i) knots[i + 1] = (l->curve[i] >> 2) & 0xfff;
    std::vector<uint16_t> curve(4096, 0);
    for (int i = 0; i < 5; ++i)
        for (uint32_t j = knots[i] + 1; j <= knots[i + 1]; ++j)
            curve[j] = (uint16_t)(curve[j - 1] + (1u << i));
    const int blocks = (int)(l->width / 32) * 2;
    std::vector<uint32_t> tables(4096 + (size_t)blocks);
    for (int i = 0; i < 4096; ++i) {
        const uint16_t c = curve[i];
        const uint16_t lo = i > 0 ? curve[i - 1] : c, up = i < 4095 ? curve[i + 1] : c;
        const uint16_t spread = (uint16_t)(up - lo);
        const uint16_t base = c == 0 ? 0 : (uint16_t)(c - (uint16_t)((uint16_t)(spread + 2) / 4));
        tables[i] = (uint32_t)base | (uint32_t)spread << 16;
    }
    const uint64_t m = 15700ull * 65536ull - 1ull;
    uint64_t step16 = 1;
    for (int k = 0; k < 16; ++k) step16 = step16 * 15700ull % m;
    uint64_t jump = 1;
    for (int b = 0; b < blocks; ++b) {
        tables[4096 + b] = (uint32_t)jump;
        jump = jump * step16 % m;
    }

    (void)hipSetDevice(ctx->device);
    const uint64_t strip_bytes = (l->bytes + 255) & ~255ull;
    const uint64_t table_bytes = (tables.size() * sizeof(uint32_t) + 255) & ~255ull;
    rhd_status st =
        ensure_scratch(ctx, L, L.raw, strip_bytes + table_bytes + 4ull * l->height);
    if (st != RHD_OK) return st;
    void* mosaic = nullptr;
    if ((st = device_alloc(ctx, &L, px * sizeof(uint16_t), &mosaic)) != RHD_OK) return st;
    uint8_t* d_strip = (uint8_t*)L.raw.ptr;
    uint32_t* d_tables = (uint32_t*)(d_strip + strip_bytes);
    uint32_t* d_first_bad = (uint32_t*)(d_strip + strip_bytes + table_bytes);
    hipError_t e = hipMemcpyAsync(d_strip, file + l->offset, l->bytes, hipMemcpyHostToDevice,
                                  L.stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_tables, tables.data(), tables.size() * sizeof(uint32_t),
                           hipMemcpyHostToDevice, L.stream);
    if (e == hipSuccess) e = hipEventRecord(L.ev[3], L.stream);
    if (e == hipSuccess)
        e = launch_arw2(d_strip, l->bytes, (int)l->width, (int)l->height, d_tables,
                        d_tables + 4096, d_first_bad, (uint16_t*)mosaic, L.stream, PROF(L));
    if (e == hipSuccess) e = hipEventRecord(L.ev[4], L.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(L.stream);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(L.stream);
        device_free(ctx, mosaic, px * sizeof(uint16_t));
        return hip_fail(ctx, e, "compressed ARW decode");
    }
    if (gpu_ms) *gpu_ms = elapsed(L.ev[3], L.ev[4]);
    ctx->refs.fetch_add(1);
    *out = new rhd_mosaic{ctx, (uint16_t*)mosaic, l->width, l->width, l->height,
                          px * sizeof(uint16_t)};
    return RHD_OK;
}


}  // namespace rhd

using namespace rhd;

extern "C" {

rhd_status rhd_arw_probe(const uint8_t* file, uint64_t size, rhd_arw_layout* out) {
    if (!file || !out) return RHD_INVALID_INPUT;
    *out = rhd_arw_layout{};
    if (size < 8 || !((file[0] == 'I' && file[1] == 'I') || (file[0] == 'M' && file[1] == 'M')))
        return RHD_RAW_UNSUPPORTED;
    const Tiff t{file, size, file[0] == 'I'};
    if (t.u16(2) != 42) return RHD_RAW_UNSUPPORTED;
    // IFD0, its SubIFDs depth first, then the next IFD in the chain.
    std::vector<uint64_t> pending{t.u32(4)}, visited;
    std::vector<Field> fields;
    while (!pending.empty()) {
        const uint64_t at = pending.back();
        pending.pop_back();
        if (at == 0) continue;
        for (uint64_t v : visited)
            if (v == at) return RHD_RAW_CORRUPT;  // a loop
        if (visited.size() >= 64) return RHD_RAW_CORRUPT;
        visited.push_back(at);
        uint64_t next = 0;
        rhd_status st = read_fields(t, at, fields, next);
        if (st != RHD_OK) return st;
        rhd_arw_layout l{};
        bool one_strip = false, one_count = false, curve = false;
        std::vector<uint64_t> subs;
        for (const Field& f : fields) {
            switch (f.tag) {
                case 0x0100: l.width = f.get(t, 0); break;
                case 0x0101: l.height = f.get(t, 0); break;
                case 0x0102: l.bits_per_sample = f.get(t, 0); break;
                case 0x0103: l.compression = f.get(t, 0); break;
                case 0x0111: l.offset = f.get(t, 0); one_strip = f.count == 1; break;
                case 0x0117: l.bytes = f.get(t, 0); one_count = f.count == 1; break;
                case 0x7010:  // SonyCurve: four knots
                    if (f.count >= 4 && field_size(f.type) == 2) {
                        for (uint32_t k = 0; k < 4; ++k) l.curve[k] = (uint16_t)f.get(t, k);
                        curve = true;
                    }
                    break;
                case 0x014a:
                    for (uint32_t k = 0; k < f.count && k < 16; ++k) subs.push_back(f.get(t, k));
                    break;
            }
        }
        const uint64_t px = (uint64_t)l.width * l.height;
        if (one_strip && one_count && px > 0) {
            if (l.compression == 32767 && l.bytes == px && curve)
                l.kind = RHD_ARW_COMPRESSED;
            else if (l.compression == 1 && l.bits_per_sample == 16 && l.bytes == 2 * px && t.le)
                l.kind = RHD_ARW_UNCOMPRESSED16;
            if (l.kind) {
                if (!t.fits(l.offset, l.bytes)) return RHD_RAW_CORRUPT;
                *out = l;
                return RHD_OK;
            }
        }
        pending.push_back(next);
        for (size_t k = subs.size(); k-- > 0;) pending.push_back(subs[k]);
    }
    return RHD_RAW_UNSUPPORTED;
}

rhd_status rhd_mosaic_decode_arw(rhd_context* ctx, const uint8_t* file, uint64_t size,
                                 const rhd_arw_layout* l, rhd_mosaic** out, float* gpu_ms) {
    if (!ctx || !file || !l || !out) return RHD_INVALID_INPUT;
    return decode_arw_on(ctx, ctx->lanes[0], file, size, l, out, gpu_ms);
}

rhd_status rhd_mosaic_read(rhd_context* ctx, const rhd_mosaic* m, uint16_t* dst, uint64_t count) {
    if (!ctx || !m || !dst) return RHD_INVALID_INPUT;
    rhd_status st = check_handle_ctx(ctx, m->ctx);
    if (st != RHD_OK) return st;
    if (count != m->row_stride * m->height)
        return fail(ctx, RHD_INVALID_INPUT, "destination must hold row_stride * height samples");
    Lane& L = ctx->lanes[0];
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipMemcpyAsync(dst, m->d, count * sizeof(uint16_t), hipMemcpyDeviceToHost,
                                  L.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(L.stream);
    return e == hipSuccess ? RHD_OK : hip_fail(ctx, e, "mosaic readback");
}

}  // extern "C"
