"""Pins the ARW decode oracle (oracle/arw.hpp) and checks the library's host
container walk against it.

rawler 0.7.2, the reference's decoder (crates/raw/src/rawler_adapter.rs), is
not vendored and no decoded pixels ship with the reference, so the format is
pinned here by independent computations: the tone curve and its dithering
lookup in closed form, the multiply-with-carry stream against modular
powers, hand-packed blocks, and a vectorised decode of encoded rows. The
reference's own corrupt fixtures (tests/golden/raw, from
crates/raw/tests/malformed_inputs.rs:11-30) must be rejected."""
import struct
from pathlib import Path

import numpy as np
import pytest

from rohditor_amd import _lib, sensor, synth

GOLDEN = Path(__file__).resolve().parent / "golden" / "raw"
M = 15700 * 65536 - 1


def _table_closed_form(curve):
    k = [0] + [(v >> 2) & 0xFFF for v in curve] + [4095]
    j = np.arange(4096)
    t = sum((1 << i) * np.clip(j - k[i], 0, max(k[i + 1] - k[i], 0)) for i in range(5))
    return t.astype(np.uint16)


@pytest.mark.parametrize("curve", [synth.SONY_CURVE, (0, 0, 0, 0), (8000, 10400, 12900, 14100),
                                   (16380, 16380, 16380, 16380), (9000, 4000, 12000, 400)])
def test_curve_and_dither_lookup_closed_form(oracle, curve):
    t, base, delta = oracle.arw_curve(curve)
    knots = [(v >> 2) & 0xFFF for v in curve]
    if knots == sorted(knots):
        assert np.array_equal(t, _table_closed_form(curve))
    else:
        # Out-of-order knots: later segments overwrite earlier ones, each
        # entry climbing from its left neighbour (u16 wrap).
        k = [0] + knots + [4095]
        ref = [0] * 4096
        for i in range(5):
            for j in range(k[i] + 1, k[i + 1] + 1):
                ref[j] = (ref[j - 1] + (1 << i)) & 0xFFFF
        assert np.array_equal(t, np.array(ref, np.uint16))
    lo = np.concatenate([t[:1], t[:-1]]).astype(np.int64)
    up = np.concatenate([t[1:], t[-1:]]).astype(np.int64)
    spread = (up - lo) & 0xFFFF
    exp_base = np.where(t == 0, 0, (t.astype(np.int64) - ((spread + 2) & 0xFFFF) // 4) & 0xFFFF)
    assert np.array_equal(delta, spread.astype(np.uint16))
    assert np.array_equal(base, exp_base.astype(np.uint16))


@pytest.mark.parametrize("seed", [0, 1, 2, 12345, 65535, 40000])
def test_dither_stream_is_a_modular_power(seed):
    """r' = 15700 (r & 0xffff) + (r >> 16) from a 16-bit seed equals
    15700^n r0 mod (15700 * 2^16 - 1) and stays below the modulus: the decoder
    jumps to any block with one multiplication."""
    r = seed
    for n in range(1, 3000):
        r = 15700 * (r & 65535) + (r >> 16)
        assert r < M
        assert r == pow(15700, n, M) * seed % M


def _pack(fields):
    """fields: [(value, bits)] LSB first into 16 bytes."""
    v, pos = 0, 0
    for value, bits in fields:
        v |= (value & ((1 << bits) - 1)) << pos
        pos += bits
    assert pos == 128
    return v.to_bytes(16, "little")


def _dither_expect(curve, vals, seed, start=0):
    _, base, delta = (np.asarray(a, np.int64) for a in _curve(curve))
    out = []
    for n, v in enumerate(vals):
        r = pow(15700, start + n, M) * seed % M
        idx = (v << 1) & 0xFFFF
        out.append((base[idx] + ((delta[idx] * (r & 2047) + 1024) >> 12)) & 0xFFFF)
    return out


_CURVES = {}


def _curve(curve):
    from oracle import oracle
    key = tuple(curve)
    if key not in _CURVES:
        _CURVES[key] = oracle.arw_curve(curve)
    return _CURVES[key]


def test_hand_packed_blocks(oracle):
    # Block A (even columns): max 1500 at 3, min 1400 at 9: spread 100 < 128,
    # no shift; deltas 0..13 for the other positions in order.
    deltas_a = list(range(14))
    a = _pack([(1500, 11), (1400, 11), (3, 4), (9, 4)] + [(d, 7) for d in deltas_a])
    # Block B (odd columns): max 2047 at 0, min 100 at 15: spread 1947 (11
    # bits, shift 4); deltas 127 (2132, clamped to 0x7ff) then 1, 2, ...
    deltas_b = [127] + list(range(1, 14))
    b = _pack([(2047, 11), (100, 11), (0, 4), (15, 4)] + [(d, 7) for d in deltas_b])
    row = a + b
    vals_a, k = [], 0
    for i in range(16):
        if i == 3:
            vals_a.append(1500)
        elif i == 9:
            vals_a.append(1400)
        else:
            vals_a.append(deltas_a[k] + 1400)
            k += 1
    vals_b, k = [], 0
    for i in range(16):
        if i == 0:
            vals_b.append(2047)
        elif i == 15:
            vals_b.append(100)
        else:
            vals_b.append(min(0x7FF, (deltas_b[k] << 4) + 100))
            k += 1
    assert vals_b[1] == 0x7FF
    seed = row[0] | row[1] << 8
    exp = np.zeros(32, np.int64)
    exp[0::2] = _dither_expect(synth.SONY_CURVE, vals_a, seed, 0)
    exp[1::2] = _dither_expect(synth.SONY_CURVE, vals_b, seed, 16)
    st, out = oracle.decode_arw2(row, 32, 1, synth.SONY_CURVE)
    assert st == 0 and np.array_equal(out[0].astype(np.int64), exp)


def _vector_decode(codes_dec, strip, width, curve):
    """Expected output for rows whose decoded 11-bit values are `codes_dec`:
    curve lookup at value * 2 dithered by the row's stream in block order."""
    h = codes_dec.shape[0]
    _, base, delta = (np.asarray(a, np.int64) for a in _curve(curve))
    raw = np.frombuffer(strip, np.uint8).reshape(h, width).astype(np.int64)
    seeds = raw[:, 0] | raw[:, 1] << 8
    n = np.arange(width)
    # pixel column c = 32 g + j + 2 i has stream index 32 g + 16 j + i
    g, rem = n // 32, n % 32
    order = 32 * g + 16 * (rem % 2) + rem // 2
    powers = np.array([pow(15700, int(k), M) for k in order], dtype=np.int64)
    r = (powers[None, :] * seeds[:, None]) % M
    idx = (codes_dec.astype(np.int64) << 1) & 0xFFFF
    return ((base[idx] + ((delta[idx] * (r & 2047) + 1024) >> 12)) & 0xFFFF).astype(np.uint16)


def _quantised(codes):
    """The values the decoder returns for what encode_arw2 packs."""
    h, w = codes.shape
    v = codes.astype(np.int64).reshape(h, w // 32, 16, 2).transpose(0, 1, 3, 2)
    vmax, vmin = v.max(-1, keepdims=True), v.min(-1, keepdims=True)
    bl = np.zeros_like(vmax)
    d = vmax - vmin
    while np.any(d):
        bl += d > 0
        d >>= 1
    shift = np.maximum(bl - 7, 0)
    q = np.clip((v - vmin) >> shift, 0, 127)
    dec = np.minimum((q << shift) + vmin, 0x7FF)
    dec = np.where(v == vmax, vmax, dec)
    dec = np.where(v == vmin, vmin, dec)
    return dec.transpose(0, 1, 3, 2).reshape(h, w)


@pytest.mark.parametrize("spread", [100, 2047])
def test_encoded_rows_decode_as_the_vectorised_stream(oracle, spread):
    rng = np.random.default_rng(spread)
    h, w = 5, 256
    base = rng.integers(0, 2048 - min(spread, 2047), (h, w // 32, 1, 2))
    codes = (base + rng.integers(0, spread + 1, (h, w // 32, 16, 2))).clip(0, 2047)
    codes = codes.transpose(0, 1, 3, 2).reshape(h, w // 32, 2, 16)
    codes = codes.transpose(0, 1, 3, 2).reshape(h, w).astype(np.uint16)
    strip = synth.encode_arw2(codes)
    st, out = oracle.decode_arw2(strip, w, h, synth.SONY_CURVE)
    assert st == 0
    exp = _vector_decode(_quantised(codes), strip, w, synth.SONY_CURVE)
    assert np.array_equal(out, exp)


def test_tail_columns_stay_zero_and_short_strips_fail(oracle):
    rng = np.random.default_rng(3)
    strip = rng.integers(0, 256, 40 * 3, dtype=np.uint8).tobytes()
    st, out = oracle.decode_arw2(strip, 40, 3, synth.SONY_CURVE)
    assert st == 0 and not out[:, 32:].any() and out[:, :32].any()
    st, out = oracle.decode_arw2(strip[:16], 16, 1, synth.SONY_CURVE)
    assert st == 0 and not out.any()
    assert oracle.decode_arw2(strip[:100], 40, 3, synth.SONY_CURVE)[0] == 1


def _tiff(entries, le=True, next_ifd=0, tail=b""):
    e = "<" if le else ">"
    head = (b"II*\0" if le else b"MM\0*") + struct.pack(e + "I", 8)
    body = struct.pack(e + "H", len(entries))
    for tag, typ, count, val in entries:
        body += struct.pack(e + "HHI", tag, typ, count) + val.ljust(4, b"\0")
    return head + body + struct.pack(e + "I", next_ifd) + tail


def test_probe_finds_the_raw_strip_and_rejects_damaged_files(oracle):
    rng = np.random.default_rng(9)
    w, h = 96, 4
    strip = synth.encode_arw2(rng.integers(0, 2048, (h, w)).astype(np.uint16))
    f = synth.arw_file(strip, w, h)
    st, lay = oracle.arw_probe(f)
    assert st == 0 and lay["kind"] == 1 and (lay["width"], lay["height"]) == (w, h)
    assert f[lay["offset"]:lay["offset"] + lay["bytes"]] == strip
    assert lay["curve"] == list(synth.SONY_CURVE)
    # uncompressed 16-bit little-endian
    u16 = rng.integers(0, 16384, (h, w)).astype("<u2").tobytes()
    g = synth.arw_file(u16, w, h, compression=1, bits_per_sample=16)
    st, lay = oracle.arw_probe(g)
    assert st == 0 and lay["kind"] == 2 and g[lay["offset"]:lay["offset"] + lay["bytes"]] == u16
    # the reference's corrupt fixtures: not TIFF containers at all
    for name in ("not-a-raw.ARW", "truncated-tiff.ARW"):
        assert oracle.arw_probe((GOLDEN / name).read_bytes())[0] == 1
    assert oracle.arw_probe(b"")[0] == 1
    # truncated strip, strip past the end, IFD outside the file, IFD loop
    assert oracle.arw_probe(f[:-1])[0] == 2
    assert oracle.arw_probe(_tiff([], next_ifd=1 << 20))[0] == 2
    assert oracle.arw_probe(b"II*\0" + struct.pack("<I", 4096))[0] == 2
    assert oracle.arw_probe(_tiff([], next_ifd=8))[0] == 2
    # a TIFF with no raw IFD, and a big-endian uncompressed strip
    assert oracle.arw_probe(_tiff([(0x0100, 4, 1, struct.pack("<I", 5))]))[0] == 3
    be = _tiff([(0x0100, 4, 1, struct.pack(">I", 2)), (0x0101, 4, 1, struct.pack(">I", 1)),
                (0x0102, 3, 1, struct.pack(">H", 16)), (0x0103, 3, 1, struct.pack(">H", 1)),
                (0x0111, 4, 1, struct.pack(">I", 8)), (0x0117, 4, 1, struct.pack(">I", 4))],
               le=False)
    assert oracle.arw_probe(be)[0] == 3


def test_library_probe_matches_the_oracle(oracle):
    """rhd_arw_probe is host code: it runs without a GPU."""
    rng = np.random.default_rng(11)
    w, h = 64, 6
    strip = synth.encode_arw2(rng.integers(0, 2048, (h, w)).astype(np.uint16))
    f = synth.arw_file(strip, w, h, curve=(8000, 10400, 12900, 14100))
    lay = sensor.arw_probe(f)
    st, exp = oracle.arw_probe(f)
    assert st == 0
    got = {"width": lay.width, "height": lay.height, "bits_per_sample": lay.bits_per_sample,
           "compression": lay.compression, "offset": lay.offset, "bytes": lay.bytes,
           "curve": list(lay.curve), "kind": lay.kind}
    assert got == {k: v for k, v in exp.items() if k != "has_curve"}
    status = {1: _lib.RAW_UNSUPPORTED, 2: _lib.RAW_CORRUPT, 3: _lib.RAW_UNSUPPORTED}
    for bad in [(GOLDEN / "not-a-raw.ARW").read_bytes(), (GOLDEN / "truncated-tiff.ARW").read_bytes(),
                f[:-1], _tiff([], next_ifd=8), _tiff([(0x0100, 4, 1, struct.pack("<I", 5))]), b""]:
        st, _ = oracle.arw_probe(bad)
        with pytest.raises(sensor.GpuPreviewError) as e:
            sensor.arw_probe(bad)
        assert e.value.status == status[st]


def _stream_decode(strip, width, height, curve):
    """Bit-by-bit reference reader: one continuous LSB-first stream per row
    starting at the row's first byte and running on past it (zeros past the
    strip's end); 15 deltas when the max and min positions coincide."""
    _, base, delta = (np.asarray(a, np.int64) for a in _curve(curve))
    data = int.from_bytes(strip, "little")
    total = len(strip) * 8
    out = np.zeros((height, width), np.uint16)

    def take(pos, n):
        if pos >= total:
            return 0
        return (data >> pos) & ((1 << n) - 1) if pos + n <= total else (data >> pos)

    for row in range(height):
        pos = row * width * 8
        r = take(pos, 16)
        for g in range(width // 32):
            for j in range(2):
                mx, mn = take(pos, 11), take(pos + 11, 11)
                imax, imin = take(pos + 22, 4), take(pos + 26, 4)
                pos += 30
                sh = max(((mx - mn) & 0xFFFFFFFF).bit_length() - 7, 0)
                for i in range(16):
                    if i == imax:
                        v = mx
                    elif i == imin:
                        v = mn
                    else:
                        v = min(0x7FF, (((take(pos, 7)) << sh) + mn) & 0xFFFFFFFF)
                        pos += 7
                    idx = (v << 1) & 0xFFFF
                    out[row, 32 * g + j + 2 * i] = (base[idx] + ((delta[idx] * (r & 2047) + 1024)
                                                                 >> 12)) & 0xFFFF
                    r = (15700 * (r & 65535) + (r >> 16)) & 0xFFFFFFFF
    return out


def test_coinciding_positions_misalign_the_rest_of_the_row(oracle):
    """Random bytes: blocks whose max and min positions coincide carry a 15th
    delta, and the row's later blocks start 7 bits later."""
    rng = np.random.default_rng(17)
    w, h = 96, 6
    strip = rng.integers(0, 256, w * h, dtype=np.uint8).tobytes()
    st, out = oracle.decode_arw2(strip, w, h, synth.SONY_CURVE)
    assert st == 0
    assert np.array_equal(out, _stream_decode(strip, w, h, synth.SONY_CURVE))
