"""Seeded synthetic Bayer mosaics for the benchmark workloads.

Scene model (linear camera RGB, one value per CFA site): smooth per-channel
gradients, random hard-edged rectangles (the brightest scaled so a target
fraction of sites exceeds white), optional saturated Gaussian highlight blobs
(some green-only, so Opposed has opposing support), a WB-like channel scale,
then Poisson shot noise in electrons plus Gaussian read noise, rounded and
clamped to the sensor range. numpy's PCG64 stream makes every mosaic a pure
function of its seed.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .sensor import PATTERNS, ImageRect, LevelPattern, RawFileInfo, RawFrame

# Channel scale standing in for a camera's native response (as-shot WB 2.0/1.0/1.43).
CHANNEL_SCALE = np.array([0.5, 1.0, 0.7])
AS_SHOT = (2.0, 1.0, 1.0 / 0.7)


@dataclass
class SceneSpec:
    width: int
    height: int
    seed: int
    cfa: str = "RGGB"
    black: float = 512.0
    white: float = 16383.0
    clipped_fraction: float = 0.02
    blob_fraction: float = 0.0
    rectangles: int = 40
    gain: float = 0.5  # DN per electron
    read_sigma: float = 3.0


def _channel_map(pattern: int, width: int, height: int) -> np.ndarray:
    table = {0: (0, 1, 1, 2), 1: (2, 1, 1, 0), 2: (1, 0, 2, 1), 3: (1, 2, 0, 1)}[pattern]
    ch = np.empty((2, 2), dtype=np.int8)
    for y in range(2):
        for x in range(2):
            ch[y, x] = table[(y << 1) | x]
    return np.tile(ch, ((height + 1) // 2, (width + 1) // 2))[:height, :width]


def scene(spec: SceneSpec) -> np.ndarray:
    """Linear scene value at every CFA site (relative to white = 1)."""
    rng = np.random.default_rng(spec.seed)
    w, h = spec.width, spec.height
    ch = _channel_map(PATTERNS[spec.cfa], w, h)
    a = rng.uniform(0.5, 1.5, 3)
    b = rng.uniform(0.5, 1.5, 3)
    xs = (np.arange(w, dtype=np.float64) / w)[None, :]
    ys = (np.arange(h, dtype=np.float64) / h)[:, None]
    out = 0.15 + 0.35 * a[ch] * xs + 0.25 * b[ch] * ys
    rects = []
    for _ in range(spec.rectangles):
        rw = int(rng.uniform(w / 40, w / 6))
        rh = int(rng.uniform(h / 40, h / 6))
        x0 = int(rng.integers(0, max(1, w - rw)))
        y0 = int(rng.integers(0, max(1, h - rh)))
        colour = rng.uniform(0.05, 0.9, 3)
        rects.append([x0, y0, rw, rh, colour])
    # Scale the brightest rectangles until they cover the clipped fraction.
    order = sorted(range(len(rects)), key=lambda k: -float(np.max(rects[k][4] * CHANNEL_SCALE)))
    covered, target = 0.0, spec.clipped_fraction * w * h
    for k in order:
        if covered >= target:
            break
        r = rects[k]
        r[4] = r[4] * (1.3 / float(np.max(r[4] * CHANNEL_SCALE)))
        clips = (r[4] * CHANNEL_SCALE) > 1.0
        covered += r[2] * r[3] * (0.25 * clips[0] + 0.5 * clips[1] + 0.25 * clips[2])
    for x0, y0, rw, rh, colour in rects:
        out[y0:y0 + rh, x0:x0 + rw] = colour[ch[y0:y0 + rh, x0:x0 + rw]] / CHANNEL_SCALE[
            ch[y0:y0 + rh, x0:x0 + rw]]
    # Saturated Gaussian highlight blobs.
    area, target = 0.0, spec.blob_fraction * w * h
    while area < target:
        radius = rng.uniform(20, 200)
        sigma = radius / 2.0
        peak = rng.uniform(1.5, 4.0)
        green_only = rng.random() < 0.35
        tint = np.array([0.0, 1.0, 0.0]) if green_only else np.array([1.0, 1.0, 1.0])
        cx, cy = rng.uniform(0, w), rng.uniform(0, h)
        x0, x1 = int(max(0, cx - 3 * sigma)), int(min(w, cx + 3 * sigma + 1))
        y0, y1 = int(max(0, cy - 3 * sigma)), int(min(h, cy + 3 * sigma + 1))
        if x1 <= x0 or y1 <= y0:
            continue
        gx = np.exp(-0.5 * ((np.arange(x0, x1) - cx) / sigma) ** 2)[None, :]
        gy = np.exp(-0.5 * ((np.arange(y0, y1) - cy) / sigma) ** 2)[:, None]
        sub = ch[y0:y1, x0:x1]
        out[y0:y1, x0:x1] += peak * (gy * gx) * tint[sub] / CHANNEL_SCALE[sub]
        area += np.pi * sigma * sigma * 2.0 * np.log(peak)
    return out * CHANNEL_SCALE[ch]


def mosaic(spec: SceneSpec) -> np.ndarray:
    """u16 mosaic (height, width) with shot and read noise."""
    rng = np.random.default_rng(spec.seed ^ 0x5EED)
    s = scene(spec)
    span = spec.white - spec.black
    electrons = np.clip(s * span, 0.0, 4.0 * span) / spec.gain
    dn = spec.black + rng.poisson(electrons).astype(np.float64) * spec.gain
    dn += rng.normal(0.0, spec.read_sigma, dn.shape)
    return np.clip(np.rint(dn), 0, spec.white).astype(np.uint16)


def frame(spec: SceneSpec) -> RawFrame:
    info = RawFileInfo(width=spec.width, height=spec.height, cfa=spec.cfa,
                       black_levels=LevelPattern([spec.black]), white_levels=(spec.white,),
                       as_shot_white_balance=AS_SHOT,
                       active_area=ImageRect(0, 0, spec.width, spec.height))
    return RawFrame(info=info, mosaic=mosaic(spec))


def frame_from_mosaic(samples: np.ndarray, cfa: str, black: float = 512.0,
                      white: float = 16383.0) -> RawFrame:
    """A RawFrame around existing u16 samples, labelled with `cfa`."""
    h, w = samples.shape
    info = RawFileInfo(width=w, height=h, cfa=cfa, black_levels=LevelPattern([black]),
                       white_levels=(white,), as_shot_white_balance=AS_SHOT,
                       active_area=ImageRect(0, 0, w, h))
    return RawFrame(info=info, mosaic=samples)


# Workloads (seeds and sizes of the benchmark configurations).
def cpu_case() -> SceneSpec:
    """1536x1024 RGGB, 14-bit, 2 % clipped; normalize + Clip + AMaZE."""
    return SceneSpec(1536, 1024, 0x52480001)


def full_sensor_24mp() -> SceneSpec:
    """6048x4024 RGGB, 14-bit, 2 % clipped rectangles plus 5 % saturated blobs."""
    return SceneSpec(6048, 4024, 0x52480002, blob_fraction=0.05)


# Sony compressed ARW (TIFF compression 32767): tone-curve knots of the tag
# 0x7010 (each value >> 2 is a 12-bit knot) whose table ends at 15352.
SONY_CURVE = (4096, 8192, 12288, 16380)


def sony_curve_table(curve=SONY_CURVE) -> np.ndarray:
    """The 4096-entry tone curve the four knots define (segment i climbs 2^i)."""
    knots = [0] + [(int(v) >> 2) & 0xFFF for v in curve] + [4095]
    out = np.zeros(4096, np.int64)
    for i in range(5):
        for j in range(knots[i] + 1, knots[i + 1] + 1):
            out[j] = out[j - 1] + (1 << i)
    return out.astype(np.uint16)


def sony_codes(samples: np.ndarray, curve=SONY_CURVE) -> np.ndarray:
    """11-bit codes whose curve value (at index code * 2) is nearest below or
    at each sample."""
    table = sony_curve_table(curve)[0::2].astype(np.int64)
    idx = np.searchsorted(table, samples.astype(np.int64), side="right") - 1
    return np.clip(idx, 0, 2047).astype(np.uint16)


def encode_arw2(codes: np.ndarray) -> bytes:
    """Packs (H, W) 11-bit codes, W a multiple of 32, into Sony's compressed
    rows: per 32 pixels two 16-byte blocks (even columns, then odd), each an
    11-bit max, 11-bit min, 4-bit positions of both and fourteen 7-bit deltas
    scaled by 2^max(0, bit_length(max - min) - 7), read LSB first."""
    h, w = codes.shape
    if w % 32:
        raise ValueError("width must be a multiple of 32")
    v = codes.astype(np.int64).reshape(h, w // 32, 16, 2).transpose(0, 1, 3, 2).reshape(-1, 16)
    vmax, vmin = v.max(1), v.min(1)
    imax = v.argmax(1)
    imin = v.argmin(1)
    same = imax == imin  # all 16 equal: any two distinct positions
    imax = np.where(same, 0, imax)
    imin = np.where(same, 1, imin)
    delta = vmax - vmin
    bl = np.zeros_like(delta)
    d = delta.copy()
    while np.any(d):
        bl += d > 0
        d >>= 1
    shift = np.maximum(bl - 7, 0)
    q = np.clip((v - vmin[:, None]) >> shift[:, None], 0, 127)
    bits = np.zeros((v.shape[0], 128), np.uint8)

    def put(start, width, value):
        for k in range(width):
            bits[:, start + k] = (value >> k) & 1

    put(0, 11, vmax)
    put(11, 11, vmin)
    put(22, 4, imax)
    put(26, 4, imin)
    pos = np.full(v.shape[0], 30)
    rows = np.arange(v.shape[0])
    for i in range(16):
        keep = (imax != i) & (imin != i)
        for k in range(7):
            bits[rows[keep], pos[keep] + k] = (q[keep, i] >> k) & 1
        pos = pos + 7 * keep
    return np.packbits(bits, axis=1, bitorder="little").tobytes()


def arw_file(strip: bytes, width: int, height: int, curve=SONY_CURVE, compression: int = 32767,
             bits_per_sample: int = 14) -> bytes:
    """A minimal little-endian ARW-like TIFF: IFD0 (a tiny placeholder) whose
    SubIFD holds the raw strip with its width, height, bits per sample,
    compression, strip offset and byte count, and the Sony curve (0x7010)."""
    import struct

    def ifd(entries, at, next_ifd=0):
        # entries: (tag, type, count, value bytes); values over 4 bytes go after.
        out = struct.pack("<H", len(entries))
        extra = b""
        data_at = at + 2 + 12 * len(entries) + 4
        for tag, typ, count, val in entries:
            if len(val) <= 4:
                out += struct.pack("<HHI", tag, typ, count) + val.ljust(4, b"\0")
            else:
                out += struct.pack("<HHII", tag, typ, count, data_at + len(extra))
                extra += val + (b"\0" if len(val) % 2 else b"")
        return out + struct.pack("<I", next_ifd) + extra

    ifd0_at = 8
    ifd0_size = 2 + 12 * 2 + 4
    raw_at = ifd0_at + ifd0_size
    raw_entries = [
        (0x0100, 4, 1, struct.pack("<I", width)),
        (0x0101, 4, 1, struct.pack("<I", height)),
        (0x0102, 3, 1, struct.pack("<H", bits_per_sample)),
        (0x0103, 3, 1, struct.pack("<H", compression)),
        (0x0111, 4, 1, b"\0\0\0\0"),
        (0x0117, 4, 1, struct.pack("<I", len(strip))),
        (0x7010, 3, 4, struct.pack("<4H", *curve)),
    ]
    raw_size = len(ifd(raw_entries, raw_at))
    strip_at = raw_at + raw_size
    strip_at += (-strip_at) % 16
    raw_entries[4] = (0x0111, 4, 1, struct.pack("<I", strip_at))
    raw = ifd(raw_entries, raw_at)
    ifd0 = ifd([(0x00FE, 4, 1, struct.pack("<I", 1)), (0x014A, 4, 1, struct.pack("<I", raw_at))],
               ifd0_at)
    head = b"II*\0" + struct.pack("<I", ifd0_at) + ifd0 + raw
    return head + b"\0" * (strip_at - len(head)) + strip
