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
