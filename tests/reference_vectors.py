"""Inputs and expected values held by the reference's own tests.

Each generator restates an input builder from a reference test (integer
arithmetic and IEEE f32 operations only, so the values are bit-identical), and
each constant is an expected value copied from that test as data.
"""
from __future__ import annotations

import numpy as np

F = np.float32

# crates/demosaic/tests/amaze_golden.rs:74-83
AMAZE_GOLDEN_HASHES = {
    ("RGGB", "identity"): 0xc60f3fec92045b18,
    ("RGGB", "balanced"): 0xd57793d2c96ccba7,
    ("BGGR", "identity"): 0x4ee5cebe694421fc,
    ("BGGR", "balanced"): 0xb9c80f0cf7e9638e,
    ("GRBG", "identity"): 0x2a4e9926ffacb7e2,
    ("GRBG", "balanced"): 0x269fc286e320f8b8,
    ("GBRG", "identity"): 0xaa4084ee82afc286,
    ("GBRG", "balanced"): 0x6fa2848f15978563,
}
# crates/demosaic/tests/amaze_golden.rs:57-61
AMAZE_GOLDEN_BALANCE = (2.1, 1.0, 2.6)


def amaze_golden_mosaic() -> np.ndarray:
    """crates/demosaic/tests/amaze_golden.rs:18-41: 513x389 LCG mosaic."""
    width, height = 513, 389
    n = width * height
    states = np.empty(n, dtype=np.uint32)
    state = 12345
    for i in range(n):
        state = (state * 1664525 + 1013904223) & 0xFFFFFFFF
        states[i] = state
    noise = (states >> np.uint32(8)).astype(F) / F(16777216.0)
    idx = np.arange(n)
    x = idx % width
    y = idx // width
    stripes = ((x * 3 + y * 5) % 7).astype(F) / F(7.0)
    gradient = (x + 2 * y).astype(F) / (F(3.0) * F(width))
    value = F(0.1) + F(0.5) * stripes + F(0.3) * gradient + F(0.2) * (noise - F(0.5))
    value = np.maximum(value, F(-0.05))
    clipped = ((y // 40 + x // 60) % 5) == 0
    dim = ((x // 90 + y // 70) % 4) == 1
    out = np.where(clipped, F(1.1), np.where(dim, F(0.35) + F(0.1) * noise, value)).astype(F)
    return out.reshape(height, width)


def fnv1a_f32(values: np.ndarray) -> int:
    """crates/demosaic/tests/amaze_golden.rs:44-48: FNV-1a over f32 bit patterns."""
    bits = np.ascontiguousarray(values, dtype=np.float32).view(np.uint32).ravel()
    h = 0xcbf29ce484222325
    prime = 0x100000001b3
    mask = 0xFFFFFFFFFFFFFFFF
    for b in bits.tolist():
        h = ((h ^ b) * prime) & mask
    return h


def amaze_stage_fixture() -> np.ndarray:
    """crates/demosaic/src/amaze/mod.rs:280-299: 53x47 signed fixture (RGGB)."""
    width, height = 53, 47
    data = np.empty(width * height, dtype=F)
    for i in range(width * height):
        x, y = i % width, i // width
        k = (x + 3 * y) % 11
        if k == 0:
            v = F(-0.0)
        elif k == 1:
            v = F(1.0e-6)
        elif k == 2:
            v = F(-1.0e-6)
        elif k == 3:
            v = F(1.5)
        elif k == 4:
            v = F(-0.2)
        elif k == 5:
            v = F(2.0)
        else:
            v = F((x * 37 + y * 19 + x * y) % 101) / F(63.0) - F(0.4)
        data[i] = v
    return data.reshape(height, width)


# crates/demosaic/src/amaze/mod.rs:368-424: (plane, tile row, tile col, value);
# `half` planes are indexed by (row * 128 + col / 2).
AMAZE_STAGE_PROBES = [
    ("nyqutest", 25, 57, 0.08931726), ("nyquist", 25, 57, 1.0), ("nyquist2", 25, 57, 1.0),
    ("hvwt", 25, 57, 0.6249956), ("dgrb2_h", 25, 57, 0.087723866),
    ("dgrb2_v", 25, 57, 0.13902828), ("dgrb0", 25, 57, 0.7487581), ("dgrb1", 25, 57, 1.1797017),
    ("green", 25, 57, 1.1797007), ("pmwt", 25, 57, 0.59881574), ("rbint", 25, 57, -0.018249098),
    ("nyquist2", 26, 56, 1.0), ("hvwt", 26, 56, 0.77032536), ("dgrb0", 26, 56, 0.043839097),
    ("dgrb1", 26, 56, 0.51848173), ("green", 26, 56, -0.1561609),
    ("nyqutest", 32, 32, 0.051671088), ("nyquist", 32, 32, 1.0), ("nyquist2", 32, 32, 0.0),
    ("hvwt", 32, 32, 0.22237617), ("dgrb2_h", 32, 32, 0.0), ("green", 32, 32, 0.7930795),
    ("cfa", 31, 33, 0.5206349), ("vcd", 31, 33, 0.1482284), ("hcd", 31, 33, -0.8253968),
    ("vcdalt", 31, 33, 0.1482284), ("hcdalt", 31, 33, -0.8630527),
    ("cddiffsq", 31, 33, 0.947946), ("dgintv", 31, 33, 1.6694309),
    ("dginth", 31, 33, 0.58292013), ("delhvsqsum", 31, 33, 1.671696),
    ("dirwts0", 31, 33, 2.3333435), ("dirwts1", 31, 33, 1.5746133), ("delp", 31, 33, 1.7492063),
    ("delm", 31, 33, 1.5), ("sq1m", 31, 33, 2.0440006), ("sq1p", 31, 33, 1.7466767),
    ("hvwt", 31, 33, 0.16582216), ("pmwt", 31, 33, 0.5514841), ("rbm", 31, 33, 0.69720316),
    ("rbp", 31, 33, 0.85232896), ("rbint", 31, 33, 0.6516937), ("dgrb0", 31, 33, -0.929778),
    ("dgrb1", 31, 33, -0.6639482), ("green", 31, 33, -0.14331329), ("vcd", 33, 35, 0.4021777),
    ("hcd", 33, 35, 1.7952101), ("cddiffsq", 33, 35, 1.9405394), ("delp", 33, 35, 0.098411694),
    ("delm", 33, 35, 0.04444444), ("hvwt", 33, 35, 0.1975245), ("pmwt", 33, 35, 0.60838723),
    ("rbint", 33, 35, -0.12558362), ("dgrb0", 33, 35, 1.3346572), ("dgrb1", 33, 35, 1.5200521),
    ("green", 33, 35, 1.320052),
]
# crates/demosaic/src/amaze/mod.rs:427-428: (nyquist sites, voted sites)
AMAZE_STAGE_NYQUIST_COUNTS = (1428, 643)
HALF_PLANES = {"hvwt", "dgrb0", "dgrb1", "dgrb2_h", "dgrb2_v", "delp", "delm", "sq1m", "sq1p",
               "nyqutest", "pmwt", "rbm", "rbp", "rbint", "nyquist", "nyquist2"}
