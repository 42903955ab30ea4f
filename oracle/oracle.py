"""Oracle (test infrastructure only): ctypes bindings to the C++ restatement.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import
this module, and only as the checker; the product path never does.
"""
from __future__ import annotations

import contextlib
import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "_build" / "librhd_oracle.so"
_lib = None

ALGORITHMS = {"bilinear": 0, "mhc": 1, "rcd": 2, "amaze": 3}
PATTERNS = {"RGGB": 0, "BGGR": 1, "GRBG": 2, "GBRG": 3}
AMAZE_PLANES = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)


def lib():
    global _lib, AMAZE_PLANES
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(_LIB_PATH))
        _lib.rhdo_amaze_plane_names.restype = ctypes.c_char_p
        AMAZE_PLANES = _lib.rhdo_amaze_plane_names().decode().split(",")
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _sz(v) -> ctypes.c_size_t:
    return ctypes.c_size_t(int(v))


def normalize(mosaic: np.ndarray, sensor_pattern: int, black: np.ndarray, repeat_w: int,
              repeat_h: int, white: np.ndarray, crop) -> np.ndarray:
    """crates/core/src/pipeline/sensor.rs:221-277 at full crop size."""
    mosaic = np.ascontiguousarray(mosaic, dtype=np.uint16)
    black = np.ascontiguousarray(black, dtype=np.float32)
    white = np.ascontiguousarray(white, dtype=np.float32)
    cx, cy, cw, ch = crop
    out = np.empty((ch, cw), dtype=np.float32)
    lib().rhdo_normalize(_p(mosaic), _sz(mosaic.shape[1]), ctypes.c_int(sensor_pattern), _p(black),
                         _sz(repeat_w), _sz(repeat_h), _p(white), _sz(white.size), _sz(cx),
                         _sz(cy), _sz(cw), _sz(ch), _p(out))
    return out


def normalize_f64(mosaic: np.ndarray, sensor_pattern: int, black: np.ndarray, repeat_w: int,
                  repeat_h: int, white: np.ndarray, crop) -> np.ndarray:
    """The FP64 twin of normalize."""
    mosaic = np.ascontiguousarray(mosaic, dtype=np.uint16)
    black = np.ascontiguousarray(black, dtype=np.float32)
    white = np.ascontiguousarray(white, dtype=np.float32)
    cx, cy, cw, ch = crop
    out = np.empty((ch, cw), dtype=np.float64)
    lib().rhdo_normalize_f64(_p(mosaic), _sz(mosaic.shape[1]), ctypes.c_int(sensor_pattern),
                             _p(black), _sz(repeat_w), _sz(repeat_h), _p(white), _sz(white.size),
                             _sz(cx), _sz(cy), _sz(cw), _sz(ch), _p(out))
    return out


def opposed_f64(mosaic: np.ndarray, pattern: int, detection, neutral_ceiling=None):
    """The FP64 twin of opposed. Returns (status, out, stats[9])."""
    m = np.ascontiguousarray(mosaic, dtype=np.float64).copy()
    h, w = m.shape
    det = np.asarray(detection, dtype=np.float64)
    stats = np.zeros(9, dtype=np.uint64)
    ceil_ptr = ctypes.c_void_p(0)
    if neutral_ceiling is not None:
        ceil = np.asarray(neutral_ceiling, dtype=np.float64)
        ceil_ptr = _p(ceil)
    st = lib().rhdo_opposed_f64(_p(m), _sz(w), _sz(h), ctypes.c_int(pattern), _p(det), ceil_ptr,
                                _p(stats))
    return st, m, stats


def clip(mosaic: np.ndarray, pattern: int, limits) -> tuple[int, np.ndarray, np.ndarray]:
    """crates/highlight/src/clip.rs:24-72. Returns (status, clipped, stats[6])."""
    m = np.ascontiguousarray(mosaic, dtype=np.float32).copy()
    lim = np.asarray(limits, dtype=np.float32)
    stats = np.zeros(6, dtype=np.uint64)
    st = lib().rhdo_clip(_p(m), _sz(m.shape[1]), _sz(m.shape[0]), ctypes.c_int(pattern), _p(lim),
                         _p(stats))
    return st, m, stats


def demosaic(mosaic: np.ndarray, pattern: int, balance, algorithm: str, threads: int = 1,
             poison: bool = False, fp64: bool = False) -> tuple[int, np.ndarray]:
    """crates/demosaic/src/lib.rs:302-375: interleaved RGB (H, W, 3)."""
    dt = np.float64 if fp64 else np.float32
    m = np.ascontiguousarray(mosaic, dtype=dt)
    bal = np.asarray(balance, dtype=dt)
    out = np.zeros((m.shape[0], m.shape[1], 3), dtype=dt)
    if fp64:
        st = lib().rhdo_demosaic_f64(_p(m), _sz(m.shape[1]), _sz(m.shape[0]),
                                     ctypes.c_int(pattern), _p(bal),
                                     ctypes.c_int(ALGORITHMS[algorithm]), _p(out),
                                     ctypes.c_int(threads))
    else:
        st = lib().rhdo_demosaic_f32(_p(m), _sz(m.shape[1]), _sz(m.shape[0]),
                                     ctypes.c_int(pattern), _p(bal),
                                     ctypes.c_int(ALGORITHMS[algorithm]), _p(out),
                                     ctypes.c_int(threads), ctypes.c_int(int(poison)))
    return st, out


def amaze_tile_planes(mosaic: np.ndarray, pattern: int, origin=(0, 0),
                      clip_point: float = 1.0) -> dict[str, np.ndarray]:
    """Every scratch plane of one AMaZE tile (crates/demosaic/src/amaze/mod.rs:355-430)."""
    m = np.ascontiguousarray(mosaic, dtype=np.float32)
    lib()
    planes = np.zeros((len(AMAZE_PLANES), 256 * 256), dtype=np.float32)
    lib().rhdo_amaze_tile_planes(_p(m), _sz(m.shape[1]), _sz(m.shape[0]), ctypes.c_int(pattern),
                                 _sz(origin[0]), _sz(origin[1]), ctypes.c_float(clip_point),
                                 _p(planes))
    return {name: planes[k] for k, name in enumerate(AMAZE_PLANES)}


def capture(rgb: np.ndarray, amount: float, sigma: float, noise_protection: float, ceilings,
            iterations: int = 8, threads: int = 1, fp64: bool = False, planes: bool = False):
    """crates/core/src/sharpening.rs:87-310 on interleaved RGB; returns a copy."""
    h, w = rgb.shape[:2]
    if fp64:
        out = np.ascontiguousarray(rgb, dtype=np.float64).copy()
        ceil = np.asarray(ceilings, dtype=np.float64)
        st = lib().rhdo_capture_f64(_p(out), _sz(w), _sz(h), ctypes.c_double(amount),
                                    ctypes.c_double(sigma), ctypes.c_double(noise_protection),
                                    _p(ceil), ctypes.c_int(iterations), ctypes.c_int(threads))
        return st, out
    out = np.ascontiguousarray(rgb, dtype=np.float32).copy()
    ceil = np.asarray(ceilings, dtype=np.float32)
    guide = np.zeros((h, w), np.float32) if planes else None
    mask = np.zeros((h, w), np.float32) if planes else None
    est = np.zeros((h, w), np.float32) if planes else None
    null = ctypes.c_void_p(0)
    st = lib().rhdo_capture_f32(_p(out), _sz(w), _sz(h), ctypes.c_float(amount),
                                ctypes.c_float(sigma), ctypes.c_float(noise_protection), _p(ceil),
                                ctypes.c_int(iterations), _p(guide) if planes else null,
                                _p(mask) if planes else null, _p(est) if planes else null,
                                ctypes.c_int(threads))
    if planes:
        return st, out, (guide, mask, est)
    return st, out


def blur(plane: np.ndarray, sigma: float) -> np.ndarray:
    """The capture blur (crates/core/src/sharpening.rs:328-384) of one plane."""
    x = np.ascontiguousarray(plane, dtype=np.float32)
    out = np.empty_like(x)
    st = lib().rhdo_blur_f32(_p(x), _sz(x.shape[1]), _sz(x.shape[0]), ctypes.c_float(sigma),
                             _p(out))
    assert st == 0
    return out


def gaussian_kernel(sigma: float) -> np.ndarray:
    """crates/detail/src/gaussian.rs:9-17 with the platform expf."""
    buf = np.zeros(64, dtype=np.float32)
    n = lib().rhdo_gaussian_kernel(ctypes.c_float(sigma), _p(buf), ctypes.c_int(64))
    return buf[:n].copy()


def opposed(mosaic: np.ndarray, pattern: int, detection, neutral_ceiling=None):
    """crates/highlight/src/opposed.rs:63-180. Returns (status, out, stats[9], labels, count)."""
    m = np.ascontiguousarray(mosaic, dtype=np.float32).copy()
    h, w = m.shape
    det = np.asarray(detection, dtype=np.float32)
    stats = np.zeros(9, dtype=np.uint64)
    labels = np.zeros(((h + 1) // 2, (w + 1) // 2), dtype=np.uint32)
    count = ctypes.c_uint32(0)
    if neutral_ceiling is None:
        ceil_ptr = ctypes.c_void_p(0)
    else:
        ceil = np.asarray(neutral_ceiling, dtype=np.float32)
        ceil_ptr = _p(ceil)
    st = lib().rhdo_opposed(_p(m), _sz(w), _sz(h), ctypes.c_int(pattern), _p(det), ceil_ptr,
                            _p(stats), _p(labels), ctypes.byref(count))
    return st, m, stats, labels, int(count.value)


def rcd_probe(mosaic: np.ndarray, pattern: int, red: int, green: int) -> np.ndarray:
    """Stage values of the first RCD tile (crates/demosaic/src/rcd.rs:355-466)."""
    m = np.ascontiguousarray(mosaic, dtype=np.float32)
    out = np.zeros(9, dtype=np.float32)
    lib().rhdo_rcd_probe(_p(m), _sz(m.shape[1]), _sz(m.shape[0]), ctypes.c_int(pattern),
                         _sz(red), _sz(green), _p(out))
    return out


def label_components(mask: np.ndarray) -> tuple[np.ndarray, int]:
    """Compact raster-order 8-connected labels (crates/highlight/src/regions.rs:188-271)."""
    lab = np.ascontiguousarray(mask != 0, dtype=np.uint32)
    count = ctypes.c_uint32(0)
    lib().rhdo_label_components(_p(lab), _sz(lab.shape[1]), _sz(lab.shape[0]), ctypes.byref(count))
    return lab, int(count.value)


# Optics and area reduction (oracle/optics.hpp). `desc` is any ctypes
# structure with the rhd_optics_desc layout (include/rhd.h).

def optics_apply(rgb: np.ndarray, desc) -> tuple[int, np.ndarray]:
    """crates/optics/src/correction.rs:69-262 (authority path) on (H, W, 3) RGB."""
    v = np.ascontiguousarray(rgb, dtype=np.float32).copy()
    out = np.zeros_like(v)
    st = lib().rhdo_optics_apply(_p(v), ctypes.byref(desc), _p(out))
    return st, out


def optics_map(desc, scale: float, x: int, y: int) -> tuple[int, np.ndarray]:
    """Source (x, y) per channel of output pixel (x, y) (plan.rs:465-496)."""
    out = np.zeros(6, dtype=np.float32)
    st = lib().rhdo_optics_map(ctypes.byref(desc), ctypes.c_float(scale), _sz(x), _sz(y),
                               _p(out))
    return st, out


def optics_find_scale(desc) -> tuple[int, float]:
    out = ctypes.c_float()
    st = lib().rhdo_optics_find_scale(ctypes.byref(desc), ctypes.byref(out))
    return st, float(out.value)


def inverse_model(model: int, k, x: float, y: float) -> np.ndarray:
    """1 Poly3, 2 Poly5, 3 PTLens inverse distortion; 4 reverse TCA Poly3 [v, c, b]."""
    kk = np.zeros(3, dtype=np.float32)
    kk[:len(k)] = k
    out = np.zeros(2, dtype=np.float32)
    lib().rhdo_inverse_model(ctypes.c_int(model), _p(kk), ctypes.c_float(x), ctypes.c_float(y),
                             _p(out))
    return out


def cubic_sample(rgb: np.ndarray, channel: int, x: float, y: float) -> tuple[int, float]:
    """crates/optics/src/resample.rs:3-42."""
    v = np.ascontiguousarray(rgb, dtype=np.float32)
    out = ctypes.c_float()
    st = lib().rhdo_cubic_sample(_p(v), _sz(v.shape[1]), _sz(v.shape[0]), ctypes.c_int(channel),
                                 ctypes.c_float(x), ctypes.c_float(y), ctypes.byref(out))
    return st, float(out.value)


def cubic_weight(distance: float) -> float:
    lib().rhdo_cubic_weight.restype = ctypes.c_float
    return float(lib().rhdo_cubic_weight(ctypes.c_float(distance)))


def area_axis(source: int, target: int):
    """crates/core/src/resample.rs:46-112: (first, offset, count, weights)."""
    first = np.zeros(target, np.uint64)
    offset = np.zeros(target, np.uint64)
    count = np.zeros(target, np.uint64)
    weights = np.zeros(source + 2 * target, np.float32)
    lib().rhdo_area_axis.restype = ctypes.c_long
    n = lib().rhdo_area_axis(_sz(source), _sz(target), _p(first), _p(offset), _p(count),
                             _p(weights), _sz(weights.size))
    if n < 0:
        raise ValueError("invalid area axis")
    return first, offset, count, weights[:n].copy()


def resize_area(rgb: np.ndarray, target_width: int, target_height: int) -> tuple[int, np.ndarray]:
    """crates/core/src/resample.rs:235-316 on (H, W, 3) RGB."""
    v = np.ascontiguousarray(rgb, dtype=np.float32)
    out = np.zeros((target_height, target_width, 3), np.float32)
    st = lib().rhdo_resize_area(_p(v), _sz(v.shape[1]), _sz(v.shape[0]), _p(out),
                                _sz(target_width), _sz(target_height))
    return st, out


def preview_dimensions(width: int, height: int, max_long_edge: int) -> tuple[int, int, int]:
    tw, th = ctypes.c_size_t(), ctypes.c_size_t()
    st = lib().rhdo_preview_dimensions(_sz(width), _sz(height), _sz(max_long_edge),
                                       ctypes.byref(tw), ctypes.byref(th))
    return st, int(tw.value), int(th.value)


def color(rgb: np.ndarray, gains, matrix) -> tuple[int, np.ndarray]:
    """White balance then the camera matrix (crates/core/src/cpu/stages.rs:250-308)."""
    v = np.ascontiguousarray(rgb, dtype=np.float32).copy()
    g = np.asarray(gains, dtype=np.float32)
    m = np.ascontiguousarray(matrix, dtype=np.float32).reshape(9)
    st = lib().rhdo_color(_p(v), _sz(v.size // 3), _p(g), _p(m))
    return st, v


def false_color(rgb: np.ndarray, steps: int, balance, ceilings) -> tuple[int, np.ndarray]:
    """crates/detail/src/false_color.rs:45-135 on (H, W, 3) RGB; returns a copy."""
    v = np.ascontiguousarray(rgb, dtype=np.float32).copy()
    b = np.asarray(balance, dtype=np.float32)
    c = np.asarray(ceilings, dtype=np.float32)
    st = lib().rhdo_false_color(_p(v), _sz(v.shape[1]), _sz(v.shape[0]), ctypes.c_int(steps),
                                _p(b), _p(c))
    return st, v


def cpu_threads() -> int:
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)


def noise_profiles(rgb: np.ndarray, balance):
    """crates/detail/src/denoise/mod.rs:60-110 on (H, W, 3) RGB. Returns
    (status, [(a, b, bins)] for luma, C1, C2); status 1 = too small to profile."""
    v = np.ascontiguousarray(rgb, dtype=np.float32)
    b = np.asarray(balance, dtype=np.float32)
    out = np.zeros(9, np.float64)
    st = lib().rhdo_noise_profiles(_p(v), _sz(v.shape[1]), _sz(v.shape[0]), _p(b), _p(out))
    return st, [(float(out[3 * c]), float(out[3 * c + 1]), int(out[3 * c + 2])) for c in range(3)]


def estimate_noise(plane: np.ndarray, brightness: np.ndarray):
    """crates/detail/src/noise.rs:48-125 on one (H, W) plane."""
    p = np.ascontiguousarray(plane, dtype=np.float32)
    b = np.ascontiguousarray(brightness, dtype=np.float32)
    out = np.zeros(3, np.float64)
    st = lib().rhdo_estimate_noise(_p(p), _p(b), _sz(p.shape[1]), _sz(p.shape[0]), _p(out))
    return st, (float(out[0]), float(out[1]), int(out[2]))


def chroma_denoise(rgb: np.ndarray, strength: float, balance, ceilings, profiles):
    """crates/detail/src/denoise/chroma.rs:88-148; profiles = [(a, b), ...] for
    luma, C1, C2. Returns (status, copy)."""
    v = np.ascontiguousarray(rgb, dtype=np.float32).copy()
    b = np.asarray(balance, dtype=np.float32)
    c = np.asarray(ceilings, dtype=np.float32)
    p = np.asarray([x for pr in profiles for x in pr[:2]], dtype=np.float32)
    st = lib().rhdo_chroma_denoise(_p(v), _sz(v.shape[1]), _sz(v.shape[0]),
                                   ctypes.c_float(strength), _p(b), _p(c), _p(p))
    return st, v


def luma_denoise(rgb: np.ndarray, strength: float, detail: float, balance, ceilings, profile):
    """crates/detail/src/denoise/luma.rs:128-176; profile = (a, b)."""
    v = np.ascontiguousarray(rgb, dtype=np.float32).copy()
    b = np.asarray(balance, dtype=np.float32)
    c = np.asarray(ceilings, dtype=np.float32)
    p = np.asarray(profile[:2], dtype=np.float32)
    st = lib().rhdo_luma_denoise(_p(v), _sz(v.shape[1]), _sz(v.shape[0]),
                                 ctypes.c_float(strength), ctypes.c_float(detail), _p(b), _p(c),
                                 _p(p))
    return st, v


def b3_smooth(plane: np.ndarray, step: int) -> np.ndarray:
    """crates/detail/src/atrous.rs:63-120."""
    p = np.ascontiguousarray(plane, dtype=np.float32)
    out = np.empty_like(p)
    lib().rhdo_b3_smooth(_p(p), _sz(p.shape[1]), _sz(p.shape[0]), _sz(step), _p(out))
    return out


def local_mean_square(band: np.ndarray) -> np.ndarray:
    """crates/detail/src/denoise/luma.rs:279-321."""
    p = np.ascontiguousarray(band, dtype=np.float32)
    out = np.empty_like(p)
    lib().rhdo_local_mean_square(_p(p), _sz(p.shape[1]), _sz(p.shape[0]), _p(out))
    return out


def mad(values) -> float:
    """crates/detail/src/noise.rs:127-139."""
    v = np.ascontiguousarray(values, dtype=np.float32).copy()
    lib().rhdo_mad.restype = ctypes.c_float
    return float(lib().rhdo_mad(_p(v), _sz(v.size)))


def detail_stage(rgb: np.ndarray, *, false_colour=0, chroma=0.0, luminance=0.0,
                 luminance_detail=0.5, balance=(1.0, 1.0, 1.0), ceilings=(np.inf,) * 3,
                 capture=None):
    """crates/core/src/pipeline/detail.rs:210-307: profiles of the input, false
    colour, chroma NR, luminance NR, capture sharpening. capture = (amount,
    sigma, noise_protection, iterations) or None. Returns (status, copy,
    profiles or None)."""
    v = np.ascontiguousarray(rgb, dtype=np.float32).copy()
    b = np.asarray(balance, dtype=np.float32)
    c = np.asarray(ceilings, dtype=np.float32)
    prof = np.zeros(9, np.float64)
    have = ctypes.c_int(0)
    cap = capture or (0.0, 0.7, 0.0, 0)
    st = lib().rhdo_detail_stage(
        _p(v), _sz(v.shape[1]), _sz(v.shape[0]), ctypes.c_int(false_colour),
        ctypes.c_float(chroma), ctypes.c_float(luminance), ctypes.c_float(luminance_detail),
        _p(b), _p(c), ctypes.c_int(1 if capture else 0), ctypes.c_float(cap[0]),
        ctypes.c_float(cap[1]), ctypes.c_float(cap[2]), ctypes.c_int(cap[3]), _p(prof),
        ctypes.byref(have))
    profiles = ([(float(prof[3 * k]), float(prof[3 * k + 1]), int(prof[3 * k + 2]))
                 for k in range(3)] if have.value else None)
    return st, v, profiles


def decode_arw2(strip: bytes, width: int, height: int, curve) -> tuple[int, np.ndarray]:
    """Sony compressed ARW (rawler 0.7.2's published algorithm): u16 mosaic."""
    buf = np.frombuffer(bytes(strip), dtype=np.uint8).copy()
    out = np.zeros((height, width), np.uint16)
    c = np.asarray(curve, dtype=np.uint16)
    st = lib().rhdo_decode_arw2(_p(buf), _sz(buf.size), _sz(width), _sz(height), _p(c), _p(out))
    return st, out


def arw_curve(curve):
    """(table, base, delta) of the tone curve and its dithering lookup."""
    c = np.asarray(curve, dtype=np.uint16)
    t, b, d = (np.zeros(4096, np.uint16) for _ in range(3))
    lib().rhdo_arw_curve(_p(c), _p(t), _p(b), _p(d))
    return t, b, d


def arw_probe(data: bytes):
    """(status, layout dict or None); status 1 not TIFF, 2 corrupt, 3 no raw IFD."""
    buf = np.frombuffer(bytes(data), dtype=np.uint8).copy() if len(data) else np.zeros(1, np.uint8)
    out = np.zeros(12, np.uint64)
    st = lib().rhdo_arw_probe(_p(buf), _sz(len(data)), _p(out))
    if st:
        return st, None
    keys = ["width", "height", "bits_per_sample", "compression", "offset", "bytes"]
    d = {k: int(out[i]) for i, k in enumerate(keys)}
    d["curve"] = [int(v) for v in out[6:10]]
    d["has_curve"], d["kind"] = int(out[10]), int(out[11])
    return 0, d


def libm(fn: str, values: np.ndarray) -> np.ndarray:
    """The host libm's cbrtf ("cbrt"), log2f ("log2") or cosf ("cos") over f32
    values, as the reference's f32 methods compute them on Linux."""
    x = np.ascontiguousarray(values, dtype=np.float32)
    out = np.empty_like(x)
    st = lib().rhdo_libm(ctypes.c_int({"cbrt": 0, "log2": 1, "cos": 2}[fn]), _p(x), _p(out),
                         _sz(x.size))
    assert st == 0
    return out


MUTATIONS = {
    1: "Opposed: 3-sigma trim window for the region corrections (2 sigma)",
    2: "Opposed: local/region blend weights swapped",
    3: "Opposed: 1.5 EV chromaticity gate (0.75 EV)",
    4: "Opposed: local median takes one middle element",
    5: "capture: highlight mask expanded with blur^2 (blur^4)",
    6: "capture: upper highlight edge 0.95 (0.98)",
    7: "capture: estimate clamped to [0.25, 4] x guide ([0.5, 2])",
    8: "capture: lower highlight edge 0.80 (0.85)",
    9: "Opposed: arithmetic mean as the opposing reference (cube-root mean)",
    10: "Opposed: ring evidence from 0.25 x level (0.5 x level)",
    11: "Opposed: estimates bounded at 2 x level (4 x level)",
}


_mutants_built = False


def build_mutants() -> None:
    """Builds every planted-mistake oracle (make mutants, in parallel)."""
    jobs = str(min(8, os.cpu_count() or 1))
    subprocess.run(["make", "-s", "-j", jobs, "-C", str(_HERE), "mutants"], check=True)


@contextlib.contextmanager
def using(path: Path):
    """Within the block every oracle function runs the build at `path`."""
    global _lib
    saved = lib()
    _lib = ctypes.CDLL(str(path))
    _lib.rhdo_amaze_plane_names.restype = ctypes.c_char_p
    try:
        yield
    finally:
        _lib = saved


@contextlib.contextmanager
def mutant(k: int):
    """Within the block every oracle function runs the build that carries
    planted mistake `k` (MUTATIONS[k]); the checker is restored afterwards."""
    global _mutants_built
    if not _mutants_built:
        build_mutants()
        _mutants_built = True
    with using(_HERE / "_build" / f"librhd_oracle_m{k}.so"):
        yield


def native_build() -> Path:
    """The -O3 -march=native build of the same sources (make native), for the
    CPU baseline on the host it runs on."""
    subprocess.run(["make", "-s", "-C", str(_HERE), "native"], check=True)
    return _HERE / "_build" / "librhd_oracle_native.so"
