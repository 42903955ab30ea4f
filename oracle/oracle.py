"""Oracle (test infrastructure only): ctypes bindings to the C++ restatement.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import
this module, and only as the checker; the product path never does.
"""
from __future__ import annotations

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


def cpu_threads() -> int:
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
