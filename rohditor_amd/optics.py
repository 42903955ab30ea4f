"""Lens-correction plans for the spatial step after the sensor stage.

Host-side mirror of crates/optics/src/plan.rs: the model enums (:13-29), the
resolved ``OpticsExecution`` (:36-48), build_plan's normalisation and
coefficient rescaling (:173-356), and the automatic scale (find_scale,
:521-559, run natively by ``rhd_optics_find_scale``). Lens matching and the
interpolation of calibrations across focal length, aperture and distance
(the lensfun database, crates/optics/src/catalog.rs and matching.rs) happen
before this and are not part of it: callers pass the calibration that
applies. The correction itself runs in librhd (``rhd_reduce_preview``,
``rhd_optics_full``).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Optional, Union

import numpy as np

from . import _lib


@dataclass(frozen=True)
class Poly3:
    k1: float


@dataclass(frozen=True)
class Poly5:
    k1: float
    k2: float


@dataclass(frozen=True)
class Ptlens:
    a: float
    b: float
    c: float


@dataclass(frozen=True)
class TcaLinear:
    kr: float
    kb: float


@dataclass(frozen=True)
class TcaPoly3:
    red: tuple   # (v, c, b)
    blue: tuple


@dataclass(frozen=True)
class Pa:
    k1: float
    k2: float
    k3: float


Distortion = Union[Poly3, Poly5, Ptlens]
Tca = Union[TcaLinear, TcaPoly3]


def _f32(v) -> float:
    return float(np.float32(v))


@dataclass(frozen=True)
class OpticsExecution:
    """crates/optics/src/plan.rs:36-48; f32 fields hold f32-representable values."""
    width: int
    height: int
    norm_scale: float
    norm_unscale: float
    center_x: float
    center_y: float
    scale: float = 1.0
    distortion: Optional[Distortion] = None
    tca: Optional[Tca] = None
    vignetting: Optional[Pa] = None

    def to_c(self) -> "_lib.OpticsDesc":
        d = _lib.OpticsDesc()
        d.width, d.height = self.width, self.height
        d.norm_scale, d.norm_unscale = self.norm_scale, self.norm_unscale
        d.center_x, d.center_y = self.center_x, self.center_y
        d.scale = self.scale
        m = self.distortion
        if isinstance(m, Poly3):
            d.distortion, k = 1, (m.k1, 0.0, 0.0)
        elif isinstance(m, Poly5):
            d.distortion, k = 2, (m.k1, m.k2, 0.0)
        elif isinstance(m, Ptlens):
            d.distortion, k = 3, (m.a, m.b, m.c)
        else:
            d.distortion, k = 0, (0.0, 0.0, 0.0)
        d.distortion_k[:] = k
        t = self.tca
        if isinstance(t, TcaLinear):
            d.tca = 1
            d.tca_red[:] = (t.kr, 0.0, 0.0)
            d.tca_blue[:] = (t.kb, 0.0, 0.0)
        elif isinstance(t, TcaPoly3):
            d.tca = 2
            d.tca_red[:] = tuple(t.red)
            d.tca_blue[:] = tuple(t.blue)
        if self.vignetting is not None:
            d.vignetting = 1
            d.vignetting_k[:] = (self.vignetting.k1, self.vignetting.k2, self.vignetting.k3)
        return d

    @property
    def remaps(self) -> bool:
        return self.distortion is not None or self.tca is not None

    def with_scale(self, scale: float) -> "OpticsExecution":
        return OpticsExecution(self.width, self.height, self.norm_scale, self.norm_unscale,
                               self.center_x, self.center_y, _f32(scale), self.distortion,
                               self.tca, self.vignetting)


def find_scale(execution: OpticsExecution) -> float:
    """plan.rs:521-559: smallest 1.025^k (k < 160) whose perimeter maps inside
    the cubic footprint. Raises GpuPreviewError(InvalidInput) when none does."""
    from .sensor import GpuPreviewError
    lib = _lib.load()
    d = execution.to_c()
    out = ctypes.c_float()
    st = lib.rhd_optics_find_scale(ctypes.byref(d), ctypes.byref(out))
    if st != _lib.OK:
        raise GpuPreviewError(st, "no automatic scale keeps the cubic footprint inside the image")
    return float(out.value)


def preview_dimensions(width: int, height: int, max_long_edge: int) -> tuple[int, int]:
    """crates/core/src/cpu/stages.rs:112-147."""
    from .sensor import GpuPreviewError
    lib = _lib.load()
    tw, th = ctypes.c_uint32(), ctypes.c_uint32()
    st = lib.rhd_preview_dimensions(width, height, max_long_edge, ctypes.byref(tw),
                                    ctypes.byref(th))
    if st != _lib.OK:
        raise GpuPreviewError(st, "preview dimensions need at least 2x2 pixels and a 2 px edge")
    return int(tw.value), int(th.value)


def _hugin_hs(real_focal: float, crop: float, aspect_ratio: Optional[float]) -> float:
    scale = math.hypot(36.0, 24.0) / float(np.float32(crop))
    if aspect_ratio is not None:
        scale = scale / math.hypot(float(np.float32(aspect_ratio)), 1.0)
    scale = scale / 2.0
    return float(np.float32(real_focal)) / scale


def resolve_execution(width: int, height: int, *, camera_crop: float, focal: float,
                      real_focal: Optional[float] = None, lens_center=(0.0, 0.0),
                      lens_aspect_ratio: float = 1.5, lens_crop: float = 1.0,
                      distortion: Optional[Distortion] = None, tca: Optional[Tca] = None,
                      vignetting: Optional[Pa] = None) -> OpticsExecution:
    """Numeric part of build_plan (plan.rs:196-326) from lensfun-calibration
    coefficients already interpolated for the shot: normalisation, lens
    centre, Hugin-scale rescaling of each model (rescale_distortion :358-389,
    rescale_tca :391-424, rescale_vignetting :426-442) and the automatic scale."""
    if width < 4 or height < 4:
        from .sensor import GpuPreviewError
        raise GpuPreviewError(_lib.INVALID_DIMENSIONS,
                              "cubic correction requires at least four pixels per axis")
    rf = real_focal if (real_focal is not None and math.isfinite(real_focal)
                        and real_focal > 0) else focal
    rf32 = float(np.float32(rf))
    width_f = float(width - 1) if width >= 2 else 1.0
    height_f = float(height - 1) if height >= 2 else 1.0
    norm_scale = (math.hypot(36.0, 24.0) / float(np.float32(camera_crop))
                  / math.hypot(width_f + 1.0, height_f + 1.0) / rf32)
    norm_unscale = 1.0 / norm_scale
    lens_size = min(width_f, height_f)
    cx = (width_f / 2.0 + lens_size / 2.0 * float(np.float32(lens_center[0]))) * norm_scale
    cy = (height_f / 2.0 + lens_size / 2.0 * float(np.float32(lens_center[1]))) * norm_scale

    hs = _hugin_hs(rf32, lens_crop, lens_aspect_ratio)
    hs2 = hs * hs
    hs3 = hs2 * hs
    hs4 = hs2 * hs2
    dist = None
    if isinstance(distortion, Poly3) and distortion.k1 != 0.0:
        k1 = float(np.float32(distortion.k1))
        dd = 1.0 - k1
        dist = Poly3(_f32(k1 * hs2 / (dd * dd * dd)))
    elif isinstance(distortion, Poly5):
        dist = Poly5(_f32(float(np.float32(distortion.k1)) * hs2),
                     _f32(float(np.float32(distortion.k2)) * hs4))
    elif isinstance(distortion, Ptlens):
        a, b, c = (float(np.float32(v)) for v in (distortion.a, distortion.b, distortion.c))
        dd = 1.0 - a - b - c
        d2 = dd * dd
        dist = Ptlens(_f32(a * hs3 / (d2 * d2)), _f32(b * hs2 / (d2 * dd)), _f32(c * hs / d2))
    t = None
    if isinstance(tca, TcaLinear):
        kr, kb = np.float32(tca.kr), np.float32(tca.kb)
        if math.isfinite(kr) and math.isfinite(kb) and kr != 0 and kb != 0:
            t = TcaLinear(float(np.float32(1.0) / kr), float(np.float32(1.0) / kb))
    elif isinstance(tca, TcaPoly3):
        r = [float(np.float32(v)) for v in tca.red]
        b = [float(np.float32(v)) for v in tca.blue]
        t = TcaPoly3((r[0], _f32(r[1] * hs), _f32(r[2] * hs2)),
                     (b[0], _f32(b[1] * hs), _f32(b[2] * hs2)))
    vig = None
    if isinstance(vignetting, Pa):
        vhs = _hugin_hs(rf32, lens_crop, None)
        v2 = vhs * vhs
        v4 = v2 * v2
        v6 = v4 * v2
        vig = Pa(_f32(float(np.float32(vignetting.k1)) * v2),
                 _f32(float(np.float32(vignetting.k2)) * v4),
                 _f32(float(np.float32(vignetting.k3)) * v6))
    e = OpticsExecution(width, height, norm_scale, norm_unscale, cx, cy, 1.0, dist, t, vig)
    if e.remaps:
        e = e.with_scale(find_scale(e))
    return e
