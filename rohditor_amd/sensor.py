"""Host-side mirror of the reference's sensor-stage interface over librhd.

Names and meanings follow the reference so callers and tests read the same:

* ``RawFrame`` / ``RawFileInfo`` / ``LevelPattern``: crates/raw/src/model.rs:35-210
* ``SensorDevelopmentDescription.from_frame``: crates/core/src/pipeline/sensor.rs:303-338,
  with ``highlight_execution`` / ``neutral_ceiling`` (:407-455) and the
  as-shot demosaic balance (crates/core/src/cpu/stages.rs:180-205)
* ``capture_ceilings``: crates/core/src/pipeline/detail.rs:323-337
* ``GpuSensorProcessor``: crates/gpu/src/sensor/{develop,normalize,highlight,demosaic}.rs
* ``GpuPreviewError`` with ``is_routing``: crates/gpu/src/lib.rs:55-109
* ``CancellationToken``: crates/core/src/cancel.rs:13-42

Every pixel operation runs in the HIP library; this module only resolves
metadata and moves handles.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from enum import Enum
from typing import Optional, Sequence

import numpy as np

from . import _lib

DETECTION_SAFETY_MARGIN = np.float32(0.987)  # crates/highlight/src/lib.rs:26
CAPTURE_SHARPENING_ITERATIONS = 8  # crates/core/src/sharpening.rs:13

PATTERNS = {"RGGB": 0, "BGGR": 1, "GRBG": 2, "GBRG": 3}
_COLOR_TABLE = {0: (0, 1, 1, 2), 1: (2, 1, 1, 0), 2: (1, 0, 2, 1), 3: (1, 2, 0, 1)}


def color_at(pattern: int, x: int, y: int) -> int:
    """crates/image/src/lib.rs:90-117: 0 red, 1 green, 2 blue."""
    return _COLOR_TABLE[pattern][((y & 1) << 1) | (x & 1)]


def shifted(pattern: int, x: int, y: int) -> int:
    """crates/image/src/lib.rs:119-135 (BayerPattern::shifted)."""
    for cand in range(4):
        if all(color_at(cand, xx, yy) == color_at(pattern, xx + x, yy + y)
               for yy in range(2) for xx in range(2)):
            return cand
    raise AssertionError("Bayer shifts only use parity")


class HighlightMethod(Enum):
    OFF = 0
    CLIP = 1
    OPPOSED = 2


class DemosaicAlgorithm(Enum):
    BILINEAR = 0
    MHC = 1
    RCD = 2
    AMAZE = 3


class RawCropPolicy(Enum):
    ACTIVE_AREA = 0
    RECOMMENDED = 1


class GpuPreviewError(Exception):
    """crates/gpu/src/lib.rs:55-98 with the C ABI status code."""

    def __init__(self, status: int, message: str):
        self.status = status
        self.kind = _lib.STATUS_NAMES[status] if 0 <= status < len(_lib.STATUS_NAMES) else "Unknown"
        super().__init__(f"{self.kind}: {message}")

    def is_routing(self) -> bool:
        return bool(_lib.load().rhd_status_is_routing(self.status))


@dataclass
class LevelPattern:
    values: Sequence[float]
    repeat_width: int = 1
    repeat_height: int = 1


@dataclass
class ImageRect:
    x: int
    y: int
    width: int
    height: int


@dataclass
class RawFileInfo:
    width: int
    height: int
    cfa: str = "RGGB"
    black_levels: LevelPattern = field(default_factory=lambda: LevelPattern([0.0]))
    white_levels: Sequence[float] = (65535.0,)
    as_shot_white_balance: Sequence[Optional[float]] = (1.0, 1.0, 1.0)
    active_area: Optional[ImageRect] = None
    crop_area: Optional[ImageRect] = None


@dataclass
class RawFrame:
    info: RawFileInfo
    mosaic: np.ndarray  # (height, row_stride) uint16
    row_stride: int = 0

    def __post_init__(self):
        self.mosaic = np.ascontiguousarray(self.mosaic, dtype=np.uint16)
        if not self.row_stride:
            self.row_stride = self.mosaic.shape[1]


@dataclass
class HighlightAdjustments:
    """crates/edit/src/recipe.rs:241-300: method plus the two thresholds."""
    method: HighlightMethod = HighlightMethod.CLIP
    clip_threshold: float = 1.0
    opposed_detection_threshold: float = 1.0


@dataclass
class CaptureSharpening:
    """crates/edit/src/sharpening.rs:22-48."""
    enabled: bool = False
    amount: float = 0.5
    radius: float = 0.6
    noise_protection: float = 0.5

    def is_active(self) -> bool:
        return self.enabled and self.amount > 0.0


@dataclass
class NoiseReduction:
    """crates/edit/src/noise_reduction.rs:27-61: strengths in [0, 1] (0 = off)
    and the luminance filter's texture preservation (0.5 neutral)."""
    chroma: float = 0.0
    luminance: float = 0.0
    luminance_detail: float = 0.5

    def is_active(self) -> bool:
        return self.chroma > 0.0 or self.luminance > 0.0


@dataclass
class EditRecipe:
    highlights: HighlightAdjustments = field(default_factory=HighlightAdjustments)
    capture: CaptureSharpening = field(default_factory=CaptureSharpening)
    # Resolved user white-balance gains (R, G, B); None = as shot.
    white_balance_gains: Optional[Sequence[float]] = None
    # RL iterations; the reference fixes 8.
    capture_iterations: int = CAPTURE_SHARPENING_ITERATIONS
    # FalseColorSuppression::steps (crates/edit/src/false_color.rs), 0..=5.
    false_color_steps: int = 0
    noise_reduction: "NoiseReduction" = field(default_factory=lambda: NoiseReduction())


@dataclass
class RenderOptions:
    raw_crop_policy: RawCropPolicy = RawCropPolicy.RECOMMENDED
    demosaic: DemosaicAlgorithm = DemosaicAlgorithm.AMAZE


def as_shot_gains(info: RawFileInfo) -> np.ndarray:
    """crates/core/src/cpu/stages.rs:180-205: [r/g, 1, b/g] from as-shot multipliers."""
    r, g, b = info.as_shot_white_balance[:3]
    if r is None and g is None and b is None:
        return np.ones(3, dtype=np.float32)
    vals = [np.float32(v) if v is not None else None for v in (r, g, b)]
    if any(v is None or not np.isfinite(v) or v <= 0 for v in vals):
        raise GpuPreviewError(_lib.INVALID_INPUT, "finite positive R, G, and B multipliers are required")
    return np.array([vals[0] / vals[1], np.float32(1.0), vals[2] / vals[1]], dtype=np.float32)


def _min3(g) -> np.float32:
    return np.float32(min(min(g[0], g[1]), g[2]))


@dataclass
class SensorDevelopmentDescription:
    frame_dims: tuple
    row_stride: int
    sensor_pattern: int
    crop: ImageRect
    crop_pattern: int
    black: np.ndarray
    black_repeat: tuple
    white: np.ndarray
    highlight: HighlightMethod
    clip_levels: np.ndarray
    detection_levels: np.ndarray
    neutral_ceiling: Optional[np.ndarray]
    demosaic: DemosaicAlgorithm
    demosaic_balance: np.ndarray

    @classmethod
    def from_frame(cls, frame: RawFrame, recipe: EditRecipe,
                   options: RenderOptions) -> "SensorDevelopmentDescription":
        info = frame.info
        if info.cfa.upper() not in PATTERNS:
            raise GpuPreviewError(_lib.INVALID_INPUT, f"unsupported CFA pattern {info.cfa}")
        pattern = PATTERNS[info.cfa.upper()]
        full = ImageRect(0, 0, info.width, info.height)
        active = info.active_area or full
        crop = active
        if options.raw_crop_policy is RawCropPolicy.RECOMMENDED and info.crop_area is not None:
            crop = info.crop_area
        demosaic_balance = as_shot_gains(info)
        gains = (np.asarray(recipe.white_balance_gains, dtype=np.float32)
                 if recipe.white_balance_gains is not None else demosaic_balance)
        hl = recipe.highlights
        clip_levels = np.ones(3, np.float32)
        detection = np.ones(3, np.float32)
        neutral = None
        if hl.method is HighlightMethod.CLIP:
            # sensor.rs:416-426
            common = np.float32(hl.clip_threshold) * DETECTION_SAFETY_MARGIN * _min3(gains)
            clip_levels = np.array([common / gains[0], common / gains[1], common / gains[2]],
                                   dtype=np.float32)
        elif hl.method is HighlightMethod.OPPOSED:
            # sensor.rs:427-440 and neutral_ceiling :449-455
            level = np.float32(hl.opposed_detection_threshold) * DETECTION_SAFETY_MARGIN
            detection = np.array([level] * 3, dtype=np.float32)
            common = level * _min3(demosaic_balance)
            neutral = np.array([common / demosaic_balance[0], common / demosaic_balance[1],
                                common / demosaic_balance[2]], dtype=np.float32)
        lp = info.black_levels
        return cls(
            frame_dims=(info.width, info.height), row_stride=frame.row_stride,
            sensor_pattern=pattern, crop=crop, crop_pattern=shifted(pattern, crop.x, crop.y),
            black=np.asarray(lp.values, dtype=np.float32),
            black_repeat=(lp.repeat_width, lp.repeat_height),
            white=np.asarray(info.white_levels, dtype=np.float32), highlight=hl.method,
            clip_levels=clip_levels, detection_levels=detection, neutral_ceiling=neutral,
            demosaic=options.demosaic, demosaic_balance=demosaic_balance)

    def to_c(self) -> "_lib.SensorDesc":
        d = _lib.SensorDesc()
        d.sensor_width, d.sensor_height = self.frame_dims
        d.row_stride = self.row_stride
        d.sensor_pattern = self.sensor_pattern
        d.crop_x, d.crop_y = self.crop.x, self.crop.y
        d.crop_width, d.crop_height = self.crop.width, self.crop.height
        self._keep = (np.ascontiguousarray(self.black), np.ascontiguousarray(self.white))
        d.black_levels = self._keep[0].ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        d.black_repeat_width, d.black_repeat_height = self.black_repeat
        d.white_levels = self._keep[1].ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        d.white_count = self.white.size
        d.highlight = self.highlight.value
        d.clip_levels[:] = [float(v) for v in self.clip_levels]
        d.detection_levels[:] = [float(v) for v in self.detection_levels]
        d.has_neutral_ceiling = int(self.neutral_ceiling is not None)
        if self.neutral_ceiling is not None:
            d.neutral_ceiling[:] = [float(v) for v in self.neutral_ceiling]
        d.demosaic = self.demosaic.value
        d.demosaic_balance[:] = [float(v) for v in self.demosaic_balance]
        return d


def capture_ceilings(recipe: EditRecipe, gains: Sequence[float]) -> np.ndarray:
    """crates/core/src/pipeline/detail.rs:323-337."""
    hl = recipe.highlights
    if hl.method is HighlightMethod.CLIP:
        g = np.asarray(gains, dtype=np.float32)
        ceiling = np.float32(hl.clip_threshold) * _min3(g)
        return np.array([ceiling / g[0], ceiling / g[1], ceiling / g[2]], dtype=np.float32)
    if hl.method is HighlightMethod.OPPOSED:
        return np.array([hl.opposed_detection_threshold] * 3, dtype=np.float32)
    return np.ones(3, dtype=np.float32)


def detail_desc(recipe: EditRecipe, gains: Sequence[float],
                balance: Optional[Sequence[float]] = None) -> "_lib.DetailDesc":
    """DetailContract::new (crates/core/src/pipeline/detail_contract.rs:23-40):
    ceilings from the rendering `gains`, balance = the as-shot demosaic
    balance (`gains` when not given)."""
    c = recipe.capture
    dd = _lib.DetailDesc()
    dd.capture_enabled = int(c.enabled)
    dd.amount, dd.sigma, dd.noise_protection = c.amount, c.radius, c.noise_protection
    dd.ceilings[:] = [float(v) for v in capture_ceilings(recipe, gains)]
    dd.iterations = recipe.capture_iterations
    dd.false_colour = int(recipe.false_color_steps)
    dd.balance[:] = [float(v) for v in (balance if balance is not None else gains)]
    nr = recipe.noise_reduction
    dd.nr_chroma, dd.nr_luminance = nr.chroma, nr.luminance
    dd.nr_luminance_detail = nr.luminance_detail
    return dd


class CancellationToken:
    """crates/core/src/cancel.rs:13-42: a flag the library polls between stages."""

    def __init__(self):
        self._flag = ctypes.c_uint8(0)

    def cancel(self) -> None:
        self._flag.value = 1

    def is_cancelled(self) -> bool:
        return bool(self._flag.value)

    @property
    def ptr(self):
        return ctypes.cast(ctypes.pointer(self._flag), ctypes.c_void_p)


class _Handle:
    _release = None

    def __init__(self, proc: "GpuSensorProcessor", ptr: ctypes.c_void_p):
        self.proc = proc
        self.ptr = ptr

    def release(self) -> None:
        if self.ptr:
            getattr(self.proc.lib, self._release)(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class ResidentMosaic(_Handle):
    """A decoded RawFrame mosaic kept on the GPU."""
    _release = "rhd_mosaic_release"

    def read(self, width: int, height: int, row_stride: Optional[int] = None) -> np.ndarray:
        stride = row_stride or width
        out = np.empty((height, stride), dtype=np.uint16)
        self.proc._check(self.proc.lib.rhd_mosaic_read(
            self.proc.ctx, self.ptr, out.ctypes.data_as(ctypes.c_void_p), out.size))
        return out[:, :width]


def arw_probe(data: bytes) -> "_lib.ArwLayout":
    """Finds the raw strip of a Sony ARW (TIFF) file: its size, compression,
    offset and tone-curve knots. Raises GpuPreviewError with RawUnsupported or
    RawCorrupt (RawError::Unsupported / Corrupt of crates/raw/src/lib.rs)."""
    lib = _lib.load()
    layout = _lib.ArwLayout()
    buf = np.frombuffer(data, dtype=np.uint8) if len(data) else np.zeros(1, np.uint8)
    st = lib.rhd_arw_probe(buf.ctypes.data_as(ctypes.c_void_p), len(data), ctypes.byref(layout))
    if st != _lib.OK:
        raise GpuPreviewError(st, lib.rhd_status_name(st).decode())
    return layout


class GpuMosaic(_Handle):
    """GpuNormalizedMosaic / GpuHighlightedMosaic."""
    _release = "rhd_plane_release"

    def dims(self):
        w, h, p = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_int32()
        self.proc.lib.rhd_plane_dims(self.ptr, ctypes.byref(w), ctypes.byref(h), ctypes.byref(p))
        return w.value, h.value, p.value

    def read(self) -> np.ndarray:
        w, h, _ = self.dims()
        out = np.empty((h, w), dtype=np.float32)
        self.proc._check(self.proc.lib.rhd_plane_read(self.proc.ctx, self.ptr,
                                                      out.ctypes.data_as(ctypes.c_void_p), out.size))
        return out


class GpuDevelopedSource(_Handle):
    """GpuSensorCameraSource: three resident camera-RGB planes."""
    _release = "rhd_source_release"

    def dims(self):
        w, h = ctypes.c_uint32(), ctypes.c_uint32()
        self.proc.lib.rhd_source_dims(self.ptr, ctypes.byref(w), ctypes.byref(h))
        return w.value, h.value

    def read_rgb(self) -> np.ndarray:
        """Interleaved (H, W, 3) readback, like LinearRgbImage."""
        w, h = self.dims()
        out = np.empty((h, w, 3), dtype=np.float32)
        self.proc._check(self.proc.lib.rhd_source_read_rgb(
            self.proc.ctx, self.ptr, out.ctypes.data_as(ctypes.c_void_p), out.size))
        return out


class GpuSensorProcessor:
    """crates/gpu/src/sensor/normalize.rs:75-200 and develop.rs:54-142."""

    def __init__(self, device: int = 0, budget_bytes: int = 0):
        self.lib = _lib.load()
        ctx = ctypes.c_void_p()
        st = self.lib.rhd_create(device, budget_bytes, ctypes.byref(ctx))
        if st != _lib.OK:
            raise GpuPreviewError(st, "could not create the HIP sensor processor")
        self.ctx = ctx

    def close(self) -> None:
        if getattr(self, "ctx", None):
            self.lib.rhd_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st: int) -> None:
        if st != _lib.OK:
            msg = self.lib.rhd_last_error(self.ctx)
            raise GpuPreviewError(st, msg.decode() if msg else "")

    def set_budget(self, budget_bytes: int) -> None:
        self._check(self.lib.rhd_set_budget(self.ctx, budget_bytes))

    def reserved_bytes(self) -> int:
        return int(self.lib.rhd_reserved_bytes(self.ctx))

    @staticmethod
    def supports(desc: SensorDevelopmentDescription, detail=None) -> bool:
        lib = _lib.load()
        c = desc.to_c()
        return bool(lib.rhd_supports(ctypes.byref(c), ctypes.byref(detail) if detail else None))

    def upload(self, frame: RawFrame) -> ResidentMosaic:
        ptr = ctypes.c_void_p()
        h = frame.info.height
        buf = np.ascontiguousarray(frame.mosaic[:h, :frame.row_stride])
        self._check(self.lib.rhd_mosaic_upload(self.ctx, buf.ctypes.data_as(ctypes.c_void_p),
                                               frame.row_stride, frame.info.width, h,
                                               ctypes.byref(ptr)))
        return ResidentMosaic(self, ptr)

    def decode_arw(self, data: bytes, layout: Optional["_lib.ArwLayout"] = None):
        """Uploads an ARW file's raw strip and decodes it on the device into a
        resident mosaic (row stride = width). Returns (mosaic, gpu_ms)."""
        layout = layout if layout is not None else arw_probe(data)
        buf = np.frombuffer(data, dtype=np.uint8)
        ptr = ctypes.c_void_p()
        ms = ctypes.c_float()
        self._check(self.lib.rhd_mosaic_decode_arw(self.ctx, buf.ctypes.data_as(ctypes.c_void_p),
                                                   len(data), ctypes.byref(layout),
                                                   ctypes.byref(ptr), ctypes.byref(ms)))
        return ResidentMosaic(self, ptr), float(ms.value)

    def normalize(self, mosaic: ResidentMosaic, desc: SensorDevelopmentDescription,
                  cancellation: Optional[CancellationToken] = None, metrics=None) -> GpuMosaic:
        ptr = ctypes.c_void_p()
        c = desc.to_c()
        self._check(self.lib.rhd_normalize(self.ctx, mosaic.ptr, ctypes.byref(c),
                                           cancellation.ptr if cancellation else None,
                                           ctypes.byref(ptr),
                                           ctypes.byref(metrics) if metrics else None))
        return GpuMosaic(self, ptr)

    def upload_mosaic(self, values: np.ndarray, pattern: int) -> GpuMosaic:
        v = np.ascontiguousarray(values, dtype=np.float32)
        ptr = ctypes.c_void_p()
        self._check(self.lib.rhd_plane_upload(self.ctx, v.ctypes.data_as(ctypes.c_void_p),
                                              v.shape[1], v.shape[0], pattern, ctypes.byref(ptr)))
        return GpuMosaic(self, ptr)

    def apply_highlight(self, mosaic: GpuMosaic, desc: SensorDevelopmentDescription,
                        cancellation: Optional[CancellationToken] = None) -> "_lib.Metrics":
        m = _lib.Metrics()
        c = desc.to_c()
        self._check(self.lib.rhd_highlight(self.ctx, mosaic.ptr, ctypes.byref(c),
                                           cancellation.ptr if cancellation else None,
                                           ctypes.byref(m)))
        return m

    def demosaic(self, mosaic: GpuMosaic, desc: SensorDevelopmentDescription, detail=None,
                 cancellation: Optional[CancellationToken] = None, metrics=None
                 ) -> GpuDevelopedSource:
        """Consumes `mosaic` (released even on error)."""
        ptr = ctypes.c_void_p()
        c = desc.to_c()
        plane, mosaic.ptr = mosaic.ptr, None
        self._check(self.lib.rhd_demosaic(self.ctx, plane, ctypes.byref(c),
                                          ctypes.byref(detail) if detail else None,
                                          cancellation.ptr if cancellation else None,
                                          ctypes.byref(ptr),
                                          ctypes.byref(metrics) if metrics else None))
        return GpuDevelopedSource(self, ptr)

    def capture(self, source: GpuDevelopedSource, detail,
                cancellation: Optional[CancellationToken] = None) -> None:
        self._check(self.lib.rhd_capture(self.ctx, source.ptr, ctypes.byref(detail),
                                         cancellation.ptr if cancellation else None, None))

    def capture_noalloc(self, planar_rgb: int, width: int, height: int, detail) -> None:
        """Capture sharpening in place on a caller-owned planar device buffer
        (3 * width * height floats) without allocating: rhd_capture_noalloc."""
        self._check(self.lib.rhd_capture_noalloc(self.ctx, ctypes.c_void_p(planar_rgb), width,
                                                 height, ctypes.byref(detail)))

    def source_planes(self, source: GpuDevelopedSource):
        """Device addresses of a source's R, G and B planes."""
        r, g, b = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        self._check(self.lib.rhd_source_planes(source.ptr, ctypes.byref(r), ctypes.byref(g),
                                               ctypes.byref(b)))
        return r.value, g.value, b.value

    def noise_profiles(self, source: GpuDevelopedSource, balance):
        """estimate_noise_profiles (crates/detail/src/denoise/mod.rs:60-110),
        on the device: [(a, b, bins)] for luma, C1, C2, or None when the image
        is too small to profile. Returns (profiles, gpu_ms)."""
        b = (ctypes.c_float * 3)(*[float(v) for v in balance])
        out = (_lib.NoiseProfile * 3)()
        have = ctypes.c_int32()
        ms = ctypes.c_float()
        self._check(self.lib.rhd_noise_profiles(self.ctx, source.ptr, b, out, ctypes.byref(have),
                                                ctypes.byref(ms)))
        if not have.value:
            return None, float(ms.value)
        return [(float(p.a), float(p.b), int(p.bins)) for p in out], float(ms.value)

    @staticmethod
    def _profiles(profiles):
        out = (_lib.NoiseProfile * len(profiles))()
        for o, p in zip(out, profiles):
            o.a, o.b = float(p[0]), float(p[1])
            o.bins = int(p[2]) if len(p) > 2 else 0
        return out

    def denoise_chroma(self, source: GpuDevelopedSource, strength: float, balance, ceilings,
                       profiles) -> float:
        """denoise_chroma (crates/detail/src/denoise/chroma.rs:88-148) in place;
        profiles = [(a, b), ...] for luma, C1, C2. Returns the GPU milliseconds."""
        b = (ctypes.c_float * 3)(*[float(v) for v in balance])
        c = (ctypes.c_float * 3)(*[float(v) for v in ceilings])
        ms = ctypes.c_float()
        self._check(self.lib.rhd_denoise_chroma(self.ctx, source.ptr, float(strength), b, c,
                                                self._profiles(profiles), ctypes.byref(ms)))
        return float(ms.value)

    def denoise_luma(self, source: GpuDevelopedSource, strength: float, detail: float, balance,
                     ceilings, profile) -> float:
        """denoise_luma (crates/detail/src/denoise/luma.rs:128-176) in place
        with the luma profile (a, b)."""
        b = (ctypes.c_float * 3)(*[float(v) for v in balance])
        c = (ctypes.c_float * 3)(*[float(v) for v in ceilings])
        ms = ctypes.c_float()
        self._check(self.lib.rhd_denoise_luma(self.ctx, source.ptr, float(strength),
                                              float(detail), b, c, self._profiles([profile]),
                                              ctypes.byref(ms)))
        return float(ms.value)

    def upload_source(self, rgb: np.ndarray) -> GpuDevelopedSource:
        """A resident camera source from interleaved (H, W, 3) RGB
        (GpuSpatialProcessor::upload_captured_source, crates/gpu/src/spatial/mod.rs:110)."""
        v = np.ascontiguousarray(rgb, dtype=np.float32)
        ptr = ctypes.c_void_p()
        self._check(self.lib.rhd_source_upload(self.ctx, v.ctypes.data_as(ctypes.c_void_p),
                                               v.shape[1], v.shape[0], ctypes.byref(ptr)))
        return GpuDevelopedSource(self, ptr)

    def reduce_preview(self, source: GpuDevelopedSource, optics, target: tuple,
                       cancellation: Optional[CancellationToken] = None):
        """Optics (an OpticsExecution or None) then exact area reduction to
        `target` = (width, height) (GpuSpatialProcessor::reduce_preview,
        crates/gpu/src/spatial/reduction.rs:15). Returns (source, gpu_ms)."""
        ptr = ctypes.c_void_p()
        ms = ctypes.c_float()
        d = optics.to_c() if optics is not None else None
        self._check(self.lib.rhd_reduce_preview(
            self.ctx, source.ptr, ctypes.byref(d) if d is not None else None, target[0],
            target[1], cancellation.ptr if cancellation else None, ctypes.byref(ptr),
            ctypes.byref(ms)))
        return GpuDevelopedSource(self, ptr), float(ms.value)

    def full_resolution_source(self, source: GpuDevelopedSource, optics,
                               cancellation: Optional[CancellationToken] = None):
        """The corrected image at the source's size
        (GpuSpatialProcessor::full_resolution_source, crates/gpu/src/spatial/mod.rs:200)."""
        ptr = ctypes.c_void_p()
        ms = ctypes.c_float()
        d = optics.to_c() if optics is not None else None
        self._check(self.lib.rhd_optics_full(
            self.ctx, source.ptr, ctypes.byref(d) if d is not None else None,
            cancellation.ptr if cancellation else None, ctypes.byref(ptr), ctypes.byref(ms)))
        return GpuDevelopedSource(self, ptr), float(ms.value)

    def color_transform(self, source: GpuDevelopedSource, gains, camera_to_rec2020,
                        cancellation: Optional[CancellationToken] = None) -> float:
        """White balance then the camera matrix to linear Rec.2020, in place
        (crates/core/src/cpu/stages.rs:250-308). Returns the GPU milliseconds."""
        g = (ctypes.c_float * 3)(*[float(v) for v in gains])
        m = (ctypes.c_float * 9)(*[float(v) for v in np.asarray(camera_to_rec2020).reshape(9)])
        ms = ctypes.c_float()
        self._check(self.lib.rhd_color_transform(self.ctx, source.ptr, g, m,
                                                 cancellation.ptr if cancellation else None,
                                                 ctypes.byref(ms)))
        return float(ms.value)

    def develop_resident(self, mosaic: ResidentMosaic, desc: SensorDevelopmentDescription,
                         detail=None, cancellation: Optional[CancellationToken] = None,
                         metrics: Optional["_lib.Metrics"] = None) -> GpuDevelopedSource:
        ptr = ctypes.c_void_p()
        c = desc.to_c()
        self._check(self.lib.rhd_develop_resident(
            self.ctx, mosaic.ptr, ctypes.byref(c), ctypes.byref(detail) if detail else None,
            cancellation.ptr if cancellation else None, ctypes.byref(ptr),
            ctypes.byref(metrics) if metrics is not None else None))
        return GpuDevelopedSource(self, ptr)

    def develop(self, frame: RawFrame, recipe: EditRecipe, options: RenderOptions,
                cancellation: Optional[CancellationToken] = None):
        """crates/gpu/src/sensor/develop.rs:68-142: returns (source, metrics)."""
        if cancellation is not None and cancellation.is_cancelled():
            raise GpuPreviewError(_lib.CANCELLED, "GPU processing was cancelled")
        desc = SensorDevelopmentDescription.from_frame(frame, recipe, options)
        gains = (recipe.white_balance_gains if recipe.white_balance_gains is not None
                 else desc.demosaic_balance)
        detail = detail_desc(recipe, gains, desc.demosaic_balance)
        m = _lib.Metrics()
        ptr = ctypes.c_void_p()
        c = desc.to_c()
        buf = np.ascontiguousarray(frame.mosaic[:frame.info.height, :frame.row_stride])
        self._check(self.lib.rhd_develop(
            self.ctx, buf.ctypes.data_as(ctypes.c_void_p), frame.row_stride, ctypes.byref(c),
            ctypes.byref(detail), cancellation.ptr if cancellation else None, ctypes.byref(ptr),
            ctypes.byref(m)))
        return GpuDevelopedSource(self, ptr), m


@dataclass
class ServedResult:
    """The outcome of one served request (include/rhd.h rhd_request_info)."""
    status: int
    source: Optional[GpuDevelopedSource]
    lane: int
    base_cache_hit: bool
    queued_ms: float
    run_ms: float
    latency_ms: float
    megapixels: float

    @property
    def ok(self) -> bool:
        return self.status == _lib.OK


class GpuServer:
    """Serving loop on one GPU: the counterpart of the engine worker's develop
    loop with its newest-wins mailbox (crates/engine/src/coordinator/
    worker.rs:723-800, scheduler.rs:44-173) and the resident caches of
    SpatialCache (crates/engine/src/preview_cache/spatial.rs:322-611).

    Frames are made resident with `put_frame`; `submit` queues a develop and
    returns a request id; `wait` collects it. Requests for different frames
    run concurrently on `lanes` HIP streams; a newer request for a frame
    coalesces its queued ones and cancels its running one.
    """

    def __init__(self, proc: GpuSensorProcessor, lanes: int = 3, admission_bytes: int = 0,
                 base_cache_bytes: int = 0):
        self.proc = proc
        cfg = _lib.ServerConfig(lanes, admission_bytes, base_cache_bytes)
        ptr = ctypes.c_void_p()
        proc._check(proc.lib.rhd_server_create(proc.ctx, ctypes.byref(cfg), ctypes.byref(ptr)))
        self.ptr = ptr

    def close(self) -> None:
        if getattr(self, "ptr", None):
            self.proc.lib.rhd_server_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def put_frame(self, frame_id: int, frame: RawFrame) -> None:
        h = frame.info.height
        buf = np.ascontiguousarray(frame.mosaic[:h, :frame.row_stride])
        self.proc._check(self.proc.lib.rhd_server_put_mosaic(
            self.ptr, frame_id, buf.ctypes.data_as(ctypes.c_void_p), frame.row_stride,
            frame.info.width, h))

    def put_arw(self, frame_id: int, data: bytes, layout=None) -> None:
        """Makes an ARW file's frame resident, decoded on the device."""
        layout = layout if layout is not None else arw_probe(data)
        buf = np.frombuffer(data, dtype=np.uint8)
        self.proc._check(self.proc.lib.rhd_server_put_arw(
            self.ptr, frame_id, buf.ctypes.data_as(ctypes.c_void_p), len(data),
            ctypes.byref(layout)))

    def drop_frame(self, frame_id: int) -> None:
        self.proc._check(self.proc.lib.rhd_server_drop_mosaic(self.ptr, frame_id))

    def submit(self, frame_id: int, desc: SensorDevelopmentDescription, detail=None,
               target: Optional[tuple] = None, optics=None) -> int:
        """Queues a develop; with `target` = (width, height) the request also
        runs the spatial completion (`optics`, an OpticsExecution or None,
        then area reduction) and returns the reduced preview."""
        c = desc.to_c()
        rid = ctypes.c_uint64()
        dp = ctypes.byref(detail) if detail else None
        if target is None:
            self.proc._check(self.proc.lib.rhd_server_submit(
                self.ptr, frame_id, ctypes.byref(c), dp, ctypes.byref(rid)))
        else:
            o = optics.to_c() if optics is not None else None
            self.proc._check(self.proc.lib.rhd_server_submit_spatial(
                self.ptr, frame_id, ctypes.byref(c), dp, ctypes.byref(o) if o is not None else None,
                target[0], target[1], ctypes.byref(rid)))
        return rid.value

    def wait(self, request_id: int, timeout_ms: int = -1, keep: bool = True) -> ServedResult:
        """Collects a request; with `keep` the developed source is returned
        (release it when done), otherwise it is released here."""
        ptr = ctypes.c_void_p()
        info = _lib.RequestInfo()
        st = self.proc.lib.rhd_server_wait(self.ptr, request_id, timeout_ms,
                                           ctypes.byref(ptr) if keep else None, ctypes.byref(info))
        if st == _lib.SYNCHRONIZATION and info.latency_ms == 0.0:
            raise TimeoutError(f"request {request_id} is still pending")
        src = GpuDevelopedSource(self.proc, ptr) if keep and ptr.value else None
        return ServedResult(st, src, info.lane, bool(info.base_cache_hit), info.queued_ms,
                            info.run_ms, info.latency_ms, info.megapixels)

    def stats(self) -> "_lib.ServerStats":
        out = _lib.ServerStats()
        self.proc._check(self.proc.lib.rhd_server_read_stats(self.ptr, ctypes.byref(out)))
        return out

    def reset_stats(self) -> None:
        self.proc._check(self.proc.lib.rhd_server_reset_stats(self.ptr))


class BorrowedSource:
    """A developed source owned by a batch (valid until the batch is closed)."""

    def __init__(self, proc: GpuSensorProcessor, ptr):
        self.proc = proc
        self.ptr = ptr

    def dims(self):
        w, h = ctypes.c_uint32(), ctypes.c_uint32()
        self.proc.lib.rhd_source_dims(self.ptr, ctypes.byref(w), ctypes.byref(h))
        return w.value, h.value

    def read_rgb(self) -> np.ndarray:
        w, h = self.dims()
        out = np.empty((h, w, 3), dtype=np.float32)
        self.proc._check(self.proc.lib.rhd_source_read_rgb(
            self.proc.ctx, self.ptr, out.ctypes.data_as(ctypes.c_void_p), out.size))
        return out


class GpuBatch:
    """Batched develop of resident mosaics of uneven sizes and recipes in one
    HIP graph launch (include/rhd.h rhd_batch_*). `lanes` = 0 autotunes the
    number of parallel branches."""

    def __init__(self, proc: GpuSensorProcessor, mosaics: Sequence[ResidentMosaic],
                 descs: Sequence[SensorDevelopmentDescription], details, lanes: int = 0):
        n = len(mosaics)
        self.proc = proc
        self.count = n
        self._mosaics = list(mosaics)  # must outlive the batch
        cdescs = (_lib.SensorDesc * n)()
        keep = []
        for i, d in enumerate(descs):
            c = d.to_c()
            keep.append(d._keep)
            cdescs[i] = c
        cdetails = (_lib.DetailDesc * n)(*details)
        ptrs = (ctypes.c_void_p * n)(*[m.ptr.value for m in mosaics])
        out = ctypes.c_void_p()
        proc._check(proc.lib.rhd_batch_create(proc.ctx, n, ptrs, cdescs, cdetails, lanes,
                                              ctypes.byref(out)))
        self.ptr = out

    @property
    def lanes(self) -> int:
        return int(self.proc.lib.rhd_batch_lanes(self.ptr))

    def run(self):
        """Launches the graph once; returns (per-frame statuses, device ms)."""
        statuses = (ctypes.c_int32 * self.count)()
        ms = ctypes.c_float()
        self.proc.lib.rhd_batch_run(self.ptr, statuses, None, ctypes.byref(ms))
        return list(statuses), ms.value

    def source(self, i: int) -> BorrowedSource:
        return BorrowedSource(self.proc, self.proc.lib.rhd_batch_source(self.ptr, i))

    def close(self) -> None:
        if getattr(self, "ptr", None):
            self.proc.lib.rhd_batch_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def no_capture_detail() -> "_lib.DetailDesc":
    """A detail contract with capture sharpening off (in-range settings)."""
    dd = _lib.DetailDesc()
    dd.capture_enabled = 0
    dd.amount, dd.sigma, dd.noise_protection = 0.0, 0.5, 0.0
    dd.ceilings[:] = [1.0, 1.0, 1.0]
    dd.iterations = 0
    return dd
