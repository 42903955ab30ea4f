"""rohditor_amd: MI355X-native sensor development for rohditor.

normalize -> highlight (Off / Clip / Opposed) -> demosaic (bilinear / MHC /
RCD / AMaZE) -> masked Richardson-Lucy capture sharpening, as hand-written
HIP kernels for gfx950 behind the C ABI of include/rhd.h (librhd.so).
"""
from .sensor import (  # noqa: F401
    CancellationToken, CaptureSharpening, DemosaicAlgorithm, EditRecipe, GpuDevelopedSource,
    GpuMosaic, GpuPreviewError, GpuSensorProcessor, HighlightAdjustments, HighlightMethod,
    ImageRect, LevelPattern, RawCropPolicy, RawFileInfo, RawFrame, RenderOptions,
    SensorDevelopmentDescription, capture_ceilings, detail_desc,
)
