"""Optics correction and exact area reduction on the GPU (rhd_reduce_preview,
rhd_optics_full) against the oracle, bit for bit: the mapping runs in f64 with
the same operation order, the cubic sampler and both reduction passes in f32
in the reference's tap order (crates/optics/src/correction.rs,
crates/core/src/resample.rs).

Model combinations follow the reference's GPU qualification matrix
(crates/gpu/src/spatial/tests.rs:283-444: every subset of distortion,
vignetting and TCA for Ptlens/linear, Poly3/Poly3 and Poly5/Poly3)."""
import itertools
import math

import numpy as np
import pytest

from rohditor_amd import _lib
from rohditor_amd import optics as opt
from rohditor_amd.sensor import GpuPreviewError

pytestmark = pytest.mark.gpu
F = np.float32


def _smooth_noisy(w, h, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    base = np.stack([0.4 + 0.3 * np.sin(xx / 9.0) * np.cos(yy / 7.0),
                     0.5 + 0.2 * np.cos((xx + yy) / 11.0),
                     0.3 + 0.25 * np.sin(yy / 5.0)], -1)
    edges = ((xx // 17 + yy // 13) % 2)[..., None] * 0.35
    return (base + edges + rng.normal(0, 0.02, base.shape)).astype(F)


def _plan(w, h, focal, distortion, tca, vignetting, mask):
    return opt.resolve_execution(
        w, h, camera_crop=1.0, focal=focal, lens_aspect_ratio=1.5, lens_crop=1.0,
        distortion=distortion if mask & 1 else None,
        vignetting=vignetting if mask & 2 else None,
        tca=tca if mask & 4 else None)


MODEL_SETS = {
    "ptlens-linear": (20.0, opt.Ptlens(0.001, -0.002, 0.001), opt.TcaLinear(1.001, 0.999)),
    "poly3-poly3": (35.0, opt.Poly3(0.01),
                    opt.TcaPoly3((1.0002, 0.00001, -0.00004), (0.9998, -0.00001, 0.00004))),
    "poly5-poly3": (50.0, opt.Poly5(0.005, -0.001),
                    opt.TcaPoly3((1.0001, 0.0, -0.00002), (0.9999, 0.0, 0.00002))),
    "c5": (24.0, opt.Poly5(0.05, -0.05), opt.TcaLinear(1.0005, 0.9995)),
}
VIG = opt.Pa(0.04, 0.004, 0.0004)


def _expect(oracle, rgb, plan, target):
    st, corrected = oracle.optics_apply(rgb, plan.to_c()) if plan else (0, rgb.copy())
    assert st == 0
    if target == (rgb.shape[1], rgb.shape[0]):
        return corrected, corrected
    st, reduced = oracle.resize_area(corrected, *target)
    assert st == 0
    return corrected, reduced


def _same(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("name,mask", list(itertools.product(sorted(MODEL_SETS), range(8))))
def test_optics_and_reduction_bit_exact(gpu, oracle, name, mask):
    focal, dist, tca = MODEL_SETS[name]
    w, h = 157, 113
    rgb = _smooth_noisy(w, h, 0x0B7 + mask)
    plan = _plan(w, h, focal, dist, tca, VIG, mask)
    target = opt.preview_dimensions(w, h, 61)
    corrected, reduced = _expect(oracle, rgb, plan, target)
    src = gpu.upload_source(rgb)
    full, _ = gpu.full_resolution_source(src, plan)
    assert _same(full.read_rgb(), corrected), (name, mask)
    small, _ = gpu.reduce_preview(src, plan, target)
    assert small.dims() == target
    assert _same(small.read_rgb(), reduced), (name, mask)
    # the source is not modified
    assert _same(src.read_rgb(), rgb)
    for s in (full, small, src):
        s.release()


@pytest.mark.parametrize("w,h,tw,th", [(31, 23, 13, 11), (64, 48, 17, 48), (50, 9, 50, 4),
                                       (97, 61, 23, 15), (300, 7, 1, 1), (257, 129, 256, 128),
                                       (1030, 6, 3, 6), (40, 30, 40, 30)])
def test_area_reduction_without_optics_bit_exact(gpu, oracle, w, h, tw, th):
    rng = np.random.default_rng(w * 131 + h)
    rgb = rng.uniform(-1, 3, (h, w, 3)).astype(F)
    rgb[0, 0] = (-0.0, 0.0, -0.0)
    _, exp = _expect(oracle, rgb, None, (tw, th))
    src = gpu.upload_source(rgb)
    out, _ = gpu.reduce_preview(src, None, (tw, th))
    assert _same(out.read_rgb(), exp)
    out.release()
    src.release()


def test_full_frame_c5_reduction_bit_exact(gpu, oracle):
    """11648x8736 (the 102 MP configuration) through Poly5 + TCA + vignetting
    and the area reduction to a 2560 px long edge, against the oracle."""
    w, h = 11648, 8736
    yy = np.linspace(0, 1, h, dtype=F)[:, None]
    xx = np.linspace(0, 1, w, dtype=F)[None, :]
    rgb = np.empty((h, w, 3), F)
    rgb[..., 0] = 0.3 + 0.4 * xx * yy
    rgb[..., 1] = 0.5 + 0.1 * np.sin(40 * xx) * np.cos(30 * yy)
    rgb[..., 2] = 0.2 + 0.5 * yy
    rgb[::97, :, 1] += 0.25
    rgb[:, ::89, 0] += 0.25
    plan = opt.resolve_execution(w, h, camera_crop=1.0, focal=24.0, distortion=opt.Poly5(0.05, -0.05),
                                 tca=opt.TcaLinear(1.0005, 0.9995), vignetting=VIG)
    target = opt.preview_dimensions(w, h, 2560)
    src = gpu.upload_source(rgb)
    out, ms = gpu.reduce_preview(src, plan, target)
    got = out.read_rgb()
    out.release()
    src.release()
    _, exp = _expect(oracle, rgb, plan, target)
    assert got.shape == (1920, 2560, 3)
    assert _same(got, exp)
    assert ms > 0


def test_errors_publish_nothing(gpu):
    rgb = np.ones((30, 40, 3), F)
    src = gpu.upload_source(rgb)
    before = gpu.reserved_bytes()
    bad_footprint = opt.OpticsExecution(40, 30, 0.02, 50.0, 0.39, 0.29, 1.0, opt.Poly5(-0.9, 0.0))
    with pytest.raises(GpuPreviewError) as e:
        gpu.reduce_preview(src, bad_footprint, (20, 15))
    assert e.value.status == _lib.INVALID_INPUT
    bad_gain = opt.OpticsExecution(40, 30, 0.05, 20.0, 0.975, 0.725, 1.0,
                                   vignetting=opt.Pa(3e38, 3e38, 0.0))
    with pytest.raises(GpuPreviewError) as e:
        gpu.full_resolution_source(src, bad_gain)
    assert e.value.status == _lib.INVALID_INPUT
    other = opt.OpticsExecution(41, 30, 0.05, 20.0, 0.975, 0.725, 1.0)
    with pytest.raises(GpuPreviewError) as e:
        gpu.reduce_preview(src, other, (20, 15))
    assert e.value.status == _lib.BASE_MISMATCH
    with pytest.raises(GpuPreviewError) as e:
        gpu.reduce_preview(src, None, (41, 30))
    assert e.value.status == _lib.INVALID_DIMENSIONS
    assert gpu.reserved_bytes() <= before + (1 << 20)  # only lane scratch may stay
    src.release()
