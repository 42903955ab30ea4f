"""rhd_capture_noalloc against rhd_capture, through the C ABI.

The no-allocation entry point keeps its working planes in a fixed device
array that every call on the GPU shares, so the test runs two calls in a row,
the second on a smaller image, and checks each bit for bit against the
allocating entry point on the same input.
"""
import numpy as np
import pytest

from rohditor_amd import _lib, sensor, synth

pytestmark = pytest.mark.gpu
F = np.float32


def _detail(sigma, iterations):
    dd = _lib.DetailDesc()
    dd.capture_enabled = 1
    dd.amount, dd.sigma, dd.noise_protection = 0.7, sigma, 0.4
    dd.ceilings[:] = [0.9, 1.0, 0.95]
    dd.iterations = iterations
    return dd


def _scene_rgb(w, h, seed):
    spec = synth.SceneSpec(w, h, seed, clipped_fraction=0.04)
    m = ((synth.mosaic(spec).astype(F) - F(512)) / F(16383 - 512)).astype(F)
    rng = np.random.default_rng(seed)
    rgb = np.stack([m, m * F(0.8), m * F(1.1)], axis=-1)
    return (rgb * rng.uniform(0.9, 1.1, size=(1, 1, 3)).astype(F)).astype(F)


def _pair(gpu, rgb, dd):
    h, w, _ = rgb.shape
    expected_src = gpu.upload_source(rgb)
    gpu.capture(expected_src, dd)
    expected = expected_src.read_rgb()
    target = gpu.upload_source(rgb)
    red, _, _ = gpu.source_planes(target)
    print(f"rhd_capture_noalloc {w}x{h} sigma {dd.sigma} iterations {dd.iterations}", flush=True)
    gpu.capture_noalloc(red, w, h, dd)
    return target.read_rgb(), expected


def test_noalloc_matches_capture_on_two_calls_in_a_row(gpu):
    # First call near the array's capacity, the second smaller: the second must
    # not see anything the first left in the shared array.
    first = _scene_rgb(1536, 1024, 0x4E41)
    out, expected = _pair(gpu, first, _detail(0.7, 10))
    assert out.tobytes() == expected.tobytes()
    second = _scene_rgb(517, 301, 0x4E42)
    out, expected = _pair(gpu, second, _detail(1.2, 3))
    assert out.tobytes() == expected.tobytes()
    assert not np.array_equal(out, second), "capture sharpening changed nothing"


def test_noalloc_rejects_images_beyond_its_scratch(gpu):
    rgb = _scene_rgb(64, 48, 7)
    src = gpu.upload_source(rgb)
    red, _, _ = gpu.source_planes(src)
    dd = _detail(0.7, 2)
    limit = 2 * 1024 * 1024
    with pytest.raises(sensor.GpuPreviewError) as err:
        gpu.capture_noalloc(red, limit // 1024 + 1, 1024, dd)
    assert err.value.status == _lib.INVALID_DIMENSIONS
    # Noise reduction needs more working memory than the array holds.
    dd.nr_chroma = 0.5
    dd.balance[:] = [2.0, 1.0, 1.5]
    with pytest.raises(sensor.GpuPreviewError) as err:
        gpu.capture_noalloc(red, 64, 48, dd)
    assert err.value.status == _lib.INVALID_INPUT
    # The source is untouched by the rejected calls.
    assert src.read_rgb().tobytes() == rgb.tobytes()
