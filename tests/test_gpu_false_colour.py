"""False-colour suppression on the GPU (the detail stage's first step) against
the oracle, bit for bit: staged on an uploaded source, inside a full develop
before capture sharpening, and through the serving loop's base cache."""
import numpy as np
import pytest

from rohditor_amd import _lib, sensor, synth

pytestmark = pytest.mark.gpu
F = np.float32


def _detail(steps, balance, ceilings, capture=False):
    dd = _lib.DetailDesc()
    dd.capture_enabled = int(capture)
    dd.amount, dd.sigma, dd.noise_protection = 0.5, 0.7, 0.5
    dd.ceilings[:] = list(ceilings)
    dd.iterations = 8
    dd.false_colour = steps
    dd.balance[:] = list(balance)
    return dd


@pytest.mark.parametrize("w,h,steps", [(1, 1, 1), (2, 5, 2), (17, 11, 3), (300, 200, 5),
                                       (1536, 1024, 2)])
def test_false_colour_staged_bit_exact(gpu, oracle, w, h, steps):
    rng = np.random.default_rng(w * 1000 + h)
    rgb = rng.uniform(0.0, 1.2, (h, w, 3)).astype(F)
    balance, ceilings = (2.0, 1.0, 1.43), (1.1, 1.1, 1.1)
    if w * h > 4:
        rgb[h // 2, w // 3, 0] = np.nan
        rgb[h // 3, w // 2] = (1.15, 0.5, 0.3)
    st, exp = oracle.false_color(rgb, steps, balance, ceilings)
    assert st == 0
    src = gpu.upload_source(rgb)
    gpu.capture(src, _detail(steps, balance, ceilings))
    got = src.read_rgb()
    src.release()
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))


def test_false_colour_in_develop_before_capture(gpu, oracle):
    spec = synth.SceneSpec(width=1100, height=700, seed=0x5E0FC, cfa="RGGB", blob_fraction=0.02)
    frame = synth.frame(spec)
    recipe = sensor.EditRecipe(
        highlights=sensor.HighlightAdjustments(method=sensor.HighlightMethod.CLIP),
        capture=sensor.CaptureSharpening(enabled=True, amount=0.6, radius=0.7,
                                         noise_protection=0.4),
        false_color_steps=3)
    opts = sensor.RenderOptions(demosaic=sensor.DemosaicAlgorithm.RCD)
    desc = sensor.SensorDevelopmentDescription.from_frame(frame, recipe, opts)
    detail = sensor.detail_desc(recipe, desc.demosaic_balance, desc.demosaic_balance)
    plain = sensor.detail_desc(sensor.EditRecipe(highlights=recipe.highlights),
                               desc.demosaic_balance)
    resident = gpu.upload(frame)
    full = gpu.develop_resident(resident, desc, detail)
    base = gpu.develop_resident(resident, desc, plain)
    got, rgb = full.read_rgb(), base.read_rgb()
    for hnd in (full, base, resident):
        hnd.release()
    st, fc = oracle.false_color(rgb, 3, detail.balance[:], detail.ceilings[:])
    assert st == 0
    st, exp = oracle.capture(fc, 0.6, 0.7, 0.4, detail.ceilings[:], iterations=8)
    assert st == 0
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))


def test_false_colour_through_the_base_cache(gpu):
    spec = synth.SceneSpec(width=900, height=640, seed=0x5E0FD, cfa="GRBG")
    frame = synth.frame(spec)
    recipe = sensor.EditRecipe(false_color_steps=2)
    desc = sensor.SensorDevelopmentDescription.from_frame(frame, recipe, sensor.RenderOptions())
    detail = sensor.detail_desc(recipe, desc.demosaic_balance, desc.demosaic_balance)
    resident = gpu.upload(frame)
    src = gpu.develop_resident(resident, desc, detail)
    exp = src.read_rgb()
    src.release()
    resident.release()
    proc = sensor.GpuSensorProcessor(0)
    server = sensor.GpuServer(proc, lanes=1, base_cache_bytes=1 << 30)
    server.put_frame(1, frame)
    for hit in (False, True):
        r = server.wait(server.submit(1, desc, detail))
        assert r.ok and r.base_cache_hit == hit
        np.testing.assert_array_equal(r.source.read_rgb(), exp)
        r.source.release()
    server.close()
    proc.close()


def test_invalid_false_colour_settings_are_rejected(gpu):
    src = gpu.upload_source(np.ones((4, 4, 3), F))
    for steps, bal in [(6, (1, 1, 1)), (-1, (1, 1, 1)), (1, (1, 0, 1))]:
        with pytest.raises(sensor.GpuPreviewError) as e:
            gpu.capture(src, _detail(steps, bal, (1, 1, 1)))
        assert e.value.status == _lib.INVALID_INPUT
    src.release()
