"""Noise reduction on the GPU against the oracle, bit for bit: the noise
profiles (crates/detail/src/noise.rs, sampled rows of
crates/detail/src/denoise/mod.rs) computed on the device, the chroma and
luminance filters staged with given profiles (crates/detail/src/denoise/
chroma.rs, luma.rs), the whole detail stage inside a develop in the
reference's order (crates/core/src/pipeline/detail.rs:210-307), in a
graph-captured batch and through the serving loop's base cache."""
import numpy as np
import pytest

from noise_images import CHROMA_BALANCE, colour_noise_image, grey_noise_image, noisy_ramp
from rohditor_amd import _lib, sensor, synth

pytestmark = pytest.mark.gpu
F = np.float32
INF3 = (np.inf, np.inf, np.inf)


def _same(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _textured(w, h, seed, sigma=0.02):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    base = np.stack([0.3 + 0.25 * xx / max(w, 1), 0.4 + 0.2 * yy / max(h, 1),
                     0.25 + 0.1 * np.sin(xx / 23.0)], -1)
    edges = ((xx // 41 + yy // 37) % 2)[..., None] * np.array([0.2, 0.1, 0.3], np.float32)
    noise = rng.normal(0, 1, (h, w, 3)) * (sigma * np.sqrt(0.2 + base))
    return (base + edges + noise).astype(F)


PROFILE_CASES = {
    "colour-noise-150x90": lambda: colour_noise_image(150, 90, 0.03),
    "fallback-64x64": lambda: _textured(64, 64, 1),
    "ramp-512": lambda: np.repeat(noisy_ramp(0.01, 0.0004)[..., None], 3, -1),
    "sampled-bands-96x1100": lambda: _textured(96, 1100, 2),
    "sampled-bands-2500x1300": lambda: _textured(2500, 1300, 3),
    "non-finite": lambda: _with_specials(_textured(333, 211, 4)),
}


def _with_specials(rgb):
    rgb = rgb.copy()
    rgb[10, 20, 0] = np.nan
    rgb[50, 60, 1] = np.inf
    rgb[70, 80] = (-0.5, 2.0, 3e38)
    return rgb


@pytest.mark.parametrize("name", sorted(PROFILE_CASES))
def test_noise_profiles_bit_exact(gpu, oracle, name):
    rgb = PROFILE_CASES[name]()
    balance = (2.0, 1.0, 1.5)
    st, exp = oracle.noise_profiles(rgb, balance)
    assert st == 0
    src = gpu.upload_source(rgb)
    got, ms = gpu.noise_profiles(src, balance)
    src.release()
    assert got == exp, (name, got, exp)
    assert ms > 0


def test_full_size_profile_bit_exact(gpu, oracle):
    """6048x4024 (the 24 MP frame): every fourth 64-row band, 24192 tiles."""
    rgb = _textured(6048, 4024, 5)
    st, exp = oracle.noise_profiles(rgb, (2.0, 1.0, 1.43))
    src = gpu.upload_source(rgb)
    got, _ = gpu.noise_profiles(src, (2.0, 1.0, 1.43))
    src.release()
    assert st == 0 and got == exp and exp[0][2] == 16


def test_too_small_to_profile(gpu):
    src = gpu.upload_source(np.full((15, 300, 3), 0.5, F))
    assert gpu.noise_profiles(src, (1.0, 1.0, 1.0))[0] is None
    src.release()


FILTER_CASES = [
    ("colour-150x90", lambda: colour_noise_image(150, 90, 0.03)),
    ("specials-333x211", lambda: _with_specials(_textured(333, 211, 6, 0.05))),
    ("tiny-5x3", lambda: _textured(5, 3, 7)),
    ("column-1x9", lambda: _textured(1, 9, 8)),
    ("narrow-3x70", lambda: _textured(3, 70, 9)),
    ("wide-2048x1536", lambda: _textured(2048, 1536, 10, 0.03)),
]
PROFILES = [(0.002, 0.0004), (0.001, 0.0002), (0.0015, 0.0003)]


@pytest.mark.parametrize("name,make", FILTER_CASES, ids=[c[0] for c in FILTER_CASES])
@pytest.mark.parametrize("strength", [0.3, 1.0])
def test_chroma_bit_exact(gpu, oracle, name, make, strength):
    rgb = make()
    ceilings = (1.2, 1.3, 1.25)
    st, exp = oracle.chroma_denoise(rgb, strength, CHROMA_BALANCE, ceilings, PROFILES)
    assert st == 0
    src = gpu.upload_source(rgb)
    ms = gpu.denoise_chroma(src, strength, CHROMA_BALANCE, ceilings, PROFILES)
    got = src.read_rgb()
    src.release()
    assert _same(got, exp), name
    assert ms > 0


@pytest.mark.parametrize("name,make", FILTER_CASES, ids=[c[0] for c in FILTER_CASES])
@pytest.mark.parametrize("strength,detail", [(0.4, 0.5), (1.0, 0.0), (0.7, 1.0)])
def test_luma_bit_exact(gpu, oracle, name, make, strength, detail):
    rgb = make()
    ceilings = (1.2, 1.3, 1.25)
    st, exp = oracle.luma_denoise(rgb, strength, detail, (1.8, 1.0, 1.3), ceilings, PROFILES[0])
    assert st == 0
    src = gpu.upload_source(rgb)
    gpu.denoise_luma(src, strength, detail, (1.8, 1.0, 1.3), ceilings, PROFILES[0])
    got = src.read_rgb()
    src.release()
    assert _same(got, exp), name


def test_grey_ramp_luma_with_measured_profile(gpu, oracle):
    rgb = grey_noise_image(150, 90, 0.03)
    st, prof = oracle.noise_profiles(rgb, (1.0, 1.0, 1.0))
    st, exp = oracle.luma_denoise(rgb, 0.6, 0.5, (1, 1, 1), INF3, prof[0])
    src = gpu.upload_source(rgb)
    got_prof, _ = gpu.noise_profiles(src, (1.0, 1.0, 1.0))
    gpu.denoise_luma(src, 0.6, 0.5, (1, 1, 1), INF3, got_prof[0])
    got = src.read_rgb()
    src.release()
    assert got_prof == prof and _same(got, exp)


def _nr_recipe(method=sensor.HighlightMethod.OPPOSED, capture=True, fc=2):
    return sensor.EditRecipe(
        highlights=sensor.HighlightAdjustments(method=method),
        capture=sensor.CaptureSharpening(enabled=capture, amount=0.5, radius=0.7,
                                         noise_protection=0.5),
        false_color_steps=fc,
        noise_reduction=sensor.NoiseReduction(chroma=0.6, luminance=0.4, luminance_detail=0.3))


def test_detail_stage_in_develop_bit_exact(gpu, oracle):
    """Profiles of the demosaiced input, false colour, chroma, luminance,
    capture: the whole stage inside one develop."""
    spec = synth.SceneSpec(width=1400, height=1100, seed=0x5E0D1, cfa="RGGB", blob_fraction=0.02)
    frame = synth.frame(spec)
    recipe = _nr_recipe()
    desc = sensor.SensorDevelopmentDescription.from_frame(frame, recipe, sensor.RenderOptions())
    detail = sensor.detail_desc(recipe, desc.demosaic_balance, desc.demosaic_balance)
    plain = sensor.detail_desc(sensor.EditRecipe(highlights=recipe.highlights),
                               desc.demosaic_balance)
    resident = gpu.upload(frame)
    full = gpu.develop_resident(resident, desc, detail)
    base = gpu.develop_resident(resident, desc, plain)
    got, rgb = full.read_rgb(), base.read_rgb()
    for hnd in (full, base, resident):
        hnd.release()
    st, exp, prof = oracle.detail_stage(
        rgb, false_colour=2, chroma=0.6, luminance=0.4, luminance_detail=0.3,
        balance=detail.balance[:], ceilings=detail.ceilings[:], capture=(0.5, 0.7, 0.5, 8))
    assert st == 0 and prof is not None
    assert _same(got, exp)


def test_noise_reduction_through_batch_and_base_cache(gpu):
    frames, descs, details, plain = [], [], [], []
    for i, (w, h, cfa) in enumerate([(900, 700, "GRBG"), (1300, 1030, "BGGR"), (12, 10, "RGGB")]):
        frame = synth.frame(synth.SceneSpec(width=w, height=h, seed=0x5E0D2 + i, cfa=cfa))
        recipe = _nr_recipe(method=sensor.HighlightMethod.CLIP, capture=i != 1, fc=i)
        desc = sensor.SensorDevelopmentDescription.from_frame(frame, recipe,
                                                              sensor.RenderOptions())
        frames.append(frame)
        descs.append(desc)
        details.append(sensor.detail_desc(recipe, desc.demosaic_balance, desc.demosaic_balance))
    residents = [gpu.upload(f) for f in frames]
    for res, desc, dd in zip(residents, descs, details):
        src = gpu.develop_resident(res, desc, dd)
        plain.append(src.read_rgb())
        src.release()
    batch = sensor.GpuBatch(gpu, residents, descs, details, lanes=2)
    for replay in range(3):
        statuses, _ = batch.run()
        assert statuses == [0, 0, 0]
        for i in range(3):
            np.testing.assert_array_equal(batch.source(i).read_rgb(), plain[i])
    batch.close()
    for r in residents:
        r.release()
    proc = sensor.GpuSensorProcessor(0)
    server = sensor.GpuServer(proc, lanes=1, base_cache_bytes=1 << 30)
    server.put_frame(1, frames[0])
    for hit in (False, True):
        r = server.wait(server.submit(1, descs[0], details[0]))
        assert r.ok and r.base_cache_hit == hit
        np.testing.assert_array_equal(r.source.read_rgb(), plain[0])
        r.source.release()
    server.close()
    proc.close()


def test_small_images_skip_noise_reduction(gpu, oracle):
    rgb = _textured(40, 12, 11)
    dd = _lib.DetailDesc()
    dd.amount, dd.sigma = 0.0, 0.7
    dd.ceilings[:] = [2.0, 2.0, 2.0]
    dd.false_colour = 1
    dd.balance[:] = [1.5, 1.0, 1.2]
    dd.nr_chroma, dd.nr_luminance, dd.nr_luminance_detail = 0.8, 0.8, 0.5
    src = gpu.upload_source(rgb)
    gpu.capture(src, dd)
    got = src.read_rgb()
    src.release()
    st, exp, prof = oracle.detail_stage(rgb, false_colour=1, chroma=0.8, luminance=0.8,
                                        balance=(1.5, 1.0, 1.2), ceilings=(2.0, 2.0, 2.0))
    assert prof is None and _same(got, exp)


def test_invalid_noise_reduction_settings_are_rejected(gpu):
    src = gpu.upload_source(np.full((20, 20, 3), 0.5, F))
    for field, value in [("nr_chroma", 1.5), ("nr_luminance", -0.1),
                         ("nr_luminance_detail", float("nan"))]:
        dd = _lib.DetailDesc()
        dd.sigma = 0.7
        dd.balance[:] = [1.0, 1.0, 1.0]
        setattr(dd, field, value)
        with pytest.raises(sensor.GpuPreviewError) as e:
            gpu.capture(src, dd)
        assert e.value.status == _lib.INVALID_INPUT
    for bad in (float("nan"), 1.5, -0.1):
        with pytest.raises(sensor.GpuPreviewError) as e:
            gpu.denoise_chroma(src, bad, (1, 1, 1), INF3, PROFILES)
        assert e.value.status == _lib.INVALID_INPUT
        with pytest.raises(sensor.GpuPreviewError) as e:
            gpu.denoise_luma(src, 0.5, bad, (1, 1, 1), INF3, PROFILES[0])
        assert e.value.status == _lib.INVALID_INPUT
    with pytest.raises(sensor.GpuPreviewError) as e:
        gpu.denoise_chroma(src, 0.5, (1, 0, 1), INF3, PROFILES)
    assert e.value.status == _lib.INVALID_INPUT
    before = src.read_rgb()
    gpu.denoise_chroma(src, 0.0, (1, 0, 1), INF3, PROFILES)  # strength 0: untouched
    assert _same(src.read_rgb(), before)
    src.release()
