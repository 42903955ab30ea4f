"""HIP path (through the C ABI) against the oracle: bit-exact where the
reference's arithmetic is reproduced operation for operation."""
import numpy as np
import pytest

from reference_vectors import (AMAZE_GOLDEN_BALANCE, AMAZE_GOLDEN_HASHES, amaze_golden_mosaic,
                               fnv1a_f32)
from rohditor_amd import sensor
from rohditor_amd import synth

pytestmark = pytest.mark.gpu
F = np.float32
PAT = {"RGGB": 0, "BGGR": 1, "GRBG": 2, "GBRG": 3}
ALGS = {"bilinear": sensor.DemosaicAlgorithm.BILINEAR, "mhc": sensor.DemosaicAlgorithm.MHC,
        "rcd": sensor.DemosaicAlgorithm.RCD, "amaze": sensor.DemosaicAlgorithm.AMAZE}


def bits_equal(a, b):
    return a.shape == b.shape and np.array_equal(np.asarray(a, F).view(np.uint32),
                                                 np.asarray(b, F).view(np.uint32))


def describe_mismatch(a, b):
    diff = np.argwhere(np.asarray(a, F).view(np.uint32) != np.asarray(b, F).view(np.uint32))
    if diff.size == 0:
        return "equal"
    i = tuple(diff[0])
    return f"{len(diff)} mismatches; first at {i}: gpu={a[i]!r} oracle={b[i]!r}"


def _desc(width, height, pattern=0, demosaic=sensor.DemosaicAlgorithm.AMAZE,
          balance=(1.0, 1.0, 1.0), highlight=sensor.HighlightMethod.OFF):
    """A description for a plane that is already normalized (staged API)."""
    info = sensor.RawFileInfo(width=width, height=height,
                              cfa=["RGGB", "BGGR", "GRBG", "GBRG"][pattern],
                              black_levels=sensor.LevelPattern([0.0]), white_levels=(1.0,),
                              as_shot_white_balance=balance)
    frame = sensor.RawFrame(info=info, mosaic=np.zeros((height, width), np.uint16))
    recipe = sensor.EditRecipe(highlights=sensor.HighlightAdjustments(method=highlight))
    return sensor.SensorDevelopmentDescription.from_frame(
        frame, recipe, sensor.RenderOptions(demosaic=demosaic))


def gpu_demosaic(gpu, mosaic, pattern, balance, alg, detail=None):
    h, w = mosaic.shape
    desc = _desc(w, h, pattern, ALGS[alg], balance)
    plane = gpu.upload_mosaic(mosaic, pattern)
    src = gpu.demosaic(plane, desc, detail)
    try:
        return src.read_rgb()
    finally:
        src.release()


# ---------------------------------------------------------------- normalize / clip

@pytest.mark.parametrize("pattern", ["RGGB", "BGGR", "GRBG", "GBRG"])
def test_normalization_fixture(gpu, pattern):
    # crates/core/src/pipeline/sensor.rs:767-803: 1.3 at the crop origin, 0.5 elsewhere.
    stride = 8
    samples = np.full((5, stride), 65535, np.uint16)
    for y in range(5):
        for x in range(6):
            samples[y, x] = [50, 60, 70, 80][(y & 1) * 2 + (x & 1)]
    samples[1, 1] = 160
    info = sensor.RawFileInfo(width=6, height=5, cfa=pattern,
                              black_levels=sensor.LevelPattern([0, 10, 20, 30], 2, 2),
                              white_levels=(100, 110, 120, 130),
                              active_area=sensor.ImageRect(0, 0, 6, 5),
                              crop_area=sensor.ImageRect(1, 1, 4, 4))
    frame = sensor.RawFrame(info=info, mosaic=samples, row_stride=stride)
    desc = sensor.SensorDevelopmentDescription.from_frame(frame, sensor.EditRecipe(),
                                                          sensor.RenderOptions())
    resident = gpu.upload(frame)
    plane = gpu.normalize(resident, desc)
    out = plane.read()
    _, _, crop_pattern = plane.dims()
    assert crop_pattern == sensor.shifted(PAT[pattern], 1, 1)
    assert out[0, 0] == F(1.3)
    assert np.all(out.ravel()[1:] == F(0.5))


@pytest.mark.parametrize("pattern", [0, 1, 2, 3])
@pytest.mark.parametrize("white", [(16000.0,), (15000.0, 16000.0, 16383.0),
                                   (16000.0, 16100.0, 16200.0, 16300.0)])
def test_normalize_bit_exact_with_crop_and_level_patterns(gpu, oracle, pattern, white):
    rng = np.random.default_rng(pattern * 7 + len(white))
    width, height, stride = 131, 97, 140
    raw = rng.integers(0, 16384, (height, stride), dtype=np.uint16)
    black = np.array([500.0, 510.0, 520.0, 530.0], F)
    info = sensor.RawFileInfo(width=width, height=height, cfa=["RGGB", "BGGR", "GRBG", "GBRG"][pattern],
                              black_levels=sensor.LevelPattern(black, 2, 2), white_levels=white,
                              crop_area=sensor.ImageRect(3, 2, 120, 90))
    frame = sensor.RawFrame(info=info, mosaic=raw, row_stride=stride)
    desc = sensor.SensorDevelopmentDescription.from_frame(frame, sensor.EditRecipe(),
                                                          sensor.RenderOptions())
    out = gpu.normalize(gpu.upload(frame), desc).read()
    expected = oracle.normalize(raw, pattern, black, 2, 2, np.array(white, F), (3, 2, 120, 90))
    assert bits_equal(out, expected), describe_mismatch(out, expected)


@pytest.mark.parametrize("pattern", [0, 1, 2, 3])
def test_clip_bit_exact_with_statistics(gpu, oracle, pattern):
    rng = np.random.default_rng(11 + pattern)
    m = rng.uniform(-0.1, 1.3, (83, 129)).astype(F)
    m[5, 7] = F(0.8)  # exactly at a limit
    limits = (0.8, 0.9, 1.1)
    desc = _desc(129, 83, pattern, highlight=sensor.HighlightMethod.CLIP)
    desc.clip_levels = np.array(limits, F)
    plane = gpu.upload_mosaic(m, pattern)
    metrics = gpu.apply_highlight(plane, desc)
    st, expected, stats = oracle.clip(m, pattern, limits)
    assert st == 0
    assert bits_equal(plane.read(), expected)
    c = metrics.clip
    got = [c.affected_sites, c.changed_sites, c.nominal_over_white_sites, *c.affected_by_channel]
    assert got == stats.tolist()


def test_clip_rejects_non_finite_samples(gpu):
    m = np.full((8, 8), F(0.5))
    m[3, 4] = np.nan
    desc = _desc(8, 8, highlight=sensor.HighlightMethod.CLIP)
    plane = gpu.upload_mosaic(m, 0)
    with pytest.raises(sensor.GpuPreviewError) as err:
        gpu.apply_highlight(plane, desc)
    assert err.value.kind == "InvalidInput" and not err.value.is_routing()


# ---------------------------------------------------------------- demosaic

@pytest.mark.parametrize("key", sorted(AMAZE_GOLDEN_HASHES))
def test_amaze_gpu_matches_reference_golden_hashes(gpu, key):
    # crates/demosaic/tests/amaze_golden.rs:50-86, through the HIP kernels.
    pattern, label = key
    balance = (1.0, 1.0, 1.0) if label == "identity" else AMAZE_GOLDEN_BALANCE
    out = gpu_demosaic(gpu, amaze_golden_mosaic(), PAT[pattern], balance, "amaze")
    assert fnv1a_f32(out) == AMAZE_GOLDEN_HASHES[key]


SIZES = [(33, 33), (64, 48), (240, 225), (257, 33), (450, 460), (513, 389), (25, 25), (24, 40),
         (2, 2), (5, 3), (195, 195)]


@pytest.mark.parametrize("alg", ["bilinear", "mhc", "rcd", "amaze"])
@pytest.mark.parametrize("size", SIZES)
def test_demosaic_bit_exact_sizes(gpu, oracle, alg, size):
    w, h = size
    rng = np.random.default_rng(w * 1000 + h)
    m = rng.uniform(-0.05, 1.2, (h, w)).astype(F)
    for pattern in range(4):
        balance = (2.1, 1.0, 1.7) if pattern % 2 else (1.0, 1.0, 1.0)
        out = gpu_demosaic(gpu, m, pattern, balance, alg)
        st, expected = oracle.demosaic(m, pattern, balance, alg, threads=8)
        assert st == 0
        assert bits_equal(out, expected), f"{alg} {size} pattern {pattern}: " + \
            describe_mismatch(out, expected)


@pytest.mark.parametrize("alg", ["mhc", "rcd", "amaze"])
def test_demosaic_bit_exact_on_a_natural_scene(gpu, oracle, alg):
    spec = synth.SceneSpec(1000, 700, 0x1234, clipped_fraction=0.03, blob_fraction=0.02)
    m = ((synth.mosaic(spec).astype(F) - F(512)) / F(16383 - 512)).astype(F)
    out = gpu_demosaic(gpu, m, 0, synth.AS_SHOT, alg)
    _, expected = oracle.demosaic(m, 0, synth.AS_SHOT, alg, threads=8)
    assert bits_equal(out, expected), describe_mismatch(out, expected)


# ---------------------------------------------------------------- capture sharpening

def _detail(amount=0.5, sigma=0.7, noise=0.5, ceilings=(1.0, 1.0, 1.0), iterations=8):
    from rohditor_amd import _lib
    dd = _lib.DetailDesc()
    dd.capture_enabled = 1
    dd.amount, dd.sigma, dd.noise_protection = amount, sigma, noise
    dd.ceilings[:] = list(ceilings)
    dd.iterations = iterations
    return dd


@pytest.mark.parametrize("size", [(130, 97), (64, 32), (7, 5), (300, 3), (257, 190)])
@pytest.mark.parametrize("sigma", [0.3, 0.7, 1.2])
def test_capture_bit_exact(gpu, oracle, size, sigma):
    w, h = size
    spec = synth.SceneSpec(max(w, 8), max(h, 8), 99 + w, clipped_fraction=0.05)
    m = ((synth.mosaic(spec)[:h, :w].astype(F) - F(512)) / F(16383 - 512)).astype(F)
    dd = _detail(amount=0.8, sigma=sigma, noise=0.4, ceilings=(0.9, 1.0, 0.95), iterations=8)
    out = gpu_demosaic(gpu, m, 0, (1, 1, 1), "mhc", detail=dd)
    _, base = oracle.demosaic(m, 0, (1, 1, 1), "mhc")
    st, expected = oracle.capture(base, 0.8, sigma, 0.4, (0.9, 1.0, 0.95), iterations=8, threads=8)
    assert st == 0
    assert bits_equal(out, expected), describe_mismatch(out, expected)


@pytest.mark.parametrize("size", [(63, 17), (65, 15), (128, 16), (129, 33), (40, 70), (100, 9),
                                  (136, 44), (136, 31), (136, 32), (136, 70), (140, 69)])
@pytest.mark.parametrize("sigma", [0.3, 0.6, 0.7, 1.2])
def test_capture_tile_edges_bit_exact(gpu, oracle, size, sigma):
    # Widths and heights one below, at and one above the 64x16 tiles of the
    # preparation and the first iteration and the 64x32 tiles of the later
    # ones, heights whose middle tile row is interior or just not, and
    # windows that reach across two image edges at once, for every radius:
    # the edge tiles' mirrored windows and the repaired ratio rows and
    # columns, through the prepared first iteration, a middle one and the
    # applying last one.
    w, h = size
    spec = synth.SceneSpec(max(w, 8), max(h, 8), 7 + w * h, clipped_fraction=0.05)
    m = ((synth.mosaic(spec)[:h, :w].astype(F) - F(512)) / F(16383 - 512)).astype(F)
    dd = _detail(amount=0.9, sigma=sigma, noise=0.3, ceilings=(0.9, 1.0, 0.95), iterations=3)
    out = gpu_demosaic(gpu, m, 0, (1, 1, 1), "mhc", detail=dd)
    _, base = oracle.demosaic(m, 0, (1, 1, 1), "mhc")
    st, expected = oracle.capture(base, 0.9, sigma, 0.3, (0.9, 1.0, 0.95), iterations=3, threads=8)
    assert st == 0
    assert bits_equal(out, expected), f"{size} sigma {sigma}: " + describe_mismatch(out, expected)


@pytest.mark.parametrize("size", [(301, 90), (298, 90), (300, 91), (264, 60)])
def test_capture_unaligned_widths_bit_exact(gpu, oracle, size):
    # Interior tiles load their windows as float4 rows only when every plane
    # row starts 16-byte aligned: widths 301 and 298 send every tile through
    # the element-wise staging, 300 and 264 mix both kinds.
    w, h = size
    spec = synth.SceneSpec(w, h, 11 + w, clipped_fraction=0.05)
    m = ((synth.mosaic(spec).astype(F) - F(512)) / F(16383 - 512)).astype(F)
    dd = _detail(amount=0.8, sigma=0.7, noise=0.4, ceilings=(1.0, 1.0, 1.0), iterations=4)
    out = gpu_demosaic(gpu, m, 0, (1, 1, 1), "mhc", detail=dd)
    _, base = oracle.demosaic(m, 0, (1, 1, 1), "mhc")
    _, expected = oracle.capture(base, 0.8, 0.7, 0.4, (1, 1, 1), iterations=4, threads=8)
    assert bits_equal(out, expected), f"{size}: " + describe_mismatch(out, expected)


@pytest.mark.parametrize("scale", [1e18, 1e20])
def test_capture_large_guides_bit_exact(gpu, oracle, scale):
    # A region scaled by 1e18 keeps every guide value below 2^64, where the
    # ratio takes the shortened division; 1e20 pushes guide values past it,
    # so every tile takes the full division. Both are the oracle's bits.
    spec = synth.SceneSpec(320, 96, 23)
    m = ((synth.mosaic(spec).astype(F) - F(512)) / F(16383 - 512)).astype(F)
    m[20:70, 100:230] *= F(scale)
    dd = _detail(amount=0.5, sigma=0.7, noise=0.5, ceilings=(1.0, 1.0, 1.0), iterations=3)
    out = gpu_demosaic(gpu, m, 0, (1, 1, 1), "mhc", detail=dd)
    _, base = oracle.demosaic(m, 0, (1, 1, 1), "mhc")
    _, expected = oracle.capture(base, 0.5, 0.7, 0.5, (1, 1, 1), iterations=3, threads=8)
    assert bits_equal(out, expected), f"scale {scale}: " + describe_mismatch(out, expected)


@pytest.mark.parametrize("iterations", [0, 1, 10, 20])
def test_capture_iteration_counts(gpu, oracle, iterations):
    spec = synth.SceneSpec(200, 150, 5)
    m = ((synth.mosaic(spec).astype(F) - F(512)) / F(16383 - 512)).astype(F)
    dd = _detail(iterations=iterations)
    out = gpu_demosaic(gpu, m, 0, (1, 1, 1), "mhc", detail=dd)
    _, base = oracle.demosaic(m, 0, (1, 1, 1), "mhc")
    _, expected = oracle.capture(base, 0.5, 0.7, 0.5, (1, 1, 1), iterations=iterations)
    assert bits_equal(out, expected), describe_mismatch(out, expected)


# ---------------------------------------------------------------- whole stage

def _oracle_develop(oracle, frame, desc, recipe, iterations=8):
    info = frame.info
    c = desc.crop
    norm = oracle.normalize(frame.mosaic, desc.sensor_pattern, desc.black, *desc.black_repeat,
                            desc.white, (c.x, c.y, c.width, c.height))
    stats = None
    if desc.highlight is sensor.HighlightMethod.CLIP:
        _, norm, stats = oracle.clip(norm, desc.crop_pattern, desc.clip_levels)
    elif desc.highlight is sensor.HighlightMethod.OPPOSED:
        _, norm, stats, _, _ = oracle.opposed(norm, desc.crop_pattern, desc.detection_levels,
                                              desc.neutral_ceiling)
    alg = {v: k for k, v in ALGS.items()}[desc.demosaic]
    _, rgb = oracle.demosaic(norm, desc.crop_pattern, desc.demosaic_balance, alg, threads=8)
    if recipe.capture.is_active():
        ceil = sensor.capture_ceilings(recipe, desc.demosaic_balance)
        _, rgb = oracle.capture(rgb, recipe.capture.amount, recipe.capture.radius,
                                recipe.capture.noise_protection, ceil, iterations=iterations,
                                threads=8)
    return rgb, stats


def test_develop_cpu_case_clip_amaze_bit_exact(gpu, oracle):
    # 1536x1024 RGGB, black 512 / white 16383, 2 % clipped; Clip + AMaZE.
    frame = synth.frame(synth.cpu_case())
    recipe = sensor.EditRecipe()
    src, metrics = gpu.develop(frame, recipe, sensor.RenderOptions())
    out = src.read_rgb()
    desc = sensor.SensorDevelopmentDescription.from_frame(frame, recipe, sensor.RenderOptions())
    expected, stats = _oracle_develop(oracle, frame, desc, recipe)
    assert bits_equal(out, expected), describe_mismatch(out, expected)
    c = metrics.clip
    assert [c.affected_sites, c.changed_sites, c.nominal_over_white_sites,
            *c.affected_by_channel] == stats.tolist()
    assert stats[0] > 0


def test_develop_with_capture_bit_exact(gpu, oracle):
    spec = synth.SceneSpec(900, 640, 0x77)
    frame = synth.frame(spec)
    recipe = sensor.EditRecipe(capture=sensor.CaptureSharpening(enabled=True, amount=0.5,
                                                                radius=0.7, noise_protection=0.5),
                               capture_iterations=10)
    src, _ = gpu.develop(frame, recipe, sensor.RenderOptions())
    desc = sensor.SensorDevelopmentDescription.from_frame(frame, recipe, sensor.RenderOptions())
    expected, _ = _oracle_develop(oracle, frame, desc, recipe, iterations=10)
    out = src.read_rgb()
    assert bits_equal(out, expected), describe_mismatch(out, expected)


# ---------------------------------------------------------------- error paths

def test_over_budget_is_a_routing_error_and_releases_everything(gpu):
    frame = synth.frame(synth.SceneSpec(512, 384, 3))
    before = gpu.reserved_bytes()
    gpu.set_budget(1 << 20)
    try:
        with pytest.raises(sensor.GpuPreviewError) as err:
            gpu.develop(frame, sensor.EditRecipe(), sensor.RenderOptions())
        assert err.value.kind == "OverBudget" and err.value.is_routing()
    finally:
        gpu.set_budget(0)
    assert gpu.reserved_bytes() <= before


def test_cancelled_token_publishes_nothing(gpu):
    frame = synth.frame(synth.SceneSpec(256, 200, 4))
    token = sensor.CancellationToken()
    token.cancel()
    before = gpu.reserved_bytes()
    with pytest.raises(sensor.GpuPreviewError) as err:
        gpu.develop(frame, sensor.EditRecipe(), sensor.RenderOptions(), cancellation=token)
    assert err.value.kind == "Cancelled"
    assert gpu.reserved_bytes() == before


def test_non_finite_demosaic_input_is_rejected(gpu):
    m = np.full((40, 40), F(0.5))
    m[10, 10] = np.inf
    with pytest.raises(sensor.GpuPreviewError) as err:
        gpu_demosaic(gpu, m, 0, (1, 1, 1), "mhc")
    assert err.value.kind == "InvalidInput"
