"""Opposed highlight reconstruction: the oracle against the reference's own
fixtures (crates/highlight/tests/opposed.rs), and the HIP path against the
oracle, bit for bit (the device's cbrt and log2 are the platform libm's,
which the reference's f32::cbrt and f32::log2 call; tests/test_libm.py)."""
import numpy as np
import pytest

from rohditor_amd import sensor, synth

F = np.float32
COLOR = {0: (0, 1, 1, 2), 1: (2, 1, 1, 0), 2: (1, 0, 2, 1), 3: (1, 2, 0, 1)}


def color_at(pattern, x, y):
    return COLOR[pattern][((y & 1) << 1) | (x & 1)]


def double_clipped_fixture(pattern):
    # crates/highlight/tests/opposed.rs:25-47 (visible 14x14 part)
    m = np.full((14, 14), F(0.01), F)
    for y in range(14):
        for x in range(14):
            c = color_at(pattern, x, y)
            if (x // 2, y // 2) == (3, 3):
                m[y, x] = F(0.8) if c == 2 else F(0.9)
            elif (x // 2, y // 2) in [(2, 2), (3, 2), (4, 2)]:
                m[y, x] = [F(0.8), F(0.7), F(0.4)][c]
    return m


def single_clipped_fixture():
    # crates/highlight/tests/opposed.rs:49-75
    m = np.full((14, 14), F(0.01), F)
    for y in range(14):
        for x in range(14):
            c = color_at(0, x, y)
            cell = (x // 2, y // 2)
            if cell == (3, 3):
                m[y, x] = [F(0.9), F(0.88), F(0.44)][c]
            elif cell in [(2, 2), (3, 2), (4, 2)]:
                m[y, x] = [F(0.75), F(0.5), F(0.25)][c]
    return m


def test_oracle_recovers_multiple_clipped_channels(oracle):
    # opposed.rs:78-98: exact statistics.
    st, out, stats, labels, count = oracle.opposed(double_clipped_fixture(0), 0, (0.9,) * 3)
    assert st == 0
    assert stats.tolist() == [3, 3, 3, 0, 0, 0, 1, 2, 0]
    assert out[6, 6] > 0.9 and out[6, 7] > 0.9 and out[7, 6] > 0.9
    assert count == 1


@pytest.mark.parametrize("pattern", [0, 1, 2, 3])
def test_oracle_all_bayer_phases(oracle, pattern):
    # opposed.rs:100-121
    _, _, stats, _, _ = oracle.opposed(double_clipped_fixture(pattern), pattern, (0.9,) * 3)
    assert stats[0] == 3 and stats[1] == 3


def test_oracle_one_channel_recovery(oracle):
    # opposed.rs:123-141
    _, out, stats, _, _ = oracle.opposed(single_clipped_fixture(), 0, (0.9, 1.0, 1.0))
    assert stats[0] == 1 and stats[1] == 1 and stats[2] == 1
    assert np.isfinite(out).all() and out[6, 6] > 0.9


def test_oracle_nothing_suspected_is_identity(oracle):
    m = np.full((20, 20), F(0.3), F)
    _, out, stats, _, count = oracle.opposed(m, 0, (0.9,) * 3)
    assert np.array_equal(out, m) and stats.sum() == 0 and count == 0


def test_oracle_neutral_ceiling_lowers_unraised_sites(oracle):
    # A fully blown cell has no support: it falls back and is neutralised.
    m = np.full((20, 20), F(0.3), F)
    m[8:12, 8:12] = F(1.2)
    _, out, stats, _, _ = oracle.opposed(m, 0, (0.9,) * 3, neutral_ceiling=(0.45, 0.9, 0.63))
    assert stats[3] == 16 and stats[4] == 16 and stats[5] == 16
    assert out[8, 8] == F(0.45) and out[8, 9] == F(0.9) and out[9, 9] == F(0.63)


def test_oracle_labels_are_raster_ordered(oracle):
    # Two clipped blobs: the one whose first cell comes first in raster order
    # gets label 1 (regions.rs:254-268).
    m = np.full((40, 40), F(0.3), F)
    m[30:34, 2:6] = F(1.0)    # lower-left blob, first row 15
    m[4:8, 30:34] = F(1.0)    # upper-right blob, first row 2
    _, _, _, labels, count = oracle.opposed(m, 0, (0.9,) * 3)
    assert count == 2
    assert labels[3, 16] == 1 and labels[16, 2] == 2


# ---------------------------------------------------------------- GPU

def _opposed_desc(w, h, pattern, levels, ceiling):
    info = sensor.RawFileInfo(width=w, height=h, cfa=["RGGB", "BGGR", "GRBG", "GBRG"][pattern],
                              black_levels=sensor.LevelPattern([0.0]), white_levels=(1.0,))
    frame = sensor.RawFrame(info=info, mosaic=np.zeros((h, w), np.uint16))
    recipe = sensor.EditRecipe(highlights=sensor.HighlightAdjustments(
        method=sensor.HighlightMethod.OPPOSED))
    desc = sensor.SensorDevelopmentDescription.from_frame(frame, recipe, sensor.RenderOptions())
    desc.detection_levels = np.array(levels, F)
    desc.neutral_ceiling = None if ceiling is None else np.array(ceiling, F)
    return desc


def gpu_opposed(gpu, m, pattern, levels, ceiling=None):
    h, w = m.shape
    plane = gpu.upload_mosaic(m, pattern)
    metrics = gpu.apply_highlight(plane, _opposed_desc(w, h, pattern, levels, ceiling))
    o = metrics.opposed
    stats = [o.suspected_clipped_sites, o.reconstructed_sites, o.changed_sites, o.fallback_sites,
             o.fully_unsupported_sites, o.neutralised_sites, *o.suspected_by_channel]
    return plane.read(), stats, o.regions


def bits_equal(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def describe_mismatch(a, b):
    if a.shape != b.shape:
        return f"shape {a.shape} != {b.shape}"
    bad = np.argwhere(a.view(np.uint32) != b.view(np.uint32))
    first = tuple(bad[0])
    return f"{len(bad)} samples differ; first at {first}: {a[first]!r} != {b[first]!r}"


@pytest.mark.gpu
@pytest.mark.parametrize("pattern", [0, 1, 2, 3])
def test_gpu_fixture_statistics_match(gpu, oracle, pattern):
    m = double_clipped_fixture(pattern)
    out, stats, regions = gpu_opposed(gpu, m, pattern, (0.9,) * 3)
    _, expected, ostats, _, count = oracle.opposed(m, pattern, (0.9,) * 3)
    assert stats == ostats.tolist() and regions == count
    assert bits_equal(out, expected)


@pytest.mark.gpu
def test_gpu_single_channel_fixture(gpu, oracle):
    m = single_clipped_fixture()
    out, stats, _ = gpu_opposed(gpu, m, 0, (0.9, 1.0, 1.0))
    _, expected, ostats, _, _ = oracle.opposed(m, 0, (0.9, 1.0, 1.0))
    assert stats == ostats.tolist()
    assert bits_equal(out, expected)


@pytest.mark.gpu
def test_gpu_nothing_suspected_is_identity(gpu):
    m = np.full((20, 22), F(0.3), F)
    out, stats, regions = gpu_opposed(gpu, m, 0, (0.9,) * 3)
    assert np.array_equal(out, m) and sum(stats) == 0 and regions == 0


@pytest.mark.gpu
@pytest.mark.parametrize("ceiling", [None, (0.6, 0.9, 0.7)])
@pytest.mark.parametrize("seed", [1, 2])
def test_gpu_natural_scene_matches_oracle(gpu, oracle, ceiling, seed):
    spec = synth.SceneSpec(1200, 800, 0x5248_0100 + seed, clipped_fraction=0.04,
                           blob_fraction=0.06)
    m = ((synth.mosaic(spec).astype(F) - F(512)) / F(16383 - 512)).astype(F)
    levels = (F(0.987),) * 3
    out, stats, regions = gpu_opposed(gpu, m, 0, levels, ceiling)
    _, expected, ostats, _, count = oracle.opposed(m, 0, levels, ceiling)
    assert regions == count
    assert stats == ostats.tolist()
    assert bits_equal(out, expected), describe_mismatch(out, expected)


@pytest.mark.gpu
def test_gpu_develop_full_sensor_opposed_amaze_rl(gpu, oracle):
    """The 24 MP workload end to end against the oracle chain."""
    spec = synth.full_sensor_24mp()
    frame = synth.frame(spec)
    recipe = sensor.EditRecipe(
        highlights=sensor.HighlightAdjustments(method=sensor.HighlightMethod.OPPOSED),
        capture=sensor.CaptureSharpening(enabled=True, amount=0.5, radius=0.7,
                                         noise_protection=0.5),
        capture_iterations=10)
    options = sensor.RenderOptions(demosaic=sensor.DemosaicAlgorithm.AMAZE)
    desc = sensor.SensorDevelopmentDescription.from_frame(frame, recipe, options)
    detail = sensor.detail_desc(recipe, desc.demosaic_balance)
    resident = gpu.upload(frame)
    from rohditor_amd import _lib
    m = _lib.Metrics()
    src = gpu.develop_resident(resident, desc, detail, metrics=m)
    out = src.read_rgb()
    src.release()
    c = desc.crop
    norm = oracle.normalize(frame.mosaic, desc.sensor_pattern, desc.black, *desc.black_repeat,
                            desc.white, (c.x, c.y, c.width, c.height))
    _, norm, ostats, _, count = oracle.opposed(norm, desc.crop_pattern, desc.detection_levels,
                                               desc.neutral_ceiling)
    _, rgb = oracle.demosaic(norm, desc.crop_pattern, desc.demosaic_balance, "amaze",
                             threads=oracle.cpu_threads())
    ceil = sensor.capture_ceilings(recipe, desc.demosaic_balance)
    _, rgb = oracle.capture(rgb, 0.5, 0.7, 0.5, ceil, iterations=10, threads=oracle.cpu_threads())
    o = m.opposed
    stats = [o.suspected_clipped_sites, o.reconstructed_sites, o.changed_sites, o.fallback_sites,
             o.fully_unsupported_sites, o.neutralised_sites, *o.suspected_by_channel]
    assert stats == ostats.tolist() and o.regions == count
    err = np.abs(out.astype(np.float64) - rgb) / np.maximum(np.abs(rgb), 1e-3)
    print(f"C2 24 MP develop: {np.count_nonzero(out != rgb)} samples differ, "
          f"max relative error {err.max():.3g}", flush=True)
    assert bits_equal(out, rgb), describe_mismatch(out, rgb)
