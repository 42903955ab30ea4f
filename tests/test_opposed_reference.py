"""The reference's Opposed tests, on the oracle and on the HIP path.

Ports crates/highlight/tests/opposed.rs and opposed_regions.rs (their
fixtures, closed forms and exact statistics), plus closed-form fixtures of
this project for the parts those leave open: the trimmed-mean window of the
region corrections (region_correction.rs:53-110), the local median and the
local/region blend weight (opposed.rs:258-327), the 0.75 EV chromaticity
gate (support.rs:19,116-125) and the 4 x level bound (opposed.rs:350-364). Every case runs on both engines; on the
GPU the output must also equal the oracle's bit for bit. test_mutations
shows that each planted oracle mistake fails at least one of these pins.

The reference's fixtures carry three padding columns (row stride 17); the
mosaics here are the visible samples, since a resident plane is dense.
"""
import numpy as np
import pytest

from rohditor_amd import sensor
from test_opposed import bits_equal, color_at, describe_mismatch, double_clipped_fixture, \
    gpu_opposed, single_clipped_fixture

F = np.float32
ENGINES = [pytest.param("oracle"), pytest.param("gpu", marks=pytest.mark.gpu)]
SUSPECTED, RECONSTRUCTED, CHANGED, FALLBACK, UNSUPPORTED, NEUTRALISED = range(6)


class Rejected(Exception):
    pass


def runner(engine, request):
    """run(m, levels, ceiling=None, pattern=0) -> (out, stats[9]); raises
    Rejected for inputs the reference refuses."""
    o = request.getfixturevalue("oracle")

    def run_oracle(m, levels, ceiling=None, pattern=0):
        st, out, stats, _, _ = o.opposed(m, pattern, levels, ceiling)
        if st != 0:
            raise Rejected(st)
        return out, [int(v) for v in stats]

    if engine == "oracle":
        return run_oracle
    gpu = request.getfixturevalue("gpu")

    def run_gpu(m, levels, ceiling=None, pattern=0):
        try:
            out, stats, _ = gpu_opposed(gpu, m, pattern, levels, ceiling)
        except sensor.GpuPreviewError as e:
            raise Rejected(e.status)
        expected, ostats = run_oracle(m, levels, ceiling, pattern)
        assert stats == ostats
        assert bits_equal(out, expected), describe_mismatch(out, expected)
        return out, [int(v) for v in stats]

    return run_gpu


def cbrt_reference(*values):
    """support.rs:51-72 in closed form: the cube of the mean cube root."""
    roots = [float(v) ** (1.0 / 3.0) for v in values]
    return (sum(roots) / len(roots)) ** 3


def mosaic_from_cells(cells_wide, cells_high, colour):
    """opposed_regions.rs:28-44: RGGB mosaic whose cell (cx, cy) holds
    colour(cx, cy), both greens equal."""
    m = np.empty((cells_high * 2, cells_wide * 2), F)
    for y in range(cells_high * 2):
        for x in range(cells_wide * 2):
            m[y, x] = F(colour(x // 2, y // 2)[color_at(0, x, y)])
    return m


LEVEL = 0.9
LEVELS = (F(LEVEL),) * 3
BACKGROUND = (0.02, 0.02, 0.02)
INTERIOR = (0.7, 0.95, 0.6)
SHOULDER = (0.5, 0.8, 0.4)


def large_region_colour(cx, cy):
    inside = 10 <= cx < 30 and 10 <= cy < 30
    near = 9 <= cx < 31 and 9 <= cy < 31
    return INTERIOR if inside else SHOULDER if near else BACKGROUND


def expected_green():
    # opposed_regions.rs:66-69
    return cbrt_reference(INTERIOR[0], INTERIOR[2]) + (
        SHOULDER[1] - cbrt_reference(SHOULDER[0], SHOULDER[2]))


# ---------------------------------------------------------------- pins
# Each takes run(m, levels, ceiling=None, pattern=0) and asserts.

def pin_deep_interior_from_border(run):
    # opposed_regions.rs:71-94
    out, st = run(mosaic_from_cells(40, 40, large_region_colour), LEVELS)
    assert st[SUSPECTED] == 800 and st[RECONSTRUCTED] == 800
    assert st[FALLBACK] == 0 and st[CHANGED] == 800
    e = expected_green()
    assert e > INTERIOR[1]
    for x, y in [(40, 41), (41, 40), (20, 21), (59, 58)]:
        assert abs(float(out[y, x]) - e) < 1e-5, (x, y, float(out[y, x]), e)


def _with_lone_cell(cx, cy):
    return INTERIOR if (cx, cy) == (50, 20) else large_region_colour(cx, cy)


def pin_pool_borrowing(run):
    # opposed_regions.rs:96-118
    out, st = run(mosaic_from_cells(60, 40, _with_lone_cell), LEVELS)
    e = expected_green()
    for x, y in [(100, 41), (101, 40)]:
        assert abs(float(out[y, x]) - e) < 1e-5, (x, y, float(out[y, x]), e)
    assert st[FALLBACK] == 0


def pin_isolated_region_falls_back(run):
    # opposed_regions.rs:120-143
    m = mosaic_from_cells(20, 20, lambda cx, cy: INTERIOR if (cx, cy) == (10, 10) else BACKGROUND)
    out, st = run(m, LEVELS)
    assert st[SUSPECTED] == 2 and st[RECONSTRUCTED] == 0 and st[FALLBACK] == 2
    assert bits_equal(out, m)


WARM, WARM_SHOULDER = (0.95, 0.8, 0.6), (0.8, 0.5, 0.25)
COOL, COOL_SHOULDER = (0.6, 0.8, 0.95), (0.25, 0.5, 0.8)


def _two_objects(with_cool):
    # opposed_regions.rs:152-174
    def colour(cx, cy):
        rows = 6 <= cy < 15
        if rows and 6 <= cx < 13:
            return WARM
        if with_cool and rows and 15 <= cx < 22:
            return COOL
        if 5 <= cy < 16 and 5 <= cx < 14:
            return WARM_SHOULDER
        if with_cool and 5 <= cy < 16 and 14 <= cx < 23:
            return COOL_SHOULDER
        return BACKGROUND
    return mosaic_from_cells(40, 20, colour)


def pin_touching_objects_stay_apart(run):
    # opposed_regions.rs:176-205
    both, _ = run(_two_objects(True), LEVELS)
    warm, _ = run(_two_objects(False), LEVELS)
    assert both[20, 18] > WARM[0]
    assert bits_equal(np.ascontiguousarray(both[:, :28]), np.ascontiguousarray(warm[:, :28]))
    assert color_at(0, 37, 21) == 2
    assert both[21, 37] > COOL[2]


def pin_ratio_guard(run):
    # opposed.rs:145-174
    m = single_clipped_fixture()
    for y in range(4, 6):
        for x in range(4, 6):
            if color_at(0, x, y) == 0:
                m[y, x] = F(0.9)
    for cx, cy in [(2, 2), (3, 2), (4, 2)]:
        for y in range(cy * 2, cy * 2 + 2):
            for x in range(cx * 2, cx * 2 + 2):
                m[y, x] = [F(0.8), F(0.2), F(0.6)][color_at(0, x, y)]
    out, st = run(m, LEVELS)
    assert st[SUSPECTED] == 1 and st[RECONSTRUCTED] == 0 and st[FALLBACK] == 1
    assert out[6, 6] == m[6, 6]


def _blown_cell():
    # opposed.rs:274-283
    m = np.full((5, 5), F(0.2), F)
    m[:2, :2] = F(1.0)
    return m


def pin_fully_unsupported(run):
    # opposed.rs:176-190
    _, st = run(_blown_cell(), LEVELS)
    assert st[SUSPECTED] == 4 and st[RECONSTRUCTED] == 0
    assert st[FALLBACK] == 4 and st[UNSUPPORTED] == 4


def pin_nothing_suspected_is_bit_identical(run):
    # opposed.rs:192-214 (BGGR, visible 5 x 3 part of the stride-7 fixture)
    m = np.array([[0.1, 0.2, 0.3, 0.4, 0.5],
                  [0.6, 0.7, 0.8, 0.1, 0.2],
                  [0.5, 0.6, 0.7, 0.8, 0.1]], F)
    out, st = run(m, LEVELS, pattern=1)
    assert bits_equal(out, m) and sum(st) == 0


def pin_invalid_inputs_are_rejected(run):
    # opposed.rs:216-252 and :452-478
    with pytest.raises(Rejected):
        run(double_clipped_fixture(0), (F("nan"), F(0.9), F(0.9)))
    bad = double_clipped_fixture(0)
    bad[2, 3] = F("nan")
    with pytest.raises(Rejected):
        run(bad, LEVELS)
    for red in [F("nan"), F(0.0)]:
        with pytest.raises(Rejected):
            run(_blown_cell(), LEVELS, (red, F(0.65), F(0.7)))
    with pytest.raises(Rejected):
        run(_blown_cell(), LEVELS, (F(0.6), F(0.65), F("nan")))


def pin_neutral_ceilings(run):
    # opposed.rs:285-322: no ceiling is the legacy output; fallback sites
    # above a ceiling land exactly on it; at or below it they stay.
    m = _blown_cell()
    out, st = run(m, LEVELS)
    assert bits_equal(out, m) and st[NEUTRALISED] == 0 and st[FALLBACK] == 4
    out, st = run(m, LEVELS, (F(0.6), F(0.65), F(0.7)))
    assert st[FALLBACK] == 4 and st[NEUTRALISED] == 4 and st[CHANGED] == 0
    assert out[0, 0] == F(0.6) and out[0, 1] == F(0.65)
    assert out[1, 0] == F(0.65) and out[1, 1] == F(0.7)
    out, st = run(m, LEVELS, (F(1.0),) * 3)
    assert bits_equal(out, m) and st[NEUTRALISED] == 0


def pin_raised_sites_are_not_lowered(run):
    # opposed.rs:324-341
    out, st = run(double_clipped_fixture(0), LEVELS, (F(0.85),) * 3)
    assert st[CHANGED] == 3 and st[NEUTRALISED] == 0
    assert out[6, 6] > 0.85 and out[6, 7] > 0.85 and out[7, 6] > 0.85


def pin_unsuspected_sites_are_never_touched(run):
    # opposed.rs:343-356
    m = single_clipped_fixture()
    m[13, 12] = F(0.6)
    out, st = run(m, LEVELS, (F(0.5),) * 3)
    assert out[13, 12] == m[13, 12] and st[NEUTRALISED] == 0


def pin_reconstructed_but_not_applied_is_neutralised(run):
    # opposed.rs:358-380
    m = single_clipped_fixture()
    m[6, 6] = F(4.0)
    out, st = run(m, (F(0.9), F(1.0), F(1.0)), (F(0.95), F(1.0), F(1.0)))
    assert st[RECONSTRUCTED] == 1 and st[CHANGED] == 0 and st[NEUTRALISED] == 1
    assert out[6, 6] == F(0.95)


def pin_ceiling_never_feeds_estimation(run):
    # opposed.rs:382-407
    _, low = run(_blown_cell(), LEVELS, (F(0.6), F(0.65), F(0.7)))
    _, high = run(_blown_cell(), LEVELS, (F(1.0),) * 3)
    for k in (SUSPECTED, RECONSTRUCTED, FALLBACK):
        assert low[k] == high[k]
    assert low[NEUTRALISED] == 4 and high[NEUTRALISED] == 0


def pin_neutralised_counts(run):
    # opposed.rs:409-450: odd dimensions, and a grid of blown cells
    m = np.full((7, 11), F(0.2), F)
    m[:4, :4] = F(1.0)
    _, st = run(m, LEVELS, (F(0.5),) * 3)
    assert st[NEUTRALISED] == 16
    g = np.full((60, 64), F(0.2), F)
    blown = 0
    for cy in range(0, 30, 3):
        for cx in range(0, 32, 3):
            g[cy * 2:cy * 2 + 2, cx * 2:cx * 2 + 2] = F(1.0)
            blown += 1
    _, st = run(g, LEVELS, (F(0.8),) * 3)
    assert st[SUSPECTED] == blown * 4 and st[NEUTRALISED] == blown * 4


def pin_one_channel_recovery_is_bounded(run):
    # opposed.rs:123-143
    out, st = run(single_clipped_fixture(), (F(0.9), F(1.0), F(1.0)))
    assert st[SUSPECTED] == 1 and st[RECONSTRUCTED] == 1 and st[CHANGED] == 1
    assert np.isfinite(out).all() and out[6, 6] > 0.9 and out.max() <= 3.6


# Trim window, local median and blend weight, in closed form. A 20 x 20-cell
# green-clipped block (cells 10..29) whose opposing channels keep the ring's
# R/B ratio, so every chromaticity gate passes; its one-cell ring holds the
# same R and B everywhere, and green NEAR_G[y] on the nine "near" cells at
# x = 9, y = 16..24 and GB on the other 75. Each ring cell's correction is
# its green minus the ring's cube-root reference; the region correction is
# their trimmed mean (mean +- 2 sigma + 1e-5), which here drops every near
# cell (checked below) and so equals the far cells' correction b, while a
# 3-sigma window would keep them. Interior cells at Chebyshev distance
# r = 1, 2, 3 from the near cells find only near cells as local candidates:
# the median of those at radius r, blended with weight (4 - r) / 3.
RING_RB = (0.5, 0.4)
INNER_RB = (0.8, 0.64)
GB = 0.7
NEAR_G = {y: 0.85 for y in range(16, 25)}
NEAR_G[19], NEAR_G[21] = 0.82, 0.84


def _trim_colour(cx, cy):
    if 10 <= cx < 30 and 10 <= cy < 30:
        return (INNER_RB[0], 0.91, INNER_RB[1])
    if 9 <= cx < 31 and 9 <= cy < 31:
        g = NEAR_G[cy] if cx == 9 and cy in NEAR_G else GB
        return (RING_RB[0], g, RING_RB[1])
    return BACKGROUND


def pin_trim_window_median_and_blend(run):
    m = mosaic_from_cells(40, 40, _trim_colour)
    out, st = run(m, LEVELS)
    assert st[SUSPECTED] == 800 and st[FALLBACK] == 0
    ref_ring = cbrt_reference(*RING_RB)
    near = np.array([NEAR_G[y] - ref_ring for y in range(16, 25)])
    ring = np.concatenate([near, np.full(84 - near.size, GB - ref_ring)])
    mean, sigma = ring.mean(), ring.std()
    assert np.all(np.abs(near - mean) > 2 * sigma + 1e-5)
    assert np.all(np.abs(near - mean) <= 3 * sigma)
    b = GB - ref_ring
    base = cbrt_reference(*INNER_RB)
    for r in (1, 2, 3, 11):
        if r <= 3:
            weight = (4 - r) / 3
            local = float(np.median([NEAR_G[y] - ref_ring for y in range(20 - r, 21 + r)]))
        else:
            weight, local = 0.0, 0.0
        expected = base + weight * local + (1.0 - weight) * b
        cx = 9 + r
        for x, y in [(2 * cx + 1, 40), (2 * cx, 41)]:
            assert abs(float(out[y, x]) - expected) < 1e-5, (r, x, y, float(out[y, x]), expected)


def pin_chromaticity_gate(run):
    # The ratio-guard fixture of opposed.rs:145-174 with candidates whose
    # G/B ratio is 1 EV from the target's (0.88 / 0.44 against 0.5 / 0.5):
    # beyond the 0.75 EV tolerance (support.rs:19), so the site falls back.
    m = single_clipped_fixture()
    for cx, cy in [(2, 2), (3, 2), (4, 2)]:
        for y in range(cy * 2, cy * 2 + 2):
            for x in range(cx * 2, cx * 2 + 2):
                m[y, x] = [F(0.8), F(0.5), F(0.5)][color_at(0, x, y)]
    out, st = run(m, LEVELS)
    assert st[SUSPECTED] == 1 and st[RECONSTRUCTED] == 0 and st[FALLBACK] == 1
    assert out[6, 6] == m[6, 6]


def pin_estimate_bound(run):
    # opposed.rs:350-364: an estimate is applied up to 4 x the channel's
    # level. With opposing levels of 3.0 and a green level of 0.9 the deep
    # interior estimate, ref(2.9, 2.8) + (0.8 - ref(1.5, 1.2)) ~ 2.30, lies
    # between 2 and 4 x 0.9 and lands unbounded.
    levels = (F(3.0), F(0.9), F(3.0))

    def colour(cx, cy):
        if 10 <= cx < 30 and 10 <= cy < 30:
            return (2.9, 0.95, 2.8)
        if 9 <= cx < 31 and 9 <= cy < 31:
            return (1.5, 0.8, 1.2)
        return BACKGROUND
    out, st = run(mosaic_from_cells(40, 40, colour), levels)
    expected = cbrt_reference(2.9, 2.8) + (0.8 - cbrt_reference(1.5, 1.2))
    assert 2 * 0.9 < expected < 4 * 0.9
    assert st[FALLBACK] == 0
    for x, y in [(40, 41), (41, 40)]:
        assert abs(float(out[y, x]) - expected) < 1e-5, (x, y, float(out[y, x]), expected)


PINS = [pin_deep_interior_from_border, pin_pool_borrowing, pin_isolated_region_falls_back,
        pin_touching_objects_stay_apart, pin_ratio_guard, pin_fully_unsupported,
        pin_nothing_suspected_is_bit_identical, pin_invalid_inputs_are_rejected,
        pin_neutral_ceilings, pin_raised_sites_are_not_lowered,
        pin_unsuspected_sites_are_never_touched,
        pin_reconstructed_but_not_applied_is_neutralised, pin_ceiling_never_feeds_estimation,
        pin_neutralised_counts, pin_one_channel_recovery_is_bounded,
        pin_trim_window_median_and_blend, pin_chromaticity_gate, pin_estimate_bound]


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("pin", PINS, ids=lambda f: f.__name__[4:])
def test_reference_pin(engine, pin, request):
    pin(runner(engine, request))
