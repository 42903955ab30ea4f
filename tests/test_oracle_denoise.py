"""Pins the noise-reduction oracle (oracle/denoise.hpp).

Each check is one the reference's own tests make (crates/detail/src/atrous.rs
tests, noise.rs tests, denoise/chroma.rs tests, denoise/luma.rs tests), or a
closed form computed a different way here (f64 numpy convolutions with
mirrored indices, brute-force sorts), or the stage order of
crates/core/src/pipeline/detail.rs:210-307.
"""
import math

import numpy as np
import pytest

from noise_images import (CHROMA_BALANCE, F, balanced_split, colour_noise_image, gauss,
                          grey_noise_image, noisy_ramp)

BAND_NOISE_GAIN = [0.8908, 0.2006, 0.0856, 0.0413, 0.0203]
NO_CEILING = (np.inf, np.inf, np.inf)


def _mirror(i, n):
    period = 2 * n
    f = np.mod(i, period)
    return np.where(f < n, f, 2 * n - 1 - f)


def _b3_f64(p, step):
    """Separable [1 4 6 4 1]/16 at dilation `step`, mirrored, in f64."""
    h, w = p.shape
    k = np.array([1, 4, 6, 4, 1], np.float64) / 16.0
    x = np.arange(w)
    y = np.arange(h)
    hor = sum(k[t] * p.astype(np.float64)[:, _mirror(x + (t - 2) * step, w)] for t in range(5))
    return sum(k[t] * hor[_mirror(y + (t - 2) * step, h), :] for t in range(5))


@pytest.mark.parametrize("w,h", [(67, 41), (5, 3), (1, 9), (33, 2)])
@pytest.mark.parametrize("step", [1, 2, 4, 8, 16])
def test_b3_smoothing_matches_a_mirrored_convolution(oracle, w, h, step):
    rng = np.random.default_rng(w * 100 + h + step)
    p = rng.uniform(-1, 1, (h, w)).astype(F)
    got = oracle.b3_smooth(p, step)
    assert np.allclose(got, _b3_f64(p, step), rtol=0, atol=2e-6)


def test_b3_constant_and_reconstruction(oracle):
    """atrous.rs:233-252: constant planes have zero detail; the bands plus the
    residual rebuild the input."""
    flat = np.full((11, 20), 0.4, F)
    for step in (1, 2, 4, 8):
        assert np.array_equal(oracle.b3_smooth(flat, step), flat)
    for w, h in [(67, 41), (5, 3), (1, 9)]:
        p = (((np.arange(w * h) * 7919) % 251) / 251.0).astype(F).reshape(h, w)
        cur, back = p, np.zeros_like(p)
        for level in range(5):
            s = oracle.b3_smooth(cur, 1 << level)
            back = back + (cur - s)
            cur = s
        assert np.abs(back + cur - p).max() < 1e-6


def test_band_gains_match_white_noise(oracle):
    """luma.rs:386-408: the per-level noise gains of the luma filter."""
    n = gauss(21, 512 * 512).reshape(512, 512)
    cur = n
    for level in range(5):
        s = oracle.b3_smooth(cur, 1 << level)
        band = (cur - s).astype(np.float64)
        std = math.sqrt(float(np.mean(band * band)))
        assert abs(std - BAND_NOISE_GAIN[level]) / BAND_NOISE_GAIN[level] < 0.08, (level, std)
        cur = s


def test_local_mean_square(oracle):
    """luma.rs:410-423 (constant and spike), plus a brute-force 7x7 window."""
    e = oracle.local_mean_square(np.full((7, 9), 3.0, F))
    assert np.all(np.abs(e - 9.0) < 1e-5)
    spike = np.zeros((11, 20), F)
    spike[5, 8] = 7.0
    e = oracle.local_mean_square(spike)
    yy, xx = np.mgrid[0:11, 0:20]
    inside = (np.abs(xx - 8) <= 3) & (np.abs(yy - 5) <= 3)
    assert np.all(np.abs(e - np.where(inside, 1.0, 0.0)) < 1e-5)
    rng = np.random.default_rng(3)
    p = rng.normal(0, 1, (13, 6)).astype(F)
    sq = p.astype(np.float64) ** 2
    ref = np.zeros_like(sq)
    for dy in range(-3, 4):
        for dx in range(-3, 4):
            ref += sq[_mirror(np.arange(13) + dy, 13)][:, _mirror(np.arange(6) + dx, 6)]
    assert np.allclose(oracle.local_mean_square(p), ref / 49.0, rtol=1e-5, atol=0)


@pytest.mark.parametrize("n", [1, 2, 7, 256, 1001])
def test_mad_matches_a_full_sort(oracle, n):
    rng = np.random.default_rng(n)
    v = rng.normal(0, 1, n).astype(F)
    s = np.sort(v)
    med = s[n // 2]
    d = np.sort(np.abs(v - med))
    assert oracle.mad(v) == d[n // 2]


def test_mad_orders_signed_zeros_totally(oracle):
    # total order puts -0 below +0: the median of {-0, +0, +0} is +0 and
    # every deviation is |x - (+0)|.
    assert math.copysign(1.0, oracle.mad(np.array([-0.0, 0.0, 0.0], F))) == 1.0
    # upper median 2 (index len/2 of [-1, 1, 2, 5]); deviations [1, 3, 3, 0] -> 3
    assert oracle.mad(np.array([1.0, -1.0, 5.0, 2.0], F)) == 3.0


@pytest.mark.parametrize("a,b", [(0.002, 0.0001), (0.01, 0.0004), (0.03, 0.001)])
def test_recovers_known_poisson_gaussian_parameters(oracle, a, b):
    """noise.rs:193-215."""
    plane = noisy_ramp(a, b)
    i = np.arange(512 * 512)
    bright = (F(0.05) + F(0.9) * (i % 512).astype(F) / F(511.0)).astype(F).reshape(512, 512)
    st, (pa, pb, bins) = oracle.estimate_noise(plane, bright)
    assert st == 0 and bins >= 8
    for x in (0.1, 0.5, 0.9):
        truth = math.sqrt(a * x + b)
        sigma = math.sqrt(pa * max(x, 0.0) + pb)
        assert abs(sigma - truth) / truth < 0.15, (a, b, x, pa, pb)


def test_small_planes_fall_back_and_tiny_ones_are_rejected(oracle):
    """noise.rs:217-232."""
    p = (F(0.5) + F(0.02) * gauss(3, 128 * 128)).reshape(128, 128)
    st, (a, b, bins) = oracle.estimate_noise(p, p)
    assert st == 0 and a == 0.0 and bins == 0
    assert abs(math.sqrt(b) - 0.02) / 0.02 < 0.15
    flat = np.full((8, 8), 0.5, F)
    assert oracle.estimate_noise(flat, flat)[0] == 1


def test_profiles_read_every_fourth_64_row_band(oracle):
    """mod.rs:112-150: tall images are profiled from rows 0..64, 256..320, ...
    stacked; the luma profile bins against itself, C1 and C2 against luma."""
    rng = np.random.default_rng(5)
    h, w = 1100, 48
    rgb = (0.3 + 0.05 * rng.normal(size=(h, w, 3))).astype(F)
    rgb[:, :, 0] += np.linspace(0, 0.5, w, dtype=F)[None, :]
    rows = [y for s in range(0, h, 256) for y in range(s, min(s + 64, h))]
    y, c1, c2 = balanced_split(rgb[rows], (1.5, 1.0, 2.0))
    st, prof = oracle.noise_profiles(rgb, (1.5, 1.0, 2.0))
    assert st == 0
    for plane, got in zip((y, c1, c2), prof):
        assert oracle.estimate_noise(plane, y) == (0, got)
    # whole image below 1024 rows
    small = rgb[:300]
    y, c1, c2 = balanced_split(small, (1.5, 1.0, 2.0))
    assert oracle.noise_profiles(small, (1.5, 1.0, 2.0))[1][1] == oracle.estimate_noise(c1, y)[1]
    assert oracle.noise_profiles(rgb[:10, :10], (1.5, 1.0, 2.0))[0] == 1
    assert oracle.noise_profiles(rgb, (0.0, 1.0, 1.0))[0] == 2


def _chroma_noise(oracle, image, balance):
    st, prof = oracle.noise_profiles(image, balance)
    if st == 1:
        return [(0.0, 1e-4)] * 3
    return [p[:2] for p in prof]


def _chroma_rms(image):
    _, c1, c2 = balanced_split(image, CHROMA_BALANCE)
    v = np.concatenate([c1.ravel(), c2.ravel()]).astype(np.float64)
    return math.sqrt(float(np.mean(v * v)))


def test_chroma_zero_strength_and_bad_strength(oracle):
    """chroma.rs:495-507."""
    img = colour_noise_image(64, 48, 0.02)
    prof = _chroma_noise(oracle, img, CHROMA_BALANCE)
    st, out = oracle.chroma_denoise(img, 0.0, CHROMA_BALANCE, NO_CEILING, prof)
    assert st == 0 and np.array_equal(out, img)
    for bad in (-0.1, 1.5, float("nan")):
        assert oracle.chroma_denoise(img, bad, CHROMA_BALANCE, NO_CEILING, prof)[0] == 1


def test_chroma_halves_colour_noise_and_keeps_luma(oracle):
    """chroma.rs:509-535."""
    noisy = colour_noise_image(150, 90, 0.03)
    prof = _chroma_noise(oracle, noisy, CHROMA_BALANCE)
    st, out = oracle.chroma_denoise(noisy, 0.5, CHROMA_BALANCE, NO_CEILING, prof)
    assert st == 0
    assert _chroma_rms(out) < 0.5 * _chroma_rms(noisy)
    ya, _, _ = balanced_split(noisy, CHROMA_BALANCE)
    yb, _, _ = balanced_split(out, CHROMA_BALANCE)
    assert np.abs(ya - yb).max() < 1e-5


def test_chroma_does_not_bleed_across_a_luma_edge(oracle):
    """chroma.rs:537-570."""
    w, h = 128, 64
    n = F(0.01) * gauss(5, w * h).reshape(h, w)
    left = np.arange(w)[None, :] < w // 2
    img = np.empty((h, w, 3), F)
    img[..., 0] = np.where(left, F(0.30), F(0.15)) + n
    img[..., 1] = np.where(left, F(0.10), F(0.50))
    img[..., 2] = np.where(left, F(0.05), F(0.80)) - n
    prof = _chroma_noise(oracle, img, (1.0, 1.0, 1.0))
    st, out = oracle.chroma_denoise(img, 1.0, (1.0, 1.0, 1.0), NO_CEILING, prof)
    assert st == 0
    for x in (w // 2 - 2, w // 2 + 1):
        assert abs(out[32, x, 0] - img[32, x, 0]) < 0.05
        assert abs(out[32, x, 2] - img[32, x, 2]) < 0.05


def test_chroma_non_finite_and_blown_pixels_pass_through(oracle):
    """chroma.rs:572-610."""
    img = colour_noise_image(96, 64, 0.03)
    img[20, 40, 0] = np.nan
    img[20, 40, 1] = np.inf
    img[30, 10] = (0.9, 0.2, 0.1)
    prof = _chroma_noise(oracle, img, CHROMA_BALANCE)
    st, out = oracle.chroma_denoise(img, 0.8, CHROMA_BALANCE, (0.8, 0.8, 0.8), prof)
    assert st == 0
    assert np.isnan(out[20, 40, 0]) and out[20, 40, 1] == np.inf
    assert np.array_equal(out[30, 10], img[30, 10])
    assert np.isfinite(out).sum() >= np.isfinite(img).sum()
    assert np.all(np.isfinite(out[20, 41]))
    assert oracle.noise_profiles(colour_noise_image(10, 10, 0.05), CHROMA_BALANCE)[0] == 1


def _luma_rms_error(image, w):
    y = (image[..., 0] + F(2.0) * image[..., 1] + image[..., 2]) * F(0.25)
    truth = (F(0.2) + F(0.6) * np.arange(w, dtype=F) / F(w))[None, :]
    d = (y - truth).astype(np.float64)
    return math.sqrt(float(np.mean(d * d)))


def test_luma_zero_strength_and_bad_parameters(oracle):
    """luma.rs:483-496."""
    img = grey_noise_image(64, 48, 0.03)
    prof = (0.0, 0.03 * 0.03)
    st, out = oracle.luma_denoise(img, 0.0, 0.5, (1, 1, 1), NO_CEILING, prof)
    assert st == 0 and np.array_equal(out, img)
    for bad in (-0.1, 1.5, float("nan")):
        assert oracle.luma_denoise(img, bad, 0.5, (1, 1, 1), NO_CEILING, prof)[0] == 1
        assert oracle.luma_denoise(img, 0.5, bad, (1, 1, 1), NO_CEILING, prof)[0] == 1


def test_luma_reduces_noise_and_keeps_neutral_colour(oracle):
    """luma.rs:498-517."""
    w, h, sigma = 150, 90, 0.03
    noisy = grey_noise_image(w, h, sigma)
    st, out = oracle.luma_denoise(noisy, 0.6, 0.5, (1, 1, 1), NO_CEILING, (0.0, sigma * sigma))
    assert st == 0
    assert _luma_rms_error(out, w) < 0.5 * _luma_rms_error(noisy, w)
    assert np.all(out[..., 0] == out[..., 1]) and np.all(out[..., 1] == out[..., 2])


def test_luma_keeps_a_strong_edge(oracle):
    """luma.rs:519-538."""
    w, h, sigma = 128, 64, 0.01
    n = F(sigma) * gauss(9, w * h).reshape(h, w)
    v = np.where(np.arange(w)[None, :] < w // 2, F(0.2), F(0.8)) + n
    img = np.repeat(v[..., None], 3, -1).astype(F)
    st, out = oracle.luma_denoise(img, 1.0, 0.5, (1, 1, 1), NO_CEILING, (0.0, sigma * sigma))
    assert st == 0
    assert out[32, w // 2 + 1, 0] - out[32, w // 2 - 2, 0] > 0.5


def test_luma_non_finite_blown_and_dark_pixels_pass_through(oracle):
    """luma.rs:540-565."""
    img = grey_noise_image(96, 64, 0.03)
    img[20, 40, 0] = np.nan
    img[30, 10] = (0.9, 0.9, 0.9)
    img[40, 50] = (-0.01, 0.0, 0.0)
    st, out = oracle.luma_denoise(img, 0.8, 0.5, (1, 1, 1), (0.88, 0.88, 0.88),
                                  (0.0, 0.03 * 0.03))
    assert st == 0
    assert np.isnan(out[20, 40, 0])
    assert np.array_equal(out[30, 10], img[30, 10])
    assert np.array_equal(out[40, 50], img[40, 50])
    assert np.all(np.isfinite(out[20, 41]))


def test_detail_stage_profiles_the_input_before_false_colour(oracle):
    """detail.rs:210-261: the profile comes from the stage input, then false
    colour, chroma NR and luminance NR run in that order."""
    img = colour_noise_image(80, 70, 0.03)
    img[::7, ::5, 0] *= 1.4
    bal, ceil = CHROMA_BALANCE, (3.0, 3.0, 3.0)
    st, out, prof = oracle.detail_stage(img, false_colour=2, chroma=0.6, luminance=0.4,
                                        luminance_detail=0.3, balance=bal, ceilings=ceil)
    assert st == 0 and prof is not None
    s0, p0 = oracle.noise_profiles(img, bal)
    assert s0 == 0 and p0 == prof
    _, x = oracle.false_color(img, 2, bal, ceil)
    _, x = oracle.chroma_denoise(x, 0.6, bal, ceil, [p[:2] for p in p0])
    _, x = oracle.luma_denoise(x, 0.4, 0.3, bal, ceil, p0[0][:2])
    assert np.array_equal(out.view(np.uint32), x.view(np.uint32))
    # too small to profile: noise reduction is skipped, false colour still runs
    tiny = img[:12, :12]
    st, out, prof = oracle.detail_stage(tiny, false_colour=1, chroma=0.6, balance=bal,
                                        ceilings=ceil)
    assert st == 0 and prof is None
    assert np.array_equal(out, oracle.false_color(tiny, 1, bal, ceil)[1])


def _guided_residual_f64(y, c, inv):
    """Five guided 3x3 binomial smoothings (taps 1, 2, 4, 8, 16 apart) of c
    with weights K_j K_k (max(1 - d^2 inv / 8, 0))^8, d the guide difference."""
    h, w = y.shape
    k = np.array([0.25, 0.5, 0.25])
    yy = y.astype(np.float64)
    cur = c.astype(np.float64)
    for level in range(5):
        s = 1 << level
        num = np.zeros_like(cur)
        den = np.zeros_like(cur)
        for j in range(3):
            rows = _mirror(np.arange(h) + (j - 1) * s, h)
            for i in range(3):
                cols = _mirror(np.arange(w) + (i - 1) * s, w)
                d = yy[rows][:, cols] - yy
                wt = k[j] * k[i] * np.maximum(1.0 - d * d * inv / 8.0, 0.0) ** 8
                num += wt * cur[rows][:, cols]
                den += wt
        cur = num / den
    return cur


def test_chroma_limits_pin_the_guided_smoothing(oracle):
    """Thresholds far above every detail leave only the residual, the five
    guided smoothings; zero thresholds keep every detail, so the planes sum
    back to the input."""
    img = colour_noise_image(53, 37, 0.05)
    img[:, 30:] *= 1.7  # a luma edge the weights must respect
    bal = CHROMA_BALANCE
    y, c1, c2 = balanced_split(img, bal)
    luma_b = 0.004
    inv = 1.0 / (math.sqrt(np.float32(luma_b)) * 4.0) ** 2
    st, out = oracle.chroma_denoise(img, 1.0, bal, NO_CEILING,
                                    [(0.0, luma_b), (0.0, 1e30), (0.0, 1e30)])
    assert st == 0
    _, o1, o2 = balanced_split(out, bal)
    assert np.allclose(o1, _guided_residual_f64(y, c1, inv), atol=2e-6)
    assert np.allclose(o2, _guided_residual_f64(y, c2, inv), atol=2e-6)
    st, out = oracle.chroma_denoise(img, 1.0, bal, NO_CEILING,
                                    [(0.0, luma_b), (0.0, 0.0), (0.0, 0.0)])
    assert st == 0 and np.allclose(out, img, atol=2e-6)


def test_luma_limits_pin_the_b3_chain(oracle):
    """A profile far above the signal shrinks every band to zero, leaving the
    five-level B3 residual as the new luma (applied as a common gain); a zero
    profile keeps every band, so the luma sums back to the input."""
    rng = np.random.default_rng(12)
    h, w = 41, 58
    base = (0.3 + 0.2 * rng.random((h, w))).astype(F)
    img = np.stack([base * 1.1, base, base * 0.9], -1).astype(F)
    bal = (1.0, 1.0, 1.0)
    y = (img[..., 0] + F(2.0) * img[..., 1] + img[..., 2]) * F(0.25)
    res = y.astype(np.float64)
    for level in range(5):
        res = _b3_f64(res, 1 << level)
    st, out = oracle.luma_denoise(img, 1.0, 0.0, bal, NO_CEILING, (0.0, 1e30))
    assert st == 0
    gain = np.clip(res / y, 0.0, 4.0)
    assert np.allclose(out, img * gain[..., None], rtol=2e-6, atol=0)
    st, out = oracle.luma_denoise(img, 1.0, 0.5, bal, NO_CEILING, (0.0, 0.0))
    assert st == 0 and np.allclose(out, img, rtol=2e-6, atol=0)
