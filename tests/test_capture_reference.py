"""The reference's capture-sharpening tests, on the oracle and on the HIP path.

Ports crates/core/src/sharpening/tests.rs (disabled and zero amount are
bit-exact, constants and tiny images unchanged, known-blur recovery with
bounded overshoot, flat-noise protection, clipped highlights untouched,
channel ratios kept), plus closed forms of this project for what those leave
open, read off the solver's planes (guide, mask, estimate; sharpening.rs:
180-310):
- the highlight mask 1 - smoothstep(0.85, 0.98, peak) at a lone pixel;
- its expansion min(m, blur(m)^4) around a lone blown pixel, where the blur
  of the indicator is 1 - w(dx) w(dy) with the Gaussian weights of
  crates/detail/src/gaussian.rs:9-17;
- the estimate clamp to [0.5, 2] x guide, which binds at both ends on
  blurred point sources after the reference's 8 iterations.
The behaviour pins run on both engines (on the GPU the output must equal the
oracle's bit for bit); the plane pins read the oracle's planes.
test_mutations shows that each planted oracle mistake fails one of them.
"""
import math

import numpy as np
import pytest

F = np.float32
ENGINES = [pytest.param("oracle"), pytest.param("gpu", marks=pytest.mark.gpu)]
# CaptureSharpening::default() with enabled = true (crates/edit/src/sharpening.rs:35-43).
ENABLED = dict(amount=0.5, sigma=0.6, noise=0.5)
ITERATIONS = 8  # sharpening.rs:13


def runner(engine, request):
    """run(rgb, amount, sigma, noise, ceilings, iterations=8) -> RGB (H, W, 3)."""
    o = request.getfixturevalue("oracle")

    def run_oracle(rgb, amount, sigma, noise, ceilings, iterations=ITERATIONS):
        st, out = o.capture(rgb, amount, sigma, noise, ceilings, iterations=iterations)
        assert st == 0
        return out

    if engine == "oracle":
        return run_oracle
    gpu = request.getfixturevalue("gpu")
    from rohditor_amd import _lib

    def run_gpu(rgb, amount, sigma, noise, ceilings, iterations=ITERATIONS):
        dd = _lib.DetailDesc()
        dd.capture_enabled = 1
        dd.amount, dd.sigma, dd.noise_protection = amount, sigma, noise
        dd.ceilings[:] = [float(c) for c in ceilings]
        dd.iterations = iterations
        src = gpu.upload_source(np.ascontiguousarray(rgb, F))
        gpu.capture(src, dd)
        out = src.read_rgb()
        src.release()
        expected = run_oracle(rgb, amount, sigma, noise, ceilings, iterations)
        assert out.tobytes() == expected.tobytes()
        return out

    return run_gpu


def grey(plane):
    return np.repeat(np.asarray(plane, F)[..., None], 3, axis=2).copy()


def fixture(width, height):
    # sharpening/tests.rs:68-80
    y, x = np.mgrid[0:height, 0:width]
    on = ((x > width // 3) & (x < width * 2 // 3) & (y > height // 4)) | ((x == 3) & (y == 5))
    return np.where(on, F(0.65), F(0.2)).astype(F)


def mse(a, b):
    # sharpening/tests.rs:95-97 (f32 sum in index order)
    d = (np.asarray(a, F) - np.asarray(b, F)).astype(F)
    total = F(0.0)
    for v in (d * d).astype(F).ravel():
        total = F(total + v)
    return float(total / F(d.size))


# ---------------------------------------------------------------- behaviour pins

def pin_zero_amount_is_bit_exact(run, blur):
    # sharpening/tests.rs:22-48 (signed HDR values; disabled is the host not
    # calling the stage at all)
    rgb = np.array([[[-0.0, -0.1, 2.0], [0.2, 0.3, 0.4]]], F)
    out = run(rgb, 0.0, 0.6, 0.5, (1.0, 1.0, 1.0))
    assert out.tobytes() == rgb.tobytes()


def pin_constants_and_tiny_images_unchanged(run, blur):
    # sharpening/tests.rs:50-66
    for w, h in [(1, 1), (1, 7), (9, 1), (3, 5)]:
        for value in [-0.1, 0.0, 0.2, 1.0, 2.0]:
            rgb = grey(np.full((h, w), F(value)))
            out = run(rgb, ENABLED["amount"], ENABLED["sigma"], ENABLED["noise"], (1.0,) * 3)
            assert out.tobytes() == rgb.tobytes(), (w, h, value)


def pin_known_blur_is_recovered(run, blur):
    # sharpening/tests.rs:99-127
    original = fixture(53, 37)
    observation = blur(original, 0.6)
    out = run(grey(observation), ENABLED["amount"], ENABLED["sigma"], ENABLED["noise"],
              (1.0,) * 3)
    sharpened = out[..., 0]
    off, rl = mse(observation, original), mse(sharpened, original)
    overshoot = max(0.65, float(sharpened.max())) - 0.65
    undershoot = 0.2 - min(0.2, float(sharpened.min()))
    assert rl < off * 0.9, (rl, off)
    assert overshoot < 0.025 and undershoot < 0.025, (overshoot, undershoot)


def pin_flat_noise_and_clipped_highlights_are_protected(run, blur):
    # sharpening/tests.rs:129-172
    i = np.arange(53 * 37)
    noise = (F(0.15) + (((i * 7919) % 101).astype(F) / F(100.0) - F(0.5)) * F(0.002)).astype(F)
    noise = noise.reshape(37, 53)
    out = run(grey(noise), ENABLED["amount"], ENABLED["sigma"], ENABLED["noise"], (1.0,) * 3)
    mean = np.full_like(noise, F(0.15))
    gain = math.sqrt(mse(out[..., 0], mean) / mse(noise, mean))
    assert gain <= 1.01, gain
    clipped = np.where(np.arange(63) % 9 > 3, F(1.0), F(0.1)).astype(F).reshape(7, 9)
    out = run(grey(clipped), ENABLED["amount"], ENABLED["sigma"], ENABLED["noise"], (1.0,) * 3)
    blown = (np.arange(63) % 9 > 3).reshape(7, 9)
    assert np.all(out[blown] == F(1.0))


def pin_channel_ratios_are_kept(run, blur):
    # sharpening/tests.rs:174-199
    g = blur(fixture(53, 37), 0.6)
    rgb = np.stack([g * F(-0.2), g, g * F(2.0)], axis=-1).astype(F)
    out = run(rgb, ENABLED["amount"], ENABLED["sigma"], ENABLED["noise"], (10.0,) * 3)
    assert out.tobytes() != rgb.tobytes()
    assert np.all(np.abs(out[..., 0] / out[..., 1] + 0.2) < 1e-6)
    assert np.all(np.abs(out[..., 2] / out[..., 1] - 2.0) < 1e-6)


BEHAVIOUR = [pin_zero_amount_is_bit_exact, pin_constants_and_tiny_images_unchanged,
             pin_known_blur_is_recovered, pin_flat_noise_and_clipped_highlights_are_protected,
             pin_channel_ratios_are_kept]


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("pin", BEHAVIOUR, ids=lambda f: f.__name__[4:])
def test_behaviour_pin(engine, pin, request, oracle):
    pin(runner(engine, request), oracle.blur)


# ---------------------------------------------------------------- plane pins

def gaussian_weights(sigma):
    """gaussian.rs:9-17 in f64: exp(-x^2 / (2 sigma^2)) normalised, radius ceil(3 sigma)."""
    r = math.ceil(3 * sigma)
    w = [math.exp(-0.5 * (x / sigma) ** 2) for x in range(-r, r + 1)]
    s = sum(w)
    return {x: v / s for x, v in zip(range(-r, r + 1), w)}


def smoothstep(lo, hi, v):
    t = min(max((v - lo) / (hi - lo), 0.0), 1.0)
    return t * t * (3 - 2 * t)


def pin_highlight_mask_edges(oracle):
    # A lone pixel of peak p among flat 0.5 pixels (mask 1): its own mask
    # value is m = 1 - smoothstep(0.85, 0.98, p), the blur of the mask there
    # is 1 - w0^2 (1 - m), and its detail gate is open, so the result is
    # min(m, (1 - w0^2 (1 - m))^4).
    sigma = 0.7
    w0 = gaussian_weights(sigma)[0]
    for p in [0.84, 0.86, 0.9, 0.95, 0.97, 0.99]:
        img = np.full((15, 15), F(0.5))
        img[7, 7] = F(p)
        _, _, (guide, mask, _) = oracle.capture(grey(img), 0.5, sigma, 0.0, (1.0,) * 3,
                                                iterations=0, planes=True)
        m = 1.0 - smoothstep(0.85, 0.98, float(F(p)))
        expected = min(m, (1.0 - w0 * w0 * (1.0 - m)) ** 4)
        assert abs(float(mask[7, 7]) - expected) < 2e-6, (p, float(mask[7, 7]), expected)


def pin_highlight_mask_expansion(oracle):
    # A lone blown pixel (peak >= 0.98: mask 0) among flat 0.5 pixels: the
    # mask around it is min(1, blur(m)^4) = (1 - w(dx) w(dy))^4 wherever the
    # detail gate is open (0.5 w(dx) w(dy) >= 3 x 0.0005 at noise
    # protection 0), and 0 where it is shut.
    sigma = 0.7
    w = gaussian_weights(sigma)
    img = np.full((15, 15), F(0.5))
    img[7, 7] = F(1.0)
    _, _, (_, mask, _) = oracle.capture(grey(img), 0.5, sigma, 0.0, (1.0,) * 3, iterations=0,
                                        planes=True)
    assert mask[7, 7] == 0.0
    for dy in range(-3, 4):
        for dx in range(-3, 4):
            if (dx, dy) == (0, 0):
                continue
            k = w[dx] * w[dy]
            detail = 0.5 * k
            if detail >= 0.0015 * 1.01:
                expected = (1.0 - k) ** 4
            elif detail <= 0.0005 * 0.99:
                expected = 0.0
            else:
                continue
            got = float(mask[7 + dy, 7 + dx])
            assert abs(got - expected) < 2e-6, (dx, dy, got, expected)


def pin_estimate_clamp(oracle):
    # Point sources blurred by the PSF: after the reference's 8 iterations
    # the estimate is clamped to exactly 0.5 x guide in the gaps and exactly
    # 2 x guide at some peaks, and never leaves [0.5, 2] x guide.
    y, x = np.mgrid[0:30, 0:40]
    points = np.where((y % 6 == 3) & (x % 6 == 3), F(2.0), F(0.05)).astype(F)
    obs = oracle.blur(points, 0.7)
    _, _, (guide, _, est) = oracle.capture(grey(obs), 0.5, 0.7, 0.0, (10.0,) * 3,
                                           iterations=ITERATIONS, planes=True)
    lo, hi = (guide * F(0.5)).astype(F), (guide * F(2.0)).astype(F)
    assert np.all(est >= lo) and np.all(est <= hi)
    assert np.count_nonzero(est == lo) >= 10 and np.count_nonzero(est == hi) >= 10


PLANES = [pin_highlight_mask_edges, pin_highlight_mask_expansion, pin_estimate_clamp]


@pytest.mark.parametrize("pin", PLANES, ids=lambda f: f.__name__[4:])
def test_plane_pin(pin, oracle):
    pin(oracle)
