"""Test images of the reference's noise-reduction tests, restated in numpy.

``gauss`` follows crates/detail/src/test_support.rs:5-18 (SplitMix64 then
Box-Muller in f64, narrowed to f32); the image builders follow the helpers of
crates/detail/src/denoise/chroma.rs:465-481, luma.rs:465-477 and
noise.rs:174-191. They feed statistical properties, not bit-level goldens.
"""
import numpy as np

F = np.float32
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def _uniform(seed: int, n: int) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + _GOLDEN * np.arange(1, n + 1, dtype=np.uint64)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) / float(1 << 53)


def gauss(seed: int, n: int) -> np.ndarray:
    """The first n values of Gauss(seed).next()."""
    u = _uniform(seed, 2 * n)
    a, b = np.maximum(u[0::2], 1e-12), u[1::2]
    return (np.sqrt(-2.0 * np.log(a)) * np.cos(2.0 * np.pi * b)).astype(F)


CHROMA_BALANCE = (2.0, 1.0, 1.5)


def colour_noise_image(w: int, h: int, sigma: float, seed: int = 11) -> np.ndarray:
    """Balanced grey ramp plus colour noise of `sigma` on C1/C2 only."""
    g_ = gauss(seed, 2 * w * h).reshape(h, w, 2)
    x = np.arange(w, dtype=F)
    luma = (F(0.2) + F(0.6) * x / F(w))[None, :]
    c1 = F(sigma) * g_[..., 0]
    c2 = F(sigma) * g_[..., 1]
    g = luma - (c1 + c2) * F(0.25)
    b = np.asarray(CHROMA_BALANCE, F)
    return np.stack([(g + c1) / b[0], g / b[1], (g + c2) / b[2]], -1).astype(F)


def grey_noise_image(w: int, h: int, sigma: float, seed: int = 4) -> np.ndarray:
    """Neutral grey ramp plus luma noise."""
    n = gauss(seed, w * h).reshape(h, w)
    x = np.arange(w, dtype=F)
    v = (F(0.2) + F(0.6) * x / F(w))[None, :] + F(sigma) * n
    return np.repeat(v[..., None], 3, -1).astype(F)


def noisy_ramp(a: float, b: float, w: int = 512, h: int = 512, seed: int = 7) -> np.ndarray:
    """Horizontal ramp 0.05..0.95 with Poisson-Gaussian noise; the top third
    carries fine texture the flat-tile selection must ignore."""
    i = np.arange(w * h)
    x = (F(0.05) + F(0.9) * (i % w).astype(F) / F(w - 1)).astype(F)
    texture = np.where(i // w < h // 3, F(0.05) * ((i % 5).astype(F) - F(2.0)), F(0.0)).astype(F)
    signal = x + texture
    noise = np.sqrt(F(a) * x + F(b)).astype(F) * gauss(seed, w * h)
    return (signal + noise).astype(F).reshape(h, w)


def balanced_split(rgb: np.ndarray, balance):
    """crates/detail/src/planes.rs:125-152 (Y, C1, C2)."""
    b = np.asarray(balance, F)
    r, g, bl = rgb[..., 0] * b[0], rgb[..., 1] * b[1], rgb[..., 2] * b[2]
    return (r + F(2.0) * g + bl) * F(0.25), r - g, bl - g
