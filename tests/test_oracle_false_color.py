"""False-colour suppression oracle against the reference's own tests
(crates/detail/src/false_color.rs:138-272) and against np.median for the 3x3
median network with half-sample mirrored borders."""
import numpy as np
import pytest

F = np.float32
BALANCE = (2.0, 1.0, 1.5)
NO_CEILING = (np.inf, np.inf, np.inf)


def neutral(w, h):
    """Balanced-neutral grey with a smooth luma ramp (false_color.rs:152-164)."""
    img = np.zeros((h, w, 3), F)
    for y in range(h):
        for x in range(w):
            g = F(0.2) + F(0.01) * F(x) + F(0.02) * F(y)
            img[y, x] = (g / F(BALANCE[0]), g / F(BALANCE[1]), g / F(BALANCE[2]))
    return img


def test_zero_steps_leave_the_image_bit_identical(oracle):
    img = neutral(9, 7)
    img[4, 2, 0] += F(0.3)
    st, out = oracle.false_color(img, 0, BALANCE, NO_CEILING)
    assert st == 0 and np.array_equal(out.view(np.uint32), img.view(np.uint32))


def test_removes_an_isolated_colour_speck_and_keeps_green(oracle):
    clean = neutral(9, 7)
    img = clean.copy()
    img[3, 4, 0] += F(0.4)
    img[3, 4, 2] += F(0.4)
    st, out = oracle.false_color(img, 1, BALANCE, NO_CEILING)
    assert st == 0 and np.all(np.abs(out - clean) < 1e-5)


def test_keeps_a_real_colour_edge(oracle):
    w, h = 12, 6
    img = neutral(w, h)
    img[:, w // 2:, 0] += F(0.2)
    st, out = oracle.false_color(img, 2, BALANCE, NO_CEILING)
    assert st == 0 and np.all(np.abs(out - img) < 1e-5)


def test_green_is_never_modified(oracle):
    img = neutral(10, 8)
    img += ((np.arange(img.size) * 31) % 17).reshape(img.shape).astype(F) * F(0.01)
    st, out = oracle.false_color(img, 3, BALANCE, NO_CEILING)
    assert st == 0 and not np.array_equal(out, img)
    assert np.array_equal(out[..., 1], img[..., 1])


def test_blown_and_non_finite_pixels_are_copied_through(oracle):
    img = neutral(9, 7)
    img[2, 3, 0] += F(0.5)
    img[4, 6, 0] = np.nan
    img[4, 6, 2] += F(0.5)
    ceil = (img[2, 3, 0], np.inf, np.inf)
    st, out = oracle.false_color(img, 2, BALANCE, ceil)
    assert st == 0
    assert np.array_equal(out[2, 3], img[2, 3])
    assert np.isnan(out[4, 6, 0]) and out[4, 6, 2] == img[4, 6, 2]
    assert np.isnan(out).sum() == 1


def test_rejects_invalid_balance(oracle):
    img = neutral(6, 6)
    assert oracle.false_color(img, 1, (2.0, 0.0, 1.5), NO_CEILING)[0] != 0
    assert oracle.false_color(img, 1, (2.0, np.nan, 1.5), NO_CEILING)[0] != 0


@pytest.mark.parametrize("w,h", [(1, 1), (2, 5), (3, 3), (17, 11)])
def test_median_network_matches_np_median(oracle, w, h):
    """With balance 1 and green 0 the filtered C1 is read back through red:
    one step is exactly the 3x3 median of red with half-sample mirroring."""
    rng = np.random.default_rng(w * 7 + h)
    img = np.zeros((h, w, 3), F)
    img[..., 0] = rng.permutation(w * h).reshape(h, w).astype(F) * F(0.125) - F(3)
    img[..., 2] = rng.uniform(-1, 1, (h, w)).astype(F)
    st, out = oracle.false_color(img, 1, (1, 1, 1), NO_CEILING)
    assert st == 0
    pad = np.pad(img[..., 0], 1, mode="symmetric")
    exp = np.empty((h, w), F)
    for y in range(h):
        for x in range(w):
            exp[y, x] = np.median(pad[y:y + 3, x:x + 3])
    assert np.array_equal(out[..., 0], exp)
