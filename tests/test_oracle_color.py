"""White balance + camera matrix oracle (crates/core/src/cpu/stages.rs:250-308)
pinned by the reference's colour-matrix fixtures (crates/color/src/matrix.rs
tests) and by closed forms."""
import numpy as np
import pytest

F = np.float32
# Reference constants (crates/color/src/matrix.rs:96-114): standard D65 white
# and the XYZ <-> linear Rec.2020 matrices.
D65_WHITE = [0.9504559, 1.0, 1.0890578]
XYZ_D65_TO_REC2020 = [[1.7166512, -0.35567078, -0.2533663],
                      [-0.6666843, 1.6164812, 0.015768546],
                      [0.017639857, -0.042770613, 0.94210315]]
REC2020_TO_XYZ_D65 = [[0.63695806, 0.1446169, 0.16888097],
                      [0.2627002, 0.67799807, 0.059301715],
                      [0.0, 0.028072694, 1.0609851]]


def test_rec2020_maps_d65_white_to_equal_rgb(oracle):
    # matrix.rs: rec2020_matrix_maps_d65_white_to_equal_rgb (2e-5)
    st, out = oracle.color(np.array([D65_WHITE], F), (1, 1, 1), XYZ_D65_TO_REC2020)
    assert st == 0 and np.all(np.abs(out - 1.0) <= 2e-5)


def test_forward_then_inverse_restores_a_probe(oracle):
    # an asymmetric probe, so a transposed or swapped matrix fails
    probe = np.array([[0.2, 0.7, 0.4], [0.25, 0.5, 0.75]], F)
    _, xyz = oracle.color(probe, (1, 1, 1), REC2020_TO_XYZ_D65)
    _, back = oracle.color(xyz, (1, 1, 1), XYZ_D65_TO_REC2020)
    assert np.all(np.abs(back - probe) <= 2e-6)
    _, transposed = oracle.color(probe, (1, 1, 1), np.array(REC2020_TO_XYZ_D65).T)
    assert np.max(np.abs(transposed - xyz)) > 1e-2


def test_gains_then_matrix_closed_form(oracle):
    rng = np.random.default_rng(7)
    rgb = rng.uniform(-0.2, 3.0, (1000, 3)).astype(F)
    gains = (2.1, 1.0, 1.47)
    m = rng.uniform(-1, 2, (3, 3)).astype(F)
    st, out = oracle.color(rgb, gains, m)
    assert st == 0
    exp = (rgb.astype(np.float64) * np.array(gains, F).astype(np.float64)) @ m.astype(np.float64).T
    np.testing.assert_allclose(out, exp, rtol=2e-6, atol=2e-6)
    # left-to-right f32 dot products, bit for bit
    g = rgb * np.array(gains, F)
    ref = np.stack([(m[c, 0] * g[:, 0] + m[c, 1] * g[:, 1]) + m[c, 2] * g[:, 2] for c in range(3)], 1)
    assert np.array_equal(out.view(np.uint32), ref.astype(F).view(np.uint32))


def test_identity_keeps_bits_and_invalid_gains_are_rejected(oracle):
    rgb = np.array([[-0.0, 1e-30, 5.5], [7.0, -1.0, 0.25]], F)
    st, out = oracle.color(rgb, (1, 1, 1), np.eye(3))
    assert st == 0
    # identity rows add +0 terms: -0 becomes +0, everything else is kept
    assert out[0, 0] == 0.0 and not np.signbit(out[0, 0])
    assert np.array_equal(out[:, 1:], rgb[:, 1:]) and out[1, 0] == F(7.0)
    # no finiteness check in the reference: inf * 0 in the other rows gives NaN
    _, out = oracle.color(np.array([[np.inf, 1.0, 1.0]], F), (1, 1, 1), np.eye(3))
    assert np.isinf(out[0, 0]) and np.isnan(out[0, 1]) and np.isnan(out[0, 2])
    for bad in [(0.0, 1, 1), (1, -2, 1), (1, 1, float("nan")), (1, float("inf"), 1)]:
        assert oracle.color(rgb, bad, np.eye(3))[0] != 0
