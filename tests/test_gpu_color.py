"""rhd_color_transform on the GPU against the oracle, bit for bit."""
import numpy as np
import pytest

from rohditor_amd import _lib
from rohditor_amd.sensor import GpuPreviewError

pytestmark = pytest.mark.gpu
F = np.float32


@pytest.mark.parametrize("w,h", [(1, 1), (67, 13), (1536, 1024)])
def test_color_transform_bit_exact(gpu, oracle, w, h):
    rng = np.random.default_rng(w * h)
    rgb = rng.uniform(-0.3, 4.0, (h, w, 3)).astype(F)
    gains = (2.03125, 1.0, 1.4285715)
    m = rng.uniform(-1.0, 2.0, (3, 3)).astype(F)
    st, exp = oracle.color(rgb, gains, m)
    assert st == 0
    src = gpu.upload_source(rgb)
    ms = gpu.color_transform(src, gains, m)
    got = src.read_rgb()
    src.release()
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    assert ms >= 0.0


def test_color_transform_rejects_invalid_gains(gpu):
    src = gpu.upload_source(np.ones((4, 4, 3), F))
    with pytest.raises(GpuPreviewError) as e:
        gpu.color_transform(src, (1.0, 0.0, 1.0), np.eye(3))
    assert e.value.status == _lib.INVALID_INPUT
    assert np.array_equal(src.read_rgb(), np.ones((4, 4, 3), F))
    src.release()
