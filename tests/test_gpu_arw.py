"""Sony compressed ARW decoded on the GPU (rhd_arw_probe + rhd_mosaic_decode_arw)
against the oracle, bit for bit, and the decoded mosaic through the stage:
a develop from it equals a develop from the same samples uploaded as u16."""
import numpy as np
import pytest

from rohditor_amd import _lib, sensor, synth

pytestmark = pytest.mark.gpu


def _decode(gpu, data, width, height):
    m, ms = gpu.decode_arw(data)
    out = m.read(width, height)
    m.release()
    return out, ms


@pytest.mark.parametrize("w,h,curve", [(64, 3, synth.SONY_CURVE), (40, 5, synth.SONY_CURVE),
                                       (16, 2, synth.SONY_CURVE), (1000, 7, synth.SONY_CURVE),
                                       (8000, 9, (8000, 10400, 12900, 14100)),
                                       (4096, 17, (9000, 4000, 12000, 400))])
def test_random_strips_bit_exact(gpu, oracle, w, h, curve):
    """Arbitrary bytes: every field value occurs, including min > max blocks,
    positions that coincide, clamps, and rows whose length is not a multiple
    of 16 (unaligned blocks)."""
    rng = np.random.default_rng(w * 7 + h)
    strip = rng.integers(0, 256, w * h, dtype=np.uint8).tobytes()
    data = synth.arw_file(strip, w, h, curve=curve)
    st, exp = oracle.decode_arw2(strip, w, h, curve)
    assert st == 0
    got, ms = _decode(gpu, data, w, h)
    assert np.array_equal(got, exp), (w, h)
    assert ms > 0


def test_full_sensor_frame_bit_exact(gpu, oracle):
    """The 24 MP frame's mosaic, encoded as Sony compressed ARW."""
    spec = synth.full_sensor_24mp()
    mosaic = synth.mosaic(spec)
    codes = synth.sony_codes(mosaic)
    strip = synth.encode_arw2(codes)
    data = synth.arw_file(strip, spec.width, spec.height)
    st, exp = oracle.decode_arw2(strip, spec.width, spec.height, synth.SONY_CURVE)
    got, _ = _decode(gpu, data, spec.width, spec.height)
    assert st == 0 and np.array_equal(got, exp)
    # below the curve's top (15352) the round trip lands near the source samples
    below = mosaic < 15000
    assert np.abs(got.astype(np.int64) - mosaic.astype(np.int64))[below].mean() < 40


def test_decoded_mosaic_develops_like_an_uploaded_one(gpu):
    spec = synth.SceneSpec(width=1536, height=1024, seed=0xA4A1, cfa="RGGB", blob_fraction=0.02)
    codes = synth.sony_codes(synth.mosaic(spec))
    data = synth.arw_file(synth.encode_arw2(codes), spec.width, spec.height)
    resident, _ = gpu.decode_arw(data)
    samples = resident.read(spec.width, spec.height)
    frame = synth.frame_from_mosaic(samples, "RGGB")
    recipe = sensor.EditRecipe(
        highlights=sensor.HighlightAdjustments(method=sensor.HighlightMethod.OPPOSED),
        capture=sensor.CaptureSharpening(enabled=True, amount=0.5, radius=0.7,
                                         noise_protection=0.5))
    desc = sensor.SensorDevelopmentDescription.from_frame(frame, recipe, sensor.RenderOptions())
    detail = sensor.detail_desc(recipe, desc.demosaic_balance)
    a = gpu.develop_resident(resident, desc, detail)
    uploaded = gpu.upload(frame)
    b = gpu.develop_resident(uploaded, desc, detail)
    ra, rb = a.read_rgb(), b.read_rgb()
    for hnd in (a, b, resident, uploaded):
        hnd.release()
    assert np.array_equal(ra.view(np.uint32), rb.view(np.uint32))


def test_uncompressed_strips_upload_as_they_are(gpu):
    rng = np.random.default_rng(5)
    w, h = 100, 30
    samples = rng.integers(0, 16384, (h, w)).astype("<u2")
    data = synth.arw_file(samples.tobytes(), w, h, compression=1, bits_per_sample=16)
    layout = sensor.arw_probe(data)
    assert layout.kind == _lib.ARW_UNCOMPRESSED16
    got, _ = _decode(gpu, data, w, h)
    assert np.array_equal(got, samples)


def test_damaged_layouts_are_rejected(gpu):
    rng = np.random.default_rng(6)
    strip = rng.integers(0, 256, 64 * 4, dtype=np.uint8).tobytes()
    data = synth.arw_file(strip, 64, 4)
    layout = sensor.arw_probe(data)
    before = gpu.reserved_bytes()
    for field, value, status in [("bytes", 64 * 4 - 1, _lib.RAW_CORRUPT),
                                 ("offset", len(data), _lib.RAW_CORRUPT),
                                 ("kind", 7, _lib.RAW_UNSUPPORTED),
                                 ("width", 0, _lib.INVALID_DIMENSIONS)]:
        bad = sensor.arw_probe(data)
        setattr(bad, field, value)
        with pytest.raises(sensor.GpuPreviewError) as e:
            gpu.decode_arw(data, bad)
        assert e.value.status == status, field
    assert gpu.reserved_bytes() <= before + (1 << 20)
    m, _ = gpu.decode_arw(data, layout)
    m.release()


def test_server_takes_arw_frames(gpu):
    """rhd_server_put_arw: a served develop of a frame made resident from its
    ARW file equals the plain develop of the decoded samples."""
    spec = synth.SceneSpec(width=1024, height=768, seed=0xA4A2, cfa="RGGB")
    codes = synth.sony_codes(synth.mosaic(spec))
    data = synth.arw_file(synth.encode_arw2(codes), spec.width, spec.height)
    resident, _ = gpu.decode_arw(data)
    frame = synth.frame_from_mosaic(resident.read(spec.width, spec.height), "RGGB")
    recipe = sensor.EditRecipe(capture=sensor.CaptureSharpening(enabled=True))
    desc = sensor.SensorDevelopmentDescription.from_frame(frame, recipe, sensor.RenderOptions())
    detail = sensor.detail_desc(recipe, desc.demosaic_balance)
    src = gpu.develop_resident(resident, desc, detail)
    exp = src.read_rgb()
    src.release()
    resident.release()
    proc = sensor.GpuSensorProcessor(0)
    server = sensor.GpuServer(proc, lanes=1)
    server.put_arw(7, data)
    r = server.wait(server.submit(7, desc, detail))
    assert r.ok
    np.testing.assert_array_equal(r.source.read_rgb(), exp)
    r.source.release()
    with pytest.raises(sensor.GpuPreviewError) as e:
        bad = sensor.arw_probe(data)
        bad.bytes -= 1
        server.put_arw(8, data, bad)
    assert e.value.status == _lib.RAW_CORRUPT
    server.close()
    proc.close()
