"""Batched develop in one HIP graph (include/rhd.h rhd_batch_*): frames of
uneven sizes, CFA phases, highlight methods, demosaic algorithms and RL
iteration counts give exactly the plain single-frame results, on every replay
of the graph, with eager work between replays, and with the autotuned lane
count as with a fixed one."""
import numpy as np
import pytest

from rohditor_amd import _lib, sensor, synth

pytestmark = pytest.mark.gpu

H = sensor.HighlightMethod
D = sensor.DemosaicAlgorithm

FRAMES = [
    # width, height, seed, cfa, highlight, demosaic, RL iterations (None: off)
    (1280, 960, 0xBA0001, "RGGB", H.OPPOSED, D.AMAZE, 10),
    (1000, 700, 0xBA0002, "BGGR", H.CLIP, D.RCD, 0),
    (900, 1024, 0xBA0003, "GRBG", H.OFF, D.MHC, None),
    (1536, 1024, 0xBA0004, "GBRG", H.OPPOSED, D.RCD, 20),
    (640, 480, 0xBA0005, "RGGB", H.CLIP, D.AMAZE, 3),
    (30, 28, 0xBA0006, "BGGR", H.OPPOSED, D.AMAZE, 2),  # bilinear fallback size
]


def _case(width, height, seed, cfa, method, demosaic, iterations):
    spec = synth.SceneSpec(width=width, height=height, seed=seed, cfa=cfa, blob_fraction=0.03)
    frame = synth.frame(spec)
    recipe = sensor.EditRecipe(
        highlights=sensor.HighlightAdjustments(method=method),
        capture=sensor.CaptureSharpening(enabled=iterations is not None, amount=0.5, radius=0.7,
                                         noise_protection=0.5),
        capture_iterations=iterations or 0)
    desc = sensor.SensorDevelopmentDescription.from_frame(
        frame, recipe, sensor.RenderOptions(demosaic=demosaic))
    detail = (sensor.detail_desc(recipe, desc.demosaic_balance) if iterations is not None
              else sensor.no_capture_detail())
    return frame, desc, detail


@pytest.fixture(scope="module")
def batch_case():
    proc = sensor.GpuSensorProcessor(0)
    cases = [_case(*f) for f in FRAMES]
    residents = [proc.upload(frame) for frame, _, _ in cases]
    plain = []
    for (frame, desc, detail), res in zip(cases, residents):
        src = proc.develop_resident(res, desc, detail)
        plain.append(src.read_rgb())
        src.release()
    yield proc, cases, residents, plain
    for r in residents:
        r.release()
    proc.close()


@pytest.mark.parametrize("lanes", [1, 3, 0])
def test_batch_equals_plain_on_every_replay(batch_case, lanes):
    proc, cases, residents, plain = batch_case
    batch = sensor.GpuBatch(proc, residents, [c[1] for c in cases], [c[2] for c in cases],
                            lanes=lanes)
    if lanes:
        assert batch.lanes == lanes
    else:
        assert 1 <= batch.lanes <= 6
    for replay in range(3):
        statuses, ms = batch.run()
        assert statuses == [_lib.OK] * len(cases)
        assert ms > 0.0
        for i in range(len(cases)):
            np.testing.assert_array_equal(batch.source(i).read_rgb(), plain[i],
                                          err_msg=f"replay {replay}, frame {i}")
        # Eager work on the same context between replays.
        src = proc.develop_resident(residents[replay], cases[replay][1], cases[replay][2])
        np.testing.assert_array_equal(src.read_rgb(), plain[replay])
        src.release()
    batch.close()


def test_batch_rejects_invalid_detail(batch_case):
    proc, cases, residents, plain = batch_case
    detail = sensor.detail_desc(sensor.EditRecipe(), cases[0][1].demosaic_balance)
    detail.nr_chroma = 1.5
    with pytest.raises(sensor.GpuPreviewError) as e:
        sensor.GpuBatch(proc, residents[:1], [cases[0][1]], [detail], lanes=1)
    assert e.value.status == _lib.INVALID_INPUT
