"""Serving loop (include/rhd.h rhd_server_*): results equal the plain develop,
alone and in a mix of concurrent requests; newest-wins per frame; the
demosaiced-base cache; admission by device bytes.

The reference's engine worker develops one request at a time
(crates/engine/src/coordinator/worker.rs:723-800) through a one-slot
newest-wins mailbox (scheduler.rs:43-173, tests in scheduler.rs); these tests
hold the served path to the same outputs as the plain single-request path,
which test_gpu_parity.py holds to the oracle bit for bit.
"""
import numpy as np
import pytest

from rohditor_amd import _lib, sensor, synth

pytestmark = pytest.mark.gpu


def _case(width, height, seed, cfa, method, demosaic, capture, blobs=0.03):
    spec = synth.SceneSpec(width=width, height=height, seed=seed, cfa=cfa, blob_fraction=blobs)
    frame = synth.frame(spec)
    recipe = sensor.EditRecipe(
        highlights=sensor.HighlightAdjustments(method=method),
        capture=sensor.CaptureSharpening(enabled=capture, amount=0.5, radius=0.7,
                                         noise_protection=0.5),
        capture_iterations=6)
    options = sensor.RenderOptions(demosaic=demosaic)
    desc = sensor.SensorDevelopmentDescription.from_frame(frame, recipe, options)
    detail = sensor.detail_desc(recipe, desc.demosaic_balance)
    return frame, desc, detail


H = sensor.HighlightMethod
D = sensor.DemosaicAlgorithm

CASES = [
    (1280, 960, 0x5E0001, "RGGB", H.OPPOSED, D.AMAZE, True),
    (1100, 700, 0x5E0002, "BGGR", H.CLIP, D.RCD, True),
    (900, 1020, 0x5E0003, "GRBG", H.OFF, D.MHC, False),
    (1536, 1024, 0x5E0004, "GBRG", H.OPPOSED, D.RCD, True),
]


@pytest.fixture(scope="module")
def cases():
    return [_case(*c) for c in CASES]


@pytest.fixture(scope="module")
def plain(cases):
    """Each case developed alone on the plain path (lane 0)."""
    proc = sensor.GpuSensorProcessor(0)
    out = []
    for frame, desc, detail in cases:
        resident = proc.upload(frame)
        src = proc.develop_resident(resident, desc, detail)
        out.append(src.read_rgb())
        src.release()
        resident.release()
    proc.close()
    return out


def _server(lanes=3, admission_bytes=0, base_cache_bytes=0):
    proc = sensor.GpuSensorProcessor(0)
    return proc, sensor.GpuServer(proc, lanes=lanes, admission_bytes=admission_bytes,
                                  base_cache_bytes=base_cache_bytes)


def _collect(server, rid):
    r = server.wait(rid)
    rgb = r.source.read_rgb() if r.source is not None else None
    if r.source is not None:
        r.source.release()
    return r, rgb


def test_served_alone_equals_plain(cases, plain):
    proc, server = _server(lanes=2)
    for i, (frame, desc, detail) in enumerate(cases):
        server.put_frame(i, frame)
    for i, (frame, desc, detail) in enumerate(cases):
        r, rgb = _collect(server, server.submit(i, desc, detail))
        assert r.ok, r.status
        np.testing.assert_array_equal(rgb, plain[i])
    server.close()
    proc.close()


def test_served_in_the_mix_equals_alone(cases, plain):
    """Every request of a concurrent mix (two rounds over four frames on three
    lanes, so work of different frames overlaps on the device) gives exactly
    the result it gives alone."""
    proc, server = _server(lanes=3)
    for i, (frame, desc, detail) in enumerate(cases):
        server.put_frame(i, frame)
    ids = []
    for _ in range(2):
        for i, (frame, desc, detail) in enumerate(cases):
            ids.append((i, server.submit(100 + len(ids), desc, detail)))
    # Frame ids 100.. are not resident: those requests must fail cleanly while
    # the resident ones below run in the same mix.
    mix = []
    for i, (frame, desc, detail) in enumerate(cases):
        mix.append((i, server.submit(i, desc, detail)))
    for _, rid in ids:
        r = server.wait(rid, keep=False)
        assert r.status == _lib.INVALID_INPUT
    lanes_used = set()
    for i, rid in mix:
        r, rgb = _collect(server, rid)
        assert r.ok, r.status
        lanes_used.add(r.lane)
        np.testing.assert_array_equal(rgb, plain[i])
    assert lanes_used <= {1, 2, 3}
    s = server.stats()
    assert s.completed == len(mix) and s.failed == len(ids)
    server.close()
    proc.close()


def test_newest_wins_coalesces_queued_requests(cases, plain):
    proc, server = _server(lanes=1)
    big = _case(3000, 2000, 0x5E0010, "RGGB", H.OPPOSED, D.AMAZE, True)
    server.put_frame(7, big[0])
    for i, (frame, desc, detail) in enumerate(cases):
        server.put_frame(i, frame)
    first = server.submit(7, big[1], big[2])           # occupies the only lane
    older = server.submit(0, cases[0][1], cases[0][2])  # queued behind it
    newer = server.submit(0, cases[0][1], cases[0][2])  # supersedes the queued one
    r_first = server.wait(first, keep=False)
    r_older = server.wait(older, keep=False)
    r_newer, rgb = _collect(server, newer)
    assert r_first.ok
    assert r_older.status == _lib.CANCELLED
    assert r_newer.ok
    np.testing.assert_array_equal(rgb, plain[0])
    assert server.stats().coalesced == 1
    server.close()
    proc.close()


def test_newest_wins_cancels_running_request(cases, plain):
    proc, server = _server(lanes=1)
    big = _case(3000, 2000, 0x5E0011, "RGGB", H.OPPOSED, D.AMAZE, True)
    server.put_frame(7, big[0])
    running = server.submit(7, big[1], big[2])
    # Let the first request start, then supersede it.
    import time
    time.sleep(0.002)
    newer = server.submit(7, big[1], big[2])
    r_running = server.wait(running, keep=False)
    r_newer = server.wait(newer, keep=False)
    assert r_newer.ok
    # Cancellation is polled between stages: a request that already passed
    # its last check completes.
    assert r_running.status in (_lib.CANCELLED, _lib.OK)
    server.close()
    proc.close()


def test_base_cache_reuses_demosaiced_base(cases, plain):
    proc, server = _server(lanes=2, base_cache_bytes=1 << 30)
    frame, desc, detail = cases[0]
    server.put_frame(0, frame)
    r1, rgb1 = _collect(server, server.submit(0, desc, detail))
    r2, rgb2 = _collect(server, server.submit(0, desc, detail))
    assert r1.ok and r2.ok and not r1.base_cache_hit and r2.base_cache_hit
    np.testing.assert_array_equal(rgb1, plain[0])
    np.testing.assert_array_equal(rgb2, plain[0])
    # Same sensor recipe, other capture settings: the base is reused and only
    # capture sharpening runs; the result equals a plain develop with them.
    recipe = sensor.EditRecipe(
        highlights=sensor.HighlightAdjustments(method=H.OPPOSED),
        capture=sensor.CaptureSharpening(enabled=True, amount=0.9, radius=0.5,
                                         noise_protection=0.2),
        capture_iterations=4)
    detail2 = sensor.detail_desc(recipe, desc.demosaic_balance)
    r3, rgb3 = _collect(server, server.submit(0, desc, detail2))
    assert r3.ok and r3.base_cache_hit
    resident = proc.upload(frame)
    src = proc.develop_resident(resident, desc, detail2)
    np.testing.assert_array_equal(rgb3, src.read_rgb())
    src.release()
    resident.release()
    assert server.stats().base_hits == 2
    server.close()
    proc.close()


def test_admission_limit_runs_one_at_a_time(cases, plain):
    proc, server = _server(lanes=3, admission_bytes=1)
    for i, (frame, desc, detail) in enumerate(cases):
        server.put_frame(i, frame)
    rids = [server.submit(i, desc, detail) for i, (_, desc, detail) in enumerate(cases)]
    spans = []
    for i, rid in enumerate(rids):
        r, rgb = _collect(server, rid)
        assert r.ok
        np.testing.assert_array_equal(rgb, plain[i])
        spans.append((r.queued_ms, r.queued_ms + r.run_ms))
    spans.sort()
    for (s0, e0), (s1, e1) in zip(spans, spans[1:]):
        assert s1 >= e0 - 0.5  # a request starts only after the previous one ended
    server.close()
    proc.close()


def test_served_spatial_completion_equals_plain(cases, plain):
    """Develop + optics + area reduction served on lanes equals the plain
    develop followed by rhd_reduce_preview, alone and in a concurrent mix."""
    from rohditor_amd import optics as opt
    proc, server = _server(lanes=3, base_cache_bytes=1 << 30)
    plans, expected = [], []
    for i, (frame, desc, detail) in enumerate(cases):
        server.put_frame(i, frame)
        w, h = desc.crop.width, desc.crop.height
        plan = opt.resolve_execution(w, h, camera_crop=1.0, focal=24.0,
                                     distortion=opt.Poly5(0.02 * (i - 1.5), 0.01),
                                     tca=opt.TcaLinear(1.0004, 0.9996),
                                     vignetting=opt.Pa(-0.2, 0.05, 0.0))
        target = opt.preview_dimensions(w, h, 500)
        resident = proc.upload(frame)
        src = proc.develop_resident(resident, desc, detail)
        small, _ = proc.reduce_preview(src, plan, target)
        expected.append(small.read_rgb())
        for hnd in (small, src, resident):
            hnd.release()
        plans.append((plan, target))
    for _ in range(2):  # the second round hits the base cache
        rids = [server.submit(i, desc, detail, target=plans[i][1], optics=plans[i][0])
                for i, (_, desc, detail) in enumerate(cases)]
        for i, rid in enumerate(rids):
            r, rgb = _collect(server, rid)
            assert r.ok, r.status
            np.testing.assert_array_equal(rgb, expected[i])
            assert rgb.shape[:2] == (plans[i][1][1], plans[i][1][0])
    assert server.stats().base_hits == len(cases)
    server.close()
    proc.close()
