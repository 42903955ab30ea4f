"""The batched, served and large-sensor workloads at their real sizes, on the
GPU, against the oracle.

- Batched (one HIP graph, autotuned lanes): frames of 12, 24, 33, 45 and
  60 MP with mixed CFA phases, highlight methods, AMaZE / RCD / MHC and RL
  iteration counts 0-20; every frame of two replays equals the oracle chain.
- Served: a Poisson schedule (40 requests/s) over resident 24-45 MP frames
  with newest-wins and the base cache, re-develops changing the highlight or
  demosaic method; every completed request equals the plain develop, and one
  request per demosaic algorithm equals the oracle chain. A frame replaced
  under its id is never served from a base of its earlier samples.
- Large sensor: the 11648 x 8736 16-bit mosaic (black 1024, white 65535)
  through Opposed + AMaZE + RL(10), then Poly5 distortion, quartic
  vignetting, linear TCA and the area reduction to a 2560 px long edge,
  against the oracle chain, for both the developed base and the preview.
"""
import numpy as np
import pytest

from rohditor_amd import _lib, sensor, synth

pytestmark = pytest.mark.gpu

H = sensor.HighlightMethod
D = sensor.DemosaicAlgorithm
ALG = {D.BILINEAR: "bilinear", D.MHC: "mhc", D.RCD: "rcd", D.AMAZE: "amaze"}


def make_case(width, height, seed, cfa, method, demosaic, iterations, black=512.0,
              white=16383.0):
    spec = synth.SceneSpec(width, height, seed, cfa=cfa, black=black, white=white,
                           blob_fraction=0.05)
    return describe(synth.frame(spec), method, demosaic, iterations)


def describe(frame, method, demosaic, iterations):
    recipe = sensor.EditRecipe(
        highlights=sensor.HighlightAdjustments(method=method),
        capture=sensor.CaptureSharpening(enabled=iterations is not None, amount=0.5, radius=0.7,
                                         noise_protection=0.5),
        capture_iterations=iterations or 0)
    desc = sensor.SensorDevelopmentDescription.from_frame(
        frame, recipe, sensor.RenderOptions(demosaic=demosaic))
    detail = (sensor.detail_desc(recipe, desc.demosaic_balance) if iterations is not None
              else sensor.no_capture_detail())
    return frame, recipe, desc, detail


def oracle_chain(o, frame, recipe, desc):
    """normalize -> highlight -> demosaic -> capture on the oracle."""
    c = desc.crop
    threads = o.cpu_threads()
    m = o.normalize(frame.mosaic, desc.sensor_pattern, desc.black, *desc.black_repeat,
                    desc.white, (c.x, c.y, c.width, c.height))
    if desc.highlight is H.CLIP:
        _, m, _ = o.clip(m, desc.crop_pattern, desc.clip_levels)
    elif desc.highlight is H.OPPOSED:
        _, m, _, _, _ = o.opposed(m, desc.crop_pattern, desc.detection_levels,
                                  desc.neutral_ceiling)
    st, rgb = o.demosaic(m, desc.crop_pattern, desc.demosaic_balance, ALG[desc.demosaic],
                         threads=threads)
    assert st == 0
    if recipe.capture.is_active():
        ceil = sensor.capture_ceilings(recipe, desc.demosaic_balance)
        st, rgb = o.capture(rgb, recipe.capture.amount, recipe.capture.radius,
                            recipe.capture.noise_protection, ceil,
                            iterations=recipe.capture_iterations, threads=threads)
        assert st == 0
    return rgb


def same(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def mismatch(a, b):
    bad = np.argwhere(a.view(np.uint32) != b.view(np.uint32))
    return f"{len(bad)} samples differ, first at {tuple(bad[0])}" if len(bad) else "equal"


# ---------------------------------------------------------------- batched

BATCH = [
    (4240, 2832, 0x52480300, "RGGB", H.OPPOSED, D.AMAZE, 20),
    (9504, 6336, 0x52480301, "BGGR", H.CLIP, D.RCD, 0),
    (6048, 4024, 0x52480302, "GRBG", H.OPPOSED, D.MHC, 7),
    (7008, 4672, 0x52480303, "RGGB", H.OFF, D.AMAZE, 13),
    (8256, 5504, 0x52480304, "BGGR", H.OPPOSED, D.RCD, None),
]


def test_batched_uneven_frames_equal_the_oracle(oracle):
    proc = sensor.GpuSensorProcessor(0, budget_bytes=200 << 30)
    cases = [make_case(*b) for b in BATCH]
    residents = [proc.upload(c[0]) for c in cases]
    print("batch: frames resident, tuning lanes", flush=True)
    batch = sensor.GpuBatch(proc, residents, [c[2] for c in cases], [c[3] for c in cases],
                            lanes=0)
    assert 1 <= batch.lanes <= 6
    outputs = []
    for replay in range(2):
        statuses, ms = batch.run()
        assert statuses == [_lib.OK] * len(cases) and ms > 0
        outputs.append([batch.source(i).read_rgb() for i in range(len(cases))])
        print(f"batch: replay {replay} {ms:.1f} ms on {batch.lanes} lanes", flush=True)
    batch.close()
    for r in residents:
        r.release()
    proc.close()
    for i, (frame, recipe, desc, _) in enumerate(cases):
        expected = oracle_chain(oracle, frame, recipe, desc)
        for replay in range(2):
            got = outputs[replay][i]
            assert same(got, expected), (BATCH[i][:2], replay, mismatch(got, expected))
        print(f"batch: frame {i} {BATCH[i][0]}x{BATCH[i][1]} equals the oracle", flush=True)


# ---------------------------------------------------------------- served

SERVED_FRAMES = [
    (6048, 4024, 0x52480400, "RGGB"),
    (7008, 4672, 0x52480401, "BGGR"),
    (8256, 5504, 0x52480402, "GRBG"),
    (6048, 4024, 0x52480403, "GBRG"),
]
RECIPES = [(H.OPPOSED, D.AMAZE), (H.CLIP, D.AMAZE), (H.OPPOSED, D.RCD), (H.CLIP, D.MHC),
           (H.OPPOSED, D.MHC)]


def test_poisson_served_requests_equal_plain_and_oracle(oracle):
    rng = np.random.default_rng(0x52480444)
    proc = sensor.GpuSensorProcessor(0, budget_bytes=200 << 30)
    server = sensor.GpuServer(proc, lanes=4, admission_bytes=48 << 30,
                              base_cache_bytes=16 << 30)
    cases = {}
    for fid, (w, h, seed, cfa) in enumerate(SERVED_FRAMES):
        frame = synth.frame(synth.SceneSpec(w, h, seed, cfa=cfa, blob_fraction=0.05))
        for k, (method, alg) in enumerate(RECIPES):
            cases[fid, k] = describe(frame, method, alg, 10)
        server.put_frame(fid, frame)
    print("served: frames resident", flush=True)
    import time
    schedule = []
    t = 0.0
    for _ in range(16):
        t += rng.exponential(1.0 / 40.0)
        schedule.append((t, int(rng.integers(len(SERVED_FRAMES))), int(rng.integers(len(RECIPES)))))
    start = time.perf_counter()
    submitted = []
    for at, fid, k in schedule:
        delay = at - (time.perf_counter() - start)
        if delay > 0:
            time.sleep(delay)
        submitted.append((fid, k, server.submit(fid, cases[fid, k][2], cases[fid, k][3])))
    results = []
    for fid, k, rid in submitted:
        r = server.wait(rid)
        rgb = r.source.read_rgb() if r.source is not None else None
        if r.source is not None:
            r.source.release()
        assert r.status in (_lib.OK, _lib.CANCELLED), r.status
        results.append((fid, k, r.status, rgb, r.base_cache_hit))
    stats = server.stats()
    server.close()
    completed = [x for x in results if x[2] == _lib.OK]
    print(f"served: {len(completed)} completed, {stats.cancelled} cancelled, "
          f"{stats.base_hits} base hits, p99 {stats.p99_ms:.1f} ms", flush=True)
    # The newest request of each frame is never superseded, so it completes.
    assert len(completed) >= len({fid for fid, _, _ in submitted})
    plain = {}
    for fid, k, _, rgb, _ in completed:
        if (fid, k) not in plain:
            frame, recipe, desc, detail = cases[fid, k]
            res = proc.upload(frame)
            src = proc.develop_resident(res, desc, detail)
            plain[fid, k] = src.read_rgb()
            src.release()
            res.release()
        assert same(rgb, plain[fid, k]), (fid, k, mismatch(rgb, plain[fid, k]))
    proc.close()
    checked = set()
    for (fid, k), rgb in plain.items():
        alg = cases[fid, k][2].demosaic
        if alg in checked:
            continue
        frame, recipe, desc, _ = cases[fid, k]
        expected = oracle_chain(oracle, frame, recipe, desc)
        assert same(rgb, expected), (fid, k, mismatch(rgb, expected))
        checked.add(alg)
        print(f"served: frame {fid} recipe {k} ({ALG[alg]}) equals the oracle", flush=True)
    assert checked, "no completed request to check against the oracle"


def test_replaced_frame_is_never_served_from_an_old_base():
    proc = sensor.GpuSensorProcessor(0)
    server = sensor.GpuServer(proc, lanes=2, base_cache_bytes=4 << 30)
    first = make_case(2000, 1500, 0x52480450, "RGGB", H.OPPOSED, D.AMAZE, 20)
    second = make_case(2000, 1500, 0x52480451, "RGGB", H.OPPOSED, D.AMAZE, 20)
    plain = []
    for frame, _, desc, detail in (first, second):
        res = proc.upload(frame)
        src = proc.develop_resident(res, desc, detail)
        plain.append(src.read_rgb())
        src.release()
        res.release()
    desc, detail = first[2], first[3]
    server.put_frame(0, first[0])
    r = server.wait(server.submit(0, desc, detail), keep=False)
    assert r.ok and not r.base_cache_hit
    for _ in range(3):
        # A base hit in flight while new samples are installed under the id.
        server.drop_frame(0)
        in_flight = server.submit(0, desc, detail)
        server.put_frame(0, second[0])
        r = server.wait(in_flight)
        got = r.source.read_rgb()
        r.source.release()
        assert same(got, plain[0]) or same(got, plain[1])
        r = server.wait(server.submit(0, desc, detail))
        got = r.source.read_rgb()
        r.source.release()
        assert r.ok and same(got, plain[1]), mismatch(got, plain[1])
        server.put_frame(0, first[0])
        r = server.wait(server.submit(0, desc, detail))
        got = r.source.read_rgb()
        r.source.release()
        assert r.ok and same(got, plain[0]), mismatch(got, plain[0])
    server.close()
    proc.close()


# ---------------------------------------------------------------- large sensor

def test_large_sensor_stage_and_spatial_equal_the_oracle(oracle):
    from rohditor_amd import optics as opt
    w, h = 11648, 8736
    frame, recipe, desc, detail = make_case(w, h, 0x52480005, "RGGB", H.OPPOSED, D.AMAZE, 10,
                                            black=1024.0, white=65535.0)
    plan = opt.resolve_execution(w, h, camera_crop=1.0, focal=24.0,
                                 distortion=opt.Poly5(0.031, -0.022),
                                 vignetting=opt.Pa(-0.3, 0.05, 0.0),
                                 tca=opt.TcaLinear(1.0003, 0.9996))
    target = opt.preview_dimensions(w, h, 2560)
    proc = sensor.GpuSensorProcessor(0, budget_bytes=200 << 30)
    res = proc.upload(frame)
    src = proc.develop_resident(res, desc, detail)
    base = src.read_rgb()
    small, _ = proc.reduce_preview(src, plan, target)
    preview = small.read_rgb()
    for hnd in (small, src, res):
        hnd.release()
    proc.close()
    print("large sensor: developed on the GPU", flush=True)
    expected = oracle_chain(oracle, frame, recipe, desc)
    assert same(base, expected), mismatch(base, expected)
    del base
    st, corrected = oracle.optics_apply(expected, plan.to_c())
    assert st == 0
    del expected
    st, reduced = oracle.resize_area(corrected, *target)
    assert st == 0
    assert preview.shape == (1920, 2560, 3)
    assert same(preview, reduced), mismatch(preview, reduced)
