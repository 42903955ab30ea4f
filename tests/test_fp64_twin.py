"""The stage against an FP64 oracle of the same algorithms.

Each stage of the 24 MP workload (normalize, Opposed, AMaZE, RL capture
with 10 iterations) is run on the GPU through the staged C ABI and compared
with the oracle's FP64 twin (the same steps in double precision) applied to
the GPU's own input to that stage. Stated tolerance, per stage:

- normalize and capture sharpening: every sample within 1e-5 relative
  (max(|b|, 1e-3) in the denominator) of the FP64 twin;
- Opposed: every sample within 1e-5 relative except where one of its
  discrete decisions (the 0.75 EV gate, the trim window, the ordering of a
  median, "applied only if it raises the sample") falls the other way in
  double precision, at most 1e-6 of the samples (one site of the 24 MP
  frame, where the f32 and f64 estimates straddle such a threshold);
- AMaZE: its direction, Nyquist and highlight decisions compare f32
  quantities against thresholds, and a decision that falls the other way in
  double precision moves that pixel by a share of the local contrast. The
  tolerance is the shape of the reference's own budget for AMaZE against a
  differently rounded implementation (crates/gpu/src/sensor/amaze_tests.rs:
  28-57, demosaic_tests.rs:925-940: an outlier fraction and a maximum), with
  limits for a change of every rounding rather than of a few: RMS below
  1e-5, at most 5e-3 of samples beyond 1e-5 relative, none beyond 0.05
  absolute (measured at 24 MP: RMS 8.2e-6, 0.17 %, 0.011).

The end-to-end chain is also checked against the FP64 chain. A flipped
decision there spreads through the later stages (the one flipped Opposed
site of the 24 MP frame moves its AMaZE and RL neighbourhood by up to 0.5),
so the chain is held to outlier fractions: at most 5e-3 of samples beyond
1e-5 relative and 1e-4 beyond 1e-3 relative (measured at 24 MP: 0.17 % and
4.5e-5; 99.8 % of samples agree within 7.7e-6). On the CPU the same relations are pinned at 1536 x 1024 with
the f32 oracle standing in for the GPU, which the GPU tests prove
bit-identical to it.
"""
import numpy as np
import pytest

from rohditor_amd import sensor, synth


def rel_stats(a, b):
    d = np.abs(a.astype(np.float64) - b)
    rel = d / np.maximum(np.abs(b), 1e-3)
    return dict(max_abs=float(d.max()), rms=float(np.sqrt(np.mean(d * d))),
                max_rel=float(rel.max()), frac_over=float(np.mean(rel > 1e-5)))


def assert_exact_stage(name, a, b, flips=0.0):
    s = rel_stats(a, b)
    print(f"{name}: {s}", flush=True)
    if flips == 0.0:
        assert s["max_rel"] <= 1e-5, (name, s)
    else:
        assert s["frac_over"] <= flips, (name, s)


def assert_amaze_budget(name, a, b):
    s = rel_stats(a, b)
    print(f"{name}: {s}", flush=True)
    assert s["rms"] < 1e-5 and s["frac_over"] <= 5e-3 and s["max_abs"] <= 0.05, (name, s)


def assert_chain_budget(name, a, b):
    s = rel_stats(a, b)
    rel = np.abs(a.astype(np.float64) - b) / np.maximum(np.abs(b), 1e-3)
    s["frac_over_1e-3"] = float(np.mean(rel > 1e-3))
    print(f"{name}: {s}", flush=True)
    assert s["frac_over"] <= 5e-3 and s["frac_over_1e-3"] <= 1e-4, (name, s)


def workload(spec):
    frame = synth.frame(spec)
    recipe = sensor.EditRecipe(
        highlights=sensor.HighlightAdjustments(method=sensor.HighlightMethod.OPPOSED),
        capture=sensor.CaptureSharpening(enabled=True, amount=0.5, radius=0.7,
                                         noise_protection=0.5),
        capture_iterations=10)
    options = sensor.RenderOptions(demosaic=sensor.DemosaicAlgorithm.AMAZE)
    desc = sensor.SensorDevelopmentDescription.from_frame(frame, recipe, options)
    return frame, recipe, desc


def twin_stages(o, frame, recipe, desc, stage_inputs, threads):
    """FP64 twin of each stage, applied to the given f32 stage inputs
    (normalized, highlighted, demosaiced), plus the whole FP64 chain."""
    c = desc.crop
    crop = (c.x, c.y, c.width, c.height)
    ceil = sensor.capture_ceilings(recipe, desc.demosaic_balance)
    norm, high, demo = stage_inputs
    n64 = o.normalize_f64(frame.mosaic, desc.sensor_pattern, desc.black, *desc.black_repeat,
                          desc.white, crop)
    _, h64, _ = o.opposed_f64(norm.astype(np.float64), desc.crop_pattern, desc.detection_levels,
                              desc.neutral_ceiling)
    _, d64 = o.demosaic(high.astype(np.float64), desc.crop_pattern, desc.demosaic_balance,
                        "amaze", threads=threads, fp64=True)
    _, c64 = o.capture(demo.astype(np.float64), 0.5, 0.7, 0.5, ceil, iterations=10,
                       threads=threads, fp64=True)
    _, chain, _ = o.opposed_f64(n64, desc.crop_pattern, desc.detection_levels,
                                desc.neutral_ceiling)
    _, chain = o.demosaic(chain, desc.crop_pattern, desc.demosaic_balance, "amaze",
                          threads=threads, fp64=True)
    _, chain = o.capture(chain, 0.5, 0.7, 0.5, ceil, iterations=10, threads=threads, fp64=True)
    return n64, h64, d64, c64, chain


def test_f32_oracle_against_the_fp64_twin(oracle):
    frame, recipe, desc = workload(synth.SceneSpec(1536, 1024, 0x52480002, blob_fraction=0.05))
    c = desc.crop
    norm = oracle.normalize(frame.mosaic, desc.sensor_pattern, desc.black, *desc.black_repeat,
                            desc.white, (c.x, c.y, c.width, c.height))
    _, high, _, _, _ = oracle.opposed(norm, desc.crop_pattern, desc.detection_levels,
                                      desc.neutral_ceiling)
    _, demo = oracle.demosaic(high, desc.crop_pattern, desc.demosaic_balance, "amaze", threads=8)
    ceil = sensor.capture_ceilings(recipe, desc.demosaic_balance)
    _, rgb = oracle.capture(demo, 0.5, 0.7, 0.5, ceil, iterations=10, threads=8)
    n64, h64, d64, c64, chain = twin_stages(oracle, frame, recipe, desc, (norm, high, demo), 8)
    assert_exact_stage("normalize", norm, n64)
    assert_exact_stage("opposed", high, h64, flips=1e-6)
    assert_amaze_budget("amaze", demo, d64)
    assert_exact_stage("capture", rgb, c64)
    assert_chain_budget("chain", rgb, chain)


@pytest.mark.gpu
def test_gpu_c2_stages_against_the_fp64_twin(gpu, oracle):
    frame, recipe, desc = workload(synth.full_sensor_24mp())
    detail = sensor.detail_desc(recipe, desc.demosaic_balance)
    resident = gpu.upload(frame)
    plane = gpu.normalize(resident, desc)
    norm = plane.read()
    gpu.apply_highlight(plane, desc)
    high = plane.read()
    src = gpu.demosaic(plane, desc)
    demo = src.read_rgb()
    gpu.capture(src, detail)
    rgb = src.read_rgb()
    src.release()
    resident.release()
    n64, h64, d64, c64, chain = twin_stages(oracle, frame, recipe, desc, (norm, high, demo),
                                            oracle.cpu_threads())
    assert_exact_stage("normalize", norm, n64)
    assert_exact_stage("opposed", high, h64, flips=1e-6)
    assert_amaze_budget("amaze", demo, d64)
    assert_exact_stage("capture", rgb, c64)
    assert_chain_budget("chain", rgb, chain)
