"""Per-kernel timing (include/rhd.h rhd_profile_*): every group when enabled,
only the named group under rhd_profile_only, and nothing when disabled."""
import ctypes

import pytest

from rohditor_amd import _lib, sensor, synth


def _entries(proc):
    out = {}
    for k in range(proc.lib.rhd_profile_count(proc.ctx)):
        name = ctypes.c_char_p()
        ms = ctypes.c_double()
        launches = ctypes.c_uint64()
        nbytes = ctypes.c_uint64()
        assert proc.lib.rhd_profile_entry(proc.ctx, k, ctypes.byref(name), ctypes.byref(ms),
                                          ctypes.byref(launches), ctypes.byref(nbytes)) == _lib.OK
        out[name.value.decode()] = (ms.value, launches.value, nbytes.value)
    return out


def test_profile_calls_reject_a_missing_context():
    lib = _lib.load()
    assert lib.rhd_profile_only(None, b"cap_iterate") == _lib.INVALID_INPUT
    assert lib.rhd_profile_enable(None, 1) == _lib.INVALID_INPUT


@pytest.mark.gpu
def test_profile_only_times_one_group(gpu):
    spec = synth.SceneSpec(width=512, height=384, seed=0x5052, cfa="RGGB")
    frame = synth.frame(spec)
    recipe = sensor.EditRecipe(
        highlights=sensor.HighlightAdjustments(method=sensor.HighlightMethod.OPPOSED),
        capture=sensor.CaptureSharpening(enabled=True), capture_iterations=4)
    desc = sensor.SensorDevelopmentDescription.from_frame(frame, recipe, sensor.RenderOptions())
    detail = sensor.detail_desc(recipe, desc.demosaic_balance)
    resident = gpu.upload(frame)
    lib, ctx = gpu.lib, gpu.ctx
    try:
        lib.rhd_profile_reset(ctx)
        lib.rhd_profile_enable(ctx, 1)
        gpu.develop_resident(resident, desc, detail).release()
        every = _entries(gpu)
        assert {"normalize", "opp_sums", "amz_weights", "cap_prep", "cap_iterate_first",
                "cap_iterate", "cap_iterate_apply"} <= set(every)
        # The first of the four iterations starts from cap_prep's ratio and
        # the last applies the gain.
        assert every["cap_prep"][1] == 1 and every["cap_iterate_first"][1] == 1
        assert every["cap_iterate"][1] == 2 and every["cap_iterate"][2] > 0
        assert every["cap_iterate_apply"][1] == 1 and "cap_apply" not in every

        lib.rhd_profile_reset(ctx)
        lib.rhd_profile_only(ctx, b"cap_iterate")
        gpu.develop_resident(resident, desc, detail).release()
        only = _entries(gpu)
        assert set(only) == {"cap_iterate"} and only["cap_iterate"][1] == 2

        lib.rhd_profile_reset(ctx)
        lib.rhd_profile_only(ctx, None)
        lib.rhd_profile_enable(ctx, 0)
        gpu.develop_resident(resident, desc, detail).release()
        assert _entries(gpu) == {}
    finally:
        lib.rhd_profile_only(ctx, None)
        lib.rhd_profile_enable(ctx, 0)
        lib.rhd_profile_reset(ctx)
        resident.release()
