"""The C-ABI library loads and exports every symbol include/rhd.h declares."""
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "rhd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rhd_[a-z_]+)\s*\(", text)))


def test_header_declares_the_stage_api():
    names = declared_functions()
    for required in ["rhd_create", "rhd_supports", "rhd_normalize", "rhd_highlight", "rhd_demosaic",
                     "rhd_capture", "rhd_develop", "rhd_develop_resident", "rhd_source_read_rgb",
                     "rhd_status_is_routing"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    from rohditor_amd import _lib
    lib = _lib.load()
    names = declared_functions()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_status_helpers_without_a_gpu():
    from rohditor_amd import _lib
    lib = _lib.load()
    assert lib.rhd_status_is_routing(_lib.OVER_BUDGET) == 1
    assert lib.rhd_status_is_routing(_lib.UNSUPPORTED_EDITS) == 1
    for s in [_lib.OK, _lib.CANCELLED, _lib.SYNCHRONIZATION, _lib.INVALID_INPUT]:
        assert lib.rhd_status_is_routing(s) == 0
    assert lib.rhd_status_name(_lib.OVER_BUDGET) == b"OverBudget"


def test_supports_mirrors_the_reference_plus_opposed():
    import numpy as np
    from rohditor_amd import sensor
    info = sensor.RawFileInfo(width=64, height=48, black_levels=sensor.LevelPattern([512.0]),
                              white_levels=(16383.0,), as_shot_white_balance=(2.0, 1.0, 1.4))
    frame = sensor.RawFrame(info=info, mosaic=np.zeros((48, 64), np.uint16))
    for method in sensor.HighlightMethod:
        recipe = sensor.EditRecipe(highlights=sensor.HighlightAdjustments(method=method))
        for alg in sensor.DemosaicAlgorithm:
            desc = sensor.SensorDevelopmentDescription.from_frame(
                frame, recipe, sensor.RenderOptions(demosaic=alg))
            assert sensor.GpuSensorProcessor.supports(desc)
    recipe = sensor.EditRecipe()
    desc = sensor.SensorDevelopmentDescription.from_frame(frame, recipe, sensor.RenderOptions())
    dd = sensor.detail_desc(recipe, desc.demosaic_balance)
    dd.nr_chroma, dd.nr_luminance = 0.5, 0.5
    assert sensor.GpuSensorProcessor.supports(desc, dd)
    # capture radii above 4 taps are routed (ceil(3 sigma) > 4)
    dd.capture_enabled, dd.amount, dd.sigma = 1, 0.5, 1.5
    assert not sensor.GpuSensorProcessor.supports(desc, dd)


def test_highlight_resolution_matches_reference_fixture():
    # crates/core/src/pipeline/sensor.rs:888-940: the neutral ceiling uses the
    # as-shot balance with the same mul/mul/div order.
    import numpy as np
    from rohditor_amd import sensor
    info = sensor.RawFileInfo(width=6, height=5, black_levels=sensor.LevelPattern([0, 10, 20, 30], 2, 2),
                              white_levels=(100, 110, 120, 130),
                              as_shot_white_balance=(2.0, 1.0, 1.55))
    frame = sensor.RawFrame(info=info, mosaic=np.zeros((5, 8), np.uint16), row_stride=8)
    recipe = sensor.EditRecipe(highlights=sensor.HighlightAdjustments(
        method=sensor.HighlightMethod.OPPOSED, opposed_detection_threshold=1.2),
        white_balance_gains=(2.6, 1.0, 1.395))
    desc = sensor.SensorDevelopmentDescription.from_frame(frame, recipe, sensor.RenderOptions())
    F = np.float32
    common = (F(1.2) * F(0.987)) * min(F(1.0), F(2.0), F(1.55))
    assert desc.neutral_ceiling.tolist() == [common / F(2.0), common / F(1.0), common / F(1.55)]


def test_every_symbol_host_code_resolves_is_in_the_device_code():
    """Each array that host code resolves with hipGetSymbolAddress must be in
    the gfx950 code object of its source file: the device compiler drops an
    array that no device code reads, and asking the runtime for a symbol the
    code object lacks takes the process down."""
    import shutil
    import subprocess
    import tempfile
    bundler = Path("/opt/rocm/llvm/bin/clang-offload-bundler")
    readelf = Path("/opt/rocm/llvm/bin/llvm-readelf")
    if not (bundler.exists() and readelf.exists()):
        pytest.skip("ROCm LLVM tools not installed")
    csrc = ROOT / "rohditor_amd" / "csrc"
    build = ROOT / "rohditor_amd" / "_build"
    checked = 0
    for src in sorted(csrc.glob("*.hip")):
        names = re.findall(r"HIP_SYMBOL\((\w+)\)", src.read_text())
        if not names:
            continue
        obj = build / (src.stem + ".o")
        assert obj.exists(), f"{obj} is not built"
        with tempfile.TemporaryDirectory() as tmp:
            fatbin = Path(tmp) / "fatbin"
            dev = Path(tmp) / "dev.o"
            objcopy = shutil.which("objcopy")
            assert objcopy, "objcopy is required"
            subprocess.run([objcopy, "-O", "binary", "--only-section=.hip_fatbin", str(obj),
                            str(fatbin)], check=True)
            subprocess.run([str(bundler), "--type=o", "--targets=hipv4-amdgcn-amd-amdhsa--gfx950",
                            f"--input={fatbin}", f"--output={dev}", "--unbundle"], check=True)
            syms = subprocess.run([str(readelf), "-s", "-W", str(dev)], check=True,
                                  capture_output=True, text=True).stdout
        for name in names:
            assert re.search(rf"\b_Z\w*{len(name)}{name}E\b", syms), \
                f"{name} ({src.name}) is missing from the gfx950 code object"
            checked += 1
    assert checked >= 2
