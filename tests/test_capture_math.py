"""Capture sharpening's guide and mask computed inside the AMaZE finish.

When capture sharpening is the only detail step after AMaZE, the demosaic's
last stage writes the guide and the highlight mask
(crates/core/src/sharpening.rs:194-218) as it stores each pixel, with the
seven divisions taken as reciprocal products plus one FMA correction
(rohditor_amd/csrc/capture_math.hpp). This file proves on the host that the
fast form returns the reference arithmetic's bits: exhaustively for the
division by 3 and for the smoothstep over every peak that reaches its band,
exhaustively near the band for ceilings across [2^-20, 2^20], and on random
pixels elsewhere. The GPU parity tests then compare the fused develop with
the oracle chain bit for bit.
"""
import os
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "rohditor_amd" / "csrc"


@pytest.fixture(scope="module")
def host_check(tmp_path_factory):
    exe = tmp_path_factory.mktemp("capture") / "capture_check"
    subprocess.run([os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-ffp-contract=off",
                    "-fno-fast-math", "-pthread", f"-I{CSRC}",
                    str(ROOT / "tests" / "capture_check.cc"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {k: int(v) for k, v in (line.split() for line in out.splitlines())}


def test_division_by_three_is_correctly_rounded_for_every_f32(host_check):
    assert host_check["third_mismatch"] == 0
    # The bare reciprocal product is not: the correction step does real work.
    assert host_check["third_uncorrected"] > 1_000_000


def test_highlight_mask_smoothstep_matches_on_every_peak(host_check):
    assert host_check["mask_mismatch"] == 0


def test_ceiling_ratios_match_near_the_mask_band(host_check):
    assert host_check["ratio_mismatch"] == 0
    assert host_check["ratio_mask_sample"] == 0


def test_random_pixels_give_the_same_guide_and_mask(host_check):
    assert host_check["random_mismatch"] == 0
    assert host_check["checked"] > 2_000_000_000
