"""Each planted oracle mistake fails at least one parity pin.

The oracle is built once per planted mistake (oracle/common.hpp,
RHDO_MUTATION): a 3-sigma trim window, swapped blend weights, a wider
chromaticity gate, a one-element median, an arithmetic opposing reference,
ring evidence from too low a level and a lower estimate bound for Opposed; a
squared instead of fourth-power mask expansion, moved highlight-mask edges
and wider estimate clamps for capture sharpening. Each runs every Opposed
and capture pin of test_opposed_reference.py and test_capture_reference.py,
and the ground-truth matrix of test_highlight_ground_truth.py, on the oracle,
and at least one pin must fail, which shows the pins fix
these parts of the algorithm rather than only restate them.
"""
import pytest

import test_capture_reference as capture_pins
import test_highlight_ground_truth as ground_truth
import test_opposed_reference as opposed_pins
from oracle import oracle as oracle_module


def _failures(o):
    def run_opposed(m, levels, ceiling=None, pattern=0):
        st, out, stats, _, _ = o.opposed(m, pattern, levels, ceiling)
        if st != 0:
            raise opposed_pins.Rejected(st)
        return out, [int(v) for v in stats]

    def run_capture(rgb, amount, sigma, noise, ceilings, iterations=capture_pins.ITERATIONS):
        st, out = o.capture(rgb, amount, sigma, noise, ceilings, iterations=iterations)
        assert st == 0
        return out

    class Request:
        def getfixturevalue(self, name):
            assert name == "oracle"
            return o

    def run_matrix(method, mosaic, sat, ceilings):
        return ground_truth.run_method("oracle", Request(), method, mosaic, sat, ceilings)

    failed = []
    checks = [(p, lambda p=p: p(run_opposed)) for p in opposed_pins.PINS]
    checks += [(ground_truth.check_scene,
                lambda sc=sc: ground_truth.check_scene(sc, run_matrix))
               for sc in ground_truth.SCENES]
    checks += [(p, lambda p=p: p(run_capture, o.blur)) for p in capture_pins.BEHAVIOUR]
    checks += [(p, lambda p=p: p(o)) for p in capture_pins.PLANES]
    for pin, check in checks:
        try:
            check()
        except (AssertionError, pytest.fail.Exception):
            failed.append(pin.__name__)
    return sorted(set(failed))


def test_the_checker_passes_every_pin(oracle):
    assert _failures(oracle) == []


@pytest.mark.parametrize("k", sorted(oracle_module.MUTATIONS),
                         ids=lambda k: f"m{k}")
def test_planted_mistake_is_caught(k, oracle):
    with oracle.mutant(k):
        failed = _failures(oracle)
    assert failed, f"no pin catches: {oracle.MUTATIONS[k]}"
