"""The reference's own engine test suite against the HIP engine.

oracle/_ref/ref_engine_tests is built in the build container from the
reference's headers (include/dpd/tests/*.hpp, test_runner.hpp,
test_differential.hpp, NaiveDPDEngine<false>) and this repository's
include/pdpd/hip_dpd_engine.hpp compiled against the reference's own DPDEngine
and WorldState. It prints TAP. Two suites:
  scenarios     every registered scenario family checks its closed-form
                invariants after each Run chunk (test_runner.hpp)
  differential  each scenario is run by NaiveDPDEngine<false> and by the HIP
                engine; x and v must agree within 0.01 sqrt(steps) and f within
                0.1 sqrt(steps) (test_differential.hpp:70-127)
"""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "ref_engine_tests")


def _run(mode):
    if not os.path.exists(BIN):
        pytest.skip("oracle/_ref/ref_engine_tests not built (needs the reference tree at build time)")
    print(f"reference harness: {mode}", flush=True)
    p = subprocess.run([BIN, mode], capture_output=True, text=True, timeout=600)
    lines = [l for l in p.stdout.splitlines() if l.startswith("ok ") or l.startswith("not ok ")]
    failed = [l for l in lines if l.startswith("not ok")]
    print(f"  {len(lines)} tests, {len(failed)} failed", flush=True)
    for l in failed[:20]:
        print("  " + l, flush=True)
    return p.returncode, lines, failed, p.stderr


def test_reference_scenarios_pass_on_hip_engine():
    rc, lines, failed, err = _run("scenarios")
    assert len(lines) >= 200
    assert not failed, "\n".join(failed[:10]) + "\n" + err[-2000:]
    assert rc == 0


def test_reference_differential_vs_naive_engine():
    rc, lines, failed, err = _run("differential")
    assert len(lines) >= 200
    assert not failed, "\n".join(failed[:10]) + "\n" + err[-2000:]
    assert rc == 0
