"""The closed-form cases of tests/closed_form_cases.py on the HIP path (through
the C-ABI): XSPH, vorticity + confinement (corrected and literal kernel
argument), inner and outer sphere responses and the plane-crossing response
(the general sweep and the six-plane box's own code path).  Each must meet
the closed form's bar and equal the oracle bit for bit."""
from __future__ import annotations

import pytest

from helpers import bit_mismatches, gpu_run, oracle_run

import closed_form_cases as cf

pytestmark = pytest.mark.gpu

PASS_CASES = {c.name: c for c in cf.pass_cases()}
BOUNDARY_CASES = {c.name: c for c in cf.sphere_cases() + cf.crossing_cases()}


@pytest.mark.parametrize("name", sorted(PASS_CASES))
def test_gpu_pass_closed_form(name):
    case = PASS_CASES[name]
    print(f"case {name}", flush=True)
    off, on = gpu_run(case.scene_off), gpu_run(case.scene_on)
    err = case.check(off, on)
    print(f"{name}: normwise error {err:.2e} (bar {case.tol:.0e})", flush=True)
    assert err <= case.tol
    ref = oracle_run(case.scene_on)
    assert bit_mismatches(on["x"], ref["x"]) == 0 and bit_mismatches(on["v"], ref["v"]) == 0


@pytest.mark.parametrize("name", sorted(BOUNDARY_CASES))
def test_gpu_boundary_closed_form(name):
    case = BOUNDARY_CASES[name]
    print(f"case {name}", flush=True)
    got = gpu_run(case.scene)
    err = case.check(got)
    print(f"{name}: max abs error {err:.2e} (bar {case.tol:.0e})", flush=True)
    assert err <= case.tol
    assert got["stats"]["sweep_cap_hits"] == 0
    ref = oracle_run(case.scene)
    assert bit_mismatches(got["x"], ref["x"]) == 0 and bit_mismatches(got["v"], ref["v"]) == 0
