"""Pins the oracle's XSPH, vorticity, confinement, sphere responses and
plane-crossing response with FP64 closed forms (tests/closed_form_cases.py),
and shows that each check catches a plausible mistake: the oracle is run
again with that mistake planted (oracle/pbf_oracle.h, ORC_PLANT_*) and the
same check must then fail.

Reference semantics: pbf/solver/cuda/pbf_xsph.cu:15-29,84-113;
pbf/solver/cuda/pbf_vorticity_confinement.cu:15-31,87-138,192-222;
pbf/solver/cuda/boundary/pbf_sphere_boundary.cu:8-64;
pbf/solver/cuda/boundary/pbf_plane_boundary.cu:6-45;
pbf/solver/cuda/pbf_response_collision.cu:12-66.
"""
from __future__ import annotations

import dataclasses

import pytest

from oracle import pyoracle as orc

import closed_form_cases as cf

PASS_CASES = {c.name: c for c in cf.pass_cases()}
BOUNDARY_CASES = {c.name: c for c in cf.sphere_cases() + cf.crossing_cases()}

# the mistakes each pass check must catch
PASS_PLANTS = {
    "xsph": [orc.PLANT_XSPH_SIGN, orc.PLANT_XSPH_NO_MR, orc.PLANT_XSPH_H],
    "vorticity": [orc.PLANT_VORT_CROSS, orc.PLANT_VORT_NO_MR, orc.PLANT_CONF_CROSS, orc.PLANT_CONF_RHO,
                  orc.PLANT_LOC_SELF],
}


def run_pass(case):
    return orc.step(case.scene_off), orc.step(case.scene_on)


@pytest.mark.parametrize("name", sorted(PASS_CASES))
def test_pass_closed_form(name):
    case = PASS_CASES[name]
    off, on = run_pass(case)
    err = case.check(off, on)
    print(f"{name}: normwise error {err:.2e} (bar {case.tol:.0e})")
    assert err <= case.tol


@pytest.mark.parametrize("name,plant", [(n, p) for n in sorted(PASS_CASES) if not n.endswith("inv_h")
                                        for p in PASS_PLANTS[n.split("_")[0]]])
def test_pass_check_catches_planted_mistake(name, plant):
    case = PASS_CASES[name]
    with orc.planted(plant):
        off, on = run_pass(case)
    err = case.check(off, on)
    print(f"{name} with planted mistake {plant}: normwise error {err:.2e}")
    assert err > 100 * case.tol


def test_vorticity_check_distinguishes_h_from_inv_h():
    # The reference's literal kernel argument (h for 1/h) is the oracle's
    # vort_h_as_inv_h mode; the corrected form's closed form must reject it,
    # and the literal form's closed form must reject the corrected kernel.
    lit = PASS_CASES["vorticity_h_as_inv_h"]
    cor = PASS_CASES["vorticity_4"]
    for case, other in ((cor, lit), (lit, cor)):
        off = orc.step(dataclasses.replace(case.scene_off, params=other.scene_off.params))
        on = orc.step(dataclasses.replace(case.scene_on, params=dataclasses.replace(
            other.scene_on.params, vc_parameter=case.scene_on.params.vc_parameter)))
        assert case.check(off, on) > 100 * case.tol


@pytest.mark.parametrize("name", sorted(BOUNDARY_CASES))
def test_boundary_closed_form(name):
    case = BOUNDARY_CASES[name]
    r = orc.step(case.scene)
    err = case.check(r)
    print(f"{name}: max abs error {err:.2e} (bar {case.tol:.0e}); sweeps used {r['stats']['max_sweeps_used']}")
    assert err <= case.tol
    assert r["stats"]["sweep_cap_hits"] == 0


@pytest.mark.parametrize("name,plant", [("inner_sphere", orc.PLANT_NUDGE), ("outer_sphere", orc.PLANT_NUDGE),
                                        ("plane_crossing", orc.PLANT_CROSS_FRAC),
                                        ("box_crossing", orc.PLANT_CROSS_FRAC)])
def test_boundary_check_catches_planted_mistake(name, plant):
    case = BOUNDARY_CASES[name]
    with orc.planted(plant):
        r = orc.step(case.scene)
    err = case.check(r)
    print(f"{name} with planted mistake {plant}: max abs error {err:.2e}")
    assert err > 10 * case.tol


def test_outer_sphere_answered_as_inner_is_caught():
    # pbf_response_collision.cu:46-51 calls the inner response for outer
    # spheres; the oracle's outer_as_inner mode reproduces that, and the
    # outer sphere's closed form must reject it (the particle stays inside).
    case = BOUNDARY_CASES["outer_sphere"]
    sc = dataclasses.replace(case.scene, params=dataclasses.replace(case.scene.params, outer_as_inner=1))
    assert case.check(orc.step(sc)) > 10 * case.tol


def test_plant_switch_resets():
    # the planted-mistake switch is off again after its block
    case = BOUNDARY_CASES["inner_sphere"]
    with orc.planted(orc.PLANT_NUDGE):
        pass
    assert case.check(orc.step(case.scene)) <= case.tol
