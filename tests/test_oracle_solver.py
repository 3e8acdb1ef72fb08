"""Pins the oracle's solver stages (lambda, delta-p, collision, update, XSPH,
vorticity) with results the mathematics fixes independently of its code:
closed forms for one and two particles (from the textbook poly6 / spiky
kernels in double precision), conservation laws of the pair sums,
invariance under uniform translation, projection onto a wall, and agreement
with the FP64 twin.  Reference semantics: pbf/solver/cuda/pbf_constraint.cu:168-209
(lambda), pbf/solver/cuda/pbf_position_update.cu:131-188 (delta-p),
pbf/solver/cuda/boundary/pbf_plane_boundary.cu:6-45 (planes),
pbf/solver/cuda/pbf_update.cu:25-39 (velocity), pbf/solver/cuda/pbf_xsph.cu:84-113,
pbf/solver/cuda/pbf_vorticity_confinement.cu:87-222."""
from __future__ import annotations

import dataclasses
import math

import numpy as np
import pytest

from oracle import pyoracle as orc
from ssf_pbf_cuda_amd import scenes


def poly6(r, h):
    return 315.0 / (64.0 * math.pi * h ** 9) * (h * h - r * r) ** 3 if r <= h else 0.0


def spiky_deriv(r, h):
    return -45.0 / (math.pi * h ** 6) * (h - r) ** 2 if r <= h else 0.0


def custom_scene(x, v, h=0.1, box=(4.0, 4.0, 4.0), iters=1, **over):
    """Particles at x with velocities v in an open box [0, box] (planes on its faces)."""
    base = scenes.small_dam(4, 4, 4, h=h, box=box, iters=iters)
    p = dataclasses.replace(base.params, body_force=(0.0, 0.0, 0.0), **over)
    return dataclasses.replace(base, x=np.ascontiguousarray(x, np.float32), v=np.ascontiguousarray(v, np.float32),
                               params=p)


def lattice(n, h, origin, quarter=True):
    s = 0.5 * h
    i = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), -1).reshape(-1, 3)
    off = 0.5 * s if quarter else 0.0  # points at cell quarter positions: away from cell faces
    return (np.asarray(origin) + (i * s + off)).astype(np.float32)


def test_isolated_particle_closed_form():
    # Only the self pair: W(0), zero gradient -> s = C / eps, no displacement;
    # predict x* = fmaf(fmaf(g,dt,v),dt,x) and v = (x* - x) * (1/dt).
    h = 0.1
    x = np.array([[1.23, 2.01, 1.77]], np.float32)
    v = np.array([[0.3, -0.2, 0.1]], np.float32)
    sc = custom_scene(x, v, h=h, xsph_on=1, vort_on=1)
    sc = dataclasses.replace(sc, params=dataclasses.replace(sc.params, body_force=(0.0, -9.8, 0.0)))
    r = orc.step(sc, want_ns=True)
    p = sc.params
    dt = np.float32(p.time_step)
    g = np.array(p.body_force, np.float32)
    vs = (g.astype(np.float64) * dt + v.astype(np.float64)).astype(np.float32)
    xs = (vs.astype(np.float64) * dt + x.astype(np.float64)).astype(np.float32)
    assert np.array_equal(r["x"], xs)
    inv_dt = np.float32(1.0) / dt
    np.testing.assert_array_equal(r["v"], ((xs - x) * inv_dt).astype(np.float32))
    C = p.particle_mass / p.stable_density * poly6(0.0, h) - 1.0
    assert r["lambda_last"][0] == pytest.approx(C / p.relaxation, rel=1e-5)


@pytest.mark.parametrize("mode", [scenes.SCORR_PAPER, scenes.SCORR_REFERENCE])
def test_two_particles_lambda_and_deltap_closed_form(mode):
    h = 0.1
    r0 = 0.6 * h
    x = np.array([[2.2 * h, 2.5 * h, 2.5 * h], [2.2 * h + r0, 2.5 * h, 2.5 * h]], np.float32)
    sc = custom_scene(x, np.zeros_like(x), h=h, box=(1.0, 1.0, 1.0), scorr_mode=mode, xsph_on=0, vort_on=0)
    p = sc.params
    r = orc.step(sc, want_ns=True)
    d = x[0].astype(np.float64) - x[1].astype(np.float64)
    rr = float(np.linalg.norm(d))
    mr = p.particle_mass / p.stable_density
    C = mr * (poly6(0.0, h) + poly6(rr, h)) - 1.0
    wd = spiky_deriv(rr, h)
    G = 2.0 * wd * wd  # |sum grad|^2 + sum |grad|^2 with one non-self pair
    s = C / (mr * mr * G + p.relaxation)
    np.testing.assert_allclose(r["lambda_last"], [s, s], rtol=2e-5)
    if mode == scenes.SCORR_PAPER:
        corr = p.scorr_k * (poly6(rr, h) / poly6(p.scorr_dq_over_h * h, h)) ** p.scorr_n
    else:
        inv_k = 1.0 / poly6(0.4 * h, h) ** 4
        corr = 10.0 * (max(-s, 0.0) * 2) * poly6(rr, h) ** 4 * inv_k
    grad = wd * d / rr  # gradient of W_ij with respect to x_i
    dp0 = mr * (-(2.0 * s + corr)) * grad
    np.testing.assert_allclose(r["x"][0].astype(np.float64) - x[0], dp0, rtol=1e-4, atol=1e-9)
    np.testing.assert_allclose(r["x"][1].astype(np.float64) - x[1], -dp0, rtol=1e-4, atol=1e-9)


def _cloud(h=0.1, n=(9, 8, 7)):
    return lattice(n, h, origin=(1.0, 1.0, 1.0))


@pytest.mark.parametrize("mode", [scenes.SCORR_PAPER, scenes.SCORR_REFERENCE])
def test_deltap_conserves_momentum_on_symmetric_lists(mode):
    # With every in-range pair listed (no truncation) the pair terms are
    # antisymmetric, so sum_i delta-p_i = 0 up to rounding.
    h = 0.1
    x = _cloud(h)
    rng = np.random.default_rng(3)
    x = (x + rng.uniform(-0.05 * h, 0.05 * h, x.shape)).astype(np.float32)
    sc = custom_scene(x, np.zeros_like(x), h=h, max_pair=64, scorr_mode=mode, xsph_on=0, vort_on=0)
    r = orc.step(sc)
    assert r["stats"]["truncated"] == 0
    dp = r["x"].astype(np.float64).sum(0) - x.astype(np.float64).sum(0)
    tot = np.abs(np.sort(r["x"].astype(np.float64), 0) - np.sort(x.astype(np.float64), 0)).sum(0)
    assert tot.max() > 0
    assert np.all(np.abs(dp) <= 1e-5 * tot + 1e-9), (dp, tot)


def test_xsph_conserves_momentum():
    # XSPH adds c*m/rho0*sum_j (v_j - v_i) W_ij: zero in total when every pair
    # is seen from both sides (particles stay in their cells, no delta-p).
    h = 0.1
    x = _cloud(h)
    rng = np.random.default_rng(5)
    v = rng.uniform(-0.01, 0.01, x.shape).astype(np.float32)
    common = dict(h=h, scorr_mode=scenes.SCORR_REFERENCE, relaxation=1e20, vort_on=0, xsph_parameter=0.3)
    off = orc.step(custom_scene(x, v, xsph_on=0, **common))
    on = orc.step(custom_scene(x, v, xsph_on=1, **common))
    np.testing.assert_array_equal(on["x"], off["x"])
    change = on["v"].astype(np.float64) - off["v"].astype(np.float64)
    assert np.abs(change).sum() > 0
    assert np.all(np.abs(change.sum(0)) <= 1e-4 * np.abs(change).sum(0))


def test_vorticity_vanishes_for_uniform_translation():
    # Every particle moves with the same u: omega_i sums (v_j - v_i) x grad W,
    # which is zero, so confinement must leave the velocities alone (a form
    # that used v_j alone would push the surface particles).
    h = 0.1
    x = _cloud(h)
    u = np.array([0.01, 0.005, -0.003], np.float32)
    v = np.tile(u, (x.shape[0], 1))
    common = dict(h=h, scorr_mode=scenes.SCORR_REFERENCE, relaxation=1e20, xsph_on=0, vc_parameter=0.05)
    off = orc.step(custom_scene(x, v, vort_on=0, **common))
    on = orc.step(custom_scene(x, v, vort_on=1, **common))
    assert np.abs(on["v"] - off["v"]).max() <= 1e-3 * float(np.abs(u).max())


def test_plane_collision_projects_onto_the_wall():
    # A particle whose predicted position is below the floor (y = 0) is
    # projected onto it: x*.y = 0, the other components unchanged.
    h = 0.1
    x = np.array([[1.0, 0.01, 1.0]], np.float32)
    v = np.array([[0.2, -3.0, 0.1]], np.float32)
    sc = custom_scene(x, v, h=h, xsph_on=0, vort_on=0)
    r = orc.step(sc)
    dt = np.float32(sc.params.time_step)
    xs = (v.astype(np.float64) * dt + x.astype(np.float64)).astype(np.float32)
    assert xs[0, 1] < 0
    assert abs(float(r["x"][0, 1])) <= 1e-7
    assert r["x"][0, 0] == xs[0, 0] and r["x"][0, 2] == xs[0, 2]
    assert r["stats"]["max_sweeps_used"] >= 1


@pytest.mark.parametrize("iters", [1, 2])
def test_fp64_twin_agrees_with_fp32(iters):
    # The FP32 oracle's rounding error on a full step of a small dam, against
    # its FP64 twin (same sort and lists).  With the dam parameters (PAPER
    # s_corr k = 0.1) each Jacobi iteration amplifies the error by about an
    # order of magnitude: 2e-7 after one iteration, 7e-6 after two, and by the
    # fourth the run is ill-conditioned (particles reach walls differently).
    sc = scenes.small_dam(10, 8, 6, iters=iters)
    a = orc.step(sc)
    b = orc.step_f64(sc)
    scale = np.abs(b["x"]).max()
    ex = float(np.abs(a["x"].astype(np.float64) - b["x"]).max() / scale)
    ev = float(np.abs(a["v"].astype(np.float64) - b["v"]).max() / max(np.abs(b["v"]).max(), 1e-30))
    print(f"fp32 vs fp64, {iters} iteration(s): rel_x={ex:.2e} rel_v={ev:.2e}")
    assert ex < 1e-5
    assert ev < 1e-4


def test_density_over_lists_is_rest_density_inside_a_lattice():
    # rho_i = m * sum_j W over the (uncapped) lists at the positions the lists
    # were built on: interior particles of the exact lattice the rest density
    # was defined on (rho0 = m * lattice sum of W) get rho0 back.
    h = 0.1
    x = lattice((12, 12, 12), h, origin=(1.0, 1.0, 1.0))
    sc = custom_scene(x, np.zeros_like(x), h=h, max_pair=64)
    r = orc.step(sc, want_ns=True)
    p0 = r["pred0"].reshape(-1, 3).astype(np.float64)
    rho = orc.density(sc, p0.astype(np.float32), r["neighbors"])
    s = sc.params.stable_distance
    lo, hi = p0.min(0) + 2.5 * s, p0.max(0) - 2.5 * s
    inner = np.all((p0 > lo) & (p0 < hi), axis=1)
    assert inner.sum() > 100
    np.testing.assert_allclose(rho[inner], sc.params.stable_density, rtol=1e-5)
    assert rho[~inner].max() < sc.params.stable_density * (1 + 1e-5)  # surface particles miss neighbours
