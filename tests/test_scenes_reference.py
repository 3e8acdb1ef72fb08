"""The reference's arrangement generators and simulator set-ups
(pbf/util/pbf_arrangement.h; simulator/dam/pbf_dam_init_cond.cpp,
simulator/sphere/pbf_sphere_init_cond.cpp, pbf_sphere_sim.cpp) as restated in
ssf_pbf_cuda_amd/scenes.py, and the jet scenario on the oracle."""
from __future__ import annotations

import math

import numpy as np

from oracle import pyoracle as orc
from ssf_pbf_cuda_amd import scenes


def test_volume_counts_follow_the_epsilon_rule():
    # (5.65 - 2.25) / 0.16 = 21.25 -> 22 points per axis; an exact multiple
    # keeps its end point: (2 - 0) / 0.5 = 4 -> 5
    x = scenes.cartesian_volume_range((2.25,) * 3, (5.65,) * 3, 0.16)
    assert x.shape == (22 ** 3, 3)
    assert scenes.cartesian_volume_range((0, 0, 0), (2, 1, 0.5), 0.5).shape == (5 * 3 * 2, 3)
    # x fastest, then y, then z
    assert np.allclose(x[1] - x[0], (np.float32(0.16), 0, 0)) and np.allclose(x[22] - x[0], (0, np.float32(0.16), 0))


def test_rectangle_jet_and_axes():
    sc, jx, jv = scenes.reference_sphere_jet(0.1)
    assert jx.shape == (121, 3) and np.all(jx[:, 0] == np.float32(1.25))
    assert np.allclose(jx[:, 1].min(), 3.5) and np.allclose(jx[:, 1].max(), 4.5)
    assert np.allclose(jx[:, 2].min(), 2.5) and np.allclose(jx[:, 2].max(), 3.5)
    assert np.all(jv == np.float32([2.8, 0, 0]))
    # loop orders: "x": y fastest; "y": x fastest over (z, x); "z": y fastest over (x, y)
    r = scenes.cartesian_rectangle("y", 2.0, (0.0, 0.0), (0.2, 0.1), 0.1)
    assert r.shape == (6, 3) and np.all(r[:, 1] == 2.0) and r[1, 0] > r[0, 0] and r[1, 2] == r[0, 2]
    r = scenes.cartesian_rectangle("z", 2.0, (0.0, 0.0), (0.1, 0.2), 0.1)
    assert r.shape == (6, 3) and np.all(r[:, 2] == 2.0) and r[1, 1] > r[0, 1] and r[1, 0] == r[0, 0]


def test_cylinder_ring_closed_form():
    c = scenes.cartesian_cylinder("x", 3.0, (3.0, 3.0), 0.8, 0.2, 0.1)
    n = math.floor(2 * math.pi / (2 * math.asin(0.1 / 1.6)))
    assert c.shape[0] == 2 * n  # rings at h = 0 and 0.1
    d = np.hypot(c[:, 1] - 3.0, c[:, 2] - 3.0)
    assert np.allclose(d, 0.8, atol=1e-6)
    ring = c[:n]
    gaps = np.hypot(*(np.roll(ring[:, 1:], -1, axis=0) - ring[:, 1:]).T)
    assert np.all(gaps >= 0.1 - 1e-6)  # n = floor(2 pi / theta) keeps neighbours at least `space` apart


def test_line_inclusive():
    line = scenes.cartesian_line_inclusive((0, 0, 0), (1, 0, 0), 0.1)
    assert line.shape == (11, 3) and np.allclose(line[-1], (1, 0, 0), atol=1e-6)


def test_parameter_derivation():
    p = scenes.reference_params(0.1, 0.007, 2.0)
    # m = pow((double)space, 3) * rho0, narrowed to float
    assert p.particle_mass == float(np.float32(float(np.float32(0.1)) ** 3 * 1000.0))
    assert p.smoothing_length == float(np.float32(np.float32(1.4) * np.float32(0.1)))
    assert abs(p.relaxation - 200.0) < 1e-3
    d = scenes.reference_dam(0.16)
    assert d.n == 10648 and d.grid == (32, 32, 32) and abs(d.params.relaxation - 1 / 0.16 ** 2) < 1e-3


def test_jet_schedule():
    assert [s for s in range(1, 25) if scenes.jet_emits(s, 100)] == [8, 16, 24]
    assert not scenes.jet_emits(8, int(0.8 * scenes.JET_CAPACITY))


def run_jet_oracle(steps):
    sc, jx, jv = scenes.reference_sphere_jet(0.1)
    x, v = sc.x.copy(), sc.v.copy()
    for k in range(1, steps + 1):
        if scenes.jet_emits(k, x.shape[0]):
            x = np.concatenate([x, jx])
            v = np.concatenate([v, jv])
        r = orc.step(sc, x=x, v=v)
        x, v = r["x"], r["v"]
    return sc, x, v


def test_jet_on_the_oracle_stays_in_the_sphere():
    sc, x, v = run_jet_oracle(40)
    assert x.shape[0] == 121 * 6  # the construction batch and steps 8, 16, 24, 32, 40
    assert np.all(np.isfinite(x))
    r = np.linalg.norm(x - np.float32(3.0), axis=1)
    assert r.max() <= 3.0 * 1.001
