"""Closed-form cases for the final passes and the boundary responses.

Each case is a scene of two to four particles whose result the mathematics
fixes, computed here in FP64 from the textbook kernels
  W(r)  = 315 / (64 pi h^9) (h^2 - r^2)^3        (poly6, r <= h)
  W'(r) = -45 / (pi h^6) (h - r)^2                (spiky derivative, r <= h)
and the published formulas of the passes, not from the oracle's code:

  XSPH (pbf/solver/cuda/pbf_xsph.cu:15-29,106-111)
      v_i' = v_i + c (m / rho0) sum_j (v_j - v_i) W(|x_j - x_i|)
  vorticity (pbf/solver/cuda/pbf_vorticity_confinement.cu:15-31)
      omega_i = (m / rho0) sum_j (v_j - v_i) x (x_j - x_i)/r_ij W'(r_ij)
  confinement (pbf_vorticity_confinement.cu:121-138,192-222)
      eta_i = sum_j |omega_j| (x_j - x_i)/r_ij W'(r_ij),  N_i = eta_i / |eta_i|
      v_i' = v_i + eps (N_i x omega_i) rho0 dt
  inner sphere (pbf/solver/cuda/boundary/pbf_sphere_boundary.cu:8-34):
      a predicted point at distance d > R from the centre c moves to
      c + (x - c) (R/d - 1e-4)                     (onto the sphere, nudged inward)
  outer sphere (:36-64): a point at d < R moves to c + (x - c)(R/d + 1e-4)
  plane crossing (pbf/solver/cuda/boundary/pbf_plane_boundary.cu:21-23,38-40):
      a point that starts in front of the plane and ends behind it moves to
      the plane along the normal (the segment's intersection I lies on the
      plane, so -((x - I).n) n = -((x - P).n) n).

The XSPH and vorticity checks relate two runs of the same scene, the pass
off and on: the positions and the velocities entering the pass are the run
without it (so the check needs nothing else from the code under test).
The relaxation is set to 1e20 and the s_corr to the REFERENCE form, so the
density constraint moves no particle (s ~ 1e-20), and gravity is zero.

A case's `check(off, on)` returns the normwise relative error of the
quantity the closed form fixes; `tol` is the bar.  Used by
tests/test_oracle_closed_forms.py (the oracle, with planted mistakes that
must fail the same checks) and tests/test_gpu_closed_forms.py (the HIP
path).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from ssf_pbf_cuda_amd import scenes

H = 0.1


def poly6(r, h):
    return 315.0 / (64.0 * math.pi * h ** 9) * (h * h - r * r) ** 3 if r <= h else 0.0


def spiky_deriv(r, h, arg=None):
    """W'(r) with kernel argument `arg` (1/h normally): -45/pi arg^4 (1 - r arg)^2."""
    a = 1.0 / h if arg is None else arg
    q = r * a
    return -45.0 / math.pi * a ** 4 * (1.0 - q) ** 2 if q <= 1.0 else 0.0


def free_scene(x, v, h=H, box=(4.0, 4.0, 4.0), planes=None, inner=None, outer=None, iters=1, **over):
    """Particles x, v in a grid of `box`, with no boundary unless given:
    planes = (points (k,3), normals (k,3)); inner / outer = (k,4) spheres."""
    base = scenes.small_dam(4, 4, 4, h=h, box=box, iters=iters)
    p = dataclasses.replace(base.params, body_force=(0.0, 0.0, 0.0), **over)
    pts, nrm = (np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32)) if planes is None else planes
    return dataclasses.replace(
        base, x=np.ascontiguousarray(x, np.float32), v=np.ascontiguousarray(v, np.float32), params=p,
        plane_points=np.ascontiguousarray(pts, np.float32), plane_normals=np.ascontiguousarray(nrm, np.float32),
        inner_spheres=np.zeros((0, 4), np.float32) if inner is None else np.ascontiguousarray(inner, np.float32),
        outer_spheres=np.zeros((0, 4), np.float32) if outer is None else np.ascontiguousarray(outer, np.float32))


def _cluster(seed, k=3):
    """k particles within 0.8h of each other, not coplanar, away from cell
    faces' special values; velocities U(-1, 1)."""
    rng = np.random.default_rng(seed)
    c = np.array([1.234, 1.467, 1.611])
    dirs = rng.normal(size=(k, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    x = (c + dirs * rng.uniform(0.2, 0.38, (k, 1)) * H).astype(np.float32)
    v = rng.uniform(-1.0, 1.0, (k, 3)).astype(np.float32)
    return x, v


def _f64(a):
    return np.asarray(a, np.float64)


def _mr(p):
    return float(np.float32(p.particle_mass)) / float(np.float32(p.stable_density))


def xsph_expected(p, x, v):
    """c (m/rho0) sum_j (v_j - v_i) W(|x_j - x_i|) for every particle."""
    x, v = _f64(x), _f64(v)
    h = float(np.float32(p.smoothing_length))
    out = np.zeros_like(v)
    for i in range(len(x)):
        for j in range(len(x)):
            out[i] += (v[j] - v[i]) * poly6(float(np.linalg.norm(x[j] - x[i])), h)
    return float(np.float32(p.xsph_parameter)) * _mr(p) * out


def vorticity_expected(p, x, v):
    """Velocity change of the vorticity confinement (omega, eta, N as in the
    module docstring) for every particle."""
    x, v = _f64(x), _f64(v)
    h = float(np.float32(p.smoothing_length))
    arg = h if p.vort_h_as_inv_h else 1.0 / h
    n = len(x)
    om = np.zeros_like(v)
    for i in range(n):
        for j in range(n):
            d = x[j] - x[i]
            r = float(np.linalg.norm(d))
            if r > 0:
                om[i] += np.cross(v[j] - v[i], d / r) * spiky_deriv(r, h, arg)
    om *= _mr(p)
    dv = np.zeros_like(v)
    rho0 = float(np.float32(p.stable_density))
    dt = float(np.float32(p.time_step))
    eps = float(np.float32(p.vc_parameter))
    for i in range(n):
        eta = np.zeros(3)
        for j in range(n):
            d = x[j] - x[i]
            r = float(np.linalg.norm(d))
            if r > 0:
                eta += np.linalg.norm(om[j]) * d / r * spiky_deriv(r, h, arg)
        if np.linalg.norm(eta) > 0:
            N = eta / np.linalg.norm(eta)
            dv[i] = eps * np.cross(N, om[i]) * rho0 * dt
    return dv


def _normwise(got, ref):
    got, ref = _f64(got), _f64(ref)
    return float(np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300))


@dataclasses.dataclass
class PassCase:
    """A final pass, checked as (run with it on) - (run with it off)."""
    name: str
    scene_off: scenes.Scene
    scene_on: scenes.Scene
    expected: callable   # (params, x, v_off) -> velocity change
    tol: float

    def check(self, off, on):
        assert np.array_equal(on["x"], off["x"]), f"{self.name}: the pass moved particles"
        dv = _f64(on["v"]) - _f64(off["v"])
        want = self.expected(self.scene_on.params, off["x"], off["v"])
        assert np.abs(want).max() > 1e-3, f"{self.name}: degenerate case"
        return _normwise(dv, want)


def pass_cases():
    cases = []
    common = dict(scorr_mode=scenes.SCORR_REFERENCE, relaxation=1e20, max_pair=64)
    for seed, k in ((11, 3), (12, 4)):
        x, v = _cluster(seed, k)
        cases.append(PassCase(f"xsph_{k}", free_scene(x, v, xsph_on=0, vort_on=0, xsph_parameter=0.3, **common),
                              free_scene(x, v, xsph_on=1, vort_on=0, xsph_parameter=0.3, **common),
                              xsph_expected, 1e-5))
        cases.append(PassCase(f"vorticity_{k}", free_scene(x, v, xsph_on=0, vort_on=0, vc_parameter=0.001, **common),
                              free_scene(x, v, xsph_on=0, vort_on=1, vc_parameter=0.001, **common),
                              vorticity_expected, 1e-5))
    # the reference's literal kernel argument (h where 1/h belongs): omega is
    # ~1e-8 of the corrected form's, so eps is raised to make it visible
    x, v = _cluster(13, 4)
    cases.append(PassCase("vorticity_h_as_inv_h",
                          free_scene(x, v, xsph_on=0, vort_on=0, vc_parameter=3e5, vort_h_as_inv_h=1, **common),
                          free_scene(x, v, xsph_on=0, vort_on=1, vc_parameter=3e5, vort_h_as_inv_h=1, **common),
                          vorticity_expected, 1e-5))
    return cases


@dataclasses.dataclass
class BoundaryCase:
    """One step whose final positions the closed form fixes."""
    name: str
    scene: scenes.Scene
    expected: np.ndarray  # (n, 3) FP64 final positions in the step's output order
    tol: float            # absolute

    def check(self, out):
        got = _f64(out["x"])
        assert got.shape == self.expected.shape, (got.shape, self.expected.shape)
        return float(np.abs(got - self.expected).max())


def _predict(x, v, dt):
    return (_f64(v) * float(np.float32(dt)) + _f64(x)).astype(np.float32).astype(np.float64)


def sphere_cases():
    cases = []
    c = np.array([2.013, 2.107, 1.911])
    u = np.array([0.36, -0.48, 0.8])
    u /= np.linalg.norm(u)
    # inner sphere R = 0.5: the prediction ends 0.515 from the centre
    R = 0.5
    x = (c + 0.49 * u).astype(np.float32)[None]
    sc = free_scene(x, (3.0 * u).astype(np.float32)[None], inner=np.array([[*c, R]]), xsph_on=0, vort_on=0)
    xs = _predict(sc.x, sc.v, sc.params.time_step)[0]
    d = float(np.linalg.norm(xs - c))
    assert d > R
    cases.append(BoundaryCase("inner_sphere", sc, (c + (xs - c) * (R / d - 1e-4))[None], 1e-6))
    # outer sphere (an obstacle) R = 0.3: the prediction ends 0.285 from the centre
    R = 0.3
    x = (c + 0.31 * u).astype(np.float32)[None]
    sc = free_scene(x, (-3.0 * u).astype(np.float32)[None], outer=np.array([[*c, R]]), xsph_on=0, vort_on=0)
    xs = _predict(sc.x, sc.v, sc.params.time_step)[0]
    d = float(np.linalg.norm(xs - c))
    assert d < R
    cases.append(BoundaryCase("outer_sphere", sc, (c + (xs - c) * (R / d + 1e-4))[None], 1e-6))
    return cases


def two_particle_deltap(p, xa, xb):
    """Delta-p of particle a in a two-particle system, REFERENCE s_corr, one
    Jacobi iteration (the closed form of the density constraint)."""
    h = float(np.float32(p.smoothing_length))
    d = _f64(xa) - _f64(xb)
    rr = float(np.linalg.norm(d))
    mr = _mr(p)
    C = mr * (poly6(0.0, h) + poly6(rr, h)) - 1.0
    wd = spiky_deriv(rr, h)
    s = C / (mr * mr * 2.0 * wd * wd + float(np.float32(p.relaxation)))
    corr = 10.0 * (max(-s, 0.0) * 2) * poly6(rr, h) ** 4 / poly6(0.4 * h, h) ** 4
    return mr * (-(2.0 * s + corr)) * wd * d / rr


def crossing_cases():
    """A pair straddling the floor y = 0: the constraint pulls the upper
    particle (0.001 above the floor) through it, so its response takes the
    crossing branch and ends on the floor, with the tangential part of its
    delta-p kept.  Once with the floor alone (the general sweep) and once
    as the bottom face of a six-plane box (the box's own code path)."""
    xa = np.array([1.0031, 0.001, 1.0017])
    # at 0.65h the pair attracts (below ~0.55h the tensile term repels)
    u = np.array([0.2, -0.9, 0.05])
    xb = xa + 0.65 * H * u / np.linalg.norm(u)
    x = np.stack([xa, xb]).astype(np.float32)
    cases = []
    floor = (np.zeros((1, 3), np.float32), np.array([[0.0, 1.0, 0.0]], np.float32))
    box = scenes.box_planes((0.0, 0.0, 0.0), (2.0, 2.0, 2.0))
    for name, planes in (("plane_crossing", floor), ("box_crossing", box)):
        sc = free_scene(x, np.zeros_like(x), planes=planes, scorr_mode=scenes.SCORR_REFERENCE, xsph_on=0,
                        vort_on=0)
        dpa = two_particle_deltap(sc.params, x[0], x[1])
        assert _f64(x[0])[1] + dpa[1] < 0, "the case must cross the floor"
        want_a = _f64(x[0]) + dpa
        want_a[1] = 0.0
        # b starts behind the floor: projected onto it (the other branch)
        want_b = _f64(x[1]) - two_particle_deltap(sc.params, x[0], x[1])
        want_b[1] = 0.0
        # sorted order: the pair shares a cell column; b (y < 0) hashes to row -1,
        # which wraps to the grid's last row, so a comes first
        cases.append(BoundaryCase(name, sc, np.stack([want_a, want_b]), 3e-7))
    return cases
