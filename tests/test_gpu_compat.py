"""The reference-shaped C++ interface (include/pbf_compat.hpp): a C++ program
that uses pbf_particle / pbf_buffer / pbf_boundary / pbf::one_step exactly as
the reference's simulators do (simulator/sphere/pbf_sphere_sim.cpp:384-389)
runs through libpbfx.so, and its results match the oracle bit for bit: one
simulatee stepped as shipped (REFERENCE s_corr, no XSPH, no vorticity) and,
in the same process, one with its own options for the full step."""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np
import pytest

from ssf_pbf_cuda_amd import scenes
from tests.helpers import oracle_run, replace_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "compat_step")


def write_input(path, sc, steps):
    p = sc.params
    with open(path, "wb") as f:
        np.array([sc.n, steps, sc.iters, sc.plane_points.shape[0]], np.uint32).tofile(f)
        np.array([p.stable_density, p.stable_distance, p.particle_mass, p.smoothing_length, p.time_step,
                  p.relaxation, p.xsph_parameter, p.vc_parameter, *p.body_force, *sc.box_hi, *sc.area_lo,
                  *sc.area_hi], np.float32).tofile(f)
        np.ascontiguousarray(sc.plane_points, np.float32).tofile(f)
        np.ascontiguousarray(sc.plane_normals, np.float32).tofile(f)
        np.ascontiguousarray(sc.x, np.float32).tofile(f)
        np.ascontiguousarray(sc.v, np.float32).tofile(f)


def read_output(path):
    """The two simulatees' results: [(num, x, v) as shipped, (num, x, v) full]."""
    raw = np.fromfile(path, np.uint8)
    out, o = [], 0
    for _ in range(2):
        num = int(raw[o:o + 4].view(np.uint32)[0])
        f = raw[o + 4:o + 4 + 24 * num].view(np.float32)
        out.append((num, f[: 3 * num].reshape(num, 3), f[3 * num: 6 * num].reshape(num, 3)))
        o += 4 + 24 * num
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("steps", [1, 3])
def test_compat_one_step_matches_reference_mode_oracle(steps):
    assert os.path.exists(EXE), f"{EXE} is missing: run __graft_entry__.build()"
    sc = replace_params(scenes.small_dam(14, 10, 8), scorr_mode=scenes.SCORR_REFERENCE, xsph_on=0, vort_on=0,
                        max_pair=30, max_sweeps=32)
    with tempfile.TemporaryDirectory() as d:
        inp, out = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        write_input(inp, sc, steps)
        print(f"compat_step: {sc.name} n={sc.n} steps={steps}", flush=True)
        r = subprocess.run([EXE, inp, out], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        shipped, full = read_output(out)
    # two simulatees in one process, each with its own options: the step as
    # shipped and the full step (PAPER s_corr, XSPH, vorticity)
    full_sc = replace_params(sc, scorr_mode=scenes.SCORR_PAPER, scorr_k=0.1, scorr_n=4, scorr_dq_over_h=0.2,
                             xsph_on=1, vort_on=1)
    for (num, x, v), s in ((shipped, sc), (full, full_sc)):
        ref = oracle_run(s, steps=steps)
        assert num == ref["num"] == sc.n
        assert np.array_equal(x.view(np.uint32), ref["x"].view(np.uint32))
        assert np.array_equal(v.view(np.uint32), ref["v"].view(np.uint32))
