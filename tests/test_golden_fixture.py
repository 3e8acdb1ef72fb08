"""The committed oracle fixture (tests/golden/oracle_dam8x8x6.npz, written by
tests/golden/make_oracle_fixture.py): the oracle must still reproduce it bit
for bit, and so must the HIP solver."""
from __future__ import annotations

import os

import numpy as np
import pytest

from tests.golden.make_oracle_fixture import STEPS, run, scene

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_dam8x8x6.npz")


def load():
    d = np.load(FIX, allow_pickle=False)
    return {k: d[k] for k in d.files}


def test_fixture_inputs_are_the_scene():
    d = load()
    sc = scene()
    assert np.array_equal(d["x_in"], sc.x) and np.array_equal(d["v_in"], sc.v)


def test_oracle_reproduces_fixture():
    d = load()
    _, r = run()
    assert int(d["num"]) == r["num"]
    for k in ("x", "v", "lambda_last"):
        assert np.array_equal(d[k].view(np.uint32), r[k].view(np.uint32)), k
    for k in ("sorted_index", "cell_offset", "neighbors", "neighbor_count"):
        assert np.array_equal(d[k], r[k]), k


@pytest.mark.gpu
def test_hip_solver_reproduces_fixture():
    from tests.helpers import gpu_run

    d = load()
    sc = scene()
    print(f"fixture: {sc.name} n={sc.n} steps={STEPS}", flush=True)
    g = gpu_run(sc, steps=STEPS, with_ns=True)
    n = int(d["num"])
    assert g["num"] == n
    assert np.array_equal(g["x"].view(np.uint32), d["x"].view(np.uint32))
    assert np.array_equal(g["v"].view(np.uint32), d["v"].view(np.uint32))
    assert np.array_equal(g["sorted_index"], d["sorted_index"])
    assert np.array_equal(g["cell_offset"], d["cell_offset"])
    assert np.array_equal(g["neighbor_count"], d["neighbor_count"])
    assert np.array_equal(g["lambda_last"].view(np.uint32), d["lambda_last"].view(np.uint32))
