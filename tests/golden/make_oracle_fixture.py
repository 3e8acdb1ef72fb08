"""Writes tests/golden/oracle_dam8x8x6.npz: two full steps (PAPER s_corr,
XSPH, vorticity) of a 384-particle dam break computed by the CPU oracle,
with the last step's sort and neighbour lists.  The CPU tests check that the
oracle still reproduces it bit for bit; the GPU tests check the HIP solver
against it.  Run from the repository root: python tests/golden/make_oracle_fixture.py"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import pyoracle as orc  # noqa: E402
from ssf_pbf_cuda_amd import scenes  # noqa: E402

STEPS = 2


def scene():
    return scenes.small_dam(8, 8, 6, seed=11)


def run():
    sc = scene()
    x, v = sc.x, sc.v
    r = None
    for s in range(STEPS):
        r = orc.step(sc, x=x, v=v, want_ns=(s == STEPS - 1))
        x, v = r["x"], r["v"]
    return sc, r


if __name__ == "__main__":
    sc, r = run()
    out = os.path.join(ROOT, "tests", "golden", "oracle_dam8x8x6.npz")
    np.savez_compressed(out, x_in=sc.x, v_in=sc.v, x=r["x"], v=r["v"], num=np.uint32(r["num"]),
                        sorted_index=r["sorted_index"], cell_offset=r["cell_offset"], neighbors=r["neighbors"],
                        neighbor_count=r["neighbor_count"], lambda_last=r["lambda_last"], steps=np.uint32(STEPS))
    print("wrote", out, os.path.getsize(out), "bytes")
