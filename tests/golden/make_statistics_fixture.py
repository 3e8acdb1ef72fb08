"""Writes tests/golden/dam32k_1000steps.json: the oracle's state after 1000
steps of the 32,768-particle dam break (scenes.config0, 4 iterations) as
SHA-256 digests of its x and v bytes, its live count and its statistics
(mean / max density-constraint error, mean positive error, kinetic energy).
The GPU test compares its own 1000-step state against it.  Regenerate with
`python tests/golden/make_statistics_fixture.py`: the oracle needs several
minutes of CPU for these 1000 steps (the dam collapses into a dense pile whose
27-cell sums grow with it), too long for the CPU test suite, so no test
reruns it."""
from __future__ import annotations

import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dam32k_1000steps.json")
STEPS = 1000


def state_record(sc, x, v, num):
    from oracle import pyoracle as orc

    m = orc.metrics(sc, x, v)
    return {
        "steps": STEPS,
        "num": int(num),
        "sha256_x": hashlib.sha256(x.tobytes()).hexdigest(),
        "sha256_v": hashlib.sha256(v.tobytes()).hexdigest(),
        "metrics": {k: float(val) for k, val in m.items()},
    }


def oracle_record():
    from ssf_pbf_cuda_amd import scenes
    from tests.helpers import oracle_run

    sc = scenes.config0()
    r = oracle_run(sc, steps=STEPS)
    return state_record(sc, r["x"], r["v"], r["num"])


if __name__ == "__main__":
    rec = oracle_record()
    with open(PATH, "w") as f:
        json.dump(rec, f, indent=1)
    print(PATH, rec["metrics"])
