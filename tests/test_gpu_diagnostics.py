"""Density diagnostics (pbfx_density), order restoration (pbfx_restore_order)
and the 1000-step statistics of the plain dam break on the HIP solver, against
the oracle.  Reference: pbf/solver/cuda/pbf_density.cu:11-99,
pbf/interaction/cuda/pbf_neighbor_search_util.cu:156-170."""
from __future__ import annotations

import numpy as np
import pytest

from ssf_pbf_cuda_amd import scenes
from tests.helpers import oracle_run


def _solver_run(sc, steps):
    import torch

    from ssf_pbf_cuda_amd import pbfx

    solver = pbfx.Solver.for_scene(sc)
    x = torch.from_numpy(sc.x).cuda()
    v = torch.from_numpy(sc.v).cuda()
    x0, v0 = x.clone(), v.clone()
    num = sc.n
    for _ in range(steps):
        x0.copy_(x)
        v0.copy_(v)
        num = solver.step(x, v, num, sc.area_lo, sc.area_hi, sc.iters)
    return solver, x, v, x0, v0, num


@pytest.mark.gpu
def test_density_matches_oracle_bit_exact():
    import torch

    sc = scenes.small_dam(12, 10, 8)
    print(f"density: {sc.name} n={sc.n}", flush=True)
    solver, x, v, _, _, num = _solver_run(sc, 1)
    rho = torch.zeros(num, dtype=torch.float32, device="cuda")
    solver.density(x, rho)
    torch.cuda.synchronize()
    from oracle import pyoracle as orc

    r = orc.step(sc, want_ns=True)
    ref = orc.density(sc, r["x"], r["neighbors"])
    assert np.array_equal(rho.cpu().numpy().view(np.uint32), ref.view(np.uint32))
    solver.close()


@pytest.mark.gpu
def test_restore_order_inverts_the_sort():
    import torch

    sc = scenes.small_dam(12, 10, 8)
    print(f"restore: {sc.name} n={sc.n}", flush=True)
    solver, x, v, _, _, num = _solver_run(sc, 1)
    xo = torch.full_like(x, float("nan"))
    vo = torch.full_like(v, float("nan"))
    solver.restore_order(x, v, xo, vo)
    torch.cuda.synchronize()
    from ssf_pbf_cuda_amd import pbfx

    sidx = pbfx.d2h(solver.ns_view().sorted_index, sc.n)[:num]
    xs, vs = x.cpu().numpy()[:num], v.cpu().numpy()[:num]
    ex = np.full((sc.n, 3), np.nan, np.float32)
    ev = np.full((sc.n, 3), np.nan, np.float32)
    ex[sidx] = xs
    ev[sidx] = vs
    assert sorted(sidx.tolist()) == list(range(sc.n))
    assert np.array_equal(xo.cpu().numpy(), ex) and np.array_equal(vo.cpu().numpy(), ev)
    solver.close()


@pytest.mark.gpu
def test_plain_dam_1000_steps_statistics_match_oracle():
    # The 32,768-particle dam break (h = 0.1, 4 iterations) run for 1000
    # steps: the final state is the oracle's bit for bit (SHA-256 of the x and
    # v bytes in tests/golden/dam32k_1000steps.json, written by
    # tests/golden/make_statistics_fixture.py from the oracle, which needs
    # minutes of CPU for these 1000 steps), and its mean / max
    # density-constraint error and kinetic energy are the oracle's.
    import json

    from tests.golden.make_statistics_fixture import PATH, STEPS, state_record

    ref = json.load(open(PATH))
    sc = scenes.config0()
    print(f"statistics: {sc.name} n={sc.n} steps={STEPS}", flush=True)
    solver, x, v, _, _, num = _solver_run(sc, STEPS)
    gx, gv = x.cpu().numpy()[:num], v.cpu().numpy()[:num]
    solver.close()
    got = state_record(sc, gx, gv, num)
    print("gpu", got["metrics"], "\noracle", ref["metrics"], flush=True)
    assert got["num"] == ref["num"]
    for k in ("mean_C", "max_C", "mean_pos_C", "kinetic_energy"):
        assert got["metrics"][k] == pytest.approx(ref["metrics"][k], rel=1e-12), k
    assert got["sha256_x"] == ref["sha256_x"] and got["sha256_v"] == ref["sha256_v"]
