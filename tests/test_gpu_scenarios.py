"""The reference's own simulator scenes on the HIP solver, bit for bit
against the oracle: the application's sphere jet with emission
(simulator/sphere/pbf_sphere_sim.cpp:355-389, pbf::cuda::addParticle in
pbf/util/cuda/pbf_add.cu:12-28 through pbfx_add_particles) and the dam break
through the no-boundary step (simulator/dam/pbf_dam_sim.cpp:49-66)."""
from __future__ import annotations

import numpy as np
import pytest

from ssf_pbf_cuda_amd import scenes
from tests.helpers import oracle_run
from tests.test_scenes_reference import run_jet_oracle


@pytest.mark.gpu
def test_sphere_jet_with_emission_matches_oracle():
    import torch

    from ssf_pbf_cuda_amd import pbfx

    steps = 120
    sc, jx, jv = scenes.reference_sphere_jet(0.1)
    print(f"sphere jet: {steps} steps, batch {jx.shape[0]}", flush=True)
    solver = pbfx.Solver.for_scene(sc, capacity=scenes.JET_CAPACITY)
    x = torch.zeros((scenes.JET_CAPACITY, 3), dtype=torch.float32, device="cuda")
    v = torch.zeros_like(x)
    n = sc.n
    x[:n] = torch.from_numpy(sc.x).cuda()
    v[:n] = torch.from_numpy(sc.v).cuda()
    ax = torch.from_numpy(jx).cuda()
    av = torch.from_numpy(jv).cuda()
    for k in range(1, steps + 1):
        if scenes.jet_emits(k, n):
            n = solver.add_particles(x, v, n, ax, av, jx.shape[0])
        n = solver.step(x, v, n, sc.area_lo, sc.area_hi, sc.iters)
    gx, gv = x[:n].cpu().numpy(), v[:n].cpu().numpy()
    solver.close()
    _, rx, rv = run_jet_oracle(steps)
    # 16 batches emitted; a few particles leave the simulation area and are deleted
    assert n == rx.shape[0] and jx.shape[0] * (1 + steps // 8) - 40 < n <= jx.shape[0] * (1 + steps // 8)
    assert np.array_equal(gx.view(np.uint32), rx.view(np.uint32))
    assert np.array_equal(gv.view(np.uint32), rv.view(np.uint32))


@pytest.mark.gpu
def test_add_particles_rejects_overflow():
    import torch

    from ssf_pbf_cuda_amd import pbfx

    sc, jx, jv = scenes.reference_sphere_jet(0.1)
    solver = pbfx.Solver.for_scene(sc, capacity=200)
    x = torch.zeros((200, 3), dtype=torch.float32, device="cuda")
    v = torch.zeros_like(x)
    ax = torch.from_numpy(jx).cuda()
    av = torch.from_numpy(jv).cuda()
    n = solver.add_particles(x, v, 0, ax, av, jx.shape[0])
    assert n == 121
    with pytest.raises(pbfx.PbfxError) as e:
        solver.add_particles(x, v, n, ax, av, jx.shape[0])
    assert e.value.status == 2
    solver.close()


@pytest.mark.gpu
def test_reference_dam_no_boundary_matches_oracle():
    import torch

    from ssf_pbf_cuda_amd import pbfx

    sc = scenes.reference_dam(0.16)
    steps = 40
    print(f"reference dam: n={sc.n} steps={steps}", flush=True)
    solver = pbfx.Solver.for_scene(sc, boundary=False)
    x = torch.from_numpy(sc.x).cuda()
    v = torch.from_numpy(sc.v).cuda()
    n = sc.n
    for _ in range(steps):
        n = solver.step(x, v, n, sc.area_lo, sc.area_hi, sc.iters, boundary=False)
    ref = oracle_run(sc, steps=steps)
    solver.close()
    assert n == ref["num"]
    assert np.array_equal(x[:n].cpu().numpy().view(np.uint32), ref["x"].view(np.uint32))
    assert np.array_equal(v[:n].cpu().numpy().view(np.uint32), ref["v"].view(np.uint32))
