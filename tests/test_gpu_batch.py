"""Batched form (pbfx_batch_*): several independent scenes with their own
parameters, grids, boundaries and areas advanced by one sequence of launches
per step.  Each scene must evolve exactly as it does alone -- bit for bit
against the oracle's plain step of that scene."""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from ssf_pbf_cuda_amd import scenes
from tests.helpers import oracle_run, replace_params


def batch_scenes():
    a = scenes.small_dam(12, 10, 8, seed=3)
    b = replace_params(scenes.small_dam(10, 9, 7, seed=4), xsph_parameter=0.03, vc_parameter=0.0008)
    b = replace_params(b, stable_density=b.params.stable_density * 1.05)
    c = replace_params(scenes.small_dam(8, 6, 5, h=0.08, box=(1.28, 1.28, 0.64), seed=5), vort_on=0)
    d = replace_params(scenes.small_dam(9, 9, 9, seed=6), scorr_mode=scenes.SCORR_REFERENCE, xsph_on=0, vort_on=0)
    return [a, b, c, d]


def _run_batch(scs, steps, async_=False):
    import torch

    from ssf_pbf_cuda_amd import pbfx

    solver = pbfx.BatchSolver(scs)
    x = torch.from_numpy(np.concatenate([s.x for s in scs])).cuda()
    v = torch.from_numpy(np.concatenate([s.v for s in scs])).cuda()
    counts = [s.n for s in scs]
    iters = scs[0].iters
    if async_:
        off = torch.tensor(np.concatenate([[0], np.cumsum(counts)]).astype(np.int32), device="cuda")
        total = sum(counts)
        for _ in range(steps):
            solver.batch_step_async(x, v, off, total, iters)
            solver.batch_offsets(off)
        torch.cuda.synchronize()
        o = off.cpu().numpy().astype(np.int64)
        counts = list(np.diff(o))
    else:
        for _ in range(steps):
            counts = solver.batch_step(x, v, counts, iters)
    xs, vs = x.cpu().numpy(), v.cpu().numpy()
    solver.close()
    out, o = [], 0
    for n in counts:
        out.append((xs[o:o + n], vs[o:o + n], n))
        o += n
    return out


def _check_against_oracle(scs, res, steps):
    for k, (sc, (gx, gv, n)) in enumerate(zip(scs, res)):
        ref = oracle_run(sc, steps=steps)
        assert n == ref["num"], (k, n, ref["num"])
        assert np.array_equal(gx.view(np.uint32), ref["x"].view(np.uint32)), k
        assert np.array_equal(gv.view(np.uint32), ref["v"].view(np.uint32)), k


@pytest.mark.gpu
@pytest.mark.parametrize("steps", [1, 3])
def test_batch_matches_each_scene_alone(steps):
    scs = batch_scenes()
    print(f"batch: {[s.n for s in scs]} steps={steps}", flush=True)
    _check_against_oracle(scs, _run_batch(scs, steps), steps)


@pytest.mark.gpu
def test_batch_async_offsets_chain():
    scs = batch_scenes()
    print(f"batch async: {[s.n for s in scs]}", flush=True)
    _check_against_oracle(scs, _run_batch(scs, 2, async_=True), 2)


@pytest.mark.gpu
def test_batch_deletion_per_scene():
    # one scene's simulation area cuts through its fluid: its count shrinks,
    # the others are untouched, and the output stays concatenated in order
    scs = batch_scenes()
    b = scs[1]
    hi = np.array(b.area_hi, np.float32)
    hi[0] = np.float32(0.4)
    scs[1] = dataclasses.replace(b, area_hi=hi)
    print(f"batch deletion: {[s.n for s in scs]}", flush=True)
    res = _run_batch(scs, 1)
    assert res[1][2] < scs[1].n
    _check_against_oracle(scs, res, 1)


@pytest.mark.gpu
def test_autotune_keeps_results():
    # the autotuner only times candidates on scratch copies; the step after
    # it is still the oracle's, bit for bit
    import torch

    from ssf_pbf_cuda_amd import pbfx

    sc = scenes.small_dam(12, 10, 8)
    print(f"autotune: {sc.name} n={sc.n}", flush=True)
    solver = pbfx.Solver.for_scene(sc)
    x = torch.from_numpy(sc.x).cuda()
    v = torch.from_numpy(sc.v).cuda()
    xs, vs = torch.empty_like(x), torch.empty_like(v)
    tile, width = solver.autotune(x, v, sc.n, sc.area_lo, sc.area_hi, sc.iters, xs, vs)
    assert tile in (1024, 2048, 3072, 4096) and width in (4, 8, 16)
    assert np.array_equal(x.cpu().numpy(), sc.x)
    n = solver.step(x, v, sc.n, sc.area_lo, sc.area_hi, sc.iters)
    ref = oracle_run(sc)
    assert n == ref["num"]
    assert np.array_equal(x.cpu().numpy()[:n].view(np.uint32), ref["x"].view(np.uint32))
    solver.close()


@pytest.mark.gpu
def test_config3_batch_step():
    """One batched step of the full batched workload: 256 dam-break scenes,
    11,690,923 particles (4,352-195,112 per scene), per-scene rest density,
    XSPH c and vorticity eps.  On every scene: the count is kept (the area
    deletes nothing), the sorted index maps the scene's output slots onto a
    permutation of its own inputs, and the state is finite.  The smallest,
    the largest and four seeded other scenes equal the oracle's plain step
    bit for bit."""
    import torch

    from ssf_pbf_cuda_amd import pbfx

    scs = scenes.batch_config3()
    counts = np.array([s.n for s in scs], np.int64)
    total = int(counts.sum())
    print(f"config3 batch: {len(scs)} scenes, {total} particles", flush=True)
    solver = pbfx.BatchSolver(scs)
    x = torch.from_numpy(np.concatenate([s.x for s in scs])).cuda()
    v = torch.from_numpy(np.concatenate([s.v for s in scs])).cuda()
    out_counts = solver.batch_step(x, v, list(counts), scs[0].iters)
    assert list(out_counts) == list(counts)
    sidx = pbfx.d2h(solver.ns_view().sorted_index, total).astype(np.int64)
    xs, vs = x.cpu().numpy(), v.cpu().numpy()
    solver.close()
    assert np.isfinite(xs).all() and np.isfinite(vs).all()
    off = np.concatenate([[0], np.cumsum(counts)])
    for k in range(len(scs)):
        seg = np.sort(sidx[off[k]:off[k + 1]])
        assert np.array_equal(seg, np.arange(off[k], off[k + 1])), f"scene {k}: not a permutation of its inputs"
    rng = np.random.default_rng(2024)
    pick = sorted({int(counts.argmin()), int(counts.argmax()), *rng.choice(len(scs), 4, replace=False).tolist()})
    for k in pick:
        ref = oracle_run(scs[k])
        gx, gv = xs[off[k]:off[k + 1]], vs[off[k]:off[k + 1]]
        same = (np.array_equal(gx.view(np.uint32), ref["x"].view(np.uint32)) and
                np.array_equal(gv.view(np.uint32), ref["v"].view(np.uint32)))
        print(f"scene {k} ({scs[k].n} particles): bit-exact={same}", flush=True)
        assert ref["num"] == counts[k] and same, k


@pytest.mark.gpu
@pytest.mark.parametrize("steps_per_graph,launches", [(1, 3), (3, 1)])
def test_batch_graph_replays_chain(steps_per_graph, launches):
    """pbfx_batch_graph_capture / _launch: three chained steps recorded as one
    graph or replayed three times give each scene's oracle state after three
    steps, bit for bit; one scene's area deletes particles, so the counts the
    kernels read from the device change between replays."""
    import torch

    from ssf_pbf_cuda_amd import pbfx

    scs = batch_scenes()
    b = scs[1]
    hi = np.array(b.area_hi, np.float32)
    hi[0] = np.float32(0.4)
    scs[1] = dataclasses.replace(b, area_hi=hi)
    print(f"batch graph: {[s.n for s in scs]} steps/graph={steps_per_graph} launches={launches}", flush=True)
    solver = pbfx.BatchSolver(scs)
    x = torch.from_numpy(np.concatenate([s.x for s in scs])).cuda()
    v = torch.from_numpy(np.concatenate([s.v for s in scs])).cuda()
    counts = [s.n for s in scs]
    off = torch.tensor(np.concatenate([[0], np.cumsum(counts)]).astype(np.int32), device="cuda")
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    solver.graph_capture(x, v, off, scs[0].iters, steps_per_graph, stream)
    solver.graph_launch(launches, stream)
    stream.synchronize()
    o = off.cpu().numpy().astype(np.int64)
    xs, vs = x.cpu().numpy(), v.cpu().numpy()
    solver.close()
    res = [(xs[o[k]:o[k + 1]], vs[o[k]:o[k + 1]], int(o[k + 1] - o[k])) for k in range(len(scs))]
    assert res[1][2] < scs[1].n
    _check_against_oracle(scs, res, 3)
