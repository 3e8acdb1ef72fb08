"""pbfx_step_static, the allocation-free step, against pbfx_step_async of a
context on the same input: bit for bit, on calls in a row with inputs of
different sizes (the 32,768-particle dam, then a 960-particle dam, then the
32k dam again), since every call reuses the one static device array and must
clear what the previous call left.  A workload beyond the array's size is
rejected with PBFX_CAPACITY before any GPU work."""
from __future__ import annotations

import numpy as np
import pytest

from ssf_pbf_cuda_amd import scenes

pytestmark = pytest.mark.gpu


def _inputs(sc):
    import torch

    return torch.from_numpy(sc.x).cuda(), torch.from_numpy(sc.v).cuda()


def _context_step(sc):
    import torch

    from ssf_pbf_cuda_amd import pbfx

    x, v = _inputs(sc)
    solver = pbfx.Solver.for_scene(sc)
    solver.step_async(x, v, sc.n, sc.area_lo, sc.area_hi, sc.iters, None)
    torch.cuda.synchronize()
    solver.close()
    return x.cpu().numpy(), v.cpu().numpy()


def _static_step(sc):
    import torch

    from ssf_pbf_cuda_amd import pbfx

    x, v = _inputs(sc)
    dev = lambda a, w: torch.from_numpy(np.ascontiguousarray(a, np.float32).reshape(-1, w)).cuda() if len(a) else None
    pbfx.step_static(sc.params, sc.box_hi, x, v, sc.n, sc.area_lo, sc.area_hi, sc.iters,
                     inner=dev(sc.inner_spheres, 4), outer=dev(sc.outer_spheres, 4),
                     plane_points=dev(sc.plane_points, 3), plane_normals=dev(sc.plane_normals, 3))
    torch.cuda.synchronize()
    return x.cpu().numpy(), v.cpu().numpy()


def test_static_step_matches_context_step_on_calls_in_a_row():
    big = scenes.config0()
    small = scenes.small_dam(12, 10, 8)
    for sc in (big, small, big):
        print(f"static step: {sc.name} n={sc.n}", flush=True)
        sx, sv = _static_step(sc)
        cx, cv = _context_step(sc)
        assert np.array_equal(sx.view(np.uint32), cx.view(np.uint32)), sc.name
        assert np.array_equal(sv.view(np.uint32), cv.view(np.uint32)), sc.name


def test_static_step_rejects_oversize():
    from ssf_pbf_cuda_amd import pbfx

    sc = scenes.config4()  # 65,536 particles need ~35 MB of scratch
    print(f"static step oversize: {sc.name} n={sc.n}", flush=True)
    x, v = _inputs(sc)
    with pytest.raises(pbfx.PbfxError) as e:
        pbfx.step_static(sc.params, sc.box_hi, x, v, sc.n, sc.area_lo, sc.area_hi, sc.iters)
    assert e.value.status == pbfx.PBFX_CAPACITY


def test_static_step_with_big_cells_matches_context_step():
    """The allocation-free step on piles that the wrapping hash puts into one
    cell and a pile of identical points: the big cells' groups and boxes
    (k_big_seg) live in the static array too, after a call of another size."""
    from test_gpu_aliased_piles import aliased_scene

    small = scenes.small_dam(12, 10, 8)
    pile = aliased_scene(far=1500, near=900, point=300, seed=13)
    for sc in (small, pile):
        print(f"static step: {sc.name} n={sc.n}", flush=True)
        sx, sv = _static_step(sc)
        cx, cv = _context_step(sc)
        assert np.array_equal(sx.view(np.uint32), cx.view(np.uint32)), sc.name
        assert np.array_equal(sv.view(np.uint32), cv.view(np.uint32)), sc.name
