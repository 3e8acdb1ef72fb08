"""Shared helpers of the parity tests: run a scene on the HIP solver and on
the CPU oracle, and compare the results."""
from __future__ import annotations

import dataclasses

import numpy as np

from ssf_pbf_cuda_amd import scenes


def replace_params(scene, **kw):
    p = dataclasses.replace(scene.params, **kw)
    return dataclasses.replace(scene, params=p)


def gpu_run(scene, steps=1, tile_cap=None, unit_width=None, boundary=True, capacity=None, x=None, v=None,
            with_ns=False):
    """Runs `steps` plain steps of `scene` through the C-ABI on cuda:0."""
    import torch

    from ssf_pbf_cuda_amd import pbfx

    x = scene.x if x is None else x
    v = scene.v if v is None else v
    n = x.shape[0]
    solver = pbfx.Solver.for_scene(scene, capacity=capacity or max(n, 1))
    if tile_cap is not None or unit_width is not None:
        solver.set_tuning(2048 if tile_cap is None else tile_cap, 16 if unit_width is None else unit_width)
    cap = solver.capacity
    xt = torch.zeros((cap, 3), dtype=torch.float32, device="cuda")
    vt = torch.zeros((cap, 3), dtype=torch.float32, device="cuda")
    if n:
        xt[:n] = torch.from_numpy(np.ascontiguousarray(x, np.float32)).cuda()
        vt[:n] = torch.from_numpy(np.ascontiguousarray(v, np.float32)).cuda()
    torch.cuda.synchronize()
    num = n
    for _ in range(steps):
        num = solver.step(xt, vt, num, scene.area_lo, scene.area_hi, scene.iters, boundary=boundary)
    out = {"x": xt[:num].cpu().numpy(), "v": vt[:num].cpu().numpy(), "num": num, "stats": solver.stats()}
    if with_ns:
        out.update(ns_views(solver, n, num, scene.params.max_pair))
    solver.close()
    return out


def ns_views(solver, n_in, n_live, max_pair):
    import torch

    from ssf_pbf_cuda_amd import pbfx

    view = solver.ns_view()
    C = int(view.num_cells)
    out = {
        "sorted_index": pbfx.d2h(view.sorted_index, n_in),
        "cell_offset": pbfx.d2h(view.cell_offset, C + 2),
        "key": pbfx.d2h(view.hash, n_in),
    }
    nb = torch.zeros(max(n_live, 1) * max_pair, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(max(n_live, 1), dtype=torch.int32, device="cuda")
    if n_live:
        solver.ns_neighbors(nb, cnt)
    lam = torch.zeros(max(n_live, 1), dtype=torch.float32, device="cuda")
    solver.scaling_factor(lam)
    start = torch.zeros(C, dtype=torch.int32, device="cuda")
    end = torch.zeros(C, dtype=torch.int32, device="cuda")
    solver.ns_cell_ranges(start, end)
    torch.cuda.synchronize()
    out["neighbors"] = nb.cpu().numpy().view(np.uint32)[: n_live * max_pair].reshape(n_live, max_pair)
    out["neighbor_count"] = cnt.cpu().numpy().view(np.uint32)[:n_live]
    out["lambda_last"] = lam.cpu().numpy()[:n_live]
    out["cell_start"] = start.cpu().numpy().view(np.uint32)
    out["cell_end"] = end.cpu().numpy().view(np.uint32)
    return out


def oracle_run(scene, steps=1, boundary=True, want_ns=False, x=None, v=None):
    from oracle import pyoracle as orc

    x = scene.x if x is None else x
    v = scene.v if v is None else v
    r = None
    for s in range(steps):
        r = orc.step(scene, x=x, v=v, want_ns=want_ns and s == steps - 1, boundary=boundary)
        x, v = r["x"], r["v"]
    return r


def rel_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    if ref.size == 0:
        return 0.0
    scale = max(np.abs(ref).max(), 1e-30)
    return float(np.abs(got - ref).max() / scale)


def bit_mismatches(got, ref) -> int:
    a = np.ascontiguousarray(got, np.float32).view(np.uint32)
    b = np.ascontiguousarray(ref, np.float32).view(np.uint32)
    both_nan = np.isnan(np.asarray(got, np.float32)) & np.isnan(np.asarray(ref, np.float32))
    return int(((a != b) & ~both_nan).sum())


def compare_state(tag, got, ref, rtol=1e-5):
    """Asserts the FP32 parity bar (normwise relative error <= rtol) and
    returns a short report including the number of bitwise mismatches."""
    assert got["num"] == ref["num"], f"{tag}: live count {got['num']} != {ref['num']}"
    ex = rel_err(got["x"], ref["x"])
    ev = rel_err(got["v"], ref["v"])
    bx = bit_mismatches(got["x"], ref["x"])
    bv = bit_mismatches(got["v"], ref["v"])
    msg = f"{tag}: n={got['num']} rel_x={ex:.3e} rel_v={ev:.3e} bitdiff_x={bx} bitdiff_v={bv}"
    print(msg, flush=True)
    assert ex <= rtol and ev <= rtol, msg
    return {"rel_x": ex, "rel_v": ev, "bitdiff_x": bx, "bitdiff_v": bv}


def small(nx=12, ny=10, nz=8, **kw):
    return scenes.small_dam(nx, ny, nz, **kw)
