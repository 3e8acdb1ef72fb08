#!/usr/bin/env python3
"""Throughput of the PBF solver step on one MI355X.

A step is one full pbf::one_step over the whole particle set: predict,
hashed-grid neighbour search, I Jacobi iterations, velocity update,
vorticity confinement and XSPH.  Default workload: the 1,048,576-particle
dam break (128x128x64 lattice, spacing 0.5h, h = 0.05, 256x128x128 cells, 4
Jacobi iterations); the 8,388,608-particle double dam break (3 iterations),
the 256-scene batch and the differentiable 65,536-particle workload (40
steps forward + backward) are reported beside it.  Inputs are resident in HBM
before the timed region.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "M particle-steps/s at 1M/8M (4 solver iters); HBM GB/s; batched vs plain"
HBM_PEAK_GBS = 8000.0

# Compulsory FP32 state traffic per particle for each kernel: each logical
# array read once and written once per pass (neighbour gathers counted once).
ALG_BYTES = {
    "k_predict_hash": 48,   # x 12, v 12 -> hash 4 + (x* 12 of the reference's predict) ; rank 4, histogram read 4
    "k_scatter": 16,        # hash 4, rank 4 -> sorted index 4, sorted hash 4
    "k_reorder": 56,        # index 4, x* 12, x 12 -> x*s 12, xs 12 ; cell ranges 4
    "k_ns": 12,             # x* 12 (the lists themselves are not compulsory traffic)
    "k_lambda": 16,         # x* 12 -> lambda 4, once per Jacobi iteration
    "k_deltap": 28,         # x* 12, lambda 4 -> x*' 12
    "k_vort": 48,           # x* 12, xs 12 -> v 12, omega 12
    "k_confine": 48,        # x* 12, omega 12, v 12 -> v 12
    "k_xsph_out": 36,       # x* 12, v 12 -> v' 12
    "k_out": 36,
}
ALG_BYTES_PER_CELL = {"k_scan_cells": 12}  # count read + clear 8, offset write 4
# Kernels that carry one of the step's compulsory state passes.  The list
# builds (k_ns, k_final_list) and the final-cell regrouping (k_cur_sort) move
# no compulsory state -- lists are an implementation choice -- so the
# roofline's dominant kernel is taken among these passes; every kernel's time
# is still reported under "kernels".
STATE_PASSES = ("k_predict_hash", "k_scan_cells", "k_scatter", "k_reorder", "k_lambda", "k_deltap", "k_vort",
                "k_confine", "k_xsph_out", "k_out")


def step_bytes(n: int, cells: int, iters: int) -> float:
    """252 + 44*I bytes per particle + 12 bytes per cell."""
    return (252 + 44 * iters) * n + 12 * cells


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", type=int, default=1, choices=[0, 1, 2])
    ap.add_argument("--steps-8m", type=int, default=200, help="timed steps of the 8M workload (0: skip)")
    ap.add_argument("--kernel-steps", type=int, default=20, help="steps of the per-kernel event pass")
    ap.add_argument("--batch-steps", type=int, default=10, help="timed steps of the 256-scene batch (0: skip)")
    ap.add_argument("--adjoint-runs", type=int, default=2,
                    help="timed forward+backward runs of the differentiable workload (0: skip)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the differentiable workload (for profiling it alone)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline budget (0: skip)")
    ap.add_argument("--tile-cap", type=int, default=None, help="LDS tile (particles); default: autotuned")
    ap.add_argument("--unit-width", type=int, default=16)
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="finish with a plain device copy of this many MiB (a known-size kernel for "
                         "scaling hardware byte counters)")
    return ap.parse_args()


def traffic_from_profiles(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` measured with rocprofv3 FETCH_SIZE and
    WRITE_SIZE counters scaled by a plain copy of known size
    (profiles/hbm_traffic_pbf_step_<workload>.txt, lines '<kernel> bytes_per_launch: X')."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_pbf_step_{workload}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"{kernel} bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_gpu(scene, steps, warmup, kernel_steps, tile_cap, unit_width):
    """Times `steps` plain steps.  Every step processes its own copy of the
    workload's input (the seeded lattice at rest), resident in HBM before the
    timed region, so each timed step is the configuration's step and the
    state never drifts into the degenerate regime the workload's parameters
    reach after a few steps (the paper s_corr with k = 0.1 at epsilon = 600
    makes the Jacobi sweep diverge, and 30 neighbour slots cannot register
    compression at 0.5h spacing)."""
    import torch

    from ssf_pbf_cuda_amd import pbfx

    n = scene.n
    solver = pbfx.Solver.for_scene(scene)
    stream = torch.cuda.current_stream()
    x0 = torch.from_numpy(scene.x).cuda()
    v0 = torch.from_numpy(scene.v).cuda()
    lo, hi, iters = scene.area_lo, scene.area_hi, scene.iters
    nbuf = max(steps, 1)
    xs = [x0.clone() for _ in range(nbuf)]
    vs = [v0.clone() for _ in range(nbuf)]
    tuned_by = "given"
    if tile_cap is None:
        log(f"{scene.name}: autotuning the LDS tile and unit width")
        tile_cap, unit_width = solver.autotune(x0, v0, n, lo, hi, iters, xs[0], vs[0], stream=stream)
        log(f"{scene.name}: autotuned tile_cap={tile_cap} unit_width={unit_width}")
        # The autotuner times steps on a freshly copied (cache-warm) input, as
        # chained steps see it; the timed steps below each start from their
        # own resident copy.  Confirm the choice in that regime, untimed.
        cands = [(tile_cap, unit_width)] + [c for c in ((2048, 16), (2048, 32), (4096, 16), (4096, 32))
                                            if c != (tile_cap, unit_width)]
        best, k = None, min(30, nbuf)  # 30 x 24 MB of inputs: most of them out of the 256 MB cache
        for tc, uw in cands:
            solver.set_tuning(tc, uw)
            for i in range(k):
                xs[i].copy_(x0)
                vs[i].copy_(v0)
            torch.cuda.synchronize()
            c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            c0.record(stream)
            for i in range(k):
                solver.step_async(xs[i], vs[i], n, lo, hi, iters, stream)
            c1.record(stream)
            torch.cuda.synchronize()
            t = c0.elapsed_time(c1) / k
            if best is None or t < best[0]:
                best = (t, tc, uw)
        tile_cap, unit_width = best[1], best[2]
        for i in range(k):
            xs[i].copy_(x0)
            vs[i].copy_(v0)
        tuned_by = "autotuner, confirmed on per-step copies"
        log(f"{scene.name}: confirmed tile_cap={tile_cap} unit_width={unit_width} ({best[0]:.3f} ms/step)")
    solver.set_tuning(tile_cap, unit_width)
    for k in range(warmup):
        solver.step_async(xs[k % nbuf], vs[k % nbuf], n, lo, hi, iters, stream)
    for k in range(min(warmup, nbuf)):
        xs[k].copy_(x0)
        vs[k].copy_(v0)
    torch.cuda.synchronize()
    log(f"{scene.name}: warmup done, timing {steps} steps")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for k in range(steps):
        solver.step_async(xs[k], vs[k], n, lo, hi, iters, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1)
    live = solver.num(stream)
    stats = solver.stats(stream)
    assert live == n, f"particles were deleted ({n} -> {live})"
    log(f"{scene.name}: {ms / max(steps, 1):.4f} ms/step; per-kernel pass ({kernel_steps} steps)")
    # per-kernel durations: the same step again, with HIP events around every
    # launch on the solver's stream
    solver.set_kernel_timing(True)
    kt = {}
    for k in range(kernel_steps):
        xs[k % nbuf].copy_(x0)
        vs[k % nbuf].copy_(v0)
        solver.step_async(xs[k % nbuf], vs[k % nbuf], n, lo, hi, iters, stream)
        kt = solver.kernel_times(stream)
    solver.set_kernel_timing(False)
    solver.close()
    del xs, vs
    torch.cuda.empty_cache()
    return {"ms_total": ms, "wall_s": wall, "steps": steps, "kernels": kt, "stats": stats,
            "tuning": {"tile_cap": tile_cap, "unit_width": unit_width, "by": tuned_by}}


def run_chained(scene, steps, tile_cap, unit_width, block=100):
    """Times `steps` CHAINED steps from the workload's input (the running
    simulation, as the reference's application times frames,
    main.cpp:339-348): each step continues from the previous one's output.
    The run is timed with HIP events at block boundaries only; being
    deterministic, it is then repeated with per-launch kernel timing for the
    per-kernel totals.  Returns ms per step over all steps and per block, the
    per-kernel totals and the statistics of the last step of each block."""
    import torch

    from ssf_pbf_cuda_amd import pbfx

    n = scene.n
    solver = pbfx.Solver.for_scene(scene)
    solver.set_tuning(tile_cap, unit_width)
    stream = torch.cuda.current_stream()
    x0 = torch.from_numpy(scene.x).cuda()
    v0 = torch.from_numpy(scene.v).cuda()
    lo, hi, iters = scene.area_lo, scene.area_hi, scene.iters
    x, v = x0.clone(), v0.clone()
    # warm-up on a scratch copy: the timed run starts from the input itself
    xw, vw = x0.clone(), v0.clone()
    for _ in range(3):
        solver.step_async(xw, vw, n, lo, hi, iters, stream)
    del xw, vw
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps // block + 2)]
    log(f"{scene.name}: timing {steps} chained steps")
    t0 = time.perf_counter()
    ev[0].record(stream)
    marks = []
    for k in range(steps):
        solver.step_async(x, v, n, lo, hi, iters, stream)
        if (k + 1) % block == 0 or k + 1 == steps:
            ev[len(marks) + 1].record(stream)
            marks.append(k + 1)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = ev[0].elapsed_time(ev[len(marks)])
    blocks = []
    prev = 0
    for i, m in enumerate(marks):
        blocks.append({"steps": f"{prev}-{m}", "ms_per_step": round(ev[i].elapsed_time(ev[i + 1]) / (m - prev), 4)})
        prev = m
    live = solver.num(stream)
    assert live == n, f"particles were deleted ({n} -> {live})"
    finite = bool(torch.isfinite(x).all() and torch.isfinite(v).all())
    # the same run again, each launch timed, with the statistics at block ends
    log(f"{scene.name}: chained per-kernel pass")
    x.copy_(x0)
    v.copy_(v0)
    solver.set_kernel_timing(True)
    totals = {}
    for k in range(steps):
        solver.step_async(x, v, n, lo, hi, iters, stream)
        totals = solver.kernel_times(stream)  # running totals since timing was enabled
        if (k + 1) % block == 0 or k + 1 == steps:
            st = solver.stats(stream)
            blk = next(b for b in blocks if b["steps"].endswith(f"-{k + 1}"))
            blk.update({key: st[key] for key in ("max_tile", "list_overflow", "slow_units", "slow_particles",
                                                 "truncated", "sweep_cap_hits", "tile_overflow", "scan_tiles")})
    solver.set_kernel_timing(False)
    solver.close()
    torch.cuda.empty_cache()
    return {"ms_total": ms_total, "wall_s": wall, "steps": steps, "blocks": blocks, "finite": finite,
            "kernels": totals}


def run_batch(scs, steps, warmup):
    """The batched form against a per-scene loop, on the same GPU and input.

    batched: one pbfx_batch_step_async for all scenes, captured in a hipGraph
    together with the restore of the input (each timed step starts from the
    workload's input, as in run_gpu).  loop: one plain context per scene and
    one pbfx_step_async per scene, the same restore first.  Returns ms per
    step of each and whether one step of each gives identical bytes."""
    import torch

    from ssf_pbf_cuda_amd import pbfx

    counts = [s.n for s in scs]
    n = sum(counts)
    iters = scs[0].iters
    x0 = torch.from_numpy(np.concatenate([s.x for s in scs])).cuda()
    v0 = torch.from_numpy(np.concatenate([s.v for s in scs])).cuda()
    off = torch.tensor(np.concatenate([[0], np.cumsum(counts)]).astype(np.int32), device="cuda")
    x, v = x0.clone(), v0.clone()
    stream = torch.cuda.Stream()
    batch = pbfx.BatchSolver(scs)
    tune = batch.autotune(x0, v0, n, (0, 0, 0), (0, 0, 0), iters, x, v, in_offsets=off, stream=stream)
    log(f"batch: autotuned tile_cap={tune[0]} unit_width={tune[1]}")
    with torch.cuda.stream(stream):
        for _ in range(max(warmup, 1)):
            x.copy_(x0)
            v.copy_(v0)
            batch.batch_step_async(x, v, off, n, iters, stream=stream)
    stream.synchronize()
    xb, vb = x.clone(), v.clone()  # one batched step from the input
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        x.copy_(x0)
        v.copy_(v0)
        batch.batch_step_async(x, v, off, n, iters, stream=stream)
    torch.cuda.synchronize()
    log(f"batch of {len(scs)} scenes ({n} particles): graph captured, timing {steps} steps")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        g.replay()
    e1.record()
    torch.cuda.synchronize()
    ms_batch = e0.elapsed_time(e1) / steps
    batch.close()
    del g
    log(f"batch: {ms_batch:.3f} ms/step; per-scene loop")
    solvers = [pbfx.Solver.for_scene(sc) for sc in scs]
    offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    cur = torch.cuda.current_stream()

    def loop_step():
        x.copy_(x0)
        v.copy_(v0)
        for sv, sc, o, c in zip(solvers, scs, offs[:-1], counts):
            sv.step_async(x[o:o + c], v[o:o + c], c, sc.area_lo, sc.area_hi, iters, cur)

    for _ in range(max(warmup, 1)):
        loop_step()
    torch.cuda.synchronize()
    same = bool(torch.equal(x.view(torch.int32), xb.view(torch.int32)) and
                torch.equal(v.view(torch.int32), vb.view(torch.int32)))
    e0.record()
    for _ in range(steps):
        loop_step()
    e1.record()
    torch.cuda.synchronize()
    ms_loop = e0.elapsed_time(e1) / steps
    for sv in solvers:
        sv.close()
    log(f"loop: {ms_loop:.3f} ms/step")
    del x, v, x0, v0, xb, vb
    torch.cuda.empty_cache()
    return {"ms_batch": ms_batch, "ms_loop": ms_loop, "bit_exact": same, "particles": n,
            "tuning": {"tile_cap": tune[0], "unit_width": tune[1]}}


def run_adjoint(runs, steps=40, every=10):
    """The differentiable workload: 65,536 particles (64x32x32, h = 0.1,
    v0 ~ N(0, 0.1^2)), `steps` steps with a checkpoint every `every`; loss =
    mean squared distance by particle to the run with XSPH c = 0.02 and
    vorticity eps = 0.0008; gradients with respect to x0, v0, c and eps.
    Times pbfx_adj_forward and pbfx_adj_backward (window recomputation from
    the checkpoints included)."""
    import dataclasses

    import torch

    from ssf_pbf_cuda_amd import pbfx, scenes

    sc = scenes.config4()
    tgt = dataclasses.replace(sc, params=dataclasses.replace(sc.params, xsph_parameter=0.02, vc_parameter=0.0008))
    x0 = torch.from_numpy(sc.x).cuda()
    v0 = torch.from_numpy(sc.v).cuda()

    def forward(solver, x, v):
        n = solver.adj_forward(x, v, sc.n, sc.area_lo, sc.area_hi, sc.iters, steps)
        ident = torch.empty(n, dtype=torch.int32, device="cuda")
        solver.adj_identity(ident)
        return n, ident.long()

    st = pbfx.Solver.for_scene(tgt)
    st.adj_enable(steps, every)
    xt, vt = x0.clone(), v0.clone()
    nt, it = forward(st, xt, vt)
    target = torch.empty_like(x0)
    target[it] = xt[:nt]
    st.close()
    solver = pbfx.Solver.for_scene(sc)
    solver.adj_enable(steps, every)
    gx0 = torch.empty(x0.shape, dtype=torch.float64, device="cuda")
    gv0 = torch.empty_like(gx0)
    t_f, t_b, res = [], [], None
    for r in range(runs + 1):  # the first run warms up
        x, v = x0.clone(), v0.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n, ident = forward(solver, x, v)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        d = (x[:n].double() - target[ident].double())
        loss = float((d ** 2).sum() / sc.n)
        gx = (2.0 / sc.n) * d
        gv = torch.zeros_like(gx)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        gc, ge, e = solver.adj_backward(gx, gv, gx0, gv0)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        if r:
            t_f.append(t1 - t0)
            t_b.append(t3 - t2)
        res = (loss, gc, ge, e)
        log(f"adjoint run {r}: forward {1e3 * (t1 - t0):.1f} ms, backward {1e3 * (t3 - t2):.1f} ms")
    solver.close()
    tf, tb = min(t_f), min(t_b)
    loss, gc, ge, e = res
    return {
        "workload": f"{sc.name}: {sc.n} particles, {steps} steps, checkpoints every {every}, {sc.iters} "
                    f"iterations; loss = mean |x - x_target|^2 (target: c = 0.02, eps = 0.0008); gradients "
                    f"w.r.t. x0, v0, c, eps (FP64 adjoint at the FP32 forward's values)",
        "value": round(sc.n * steps / (tf + tb) / 1e6, 3),
        "unit": "M particle-steps/s (forward + backward)",
        "forward_ms": round(1e3 * tf, 2),
        "backward_ms": round(1e3 * tb, 2),
        "backward_over_forward": round(tb / tf, 2),
        "loss": loss,
        "grad_exponent": e,
        "dL_dc": f"{gc:.6e} * 2^{e}",
        "dL_deps": f"{ge:.6e} * 2^{e}",
        "runs": runs,
    }


def cpu_adjoint(steps):
    """The oracle's FP64 adjoint (host C++, OpenMP) on the same workload,
    over its first `steps` steps: FP32 forward with tape + FP64 backward."""
    from oracle import pyoracle as orc
    from ssf_pbf_cuda_amd import scenes

    sc = scenes.config4()
    T = orc.Tape(sc)
    t0 = time.perf_counter()
    x, _ = T.run(steps)
    gx = np.zeros((x.shape[0], 3), np.float64)
    gx[:, 0] = 1.0 / sc.n
    T.backward(gx, np.zeros_like(gx))
    el = time.perf_counter() - t0
    T.close()
    threads = orc.get_threads()
    return {"value": round(sc.n * steps / el / 1e6, 4), "unit": "M particle-steps/s (forward + backward)",
            "cores": threads, "kind": "port",
            "sample": f"{steps} steps of {sc.name} forward + FP64 adjoint, {el:.1f} s, OpenMP threads={threads}"}


def kernel_report(kt, n, cells):
    rows = {}
    for k, (tot, cnt) in kt.items():
        avg_ms = tot / cnt
        b = ALG_BYTES.get(k, 0) * n + ALG_BYTES_PER_CELL.get(k, 0) * cells
        rows[k] = {"avg_ms": round(avg_ms, 5), "launches": cnt,
                   "alg_GBps": round(b / (avg_ms * 1e-3) / 1e9, 1) if b else None}
    return rows


def cpu_baseline(scene, budget_s):
    """The oracle (host C++ restatement of the reference step, OpenMP over
    particles) on this host's cores: whole steps of the same workload until
    the time budget is spent."""
    from oracle import pyoracle as orc

    threads = orc.get_threads()
    done, t0 = 0, time.perf_counter()
    while True:
        orc.step(scene)
        done += 1
        el = time.perf_counter() - t0
        if el >= budget_s or done >= 1000:
            break
    return {"value": round(scene.n * done / el / 1e6, 4), "unit": "M particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{done} full steps of {scene.name} ({scene.n} particles, {scene.iters} iterations), "
                      f"each from the workload's input, {el:.1f} s, OpenMP threads={threads}, "
                      f"cpu={os.cpu_count()} logical"}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import torch

    from ssf_pbf_cuda_amd import pbfx, scenes

    assert torch.cuda.is_available(), "bench.py needs a ROCm device"
    pbfx.lib()
    if a.adjoint_only:
        d = run_adjoint(max(a.adjoint_runs, 1))
        print(json.dumps({"metric": "M particle-steps/s (forward + backward), differentiable workload",
                          "value": d["value"], "unit": d["unit"], "n_gpus": 1, "higher_is_better": True,
                          "dtype": "f32 forward, f64 adjoint", "data": "synthetic",
                          "config": {"workload": d["workload"]}, "differentiable": d}), flush=True)
        return
    sc = scenes.CONFIGS[a.config]()
    r = run_gpu(sc, a.steps, a.warmup, a.kernel_steps, a.tile_cap, a.unit_width)
    n, C, I = sc.n, sc.cells, sc.iters
    ms_step = r["ms_total"] / a.steps
    value = n * a.steps / (r["ms_total"] * 1e-3) / 1e6
    kern = kernel_report(r["kernels"], n, C)
    passes = {k: t for k, t in r["kernels"].items() if k in STATE_PASSES}
    dom = max(passes.items(), key=lambda kv: kv[1][0])[0] if passes else None
    dom_avg_ms = r["kernels"][dom][0] / r["kernels"][dom][1]
    dom_bytes = ALG_BYTES.get(dom, 0) * n + ALG_BYTES_PER_CELL.get(dom, 0) * C
    achieved = dom_bytes / (dom_avg_ms * 1e-3) / 1e9
    traffic = traffic_from_profiles(dom, sc.name)
    sb = step_bytes(n, C, I)
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "M particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered lattice at rest; each timed step runs on its own resident copy)",
        "config": {
            "workload": f"{sc.name}: {n} particles, grid {sc.grid[0]}x{sc.grid[1]}x{sc.grid[2]} cells, "
                        f"{I} Jacobi iterations, paper s_corr k=0.1 n=4 dq=0.2h, XSPH c=0.01, "
                        f"vorticity eps=0.0005, 30 neighbour slots",
            "particles": n,
            "cells": C,
            "iterations": I,
            "tuning": r["tuning"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "alg_bytes_per_launch": dom_bytes,
            "avg_launch_ms": round(dom_avg_ms, 5),
            "share_of_step": round(r["kernels"][dom][0] / sum(t[0] for t in r["kernels"].values()), 4),
        },
        "step_roofline": {
            "alg_bytes_per_step": sb,
            "achieved_GBps": round(sb / (ms_step * 1e-3) / 1e9, 1),
            "frac": round(sb / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        },
        "kernels": kern,
        "state": {k: r["stats"][k] for k in ("units", "slow_units", "slow_particles", "max_tile", "truncated",
                                             "sweep_cap_hits", "tile_overflow")},
    }
    if a.steps_8m and a.config != 2:
        s8 = scenes.config2()
        r8 = run_gpu(s8, a.steps_8m, min(a.warmup, 5), min(a.kernel_steps, 5), a.tile_cap, a.unit_width)
        ms8 = r8["ms_total"] / a.steps_8m
        sb8 = step_bytes(s8.n, s8.cells, s8.iters)
        out["double_dam_8M"] = {
            "value": round(s8.n * a.steps_8m / (r8["ms_total"] * 1e-3) / 1e6, 3),
            "unit": "M particle-steps/s",
            "steps": a.steps_8m,
            "ms_per_step": round(ms8, 5),
            "iterations": s8.iters,
            "step_roofline_frac": round(sb8 / (ms8 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "kernels": kernel_report(r8["kernels"], s8.n, s8.cells),
            "state": {k: r8["stats"][k] for k in ("units", "slow_units", "max_tile", "truncated")},
        }
    if a.batch_steps:
        scs = scenes.batch_config3()
        rb = run_batch(scs, a.batch_steps, min(a.warmup, 2))
        out["batched"] = {
            "workload": f"{len(scs)} dam-break scenes, {rb['particles']} particles (4,096-200,000 each, "
                        f"log-uniform), per-scene rest density / XSPH / vorticity, 4 iterations",
            "value": round(rb["particles"] / (rb["ms_batch"] * 1e-3) / 1e6, 3),
            "unit": "M particle-steps/s",
            "ms_per_step": round(rb["ms_batch"], 4),
            "loop_ms_per_step": round(rb["ms_loop"], 4),
            "speedup_vs_loop": round(rb["ms_loop"] / rb["ms_batch"], 3),
            "bit_exact_vs_loop": rb["bit_exact"],
            "tuning": rb["tuning"],
            "steps": a.batch_steps,
        }
    if a.adjoint_runs:
        out["differentiable"] = run_adjoint(a.adjoint_runs)
    if a.cpu_seconds > 0:
        log(f"cpu baseline ({a.cpu_seconds:.0f} s budget)")
        out["cpu_baseline"] = cpu_baseline(sc, a.cpu_seconds)
        if a.adjoint_runs:
            out["differentiable"]["cpu_baseline"] = cpu_adjoint(10)
    if a.calibrate_copy_mb:
        src = torch.ones(a.calibrate_copy_mb * (1 << 18), dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        out["calibration_copy_bytes"] = src.numel() * 4
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
