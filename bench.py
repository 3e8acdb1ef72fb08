#!/usr/bin/env python3
"""Throughput of the PBF solver step on one MI355X.

A step is one full pbf::one_step over the whole particle set: predict,
hashed-grid neighbour search, I Jacobi iterations, velocity update,
vorticity confinement and XSPH.  Default workload: the 1,048,576-particle
dam break (128x128x64 lattice, spacing 0.5h, h = 0.05, 256x128x128 cells, 4
Jacobi iterations), each timed step from the workload's input; the
8,388,608-particle double dam break (3 iterations), chained runs (the 32k
dam's 500 steps, the 1M and 8M dams' first steps), the 256-scene batch over
chained steps against a per-scene loop, and the differentiable
65,536-particle workload (40 steps forward + backward) are reported beside
it.  Inputs are resident in HBM before the timed region.

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

# Algorithmic bytes per particle of each kernel: the compulsory FP32 state
# traffic of the step's passes (each logical array read once and written once
# per pass, neighbour gathers counted once), assigned to the kernel that
# carries the pass.  Their sum over a step is 252 + 44*I bytes per particle.
ALG_BYTES = {
    "k_predict_hash": 48,   # x 12, v 12 -> x* 12, hash 4 (predict + hash); histogram: hash 4 -> rank 4
    "k_scatter": 16,        # hash 4, rank 4 -> sorted index 4, sorted hash 4
    "k_reorder": 56,        # index 4, x* 12, x 12 -> x*s 12, xs 12 ; cell ranges: sorted hash 4
    "k_lambda": 16,         # x* 12 -> lambda 4, once per Jacobi iteration
    "k_deltap": 28,         # x* 12, lambda 4 -> x*' 12, once per Jacobi iteration
    "k_cur_sort": 36,       # velocity update: x* 12, xs 12 -> v 12 (k_cur_hash packs it into the records)
    "k_vort": 12,           # omega 12 written (its reads are the records above)
    "k_confine": 48,        # x* 12, omega 12, v 12 -> v 12
    "k_xsph_out": 36,       # x* 12, v 12 -> v' 12
    "k_out": 36,            # (no final passes) x* 12, xs 12 -> v 12
}
STEP_BYTES_PER_PARTICLE = lambda iters: 252 + 44 * iters  # noqa: E731
# The list builds (k_ns, k_final_list) carry no pass of the step's compulsory
# traffic (lists are an implementation choice); for their per-kernel rate
# only the x* read they need is counted.
LIST_BYTES = {"k_ns": 12, "k_final_list": 12}
# Cells: the scan writes every cell's offset (4 B); the counts are read and
# cleared (8 B) only in the occupied 4096-cell tiles -- the others are skipped.
CELL_OFFSET_BYTES = 4
CELL_COUNT_BYTES = 8
SCAN_TILE = 4096


def cell_bytes(cells: int, scan_tiles: int) -> int:
    return CELL_OFFSET_BYTES * cells + CELL_COUNT_BYTES * SCAN_TILE * scan_tiles


def step_bytes(n: int, cells: int, iters: int, scan_tiles: int) -> float:
    """(252 + 44*I) bytes per particle + the cell bytes actually moved."""
    return STEP_BYTES_PER_PARTICLE(iters) * n + cell_bytes(cells, scan_tiles)


def kernel_alg_bytes(k: str, n: int, cells: int, scan_tiles: int) -> int:
    b = (ALG_BYTES.get(k, 0) + LIST_BYTES.get(k, 0)) * n
    if k == "k_scan_cells":
        b += cell_bytes(cells, scan_tiles)
    return b


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", type=int, default=1, choices=[0, 1, 2])
    ap.add_argument("--steps-8m", type=int, default=200, help="timed steps of the 8M workload (0: skip)")
    ap.add_argument("--kernel-steps", type=int, default=20, help="steps of the per-kernel event pass")
    ap.add_argument("--chain-steps-32k", type=int, default=500,
                    help="chained steps of the 32,768-particle dam from its input (0: skip)")
    ap.add_argument("--chain-steps-1m", type=int, default=100,
                    help="chained steps of the headline workload from its input (0: skip)")
    ap.add_argument("--chain-steps-8m", type=int, default=50,
                    help="chained steps of the 8M double dam from its input (0: skip)")
    ap.add_argument("--cpu-chain-32k", type=int, default=0,
                    help="chained steps of the 32k dam on the oracle (host CPU), for the CPU baseline (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=20,
                    help="chained steps of the 256-scene batch (0: skip)")
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


def valu_from_profiles(kernel: str, workload: str):
    """VALU instructions per launch of `kernel` (rocprofv3 SQ_INSTS_VALU,
    wave instructions summed over the dispatch) and the engine clock of the
    same launches (GRBM_GUI_ACTIVE per XCD over the kernel time), from
    profiles/valu_pbf_step_<workload>.txt (lines '<kernel> valu_insts_per_launch: X'
    and '<kernel> clock_mhz: Y')."""
    path = os.path.join(ROOT, "profiles", f"valu_pbf_step_{workload}.txt")
    if not os.path.exists(path):
        return None, None
    insts, mhz = None, None
    for line in open(path):
        if line.startswith(f"{kernel} valu_insts_per_launch:"):
            insts = float(line.split(":", 1)[1])
        elif line.startswith(f"{kernel} clock_mhz:"):
            mhz = float(line.split(":", 1)[1])
    return insts, mhz


# VALU issue peak: 256 CUs x 4 SIMDs, one wave64 VALU instruction per SIMD
# every 4 cycles (16 lanes per cycle).
VALU_SIMDS = 256 * 4
VALU_CYCLES_PER_INST = 4


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_gpu(scene, steps, warmup, kernel_steps, tile_cap, unit_width):
    """Times `steps` plain steps.  Every step processes its own copy of the
    workload's input (the seeded lattice at rest), resident in HBM before the
    timed region, so each timed step is the configuration's step.  Chained,
    the workload's parameters wreck the state: the paper s_corr's k = 0.1
    term is of the order of the scaling factors themselves (s ~ C/1190 at
    epsilon = 600), so the Jacobi sweep throws the particles out in the first
    steps (the neighbour-slot cap is not the cause: the oracle's first step
    gives the same mean speed with 64 slots as with 30), the dam collapses
    into piles and the 1M state holds NaN within ~300 steps; run_chained
    reports those runs beside this one."""
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
    """The batched form against a per-scene loop, on the same GPU, both over
    `steps` CHAINED steps from the workload's input.

    batched: one step recorded by pbfx_batch_graph_capture (all scenes, the
    output offsets written over the input offsets) and replayed `steps` times
    by pbfx_batch_graph_launch.  loop: one plain context per scene, one
    pbfx_step per scene per step.  Returns ms per step of each and whether
    the two end states are identical bit for bit, scene by scene."""
    import torch

    from ssf_pbf_cuda_amd import pbfx

    counts = [s.n for s in scs]
    n = sum(counts)
    iters = scs[0].iters
    x0 = torch.from_numpy(np.concatenate([s.x for s in scs])).cuda()
    v0 = torch.from_numpy(np.concatenate([s.v for s in scs])).cuda()
    off0 = torch.tensor(np.concatenate([[0], np.cumsum(counts)]).astype(np.int32), device="cuda")
    x, v, off = x0.clone(), v0.clone(), off0.clone()
    stream = torch.cuda.Stream()
    batch = pbfx.BatchSolver(scs)
    tune = batch.autotune(x0, v0, n, (0, 0, 0), (0, 0, 0), iters, x, v, in_offsets=off0, stream=stream)
    log(f"batch: autotuned tile_cap={tune[0]} unit_width={tune[1]}")
    torch.cuda.synchronize()
    batch.graph_capture(x, v, off, iters, 1, stream)

    def restore():
        with torch.cuda.stream(stream):
            x.copy_(x0)
            v.copy_(v0)
            off.copy_(off0)

    restore()
    batch.graph_launch(max(warmup, 1), stream)
    restore()
    stream.synchronize()
    log(f"batch of {len(scs)} scenes ({n} particles): graph recorded, timing {steps} chained steps")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    batch.graph_launch(steps, stream)
    e1.record(stream)
    stream.synchronize()
    ms_batch = e0.elapsed_time(e1) / steps
    ob = off.cpu().numpy().astype(np.int64)
    xb, vb = x.cpu().numpy(), v.cpu().numpy()
    batch.close()
    log(f"batch: {ms_batch:.3f} ms/step ({n - int(ob[-1])} particles deleted); per-scene loop")
    # The loop: one context per scene, stepped with pbfx_step (one_step's
    # semantics: the live count comes back each step, so a scene whose
    # particle left its area -- a velocity overflowing to infinity -- goes on
    # with the right count).  Each scene keeps its own slice.
    solvers = [pbfx.Solver.for_scene(sc) for sc in scs]
    for sv in solvers:
        sv.set_tuning(*tune)
    offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    cur = torch.cuda.current_stream()
    live = list(counts)

    def loop_steps(k):
        for _ in range(k):
            for i, (sv, sc, o) in enumerate(zip(solvers, scs, offs[:-1])):
                live[i] = sv.step(x[o:o + counts[i]], v[o:o + counts[i]], live[i], sc.area_lo, sc.area_hi, iters,
                                  stream=cur)

    x.copy_(x0)
    v.copy_(v0)
    loop_steps(max(warmup, 1))
    x.copy_(x0)
    v.copy_(v0)
    live[:] = counts
    torch.cuda.synchronize()
    e0.record()
    loop_steps(steps)
    e1.record()
    torch.cuda.synchronize()
    ms_loop = e0.elapsed_time(e1) / steps
    xl, vl = x.cpu().numpy(), v.cpu().numpy()
    same = True
    for i in range(len(scs)):
        nb_ = int(ob[i + 1] - ob[i])
        same = same and nb_ == live[i] and \
            np.array_equal(xb[ob[i]:ob[i] + nb_].view(np.uint32), xl[offs[i]:offs[i] + nb_].view(np.uint32)) and \
            np.array_equal(vb[ob[i]:ob[i] + nb_].view(np.uint32), vl[offs[i]:offs[i] + nb_].view(np.uint32))
    for sv in solvers:
        sv.close()
    log(f"loop: {ms_loop:.3f} ms/step; end states identical scene by scene: {same}")
    del x, v, x0, v0
    torch.cuda.empty_cache()
    return {"ms_batch": ms_batch, "ms_loop": ms_loop, "bit_exact": same, "particles": n,
            "deleted": n - int(ob[-1]), "tuning": {"tile_cap": tune[0], "unit_width": tune[1]}}


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


def kernel_report(kt, n, cells, scan_tiles, steps=1):
    """Per kernel: mean launch duration, launches per step and the algorithmic
    rate (its bytes per launch over its mean duration)."""
    rows = {}
    for k, (tot, cnt) in sorted(kt.items(), key=lambda kv: -kv[1][0]):
        avg_ms = tot / cnt
        b = kernel_alg_bytes(k, n, cells, scan_tiles)
        rows[k] = {"avg_ms": round(avg_ms, 5), "ms_per_step": round(tot / steps, 5),
                   "launches_per_step": round(cnt / steps, 2),
                   "alg_GBps": round(b / (avg_ms * 1e-3) / 1e9, 1) if b else None}
    return rows


def roofline(kt, n, cells, scan_tiles, workload):
    """The dominant kernel -- the largest total time over all kernels of the
    per-kernel pass -- priced against the HBM peak with its algorithmic bytes
    per launch, beside its measured HBM bytes per launch (rocprofv3 counters,
    profiles/hbm_traffic_pbf_step_<workload>.txt)."""
    dom = max(kt.items(), key=lambda kv: kv[1][0])[0]
    tot, cnt = kt[dom]
    avg_ms = tot / cnt
    b = kernel_alg_bytes(dom, n, cells, scan_tiles)
    achieved = b / (avg_ms * 1e-3) / 1e9
    out = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(dom, workload),
           "alg_bytes_per_launch": b, "avg_launch_ms": round(avg_ms, 5),
           "share_of_step": round(tot / sum(t for t, _ in kt.values()), 4)}
    insts, mhz = valu_from_profiles(dom, workload)
    if insts and mhz:
        # what the kernel actually spends its time on: VALU issue against the
        # issue rate of all SIMDs at the measured clock
        peak = VALU_SIMDS * mhz * 1e6 / VALU_CYCLES_PER_INST
        out["valu"] = {"insts_per_launch": insts, "clock_mhz": mhz,
                       "issue_frac": round(insts / (avg_ms * 1e-3) / peak, 4)}
    return out


def cpu_chained(scene, steps):
    """The oracle on `steps` CHAINED steps of `scene` from its input, on this
    host's cores (the reference's CPU-runnable case run as a simulation)."""
    from oracle import pyoracle as orc

    threads = orc.get_threads()
    x, v = scene.x, scene.v
    t0 = time.perf_counter()
    for _ in range(steps):
        r = orc.step(scene, x=x, v=v)
        x, v = r["x"], r["v"]
    el = time.perf_counter() - t0
    return {"value": round(scene.n * steps / el / 1e6, 4), "unit": "M particle-steps/s", "cores": threads,
            "kind": "port", "ms_per_step": round(1e3 * el / steps, 3),
            "sample": f"{steps} chained steps of {scene.name} ({scene.n} particles), {el:.1f} s, "
                      f"OpenMP threads={threads}, cpu={os.cpu_count()} logical"}


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


def chained_report(sc, rc):
    n, I = sc.n, sc.iters
    ms = rc["ms_total"] / rc["steps"]
    last = rc["blocks"][-1]
    return {
        "workload": f"{sc.name}: {n} particles, {I} iterations, chained from the workload's input",
        "steps": rc["steps"],
        "value": round(n * rc["steps"] / (rc["ms_total"] * 1e-3) / 1e6, 3),
        "unit": "M particle-steps/s",
        "ms_per_step": round(ms, 4),
        "finite_at_end": rc["finite"],
        "blocks": rc["blocks"],
        "kernels": kernel_report(rc["kernels"], n, sc.cells, last.get("scan_tiles", 0), rc["steps"]),
    }


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
    tiles = r["stats"]["scan_tiles"]
    sb = step_bytes(n, C, I, tiles)
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
        "data": "synthetic (seeded jittered lattice at rest; each timed step runs the workload's step on its own "
                "resident copy of the input; chained runs reported under 'chained')",
        "config": {
            "workload": f"{sc.name}: {n} particles, grid {sc.grid[0]}x{sc.grid[1]}x{sc.grid[2]} cells, "
                        f"{I} Jacobi iterations, paper s_corr k=0.1 n=4 dq=0.2h, XSPH c=0.01, "
                        f"vorticity eps=0.0005, 30 neighbour slots",
            "particles": n,
            "cells": C,
            "iterations": I,
            "tuning": r["tuning"],
        },
        "roofline": roofline(r["kernels"], n, C, tiles, sc.name),
        "step_roofline": {
            "alg_bytes_per_step": sb,
            "alg_bytes_per_particle": STEP_BYTES_PER_PARTICLE(I),
            "cell_bytes": cell_bytes(C, tiles),
            "occupied_scan_tiles": tiles,
            "achieved_GBps": round(sb / (ms_step * 1e-3) / 1e9, 1),
            "frac": round(sb / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        },
        "kernels": kernel_report(r["kernels"], n, C, tiles, a.kernel_steps),
        "state": {k: r["stats"][k] for k in ("units", "slow_units", "slow_particles", "max_tile", "truncated",
                                             "sweep_cap_hits", "list_overflow", "scan_tiles")},
    }
    tuned = (r["tuning"]["tile_cap"], r["tuning"]["unit_width"])
    if a.steps_8m and a.config != 2:
        s8 = scenes.config2()
        r8 = run_gpu(s8, a.steps_8m, min(a.warmup, 5), min(a.kernel_steps, 5), a.tile_cap, a.unit_width)
        ms8 = r8["ms_total"] / a.steps_8m
        t8 = r8["stats"]["scan_tiles"]
        sb8 = step_bytes(s8.n, s8.cells, s8.iters, t8)
        out["double_dam_8M"] = {
            "value": round(s8.n * a.steps_8m / (r8["ms_total"] * 1e-3) / 1e6, 3),
            "unit": "M particle-steps/s",
            "steps": a.steps_8m,
            "ms_per_step": round(ms8, 5),
            "iterations": s8.iters,
            "tuning": r8["tuning"],
            "roofline": roofline(r8["kernels"], s8.n, s8.cells, t8, s8.name),
            "step_roofline_frac": round(sb8 / (ms8 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "kernels": kernel_report(r8["kernels"], s8.n, s8.cells, t8, min(a.kernel_steps, 5)),
            "state": {k: r8["stats"][k] for k in ("units", "slow_units", "max_tile", "truncated", "list_overflow",
                                                  "scan_tiles")},
        }
    chained = {}
    if a.chain_steps_32k:
        s0 = scenes.config0()
        chained[s0.name] = chained_report(s0, run_chained(s0, a.chain_steps_32k, 4096, 32, block=100))
    if a.chain_steps_1m:
        chained[sc.name] = chained_report(sc, run_chained(sc, a.chain_steps_1m, *tuned, block=25))
    if a.chain_steps_8m:
        s8 = scenes.config2()
        chained[s8.name] = chained_report(s8, run_chained(s8, a.chain_steps_8m, 4096, 32, block=25))
    if chained:
        out["chained"] = chained
    if a.batch_steps:
        scs = scenes.batch_config3()
        rb = run_batch(scs, a.batch_steps, min(a.warmup, 2))
        out["batched"] = {
            "workload": f"{len(scs)} dam-break scenes, {rb['particles']} particles (4,096-200,000 each, "
                        f"log-uniform), per-scene rest density / XSPH / vorticity, 4 iterations, "
                        f"{a.batch_steps} chained steps",
            "value": round(rb["particles"] / (rb["ms_batch"] * 1e-3) / 1e6, 3),
            "unit": "M particle-steps/s",
            "ms_per_step": round(rb["ms_batch"], 4),
            "loop_ms_per_step": round(rb["ms_loop"], 4),
            "speedup_vs_loop": round(rb["ms_loop"] / rb["ms_batch"], 3),
            "bit_exact_vs_loop": rb["bit_exact"],
            "deleted": rb["deleted"],
            "loop": "one context per scene, pbfx_step per scene per step (one_step's semantics)",
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
    if a.cpu_chain_32k:
        log(f"cpu chained 32k dam ({a.cpu_chain_32k} steps)")
        out["cpu_chained_dam32k"] = cpu_chained(scenes.config0(), a.cpu_chain_32k)
    if a.calibrate_copy_mb:
        src = torch.ones(a.calibrate_copy_mb * (1 << 18), dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        out["calibration_copy_bytes"] = src.numel() * 4
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
