"""bench.py's byte accounting and roofline choice (host logic, no GPU):
the per-kernel algorithmic bytes of one step add up to the compulsory
traffic per particle-step, 252 + 44*I bytes plus the cell bytes actually
moved; the roofline kernel is the one with the largest total time over all
kernels; more than one GPU is refused."""
from __future__ import annotations

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


@pytest.mark.parametrize("iters", [1, 3, 4])
def test_kernel_bytes_add_up_to_the_step(iters):
    n, cells, tiles = 1 << 20, 1 << 22, 128
    # one step's launches: the bookkeeping passes once, lambda and delta-p
    # once per Jacobi iteration, the final passes once
    once = ["k_predict_hash", "k_scatter", "k_reorder", "k_cur_sort", "k_vort", "k_confine", "k_xsph_out",
            "k_scan_cells"]
    total = sum(bench.kernel_alg_bytes(k, n, cells, tiles) for k in once)
    total += iters * (bench.kernel_alg_bytes("k_lambda", n, cells, tiles) +
                      bench.kernel_alg_bytes("k_deltap", n, cells, tiles))
    assert total == bench.step_bytes(n, cells, iters, tiles)
    assert bench.step_bytes(n, cells, iters, tiles) == (252 + 44 * iters) * n + 4 * cells + 8 * 4096 * tiles


def test_without_final_passes_the_write_back_carries_the_velocity():
    # k_out replaces k_cur_sort + k_vort + k_confine + k_xsph_out when neither
    # XSPH nor vorticity runs: the velocity update's 36 bytes
    assert bench.ALG_BYTES["k_out"] == bench.ALG_BYTES["k_cur_sort"]


def test_roofline_takes_the_largest_kernel_over_all_kernels():
    kt = {"k_deltap": (4 * 0.070, 4), "k_final_list": (0.300, 1), "k_ns": (0.220, 1), "k_lambda": (4 * 0.05, 4)}
    r = bench.roofline(kt, 1 << 20, 1 << 22, 128, "no_such_workload")
    assert r["kernel"] == "k_final_list"
    # the list builds carry no compulsory state: only their x* read counts
    assert r["alg_bytes_per_launch"] == 12 * (1 << 20)
    assert r["traffic"] is None and "valu" not in r
    assert r["share_of_step"] == pytest.approx(0.300 / sum(t for t, _ in kt.values()), abs=1e-4)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stdout


def test_roofline_reports_valu_issue_from_the_counter_profile():
    # the 1M dam's counter pass: k_deltap issues ~4.15e7 wave64 VALU
    # instructions per launch at ~2.4 GHz; at 75 us a launch that is the
    # issue ceiling of 1024 SIMDs at one instruction per 4 cycles
    insts, mhz = bench.valu_from_profiles("k_deltap", "dam1M_h0.05")
    assert insts is not None and 3e7 < insts < 6e7
    assert 1500 < mhz < 3000
    kt = {"k_deltap": (4 * 0.0755, 4), "k_ns": (0.2, 1)}
    r = bench.roofline(kt, 1 << 20, 1 << 22, 128, "dam1M_h0.05")
    assert r["kernel"] == "k_deltap"
    want = insts / 75.5e-6 / (bench.VALU_SIMDS * mhz * 1e6 / bench.VALU_CYCLES_PER_INST)
    assert r["valu"]["issue_frac"] == pytest.approx(want, abs=1e-3)
    assert 0.8 < r["valu"]["issue_frac"] < 1.2
