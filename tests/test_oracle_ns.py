"""Pins the oracle's neighbour search: stable counting sort against numpy's
stable argsort, cell ranges against bincount, neighbour sets against brute
force, and the reference's capped z,y,x order against an independent
enumeration (pbf_neighbor_search_util.cu:361-482)."""
import dataclasses

import numpy as np
import pytest

from oracle import pyoracle as orc
from ssf_pbf_cuda_amd import scenes

f32 = np.float32


def cloud(n=600, box=0.8, h=0.1, seed=5, max_pair=128, **kw):
    """Random particles in [0, box]^3 (denser than the lattice, so lists
    truncate when capped)."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, box, (n, 3)).astype(np.float32)
    sc = scenes.small_dam(2, 2, 2, h=h, box=(box, box, box), **kw)
    sc = dataclasses.replace(sc, x=x, v=np.zeros_like(x))
    p = dataclasses.replace(sc.params, max_pair=max_pair, body_force=(0.0, 0.0, 0.0))
    return dataclasses.replace(sc, params=p)


def run_ns(sc):
    return orc.step(sc, want_ns=True, boundary=False)


def predicted(sc):
    """x* = fmaf(fmaf(g, dt, v), dt, x) in float64-exact emulation (g = 0: x* = x)."""
    assert tuple(sc.params.body_force) == (0.0, 0.0, 0.0)
    return np.asarray(sc.x, np.float32)


def test_stable_sort_and_cell_offsets():
    sc = cloud()
    r = run_ns(sc)
    key = r["key"]
    np.testing.assert_array_equal(key, orc.hash_keys(predicted(sc), sc.params.smoothing_length, sc.grid))
    np.testing.assert_array_equal(r["sorted_index"], np.argsort(key, kind="stable").astype(np.uint32))
    C = sc.cells
    counts = np.bincount(key, minlength=C + 1)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    np.testing.assert_array_equal(r["cell_offset"], off)
    # every slot of cell c hashes to c
    sk = key[r["sorted_index"]]
    assert np.all(np.diff(sk.astype(np.int64)) >= 0)


def test_neighbor_sets_match_brute_force():
    """Uncapped lists hold exactly the pairs with r^2 < h^2 (self included)."""
    sc = cloud(box=0.5, max_pair=128)
    r = run_ns(sc)
    P = r["pred0"].astype(np.float64)
    h2 = float(f32(sc.params.smoothing_length) * f32(sc.params.smoothing_length))
    d = P[:, None, :] - P[None, :, :]
    r2 = (d ** 2).sum(-1)
    n = P.shape[0]
    assert r["neighbor_count"].max() < 127
    checked = 0
    for s in range(n):
        got = set(r["neighbors"][s, : r["neighbor_count"][s]].tolist())
        inside = set(np.nonzero(r2[s] < h2 * (1 - 1e-6))[0].tolist())
        maybe = set(np.nonzero(np.abs(r2[s] - h2) <= h2 * 1e-6)[0].tolist())
        assert inside <= got <= inside | maybe, s
        assert s in got
        checked += len(got)
    assert checked > 10 * n


def fma32(a, b, c):
    """fmaf emulated in float64 (the product of two floats is exact in double)."""
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(np.float32)


def ordered_candidates(P, s, h, grid, wrapped):
    """Slots j in the reference's visiting order for particle s: cells
    (dz, dy, dx) in z-outer, x-inner order around the cell of P[s] (wrapped),
    slots ascending within a cell -- enumerated independently by scanning all
    particles once per visited cell."""
    G = np.array(grid)
    c0 = np.floor(P[s].astype(np.float32) / f32(h)).astype(np.int64)
    out = []
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                w = (c0 + np.array([dx, dy, dz])) & (G - 1)
                out.append(np.nonzero(np.all(wrapped == w, axis=1))[0])
    return np.concatenate(out)


def capped_reference_lists(P, h, grid, cap):
    G = np.array(grid)
    wrapped = np.floor(P.astype(np.float32) / f32(h)).astype(np.int64) & (G - 1)
    h2 = f32(h) * f32(h)
    out = []
    for s in range(P.shape[0]):
        js = ordered_candidates(P, s, h, grid, wrapped)
        d = (P[s][None, :] - P[js]).astype(np.float32)
        r2 = fma32(d[:, 2], d[:, 2], fma32(d[:, 1], d[:, 1], (d[:, 0] * d[:, 0]).astype(np.float32)))
        out.append(js[r2 < h2][:cap].tolist())
    return out


def test_capped_lists_follow_reference_order():
    """With the reference's 30 slots the first 29 in-range candidates in
    z,y,x cell order (ascending slots) are kept."""
    sc = cloud(n=1500, box=0.5, max_pair=30)
    r = run_ns(sc)
    P = r["pred0"]
    ref = capped_reference_lists(P, sc.params.smoothing_length, sc.grid, 29)
    assert r["stats"]["truncated"] > 100
    for s in range(P.shape[0]):
        got = r["neighbors"][s, : r["neighbor_count"][s]].tolist()
        assert got == ref[s], s


def test_order_checker_detects_a_transposed_traversal():
    """Plants a mistake in a copy of the oracle's lists (x-outer instead of
    z-outer cell order) and checks that the order test would catch it."""
    sc = cloud(n=1500, box=0.5, max_pair=30)
    r = run_ns(sc)
    P = r["pred0"]
    ref = capped_reference_lists(P, sc.params.smoothing_length, sc.grid, 29)
    G = np.array(sc.grid)
    cells = np.floor(P / f32(sc.params.smoothing_length)).astype(np.int64) & (G - 1)
    mismatches = 0
    for s in range(P.shape[0]):
        full = r["neighbors"][s, : r["neighbor_count"][s]].tolist()
        wrong = sorted(full, key=lambda j: (cells[j][0], cells[j][1], cells[j][2], j))
        mismatches += wrong != ref[s]
    assert mismatches > 0


def test_deletion_sorts_last_and_shrinks_num():
    sc = cloud(n=300)
    sc = dataclasses.replace(sc, area_lo=np.zeros(3, np.float32), area_hi=np.array([0.4, 1.0, 1.0], np.float32))
    r = run_ns(sc)
    out = sc.x[:, 0] > f32(0.4)
    assert r["num"] == int((~out).sum())
    C = sc.cells
    assert np.all(r["key"][out] == C)
    tail = r["sorted_index"][r["num"]:]
    assert set(tail.tolist()) == set(np.nonzero(out)[0].tolist())
    assert r["cell_offset"][C] == r["num"] and r["cell_offset"][C + 1] == sc.n


def test_lattice_neighbor_counts():
    """An interior particle of the 0.5h lattice has 27 neighbours within h;
    jitter brings up to 6 more (the next shell sits at exactly h)."""
    sc = scenes.dam_break("lattice", 0.1, (1.6, 1.6, 1.6), [scenes.corner_block(0.1, (12, 12, 12))], iters=1,
                          jitter=0.0)
    sc = dataclasses.replace(sc, params=dataclasses.replace(sc.params, max_pair=64, body_force=(0.0, 0.0, 0.0)))
    r = orc.step(sc, want_ns=True, boundary=False)
    P = r["pred0"]
    s = sc.params.stable_distance
    interior = np.all((P > 2.5 * s) & (P < 9.5 * s), axis=1)
    cnt = r["neighbor_count"][interior]
    # shells at 0, s, sqrt2 s, sqrt3 s are inside h = 2s; the 6 points at
    # exactly h are kept or not by FP32 rounding of the lattice coordinates
    assert interior.sum() > 200
    assert cnt.min() >= 27 and cnt.max() <= 33
