"""Piles that the wrapping hash puts into one cell, and a pile of identical
points.

The grid wraps (pbf_grid.h:54-61): a particle exactly on the far wall
x = 3.2 of a 32-cell-wide grid lies in column 32, which hashes to column 0,
the cell of the particles at x = 0.  Here one cell row holds two piles of
that kind, 3.2 apart in space, and another cell holds 500 particles at one
point.  Every such cell holds more than 64 particles, so the step groups it
by image and keeps each group's bounding box; the scans then skip the far
group, visit the near one through its own ascending slots, fill lists from a
group wholly in range without testing pairs, and decide a final list's
overflow from such groups.  The neighbour lists, the cell ranges and the
state after one step must equal the oracle's, which scans every candidate.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from ssf_pbf_cuda_amd import scenes

from helpers import bit_mismatches, gpu_run, oracle_run

pytestmark = pytest.mark.gpu


def aliased_scene(far=3000, near=2000, point=500, seed=11):
    h = 0.1
    base = scenes.small_dam(10, 8, 6, h=h, box=(3.2, 3.2, 3.2), iters=2, seed=seed)
    rng = np.random.default_rng(seed)
    yz = lambda k: rng.uniform(1.101, 1.299, (k, 2))  # cells 11..12 in y and z
    # on the far wall: column 32, hashed into column 0
    a = np.column_stack([np.full(far, 3.2), yz(far)])
    # in column 0, the same rows
    b = np.column_stack([rng.uniform(0.001, 0.099, near), yz(near)])
    # one point, and a shell of particles around it partly within h
    c = np.tile([1.55, 2.05, 1.55], (point, 1))
    d = np.array([1.55, 2.05, 1.55]) + rng.normal(0.0, 0.08, (400, 3))
    pile = np.concatenate([a, b, c, d]).astype(np.float32)
    rng.shuffle(pile)  # input order unrelated to the cells
    x = np.concatenate([base.x, pile]).astype(np.float32)
    return dataclasses.replace(base, name="aliased", x=x, v=np.zeros_like(x))


@pytest.mark.parametrize("tile_cap", [None, 0])
def test_aliased_piles_bit_exact(tile_cap):
    sc = aliased_scene()
    print(f"aliased piles: {sc.n} particles, tile_cap={tile_cap}", flush=True)
    got = gpu_run(sc, with_ns=True, tile_cap=tile_cap, unit_width=None if tile_cap is None else 16)
    st = got["stats"]
    print(f"gpu stats: {st}", flush=True)
    ref = oracle_run(sc, want_ns=True)
    off = got["cell_offset"]
    assert int(np.diff(off[:-1]).max()) > 64  # big cells exist
    np.testing.assert_array_equal(got["sorted_index"], ref["sorted_index"])
    np.testing.assert_array_equal(off, ref["cell_offset"])
    nl = got["num"]
    m = sc.params.max_pair
    np.testing.assert_array_equal(got["neighbor_count"], ref["neighbor_count"][:nl])
    np.testing.assert_array_equal(got["neighbors"], ref["neighbors"].reshape(-1, m)[:nl])
    assert got["num"] == ref["num"]
    bx, bv = bit_mismatches(got["x"], ref["x"]), bit_mismatches(got["v"], ref["v"])
    print(f"aliased step: bitdiff_x={bx} bitdiff_v={bv} list_overflow={st['list_overflow']}", flush=True)
    assert st["list_overflow"] > 0
    assert bx == 0 and bv == 0


def test_heavy_final_builds_bit_exact():
    """Particles beside the piles have search and final-list builds of
    thousands of candidates from global memory (every unit filtered from
    global memory here); those are handed to a wave each (k_ns_heavy,
    k_final_heavy).  Their lists must equal the oracle's, and their counts
    and overflow slots give the oracle's state bit for bit."""
    sc = aliased_scene(far=3000, near=2000, point=500, seed=15)
    print(f"heavy builds: {sc.n} particles", flush=True)
    got = gpu_run(sc, with_ns=True, tile_cap=0, unit_width=16)
    st = got["stats"]
    print(f"gpu stats: {st}", flush=True)
    ref = oracle_run(sc, want_ns=True)
    assert st["heavy_builds"] > 0 and st["heavy_search"] > 0
    nl = got["num"]
    m = sc.params.max_pair
    np.testing.assert_array_equal(got["neighbor_count"], ref["neighbor_count"][:nl])
    np.testing.assert_array_equal(got["neighbors"], ref["neighbors"].reshape(-1, m)[:nl])
    assert got["num"] == ref["num"]
    bx, bv = bit_mismatches(got["x"], ref["x"]), bit_mismatches(got["v"], ref["v"])
    print(f"heavy builds step: bitdiff_x={bx} bitdiff_v={bv}", flush=True)
    assert bx == 0 and bv == 0


def test_aliased_piles_two_steps_bit_exact():
    sc = aliased_scene(far=1500, near=900, point=300, seed=12)
    print(f"aliased piles, 2 steps: {sc.n} particles", flush=True)
    got = gpu_run(sc, steps=2)
    ref = oracle_run(sc, steps=2)
    assert got["num"] == ref["num"]
    bx, bv = bit_mismatches(got["x"], ref["x"]), bit_mismatches(got["v"], ref["v"])
    print(f"aliased 2 steps: bitdiff_x={bx} bitdiff_v={bv}", flush=True)
    assert bx == 0 and bv == 0


def test_aliased_piles_in_a_batch_bit_exact():
    """The same piles as one scene of a batch beside a plain dam: big cells
    are grouped per scene (each scene's own grid and key range), and each
    scene's state equals the oracle's plain step of that scene."""
    from test_gpu_batch import _check_against_oracle, _run_batch

    # a batch runs one iteration count: the pile scene's (2)
    scs = [scenes.small_dam(12, 10, 8, seed=3, iters=2), aliased_scene(far=1500, near=900, point=300, seed=14)]
    print(f"batch with aliased piles: {[s.n for s in scs]}", flush=True)
    _check_against_oracle(scs, _run_batch(scs, 1), 1)
