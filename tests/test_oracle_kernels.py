"""Pins the oracle's SPH kernels, hashing and grid rule against the values the
reference's own header maths produce (tests/golden/reference_header_kats.json)
and against closed forms."""
import json
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as orc
from ssf_pbf_cuda_amd import scenes

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_header_kats.json")
KATS = json.load(open(GOLDEN))
f32 = np.float32


@pytest.mark.parametrize("row", KATS["kernel"], ids=lambda r: f"h={r['h']}")
def test_kernel_kats_bit_exact(row):
    h = f32(row["h"])
    inv_h = f32(f32(1.0) / h)
    assert f32(orc.wpoly6(0.0, inv_h)) == f32(row["W0"])
    assert f32(orc.wpoly6(float(h * f32(0.4)), inv_h)) == f32(row["W04h"])
    assert f32(orc.wpoly6(float(h * f32(0.2)), inv_h)) == f32(row["W02h"])
    assert scenes.lattice_density_sum(float(h), 0.5 * float(h)) == f32(row["lattice_sum"])
    assert f32(f32(1.0) / f32(row["lattice_sum"])) == f32(row["m_over_rho0"])
    p = scenes.dam_params(float(h), scorr_mode=scenes.SCORR_REFERENCE)
    inv_k, _ = orc.scorr_constants(p)
    assert f32(inv_k) == f32(row["inv_k"])
    assert f32(f32(1.0) / f32(row["k"])) == f32(row["inv_k"])


def test_spiky_derivative_kat():
    row = KATS["spiky_derivative"][0]
    h = f32(row["h"])
    inv_h = f32(f32(1.0) / h)
    assert f32(orc.wspiky_deriv(float(f32(row["r_over_h"]) * h), inv_h)) == f32(row["value"])


def test_grid_pos_and_hash_kat():
    row = KATS["grid_pos"][0]
    key = orc.hash_keys(np.array([row["p"]], np.float32), row["cell_width"], row["grid"])
    assert int(key[0]) == row["hash"]
    # the cell itself: floor(p / h) with IEEE FP32 division
    cell = [int(math.floor(f32(c) / f32(row["cell_width"]))) for c in row["p"]]
    assert cell == row["cell"]


@pytest.mark.parametrize("row", KATS["grid_size"], ids=lambda r: str(r["end"]))
def test_grid_size_kat(row):
    cs = 2.0 * float(f32(row["cell_size_of_h"]))
    assert orc.calc_grid_size(cs, row["end"]) == tuple(row["grid"])
    assert scenes.calc_grid_size(cs, row["end"]) == tuple(row["grid"])


def test_grid_size_overflow():
    assert orc.calc_grid_size(2.0 * 0.001, [10.0, 1.0, 1.0]) is None
    with pytest.raises(ValueError):
        scenes.calc_grid_size(2.0 * 0.001, [10.0, 1.0, 1.0])


def test_kernel_constants_are_the_closed_forms():
    # the reference's literals carry 12 significant digits
    assert abs(1.56668147065 / (315.0 / (64.0 * math.pi)) - 1.0) < 1e-9
    assert abs(14.3239448745 / (45.0 / math.pi) - 1.0) < 1e-9


def test_poly6_closed_form_f64():
    """W(r) = 315/(64 pi h^9) (h^2 - r^2)^3 inside the support, 0 outside."""
    h = 0.1
    for r in np.linspace(0.0, 1.2 * h, 49):
        expect = 315.0 / (64.0 * math.pi * h ** 9) * (h * h - r * r) ** 3 if r <= h else 0.0
        got = orc.wpoly6_f64(r, 1.0 / h)
        assert abs(got - expect) <= 1e-9 * max(abs(expect), 1.0)


def test_spiky_derivative_is_the_derivative_of_the_spiky_kernel():
    """W'(r) of the reference equals d/dr [15/(pi h^6) (h - r)^3]."""
    h = 0.1
    spiky = lambda r: 15.0 / (math.pi * h ** 6) * (h - r) ** 3 if r <= h else 0.0  # noqa: E731
    for r in np.linspace(0.05 * h, 0.95 * h, 19):
        e = 1e-7
        fd = (spiky(r + e) - spiky(r - e)) / (2 * e)
        got = orc.wspiky_deriv_f64(r, 1.0 / h)
        assert abs(got - fd) <= 1e-6 * abs(fd)


def test_lattice_sum_closed_form():
    """1 + 6 (3/4)^3 + 12 (1/2)^3 + 8 (1/4)^3 = 5.15625 poly6 units."""
    for h in (0.1, 0.05, 0.025):
        closed = 5.15625 * 315.0 / (64.0 * math.pi * h ** 3)
        got = float(scenes.lattice_density_sum(h, 0.5 * h))
        assert abs(got - closed) <= 2e-6 * closed


def test_hash_wraps_negative_and_far_cells():
    grid = (8, 4, 2)
    x = np.array([[-0.01, -0.01, -0.01], [0.85, 0.45, 0.25], [1e9, -1e9, 0.0], [np.nan, 0.0, 0.0]], np.float32)
    keys = orc.hash_keys(x, 0.1, grid)
    # (-1,-1,-1) wraps to (7,3,1)
    assert keys[0] == 1 * 8 * 4 + 3 * 8 + 7
    # (8,4,2) wraps to (0,0,0)
    assert keys[1] == 0
    # saturating conversion: 1e10 -> INT_MAX (& 7 = 7), -1e10 -> INT_MIN (& 3 = 0); NaN -> 0
    assert keys[2] == 0 * 32 + 0 * 8 + 7
    assert keys[3] == 0


def test_ieee_division_differs_from_reciprocal_multiply_somewhere():
    """The hash must use p / cw: p * (1/cw) floors differently for some
    positions, so a kernel that used the reciprocal would break bit-exactness."""
    cw = f32(0.1)
    inv = f32(f32(1.0) / cw)
    p = np.random.default_rng(0).uniform(0, 100, 2_000_000).astype(np.float32)
    a = np.floor(p / cw)
    b = np.floor((p * inv).astype(np.float32))
    assert np.any(a != b)
    keys = orc.hash_keys(np.stack([p, p * 0, p * 0], axis=1)[a != b], float(cw), (1024, 1, 1))
    assert np.array_equal(keys, (a[a != b].astype(np.int64) & 1023).astype(np.uint32))
