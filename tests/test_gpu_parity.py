"""GPU parity: the HIP solver through the C-ABI against the CPU oracle.

Bar: bit-exact for the neighbour search (sorted index, cell offsets, neighbour
lists, hash keys); FP32 state within normwise relative error 1e-5 of the
oracle after one step (the tests also report the number of bitwise
mismatches; the kernels follow the oracle's operation order, so it is
expected to be zero).
"""
import dataclasses

import numpy as np
import pytest

from ssf_pbf_cuda_amd import scenes

from helpers import bit_mismatches, compare_state, gpu_run, oracle_run, replace_params, small

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    import torch

    assert torch.cuda.is_available(), "no ROCm device"
    from ssf_pbf_cuda_amd import pbfx

    pbfx.lib()  # raises if the HIP library is missing
    yield


def test_device_arithmetic_is_ieee():
    """Division, sqrt, fma and floor on the device are bitwise the host's IEEE results."""
    import torch

    from oracle import pyoracle as orc
    from ssf_pbf_cuda_amd import pbfx

    rng = np.random.default_rng(3)
    a = np.concatenate([rng.uniform(-10, 10, 20000), rng.uniform(0, 1e-3, 5000), 10.0 ** rng.uniform(-30, 30, 5000),
                        [0.0, 1.0, 3.2, 0.1, 1.6, 25.6, 6.4]]).astype(np.float32)
    b = np.concatenate([rng.uniform(0.01, 3, a.size - 7), [0.1, 0.1, 0.1, 0.1, 0.2, 0.05, 0.025]]).astype(np.float32)
    b[: a.size - 7] = np.where(rng.random(a.size - 7) < 0.5, b[: a.size - 7], np.float32(0.1))
    at, bt = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    out = torch.zeros(4 * a.size, dtype=torch.float32, device="cuda")
    pbfx.selftest_arith(at, bt, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(-1, 4)
    ref = orc.arith_f32(a, b)
    sq = a >= 0
    for k, name in enumerate(["div", "sqrt", "fma", "floor(div)"]):
        mask = sq if name == "sqrt" else np.ones_like(sq)
        assert bit_mismatches(got[mask, k], ref[mask, k]) == 0, name


def test_kernel_inv_sqrt_is_ieee_on_every_float():
    """The pair terms' 1/sqrt (v_sqrt_f32 and v_rcp_f32 with correction steps
    on [2^-96, FLT_MAX], the IEEE operations elsewhere) equals IEEE
    1.f / sqrtf bit for bit on every non-negative float, infinity and a NaN."""
    from ssf_pbf_cuda_amd import pbfx

    for lo, hi in ((0x00000000, 0x7F800000), (0x7FC00000, 0x7FC00000)):
        bad, first = pbfx.selftest_inv_sqrt(lo, hi)
        assert bad == 0, (hex(lo), hex(hi), bad, hex(first))


@pytest.mark.parametrize("shape", [(12, 10, 8), (7, 5, 3), (1, 1, 1), (33, 9, 4)])
def test_neighbor_search_bit_exact(shape):
    sc = small(*shape)
    g = gpu_run(sc, with_ns=True)
    r = oracle_run(sc, want_ns=True)
    n, nl = sc.n, r["num"]
    assert g["num"] == nl
    assert np.array_equal(g["key"], r["key"])
    assert np.array_equal(g["sorted_index"], r["sorted_index"])
    assert np.array_equal(g["cell_offset"], r["cell_offset"])
    assert np.array_equal(g["neighbor_count"], r["neighbor_count"][:nl])
    assert np.array_equal(g["neighbors"], r["neighbors"][:nl])
    # the reference's cell_start / cell_end view (empty cells: start = EMPTY)
    off = r["cell_offset"]
    C = off.size - 2
    nonempty = off[1:C + 1] > off[:C]
    assert np.array_equal(g["cell_start"][nonempty], off[:C][nonempty])
    assert np.all(g["cell_start"][~nonempty] == 0xFFFFFFFF)
    assert np.array_equal(g["cell_end"][nonempty], off[1:C + 1][nonempty])
    # permutation is a bijection
    assert np.array_equal(np.sort(g["sorted_index"]), np.arange(n, dtype=np.uint32))


def test_step_as_shipped_reference_mode():
    """pbf_solve.cpp:125-145 as shipped: REFERENCE s_corr, no XSPH, no vorticity."""
    sc = replace_params(small(), scorr_mode=scenes.SCORR_REFERENCE, xsph_on=0, vort_on=0)
    compare_state("as-shipped", gpu_run(sc), oracle_run(sc))


def test_step_full_paper_mode():
    sc = small()
    rep = compare_state("full", gpu_run(sc), oracle_run(sc))
    assert rep["bitdiff_x"] == 0 and rep["bitdiff_v"] == 0


def test_lambda_matches_oracle():
    sc = small()
    g = gpu_run(sc, with_ns=True)
    r = oracle_run(sc, want_ns=True)
    assert bit_mismatches(g["lambda_last"], r["lambda_last"][: r["num"]]) == 0


@pytest.mark.parametrize("variant", [
    dict(vort_on=1, xsph_on=0),
    dict(vort_on=0, xsph_on=1),
    dict(vort_h_as_inv_h=1),
    dict(max_pair=64),
    dict(max_pair=8),
    dict(scorr_mode=scenes.SCORR_PAPER, scorr_n=3, scorr_k=0.02),
    dict(max_sweeps=1),
])
def test_step_variants(variant):
    sc = replace_params(small(), **variant)
    compare_state(f"variant {variant}", gpu_run(sc), oracle_run(sc))


@pytest.mark.parametrize("tile_cap,unit_width", [(0, 16), (64, 16), (2048, 1), (2048, 2), (2048, 8), (4096, 16), (4096, 32)])
def test_tiling_does_not_change_results(tile_cap, unit_width):
    """The on-chip tile, its global-memory fallback and every unit width give
    the oracle's results."""
    sc = small(20, 12, 6)
    compare_state(f"tile {tile_cap}/{unit_width}", gpu_run(sc, tile_cap=tile_cap, unit_width=unit_width),
                  oracle_run(sc))


def test_wide_units_bit_exact():
    """32-cell units (neighbourhoods of 9 x 34 = 306 cells, more than one per
    thread of the workgroup) on a grid 64 cells long, with the lists and the
    final passes: the oracle's results, and fewer units than at width 16."""
    sc = small(96, 8, 6, box=(6.4, 0.8, 0.8))
    assert sc.grid[0] == 64
    wide = gpu_run(sc, tile_cap=4096, unit_width=32, with_ns=True)
    narrow = gpu_run(sc, tile_cap=4096, unit_width=16)
    ref = oracle_run(sc, want_ns=True)
    nl = ref["num"]
    for tag, g in (("4096/32", wide), ("4096/16", narrow)):
        r = compare_state(tag, g, ref)
        assert r["bitdiff_x"] == 0 and r["bitdiff_v"] == 0, tag
    assert np.array_equal(wide["neighbor_count"], ref["neighbor_count"][:nl])
    assert np.array_equal(wide["neighbors"], ref["neighbors"][:nl])
    sw, sn = wide["stats"], narrow["stats"]
    assert 0 < sw["units"] < sn["units"], (sw["units"], sn["units"])
    # the search's 32-cell units fit the 4096 tile; the final-list units keep
    # their tile at 1536 particles, so only they may be filtered from global memory
    assert sw["tile_overflow"] == 0 and sw["slow_units"] <= sw["final_units"], sw


def test_spheres_and_planes():
    sc = small(10, 10, 10, box=(1.6, 1.6, 1.6))
    sc = dataclasses.replace(sc, inner_spheres=np.array([[0.8, 0.8, 0.8, 0.75]], np.float32),
                             outer_spheres=np.array([[0.3, 0.3, 0.3, 0.2]], np.float32))
    compare_state("spheres", gpu_run(sc), oracle_run(sc))
    sc2 = replace_params(sc, outer_as_inner=1)
    compare_state("spheres outer-as-inner", gpu_run(sc2), oracle_run(sc2))


def test_no_boundary_overload_uses_the_fixed_sphere():
    """one_step without boundary collides with the sphere (3,3,3), r=3
    (pbf_constraint.cu:262-315)."""
    sc = small(10, 8, 6, box=(6.4, 6.4, 6.4))
    sc = dataclasses.replace(sc, x=(sc.x + np.float32(0.5)).astype(np.float32))
    sphere = dataclasses.replace(sc, inner_spheres=np.array([[3, 3, 3, 3]], np.float32),
                                 plane_points=np.zeros((0, 3), np.float32), plane_normals=np.zeros((0, 3), np.float32))
    g = gpu_run(sc, boundary=False)
    compare_state("no-boundary", g, oracle_run(sphere))


def test_deletion_matches_oracle():
    """Particles whose predicted position leaves the simulation area are
    deleted; the survivors keep their stable order."""
    sc = small(12, 10, 8)
    sc = dataclasses.replace(sc, area_lo=np.array([0.0, 0.0, 0.0], np.float32),
                             area_hi=np.array([0.4, 2.0, 2.0], np.float32))
    g = gpu_run(sc, with_ns=True)
    r = oracle_run(sc, want_ns=True)
    assert 0 < r["num"] < sc.n
    assert np.array_equal(g["sorted_index"], r["sorted_index"])
    compare_state("deletion", g, r)


def test_empty_and_single():
    sc = small(1, 1, 1)
    g = gpu_run(sc)
    compare_state("single", g, oracle_run(sc))
    empty = dataclasses.replace(sc, x=np.zeros((0, 3), np.float32), v=np.zeros((0, 3), np.float32))
    g = gpu_run(empty, capacity=16)
    assert g["num"] == 0


def test_multi_step_config0_bit_exact():
    """20 steps of the 32,768-particle dam break stay bit-identical."""
    sc = scenes.config0()
    g = gpu_run(sc, steps=20)
    r = oracle_run(sc, steps=20)
    rep = compare_state("config0 x20", g, r)
    assert rep["bitdiff_x"] == 0 and rep["bitdiff_v"] == 0


def test_config1_one_step():
    """One step of the 1,048,576-particle dam break against the oracle."""
    sc = scenes.config1()
    rep = compare_state("config1", gpu_run(sc), oracle_run(sc))
    assert rep["bitdiff_x"] == 0 and rep["bitdiff_v"] == 0


def test_config2_one_step():
    """One step of the 8,388,608-particle double dam break (3 iterations, a
    1024x256x256-cell grid) against the oracle, bit for bit."""
    sc = scenes.config2()
    rep = compare_state("config2", gpu_run(sc), oracle_run(sc))
    assert rep["bitdiff_x"] == 0 and rep["bitdiff_v"] == 0
