"""The C-ABI library loads without a GPU and exports what include/pbfx.h
declares; host-only entry points behave like the reference."""
import ctypes
import subprocess

import numpy as np
import pytest

from ssf_pbf_cuda_amd import pbfx, scenes


def test_library_exports_every_declared_function():
    L = pbfx.lib()
    names = pbfx.header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", pbfx.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(names) <= exported


@pytest.mark.parametrize("end,h,grid", [((3.2, 3.2, 1.6), 0.1, (32, 32, 16)), ((12.8, 6.4, 6.4), 0.05, (256, 128, 128)),
                                        ((25.6, 6.4, 6.4), 0.025, (1024, 256, 256)), ((7, 7, 7), 0.14, (32, 32, 32))])
def test_calc_grid_size_matches_reference_rule(end, h, grid):
    cs = 2.0 * float(np.float32(h))
    assert pbfx.calc_grid_size(cs, end) == grid
    assert scenes.calc_grid_size(cs, end) == grid


def test_calc_grid_size_rejects_oversize():
    assert pbfx.calc_grid_size(2.0 * 0.001, (10.0, 1.0, 1.0)) is None


def test_params_default_and_status_strings():
    p = pbfx.Params()
    pbfx.lib().pbfx_params_default(ctypes.byref(p))
    assert p.max_pair == 30 and p.scorr_mode == 0 and p.xsph_on == 0 and p.vort_on == 0
    assert pbfx.lib().pbfx_status_string(0) == b"ok"
    assert pbfx.lib().pbfx_status_string(7) == b"neighbourhood tile overflow"


def test_header_struct_layout_matches_binding():
    """pbfx_params / pbfx_stats sizes in the binding match the header's field lists."""
    assert ctypes.sizeof(pbfx.Params) == 4 * 19 + 4 * 2
    assert ctypes.sizeof(pbfx.Stats) == 4 * 17
