"""ctypes binding of the CPU oracle (libpbf_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The HIP solver never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libpbf_oracle.so")
EMPTY = 0xFFFFFFFF


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("stable_density", ctypes.c_float),
        ("stable_distance", ctypes.c_float),
        ("particle_mass", ctypes.c_float),
        ("smoothing_length", ctypes.c_float),
        ("time_step", ctypes.c_float),
        ("relaxation", ctypes.c_float),
        ("xsph_parameter", ctypes.c_float),
        ("vc_parameter", ctypes.c_float),
        ("body_force", ctypes.c_float * 3),
        ("scorr_mode", ctypes.c_int32),
        ("scorr_k", ctypes.c_float),
        ("scorr_n", ctypes.c_int32),
        ("scorr_dq_over_h", ctypes.c_float),
        ("max_pair", ctypes.c_int32),
        ("xsph_on", ctypes.c_int32),
        ("vort_on", ctypes.c_int32),
        ("vort_h_as_inv_h", ctypes.c_int32),
        ("outer_as_inner", ctypes.c_int32),
        ("max_sweeps", ctypes.c_int32),
        ("grid", ctypes.c_uint32 * 3),
    ]


class OrcBoundary(ctypes.Structure):
    _fields_ = [
        ("n_inner", ctypes.c_uint32),
        ("inner", ctypes.POINTER(ctypes.c_float)),
        ("n_outer", ctypes.c_uint32),
        ("outer", ctypes.POINTER(ctypes.c_float)),
        ("n_planes", ctypes.c_uint32),
        ("plane_points", ctypes.POINTER(ctypes.c_float)),
        ("plane_normals", ctypes.POINTER(ctypes.c_float)),
    ]


class OrcNsOut(ctypes.Structure):
    _fields_ = [
        ("key", ctypes.POINTER(ctypes.c_uint32)),
        ("sorted_index", ctypes.POINTER(ctypes.c_uint32)),
        ("cell_offset", ctypes.POINTER(ctypes.c_uint32)),
        ("neighbors", ctypes.POINTER(ctypes.c_uint32)),
        ("neighbor_count", ctypes.POINTER(ctypes.c_uint32)),
        ("pred0", ctypes.POINTER(ctypes.c_float)),
        ("lambda_last", ctypes.POINTER(ctypes.c_float)),
    ]


class OrcStats(ctypes.Structure):
    _fields_ = [
        ("num_in", ctypes.c_uint32),
        ("num_live", ctypes.c_uint32),
        ("num_deleted", ctypes.c_uint32),
        ("truncated", ctypes.c_uint32),
        ("sweep_cap_hits", ctypes.c_uint32),
        ("max_sweeps_used", ctypes.c_uint32),
        ("xsph_nan", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE], stdout=subprocess.DEVNULL)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(LIB_PATH)
        f32p = ctypes.POINTER(ctypes.c_float)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        f64p = ctypes.POINTER(ctypes.c_double)
        L.orc_calc_grid_size.argtypes = [ctypes.c_double, f32p, u32p]
        L.orc_calc_grid_size.restype = ctypes.c_int
        L.orc_wpoly6.argtypes = [ctypes.c_float, ctypes.c_float]
        L.orc_wpoly6.restype = ctypes.c_float
        L.orc_wspiky_deriv.argtypes = [ctypes.c_float, ctypes.c_float]
        L.orc_wspiky_deriv.restype = ctypes.c_float
        L.orc_wpoly6_f64.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_wpoly6_f64.restype = ctypes.c_double
        L.orc_wspiky_deriv_f64.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_wspiky_deriv_f64.restype = ctypes.c_double
        L.orc_hash.argtypes = [f32p, ctypes.c_uint32, ctypes.c_float, u32p, u32p]
        L.orc_hash.restype = None
        L.orc_scorr_constants.argtypes = [ctypes.POINTER(OrcParams), f32p, f32p]
        L.orc_scorr_constants.restype = None
        L.orc_step_f32.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(OrcBoundary), f32p, f32p, u32p, f32p,
                                   f32p, ctypes.c_int, ctypes.POINTER(OrcNsOut), ctypes.POINTER(OrcStats)]
        L.orc_step_f32.restype = ctypes.c_int
        L.orc_step_f64.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(OrcBoundary), f32p, f32p,
                                   ctypes.c_uint32, f32p, f32p, ctypes.c_int, f64p, f64p, u32p]
        L.orc_step_f64.restype = ctypes.c_int
        L.orc_metrics.argtypes = [ctypes.POINTER(OrcParams), f32p, f32p, ctypes.c_uint32, f64p]
        L.orc_metrics.restype = ctypes.c_int
        L.orc_density.argtypes = [ctypes.POINTER(OrcParams), f32p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32), f32p]
        L.orc_density.restype = ctypes.c_int
        L.orc_arith_f32.argtypes = [f32p, f32p, f32p, ctypes.c_uint32]
        L.orc_arith_f32.restype = None
        vp = ctypes.c_void_p
        L.orc_tape_new.argtypes = []
        L.orc_tape_new.restype = vp
        L.orc_tape_free.argtypes = [vp]
        L.orc_tape_free.restype = None
        L.orc_tape_steps.argtypes = [vp]
        L.orc_tape_steps.restype = ctypes.c_uint32
        L.orc_tape_num.argtypes = [vp, ctypes.c_uint32, ctypes.c_int]
        L.orc_tape_num.restype = ctypes.c_uint32
        L.orc_adj_step_f32.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(OrcBoundary), f32p, f32p, u32p,
                                       f32p, f32p, ctypes.c_int, vp]
        L.orc_adj_step_f32.restype = ctypes.c_int
        L.orc_adj_replay_f64.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(OrcBoundary), vp, f64p, f64p,
                                         f64p, f64p, f64p, ctypes.c_int]
        L.orc_adj_replay_f64.restype = ctypes.c_int
        L.orc_adj_backward_f64.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(OrcBoundary), vp, f64p, f64p,
                                           f64p, f64p, f64p, f64p, ctypes.POINTER(ctypes.c_int32)]
        L.orc_adj_backward_f64.restype = ctypes.c_int
        L.orc_adj_identity.argtypes = [vp, u32p]
        L.orc_adj_identity.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_threads.restype = None
        L.orc_get_threads.argtypes = []
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_set_plant.argtypes = [ctypes.c_int]
        L.orc_set_plant.restype = None
        _lib = L
    return _lib


def _fp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _up(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_params(p, grid) -> OrcParams:
    o = OrcParams()
    for name in ("stable_density", "stable_distance", "particle_mass", "smoothing_length", "time_step", "relaxation",
                 "xsph_parameter", "vc_parameter", "scorr_k", "scorr_dq_over_h"):
        setattr(o, name, float(getattr(p, name)))
    for name in ("scorr_mode", "scorr_n", "max_pair", "xsph_on", "vort_on", "vort_h_as_inv_h", "outer_as_inner",
                 "max_sweeps"):
        setattr(o, name, int(getattr(p, name)))
    o.body_force[:] = [float(c) for c in p.body_force]
    o.grid[:] = [int(g) for g in grid]
    return o


class _Boundary:
    """Keeps the numpy buffers alive for the ctypes struct."""

    def __init__(self, scene):
        self.inner = np.ascontiguousarray(scene.inner_spheres, dtype=np.float32).reshape(-1)
        self.outer = np.ascontiguousarray(scene.outer_spheres, dtype=np.float32).reshape(-1)
        self.pts = np.ascontiguousarray(scene.plane_points, dtype=np.float32).reshape(-1)
        self.nrm = np.ascontiguousarray(scene.plane_normals, dtype=np.float32).reshape(-1)
        b = OrcBoundary()
        b.n_inner = self.inner.size // 4
        b.inner = _fp(self.inner)
        b.n_outer = self.outer.size // 4
        b.outer = _fp(self.outer)
        b.n_planes = self.pts.size // 3
        b.plane_points = _fp(self.pts)
        b.plane_normals = _fp(self.nrm)
        self.struct = b


def arith_f32(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    out = np.zeros(4 * a.size, np.float32)
    lib().orc_arith_f32(_fp(a), _fp(b), _fp(out), a.size)
    return out.reshape(-1, 4)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


# Planted mistakes (pbf_oracle.h, ORC_PLANT_*): the closed-form tests switch
# one on to show that their check fails on it.
PLANT_NONE = 0
PLANT_XSPH_SIGN = 1
PLANT_XSPH_NO_MR = 2
PLANT_XSPH_H = 3
PLANT_VORT_CROSS = 4
PLANT_VORT_NO_MR = 5
PLANT_CONF_CROSS = 6
PLANT_CONF_RHO = 7
PLANT_LOC_SELF = 8
PLANT_NUDGE = 9
PLANT_CROSS_FRAC = 10


class planted:
    """Context manager: the oracle runs with planted mistake `code` inside
    the block and as the faithful restatement again after it."""

    def __init__(self, code: int):
        self.code = int(code)

    def __enter__(self):
        lib().orc_set_plant(self.code)
        return self

    def __exit__(self, *exc):
        lib().orc_set_plant(PLANT_NONE)
        return False


def calc_grid_size(cell_size: float, end) -> Optional[tuple]:
    e = np.asarray(end, dtype=np.float32)
    out = np.zeros(3, np.uint32)
    rc = lib().orc_calc_grid_size(float(cell_size), _fp(e), _up(out))
    return None if rc else tuple(int(g) for g in out)


def wpoly6(r: float, inv_h: float) -> float:
    return float(lib().orc_wpoly6(float(r), float(inv_h)))


def wspiky_deriv(r: float, inv_h: float) -> float:
    return float(lib().orc_wspiky_deriv(float(r), float(inv_h)))


def wpoly6_f64(r: float, inv_h: float) -> float:
    return float(lib().orc_wpoly6_f64(float(r), float(inv_h)))


def wspiky_deriv_f64(r: float, inv_h: float) -> float:
    return float(lib().orc_wspiky_deriv_f64(float(r), float(inv_h)))


def hash_keys(x: np.ndarray, cell_width: float, grid) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    g = np.asarray(grid, dtype=np.uint32)
    out = np.zeros(x.shape[0], np.uint32)
    lib().orc_hash(_fp(x), x.shape[0], float(cell_width), _up(g), _up(out))
    return out


def scorr_constants(params, grid=(1, 1, 1)):
    p = make_params(params, grid)
    a = ctypes.c_float()
    b = ctypes.c_float()
    lib().orc_scorr_constants(ctypes.byref(p), ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def step(scene, x=None, v=None, iters=None, params=None, want_ns=False, boundary=True, grid=None):
    """One FP32 oracle step.  Returns a dict with x, v (sorted order), num and
    stats; with want_ns also key, sorted_index, cell_offset, neighbors,
    neighbor_count, pred0 and lambda_last."""
    p = params if params is not None else scene.params
    grid = tuple(grid) if grid is not None else tuple(scene.grid)
    x = np.array(scene.x if x is None else x, dtype=np.float32, copy=True).reshape(-1, 3)
    v = np.array(scene.v if v is None else v, dtype=np.float32, copy=True).reshape(-1, 3)
    n = x.shape[0]
    iters = scene.iters if iters is None else iters
    op = make_params(p, grid)
    bnd = _Boundary(scene) if boundary else None
    num = np.array([n], np.uint32)
    lo = np.ascontiguousarray(scene.area_lo, dtype=np.float32)
    hi = np.ascontiguousarray(scene.area_hi, dtype=np.float32)
    st = OrcStats()
    ns = None
    out = {}
    if want_ns:
        C = grid[0] * grid[1] * grid[2]
        M = int(p.max_pair)
        out["key"] = np.zeros(n, np.uint32)
        out["sorted_index"] = np.zeros(n, np.uint32)
        out["cell_offset"] = np.zeros(C + 2, np.uint32)
        out["neighbors"] = np.zeros(n * M, np.uint32)
        out["neighbor_count"] = np.zeros(n, np.uint32)
        out["pred0"] = np.zeros(3 * n, np.float32)
        out["lambda_last"] = np.zeros(n, np.float32)
        ns = OrcNsOut()
        ns.key = _up(out["key"])
        ns.sorted_index = _up(out["sorted_index"])
        ns.cell_offset = _up(out["cell_offset"])
        ns.neighbors = _up(out["neighbors"])
        ns.neighbor_count = _up(out["neighbor_count"])
        ns.pred0 = _fp(out["pred0"])
        ns.lambda_last = _fp(out["lambda_last"])
    rc = lib().orc_step_f32(ctypes.byref(op), ctypes.byref(bnd.struct) if bnd else None, _fp(x), _fp(v), _up(num),
                            _fp(lo), _fp(hi), int(iters), ctypes.byref(ns) if ns is not None else None,
                            ctypes.byref(st))
    if rc:
        raise RuntimeError(f"orc_step_f32 failed ({rc})")
    nl = int(num[0])
    out.update(x=x[:nl].copy(), v=v[:nl].copy(), num=nl,
               stats={k: getattr(st, k) for k, _ in OrcStats._fields_})
    if want_ns:
        out["neighbors"] = out["neighbors"].reshape(n, int(p.max_pair))
        out["pred0"] = out["pred0"].reshape(n, 3)
    return out


def step_f64(scene, x=None, v=None, iters=None, params=None, boundary=True, grid=None):
    """FP64 twin of :func:`step` (same discrete decisions, double arithmetic)."""
    p = params if params is not None else scene.params
    grid = tuple(grid) if grid is not None else tuple(scene.grid)
    x = np.ascontiguousarray(scene.x if x is None else x, dtype=np.float32).reshape(-1, 3)
    v = np.ascontiguousarray(scene.v if v is None else v, dtype=np.float32).reshape(-1, 3)
    n = x.shape[0]
    iters = scene.iters if iters is None else iters
    op = make_params(p, grid)
    bnd = _Boundary(scene) if boundary else None
    lo = np.ascontiguousarray(scene.area_lo, dtype=np.float32)
    hi = np.ascontiguousarray(scene.area_hi, dtype=np.float32)
    xo = np.zeros((n, 3), np.float64)
    vo = np.zeros((n, 3), np.float64)
    num = np.zeros(1, np.uint32)
    rc = lib().orc_step_f64(ctypes.byref(op), ctypes.byref(bnd.struct) if bnd else None, _fp(x), _fp(v), n,
                            _fp(lo), _fp(hi), int(iters), _dp(xo), _dp(vo), _up(num))
    if rc:
        raise RuntimeError(f"orc_step_f64 failed ({rc})")
    nl = int(num[0])
    return {"x": xo[:nl], "v": vo[:nl], "num": nl}


def density(scene, x_sorted, neighbors, params=None, grid=None):
    """rho over the given lists (n x max_pair, EMPTY padded), x in the lists' order."""
    p = params if params is not None else scene.params
    grid = tuple(grid) if grid is not None else tuple(scene.grid)
    op = make_params(p, grid)
    x = np.ascontiguousarray(x_sorted, dtype=np.float32).reshape(-1, 3)
    nb = np.ascontiguousarray(neighbors, dtype=np.uint32).reshape(x.shape[0], int(p.max_pair))
    out = np.zeros(x.shape[0], np.float32)
    lib().orc_density(ctypes.byref(op), _fp(x), x.shape[0], _up(nb), _fp(out))
    return out


def metrics(scene, x, v, params=None, grid=None):
    """mean C, max C, mean max(C,0) (C = rho/rho0 - 1, uncapped) and kinetic energy."""
    p = params if params is not None else scene.params
    grid = tuple(grid) if grid is not None else tuple(scene.grid)
    op = make_params(p, grid)
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
    v = np.ascontiguousarray(v, dtype=np.float32).reshape(-1, 3)
    out = np.zeros(4, np.float64)
    lib().orc_metrics(ctypes.byref(op), _fp(x), _fp(v), x.shape[0], _dp(out))
    return {"mean_C": out[0], "max_C": out[1], "mean_pos_C": out[2], "kinetic_energy": out[3]}


class Tape:
    """The step as a map with frozen discrete decisions (pbf_adjoint.cpp):
    ``run`` records FP32 steps (bit-identical to :func:`step`), ``replay``
    re-evaluates them in FP64 under the recorded decisions (optionally with
    other XSPH / vorticity coefficients), ``backward`` is the FP64 adjoint at
    the tape's values."""

    def __init__(self, scene, params=None, boundary=True, grid=None):
        self.scene = scene
        self.params = params if params is not None else scene.params
        self.grid = tuple(grid) if grid is not None else tuple(scene.grid)
        self._bnd = _Boundary(scene) if boundary else None
        self._h = lib().orc_tape_new()

    def close(self):
        if self._h:
            lib().orc_tape_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _bp(self):
        return ctypes.byref(self._bnd.struct) if self._bnd else None

    def _op(self, params=None):
        return make_params(params if params is not None else self.params, self.grid)

    @property
    def steps(self) -> int:
        return int(lib().orc_tape_steps(self._h))

    def run(self, steps, x=None, v=None, iters=None):
        """Appends `steps` FP32 steps from (x, v); returns the final (x, v)."""
        sc = self.scene
        x = np.array(sc.x if x is None else x, dtype=np.float32, copy=True).reshape(-1, 3)
        v = np.array(sc.v if v is None else v, dtype=np.float32, copy=True).reshape(-1, 3)
        iters = sc.iters if iters is None else iters
        op = self._op()
        lo = np.ascontiguousarray(sc.area_lo, dtype=np.float32)
        hi = np.ascontiguousarray(sc.area_hi, dtype=np.float32)
        num = np.array([x.shape[0]], np.uint32)
        for _ in range(steps):
            rc = lib().orc_adj_step_f32(ctypes.byref(op), self._bp(), _fp(x), _fp(v), _up(num), _fp(lo), _fp(hi),
                                        int(iters), self._h)
            if rc:
                raise RuntimeError(f"orc_adj_step_f32 failed ({rc})")
            x, v = x[: int(num[0])].copy(), v[: int(num[0])].copy()
        return x, v

    def num(self, step, live=True) -> int:
        return int(lib().orc_tape_num(self._h, int(step), 1 if live else 0))

    def _coef(self, coef):
        if coef is None:
            return None, None
        c = np.ascontiguousarray(coef, dtype=np.float64).reshape(2)
        return c, _dp(c)

    def replay(self, x0, v0, coef=None, record=False):
        """FP64 map under the recorded decisions with coef = (XSPH c,
        vorticity eps) in double (None: the scene's); returns the final (x, v)."""
        x0 = np.ascontiguousarray(x0, dtype=np.float64).reshape(-1, 3)
        v0 = np.ascontiguousarray(v0, dtype=np.float64).reshape(-1, 3)
        nl = self.num(self.steps - 1)
        xo = np.zeros((nl, 3), np.float64)
        vo = np.zeros((nl, 3), np.float64)
        op = self._op()
        keep, cp = self._coef(coef)
        rc = lib().orc_adj_replay_f64(ctypes.byref(op), self._bp(), self._h, _dp(x0), _dp(v0), cp, _dp(xo), _dp(vo),
                                      1 if record else 0)
        if rc:
            raise RuntimeError(f"orc_adj_replay_f64 failed ({rc})")
        return xo, vo

    def backward(self, gx, gv, coef=None, with_exponent=False):
        """FP64 adjoint: returns (dL/dx0, dL/dv0, dL/dc, dL/d vorticity eps),
        each to be multiplied by 2**exponent; with_exponent appends the
        exponent, otherwise the values are returned scaled (inf if beyond
        double range)."""
        gx = np.ascontiguousarray(gx, dtype=np.float64).reshape(-1, 3)
        gv = np.ascontiguousarray(gv, dtype=np.float64).reshape(-1, 3)
        n0 = self.num(0, live=False)
        gx0 = np.zeros((n0, 3), np.float64)
        gv0 = np.zeros((n0, 3), np.float64)
        gp = np.zeros(2, np.float64)
        op = self._op()
        keep, cp = self._coef(coef)
        ex = ctypes.c_int32(0)
        rc = lib().orc_adj_backward_f64(ctypes.byref(op), self._bp(), self._h, cp, _dp(gx), _dp(gv), _dp(gx0),
                                        _dp(gv0), _dp(gp), ctypes.byref(ex))
        if rc:
            raise RuntimeError(f"orc_adj_backward_f64 failed ({rc})")
        if with_exponent:
            return gx0, gv0, float(gp[0]), float(gp[1]), int(ex.value)
        e = int(ex.value)
        with np.errstate(over="ignore"):
            return np.ldexp(gx0, e), np.ldexp(gv0, e), float(np.ldexp(gp[0], e)), float(np.ldexp(gp[1], e))

    def identity(self):
        nl = self.num(self.steps - 1)
        out = np.zeros(nl, np.uint32)
        if lib().orc_adj_identity(self._h, _up(out)):
            raise RuntimeError("orc_adj_identity failed")
        return out
