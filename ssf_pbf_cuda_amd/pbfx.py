"""ctypes binding of libpbfx.so, the HIP PBF solver step (include/pbfx.h).

Device buffers are passed as raw device pointers; the helpers accept torch
tensors on a ROCm device (torch is only used for device memory and streams).
Loading fails loudly when the HIP library has not been built: there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

from .scenes import Scene, StepParams

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libpbfx.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "pbfx.h")

PBFX_OK = 0
PBFX_CAPACITY = 2
STATUS = {
    0: "ok",
    1: "invalid argument",
    2: "capacity",
    3: "grid too large",
    4: "HIP error",
    5: "not converged",
    6: "neighbour list truncated",
    7: "tile overflow",
    8: "not supported",
}


class PbfxError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"{what}: pbfx status {status} ({STATUS.get(status, '?')})")
        self.status = status


class Params(ctypes.Structure):
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
    ]


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("params", Params),
        ("capacity", ctypes.c_uint32),
        ("domain_end", ctypes.c_float * 3),
        ("inner_spheres4", ctypes.c_void_p),
        ("n_inner", ctypes.c_uint32),
        ("outer_spheres4", ctypes.c_void_p),
        ("n_outer", ctypes.c_uint32),
        ("plane_points3", ctypes.c_void_p),
        ("plane_normals3", ctypes.c_void_p),
        ("n_planes", ctypes.c_uint32),
        ("area_lo", ctypes.c_float * 3),
        ("area_hi", ctypes.c_float * 3),
    ]


class NsView(ctypes.Structure):
    _fields_ = [
        ("grid", ctypes.c_uint32 * 3),
        ("num_cells", ctypes.c_uint32),
        ("cell_width", ctypes.c_float),
        ("max_pair", ctypes.c_uint32),
        ("sorted_index", ctypes.c_void_p),
        ("cell_offset", ctypes.c_void_p),
        ("hash", ctypes.c_void_p),
    ]


class KernelTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("total_ms", ctypes.c_double), ("launches", ctypes.c_uint32)]


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "num_in", "num_live", "truncated", "sweep_cap_hits", "max_sweeps_used", "xsph_nan", "units", "slow_units",
        "slow_particles", "tile_overflow", "max_tile", "steps", "final_units", "list_overflow", "scan_tiles",
        "heavy_builds", "heavy_search")]


_lib = None


def lib():
    """Loads libpbfx.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(python -c 'import __graft_entry__ as g; g.build()')")
        L = ctypes.CDLL(LIB_PATH)
        vp, u32, f32p = ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_float)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        sig = {
            "pbfx_params_default": ([ctypes.POINTER(Params)], None),
            "pbfx_create": ([ctypes.POINTER(Params), u32, f32p, ctypes.POINTER(vp)], ctypes.c_int),
            "pbfx_destroy": ([vp], None),
            "pbfx_set_params": ([vp, ctypes.POINTER(Params)], ctypes.c_int),
            "pbfx_set_boundary": ([vp, f32p, u32, f32p, u32, f32p, f32p, u32], ctypes.c_int),
            "pbfx_step": ([vp, vp, vp, u32p, f32p, f32p, ctypes.c_int, vp], ctypes.c_int),
            "pbfx_step_async": ([vp, vp, vp, u32, f32p, f32p, ctypes.c_int, vp], ctypes.c_int),
            "pbfx_step_no_boundary": ([vp, vp, vp, u32p, f32p, f32p, ctypes.c_int, vp], ctypes.c_int),
            "pbfx_get_num": ([vp, vp, u32p], ctypes.c_int),
            "pbfx_batch_create": ([ctypes.POINTER(SceneDesc), u32, ctypes.POINTER(vp)], ctypes.c_int),
            "pbfx_batch_step": ([vp, vp, vp, u32p, ctypes.c_int, vp], ctypes.c_int),
            "pbfx_batch_step_async": ([vp, vp, vp, vp, u32, ctypes.c_int, vp], ctypes.c_int),
            "pbfx_batch_offsets": ([vp, vp, vp], ctypes.c_int),
            "pbfx_batch_set_params": ([vp, u32, ctypes.POINTER(Params)], ctypes.c_int),
            "pbfx_batch_graph_capture": ([vp, vp, vp, vp, ctypes.c_int, u32, vp], ctypes.c_int),
            "pbfx_batch_graph_launch": ([vp, u32, vp], ctypes.c_int),
            "pbfx_get_stats": ([vp, vp, ctypes.POINTER(Stats)], ctypes.c_int),
            "pbfx_ns_view_get": ([vp, ctypes.POINTER(NsView)], ctypes.c_int),
            "pbfx_ns_neighbors": ([vp, vp, vp, vp], ctypes.c_int),
            "pbfx_ns_cell_ranges": ([vp, vp, vp, vp], ctypes.c_int),
            "pbfx_scaling_factor": ([vp, vp, vp], ctypes.c_int),
            "pbfx_density": ([vp, vp, vp, vp], ctypes.c_int),
            "pbfx_restore_order": ([vp, vp, vp, vp, vp, vp], ctypes.c_int),
            "pbfx_set_tuning": ([vp, u32, u32], ctypes.c_int),
            "pbfx_autotune": ([vp, vp, vp, u32, vp, f32p, f32p, ctypes.c_int, vp, vp, u32p, u32p, vp], ctypes.c_int),
            "pbfx_set_report_info": ([vp, ctypes.c_int], None),
            "pbfx_get_grid": ([vp, u32p], ctypes.c_int),
            "pbfx_calc_grid_size": ([ctypes.c_double, f32p, u32p], ctypes.c_int),
            "pbfx_last_hip_error": ([vp], ctypes.c_int),
            "pbfx_status_string": ([ctypes.c_int], ctypes.c_char_p),
            "pbfx_selftest_arith": ([vp, vp, vp, u32, vp], ctypes.c_int),
            "pbfx_selftest_inv_sqrt": ([u32, u32, vp, vp, vp], ctypes.c_int),
            "pbfx_step_static": ([ctypes.POINTER(Params), f32p, vp, u32, vp, u32, vp, vp, u32, vp, vp, u32, f32p, f32p,
                                  ctypes.c_int, vp], ctypes.c_int),
            "pbfx_set_kernel_timing": ([vp, ctypes.c_int], ctypes.c_int),
            "pbfx_kernel_timing_read": ([vp, vp, ctypes.POINTER(KernelTime), u32, u32p], ctypes.c_int),
            "pbfx_add_particles": ([vp, vp, vp, u32p, vp, vp, u32, vp], ctypes.c_int),
            "pbfx_adj_enable": ([vp, u32, u32], ctypes.c_int),
            "pbfx_adj_forward": ([vp, vp, vp, u32p, f32p, f32p, ctypes.c_int, u32, vp], ctypes.c_int),
            "pbfx_adj_identity": ([vp, vp, vp], ctypes.c_int),
            "pbfx_adj_backward": ([vp, vp, vp, vp, vp, ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_int32), vp], ctypes.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def header_functions() -> list:
    """Names of the functions include/pbfx.h declares."""
    import re

    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"\b(pbfx_[a-z_]+)\s*\(", text)))


def _f3(a) -> ctypes.Array:
    arr = (ctypes.c_float * 3)()
    arr[:] = [float(x) for x in a]
    return arr


def _check(st: int, what: str) -> None:
    if st != PBFX_OK:
        raise PbfxError(st, what)


def make_params(p: StepParams) -> Params:
    o = Params()
    for name in ("stable_density", "stable_distance", "particle_mass", "smoothing_length", "time_step", "relaxation",
                 "xsph_parameter", "vc_parameter", "scorr_k", "scorr_dq_over_h"):
        setattr(o, name, float(getattr(p, name)))
    for name in ("scorr_mode", "scorr_n", "max_pair", "xsph_on", "vort_on", "vort_h_as_inv_h", "outer_as_inner",
                 "max_sweeps"):
        setattr(o, name, int(getattr(p, name)))
    o.body_force[:] = [float(c) for c in p.body_force]
    return o


def step_static(params: StepParams, domain_end, x, v, num: int, lo, hi, iters: int, inner=None, outer=None,
                plane_points=None, plane_normals=None, stream=None) -> None:
    """pbfx_step_static: one step with no context and no allocation (scratch
    in the library's static device array).  x, v: device (N, 3) float32;
    the boundary arrays are device tensors ((k, 4) spheres, (k, 3) planes)
    or None."""
    def n_rows(t):
        return 0 if t is None else int(t.shape[0])

    st = lib().pbfx_step_static(ctypes.byref(make_params(params)), _f3(domain_end), _ptr(inner), n_rows(inner),
                                _ptr(outer), n_rows(outer), _ptr(plane_points), _ptr(plane_normals),
                                n_rows(plane_points), _ptr(x), _ptr(v), int(num), _f3(lo), _f3(hi), int(iters),
                                _stream(stream))
    _check(st, "pbfx_step_static")


def calc_grid_size(cell_size: float, end: Sequence[float]):
    g = (ctypes.c_uint32 * 3)()
    st = lib().pbfx_calc_grid_size(float(cell_size), _f3(end), g)
    return None if st else tuple(g)


def _ptr(t) -> int:
    if t is None:
        return 0
    if isinstance(t, int):
        return t
    return int(t.data_ptr())


def _stream(stream) -> Optional[int]:
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return int(stream.cuda_stream)


class Solver:
    """One pbfx context (the reference's pbf_particle + pbf_buffer + neighbor_search)."""

    def __init__(self, params: StepParams, capacity: int, domain_end: Sequence[float]):
        self._lib = lib()
        self.params = params
        self.capacity = int(capacity)
        self._ctx = ctypes.c_void_p()
        _check(self._lib.pbfx_create(ctypes.byref(make_params(params)), self.capacity, _f3(domain_end),
                                     ctypes.byref(self._ctx)), "pbfx_create")

    @classmethod
    def for_scene(cls, scene: Scene, capacity: Optional[int] = None, boundary: bool = True) -> "Solver":
        s = cls(scene.params, capacity or scene.n, scene.box_hi)
        if boundary:
            s.set_boundary(scene.inner_spheres, scene.outer_spheres, scene.plane_points, scene.plane_normals)
        return s

    def close(self) -> None:
        if self._ctx:
            self._lib.pbfx_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params: StepParams) -> None:
        _check(self._lib.pbfx_set_params(self._ctx, ctypes.byref(make_params(params))), "pbfx_set_params")
        self.params = params

    def set_boundary(self, inner=None, outer=None, plane_points=None, plane_normals=None) -> None:
        def arr(a, w):
            if a is None:
                return np.zeros((0, w), np.float32)
            return np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1, w))

        ai, ao, pp, pn = arr(inner, 4), arr(outer, 4), arr(plane_points, 3), arr(plane_normals, 3)
        fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))  # noqa: E731
        self._keep = (ai, ao, pp, pn)
        _check(self._lib.pbfx_set_boundary(self._ctx, fp(ai), ai.shape[0], fp(ao), ao.shape[0], fp(pp), fp(pn),
                                           pp.shape[0]), "pbfx_set_boundary")

    def autotune(self, x, v, num: int, lo, hi, iters: int, x_scratch, v_scratch, in_offsets=None, stream=None):
        """Times each (tile_cap, unit_width) candidate on scratch copies of x, v and keeps the fastest."""
        t, w = ctypes.c_uint32(), ctypes.c_uint32()
        _check(self._lib.pbfx_autotune(self._ctx, _ptr(x), _ptr(v), int(num), _ptr(in_offsets), _f3(lo), _f3(hi),
                                       int(iters), _ptr(x_scratch), _ptr(v_scratch), ctypes.byref(t),
                                       ctypes.byref(w), _stream(stream)), "pbfx_autotune")
        return int(t.value), int(w.value)

    def set_tuning(self, tile_cap: int = 2048, unit_width: int = 16) -> None:
        _check(self._lib.pbfx_set_tuning(self._ctx, int(tile_cap), int(unit_width)), "pbfx_set_tuning")

    def set_report_info(self, enable: bool) -> None:
        self._lib.pbfx_set_report_info(self._ctx, int(bool(enable)))

    def grid(self):
        g = (ctypes.c_uint32 * 3)()
        _check(self._lib.pbfx_get_grid(self._ctx, g), "pbfx_get_grid")
        return tuple(g)

    def step(self, x, v, num: int, lo, hi, iters: int, stream=None, boundary: bool = True) -> int:
        """pbf::one_step: returns the live count after deletion (synchronises)."""
        n = ctypes.c_uint32(int(num))
        fn = self._lib.pbfx_step if boundary else self._lib.pbfx_step_no_boundary
        st = fn(self._ctx, _ptr(x), _ptr(v), ctypes.byref(n), _f3(lo), _f3(hi), int(iters), _stream(stream))
        _check(st, "pbfx_step")
        return int(n.value)

    def step_status(self, x, v, num: int, lo, hi, iters: int, stream=None):
        n = ctypes.c_uint32(int(num))
        st = self._lib.pbfx_step(self._ctx, _ptr(x), _ptr(v), ctypes.byref(n), _f3(lo), _f3(hi), int(iters),
                                 _stream(stream))
        return int(st), int(n.value)

    def step_async(self, x, v, num: int, lo, hi, iters: int, stream=None) -> None:
        _check(self._lib.pbfx_step_async(self._ctx, _ptr(x), _ptr(v), int(num), _f3(lo), _f3(hi), int(iters),
                                         _stream(stream)), "pbfx_step_async")

    def add_particles(self, x, v, num: int, add_x, add_v, add_n: int, stream=None) -> int:
        """Appends add_n particles after the live ones (pbf::cuda::addParticle); returns the new count."""
        n = ctypes.c_uint32(num)
        _check(self._lib.pbfx_add_particles(self._ctx, _ptr(x), _ptr(v), ctypes.byref(n), _ptr(add_x), _ptr(add_v),
                                            int(add_n), _stream(stream)), "pbfx_add_particles")
        return int(n.value)

    def num(self, stream=None) -> int:
        n = ctypes.c_uint32()
        _check(self._lib.pbfx_get_num(self._ctx, _stream(stream), ctypes.byref(n)), "pbfx_get_num")
        return int(n.value)

    def stats(self, stream=None) -> dict:
        s = Stats()
        _check(self._lib.pbfx_get_stats(self._ctx, _stream(stream), ctypes.byref(s)), "pbfx_get_stats")
        return {k: int(getattr(s, k)) for k, _ in Stats._fields_}

    def ns_view(self) -> NsView:
        v = NsView()
        _check(self._lib.pbfx_ns_view_get(self._ctx, ctypes.byref(v)), "pbfx_ns_view_get")
        return v

    def ns_neighbors(self, out, counts, stream=None) -> None:
        _check(self._lib.pbfx_ns_neighbors(self._ctx, _ptr(out), _ptr(counts), _stream(stream)), "pbfx_ns_neighbors")

    def ns_cell_ranges(self, start, end, stream=None) -> None:
        _check(self._lib.pbfx_ns_cell_ranges(self._ctx, _ptr(start), _ptr(end), _stream(stream)),
               "pbfx_ns_cell_ranges")

    def set_kernel_timing(self, enable: bool) -> None:
        _check(self._lib.pbfx_set_kernel_timing(self._ctx, int(bool(enable))), "pbfx_set_kernel_timing")

    def kernel_times(self, stream=None) -> dict:
        """{kernel: (total_ms, launches)} accumulated since timing was enabled."""
        buf = (KernelTime * 32)()
        n = ctypes.c_uint32()
        _check(self._lib.pbfx_kernel_timing_read(self._ctx, _stream(stream), buf, 32, ctypes.byref(n)),
               "pbfx_kernel_timing_read")
        return {buf[i].name.decode(): (float(buf[i].total_ms), int(buf[i].launches)) for i in range(min(n.value, 32))}

    def scaling_factor(self, out, stream=None) -> None:
        _check(self._lib.pbfx_scaling_factor(self._ctx, _ptr(out), _stream(stream)), "pbfx_scaling_factor")

    def density(self, x_sorted, rho, stream=None) -> None:
        """rho_i over the last step's lists (device AoS x in its sorted order)."""
        _check(self._lib.pbfx_density(self._ctx, _ptr(x_sorted), _ptr(rho), _stream(stream)), "pbfx_density")

    def restore_order(self, x_sorted, v_sorted, x_out, v_out, stream=None) -> None:
        """x_out[sorted_index[s]] = x_sorted[s] (and v) for the last step's live slots."""
        _check(self._lib.pbfx_restore_order(self._ctx, _ptr(x_sorted), _ptr(v_sorted), _ptr(x_out), _ptr(v_out),
                                            _stream(stream)), "pbfx_restore_order")

    # ---- differentiable form
    def adj_enable(self, max_steps: int, every: int) -> None:
        """Checkpoints every `every` steps, up to max_steps steps."""
        _check(self._lib.pbfx_adj_enable(self._ctx, int(max_steps), int(every)), "pbfx_adj_enable")

    def adj_forward(self, x, v, num: int, lo, hi, iters: int, steps: int, stream=None) -> int:
        """`steps` steps from (x, v) in place, kept for the backward pass; returns the live count."""
        n = ctypes.c_uint32(num)
        _check(self._lib.pbfx_adj_forward(self._ctx, _ptr(x), _ptr(v), ctypes.byref(n), _f3(lo), _f3(hi), int(iters),
                                          int(steps), _stream(stream)), "pbfx_adj_forward")
        return int(n.value)

    def adj_identity(self, out, stream=None) -> None:
        """out[s] = initial index of the particle in final slot s (device uint32)."""
        _check(self._lib.pbfx_adj_identity(self._ctx, _ptr(out), _stream(stream)), "pbfx_adj_identity")

    def adj_backward(self, gx, gv, gx0, gv0, stream=None):
        """dL/d(final x, v) -> dL/d(initial x, v) into gx0, gv0 (device
        float64, AoS); returns (dL/dc, dL/d vorticity eps, exponent): every
        gradient is the value returned times 2**exponent."""
        gp = (ctypes.c_double * 2)()
        ex = ctypes.c_int32(0)
        _check(self._lib.pbfx_adj_backward(self._ctx, _ptr(gx), _ptr(gv), _ptr(gx0), _ptr(gv0), gp, ctypes.byref(ex),
                                           _stream(stream)), "pbfx_adj_backward")
        return float(gp[0]), float(gp[1]), int(ex.value)


def selftest_arith(a, b, out, stream=None) -> None:
    n = int(a.numel())
    _check(lib().pbfx_selftest_arith(_ptr(a), _ptr(b), _ptr(out), n, _stream(stream)), "pbfx_selftest_arith")


def selftest_inv_sqrt(lo_bits: int, hi_bits: int, stream=None) -> tuple[int, int]:
    """Compares the kernels' 1/sqrt with IEEE 1.f/sqrtf on every float whose
    bits lie in [lo_bits, hi_bits]; returns (mismatches, bits of one
    mismatching input or 0)."""
    import torch

    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    first = torch.zeros(1, dtype=torch.int32, device="cuda")
    _check(lib().pbfx_selftest_inv_sqrt(lo_bits, hi_bits, _ptr(cnt), _ptr(first), _stream(stream)),
           "pbfx_selftest_inv_sqrt")
    torch.cuda.synchronize()
    return int(cnt.item()), int(first.item()) & 0xFFFFFFFF


_hip = None


def _hiprt():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemcpy.restype = ctypes.c_int
    return _hip


def d2h(ptr: int, count: int, dtype=np.uint32) -> np.ndarray:
    """Copies `count` elements from a device pointer owned by a context."""
    out = np.empty(int(count), dtype=dtype)
    if count:
        rc = _hiprt().hipMemcpy(out.ctypes.data, ctypes.c_void_p(int(ptr)), out.nbytes, 2)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed ({rc})")
    return out


class BatchSolver(Solver):
    """A batched pbfx context: independent scenes advanced by one sequence of
    launches per step (pbfx_batch_*).  Particle arrays hold the scenes
    concatenated in order."""

    def __init__(self, scenes: Sequence[Scene], capacities: Optional[Sequence[int]] = None, boundary: bool = True):
        self._lib = lib()
        self.scenes = list(scenes)
        caps = [int(c) for c in (capacities or [sc.n for sc in scenes])]
        self.capacities = caps
        self.capacity = sum(caps)
        self.params = scenes[0].params
        descs = (SceneDesc * len(scenes))()
        keep = []
        for d, sc, cap in zip(descs, scenes, caps):
            d.params = make_params(sc.params)
            d.capacity = cap
            d.domain_end[:] = [float(c) for c in sc.box_hi]
            d.area_lo[:] = [float(c) for c in sc.area_lo]
            d.area_hi[:] = [float(c) for c in sc.area_hi]
            if boundary:
                arrs = [np.ascontiguousarray(a, np.float32) for a in
                        (sc.inner_spheres, sc.outer_spheres, sc.plane_points, sc.plane_normals)]
                keep.extend(arrs)
                d.inner_spheres4, d.outer_spheres4, d.plane_points3, d.plane_normals3 = [
                    a.ctypes.data if a.size else None for a in arrs]
                d.n_inner, d.n_outer, d.n_planes = arrs[0].shape[0], arrs[1].shape[0], arrs[2].shape[0]
        self._ctx = ctypes.c_void_p()
        _check(self._lib.pbfx_batch_create(descs, len(scenes), ctypes.byref(self._ctx)), "pbfx_batch_create")

    def batch_step(self, x, v, counts: Sequence[int], iters: int, stream=None):
        """One step of every scene; returns the live counts per scene."""
        c = (ctypes.c_uint32 * len(counts))(*[int(n) for n in counts])
        _check(self._lib.pbfx_batch_step(self._ctx, _ptr(x), _ptr(v), c, int(iters), _stream(stream)),
               "pbfx_batch_step")
        return list(c)

    def batch_step_async(self, x, v, in_offsets, total: int, iters: int, stream=None) -> None:
        _check(self._lib.pbfx_batch_step_async(self._ctx, _ptr(x), _ptr(v), _ptr(in_offsets), int(total), int(iters),
                                               _stream(stream)), "pbfx_batch_step_async")

    def batch_offsets(self, out, stream=None) -> None:
        _check(self._lib.pbfx_batch_offsets(self._ctx, _ptr(out), _stream(stream)), "pbfx_batch_offsets")

    def batch_set_params(self, scene: int, params: StepParams) -> None:
        _check(self._lib.pbfx_batch_set_params(self._ctx, int(scene), ctypes.byref(make_params(params))),
               "pbfx_batch_set_params")

    def graph_capture(self, x, v, offsets, iters: int, steps: int, stream) -> None:
        """pbfx_batch_graph_capture: records `steps` chained batched steps
        (offsets: device int32 [n_scenes+1], updated in place by each step)."""
        _check(self._lib.pbfx_batch_graph_capture(self._ctx, _ptr(x), _ptr(v), _ptr(offsets), int(iters),
                                                  int(steps), _stream(stream)), "pbfx_batch_graph_capture")

    def graph_launch(self, launches: int, stream) -> None:
        _check(self._lib.pbfx_batch_graph_launch(self._ctx, int(launches), _stream(stream)),
               "pbfx_batch_graph_launch")
