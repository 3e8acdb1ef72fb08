"""Scene and initial-condition generators for the PBF solver step.

Particle lattices follow the reference's ``cartesian_volume`` order
(cuda_pbf_with_ssf/src/pbf/util/pbf_arrangement.h:82-121: x fastest, then y,
then z; ``coord = origin + (space*x, space*y, space*z)`` in FP32).  Jitter and
initial velocities come from a counter-based SplitMix64 stream, so every
consumer (HIP solver, CPU checker, benchmarks) sees identical bytes.

Parameters are derived the way the reference's dam-break set-up does
(simulator/dam/pbf_dam_init_cond.cpp:7-27: ``m = space^3 * 1000``), except
that the rest density is the poly6 density of an interior lattice particle
(``rho0 = m * sum_j W(|x_i - x_j|)`` over the 27 lattice neighbours inside h)
and the step parameters are the dam-break workload's: dt = 1/120,
g = (0, -9.8, 0), relaxation 600, paper s_corr (k = 0.1, n = 4, dq = 0.2h),
XSPH c = 0.01, vorticity epsilon = 0.0005, hash cell width h.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

SCORR_REFERENCE = 0
SCORR_PAPER = 1

# Simulation-area bound of the workloads: far enough that no predicted
# position leaves it, so the particle count of a workload stays fixed (the
# box planes keep the solved positions inside the box; only the prediction
# x + v*dt can stray, by |v|*dt).
NO_DELETION = 1.0e30


def splitmix64(seed: int, counters: np.ndarray) -> np.ndarray:
    """k-th SplitMix64 output for state ``seed``: mix(seed + (k+1)*GOLDEN)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + (counters.astype(np.uint64) + np.uint64(1)) * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01_f32(seed: int, counters: np.ndarray) -> np.ndarray:
    """Uniform [0, 1) with 24 random bits, exactly representable in FP32."""
    z = splitmix64(seed, counters)
    return ((z >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24))).astype(np.float32)


def normal_f32(seed: int, n: int, sigma: float) -> np.ndarray:
    """N(0, sigma^2) via Box-Muller on the counter stream, rounded to FP32."""
    k = np.arange(2 * n, dtype=np.uint64)
    z = splitmix64(seed, k)
    u = ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / (1 << 53))
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    return (sigma * r * np.cos(2.0 * np.pi * u2)).astype(np.float32)


def wpoly6_f32(r, inv_h):
    """pbf/kernel/cuda/detail/pbf_kernel.inl:34-48 in FP32 (numpy)."""
    r = np.float32(r)
    inv_h = np.float32(inv_h)
    q = np.float32(r * inv_h)
    alpha = np.float32(np.float32(np.float32(np.float32(1.56668147065) * inv_h) * inv_h) * inv_h)
    if q <= np.float32(1.0):
        t = np.float32(np.float32(1.0) - np.float32(q * q))
        return np.float32(np.float32(np.float32(alpha * t) * t) * t)
    return np.float32(0.0)


def lattice_density_sum(h: float, space: float) -> np.float32:
    """Sum of poly6 weights over the ideal cubic lattice around an interior
    particle (offsets in z, y, x order, FP32 accumulation)."""
    h = np.float32(h)
    s = np.float32(space)
    inv_h = np.float32(np.float32(1.0) / h)
    total = np.float32(0.0)
    reach = int(math.ceil(float(h) / float(s))) + 1
    for c in range(-reach, reach + 1):
        for b in range(-reach, reach + 1):
            for a in range(-reach, reach + 1):
                r = np.float32(s * np.sqrt(np.float32(a * a + b * b + c * c)))
                if r < h:
                    total = np.float32(total + wpoly6_f32(r, inv_h))
    return total


def calc_grid_size(cell_size: float, domain_end: Sequence[float]) -> Tuple[int, int, int]:
    """pbf/pbf_type.cpp:141-162 (calcGridSize): smallest power of two per axis
    with g*cell_size > end (double comparisons); fails past 1028."""
    g = [1, 1, 1]
    end = [float(np.float32(e)) for e in domain_end]
    while True:
        if all(g[a] * cell_size > end[a] for a in range(3)):
            break
        g = [g[a] if g[a] * cell_size > end[a] else g[a] * 2 for a in range(3)]
        if any(x > 1028 for x in g):
            raise ValueError("grid_size is too large")
    return tuple(g)


def cartesian_volume(origin, n: Sequence[int], space: float) -> np.ndarray:
    """Lattice of n[0] x n[1] x n[2] points, x fastest (pbf_arrangement.h:82-100)."""
    nx, ny, nz = n
    s = np.float32(space)
    i = np.arange(nx, dtype=np.float32)
    j = np.arange(ny, dtype=np.float32)
    k = np.arange(nz, dtype=np.float32)
    K, J, I = np.meshgrid(k, j, i, indexing="ij")
    o = np.asarray(origin, dtype=np.float32)
    pts = np.stack(
        [
            np.float32(o[0]) + (s * I).astype(np.float32),
            np.float32(o[1]) + (s * J).astype(np.float32),
            np.float32(o[2]) + (s * K).astype(np.float32),
        ],
        axis=-1,
    ).reshape(-1, 3)
    return pts.astype(np.float32)


def box_planes(lo, hi) -> Tuple[np.ndarray, np.ndarray]:
    """Six inward planes in the order bottom, top, left, right, front, back
    (the order of the plane block in pbf_constraint.cu:283-296)."""
    lo = np.asarray(lo, dtype=np.float32)
    hi = np.asarray(hi, dtype=np.float32)
    pts = np.stack([lo, hi, lo, hi, lo, hi]).astype(np.float32)
    nrm = np.array(
        [[0, 1, 0], [0, -1, 0], [1, 0, 0], [-1, 0, 0], [0, 0, 1], [0, 0, -1]], dtype=np.float32
    )
    return pts, nrm


@dataclasses.dataclass
class StepParams:
    """Mirror of pbf_parameter + pbf_external (pbf/pbf_type.h:49-72) plus the
    switches the rebuilt step adds."""

    stable_density: float
    stable_distance: float
    particle_mass: float
    smoothing_length: float
    time_step: float
    relaxation: float
    xsph_parameter: float
    vc_parameter: float
    body_force: Tuple[float, float, float] = (0.0, -9.8, 0.0)
    scorr_mode: int = SCORR_PAPER
    scorr_k: float = 0.1
    scorr_n: int = 4
    scorr_dq_over_h: float = 0.2
    max_pair: int = 30
    xsph_on: int = 1
    vort_on: int = 1
    vort_h_as_inv_h: int = 0
    outer_as_inner: int = 0
    max_sweeps: int = 32


@dataclasses.dataclass
class Scene:
    name: str
    x: np.ndarray  # (N, 3) float32, AoS
    v: np.ndarray  # (N, 3) float32
    params: StepParams
    iters: int
    box_lo: np.ndarray
    box_hi: np.ndarray
    area_lo: np.ndarray  # deletion bounds (pbf_delete.cu:22-26)
    area_hi: np.ndarray
    grid: Tuple[int, int, int]
    plane_points: np.ndarray
    plane_normals: np.ndarray
    inner_spheres: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 4), np.float32))
    outer_spheres: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 4), np.float32))

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    @property
    def cells(self) -> int:
        return int(self.grid[0] * self.grid[1] * self.grid[2])


def dam_params(h: float, **overrides) -> StepParams:
    s = np.float32(0.5 * h)
    m = np.float32(float(s) ** 3 * 1000.0)
    sw = lattice_density_sum(h, float(s))
    rho0 = np.float32(m * sw)
    p = StepParams(
        stable_density=float(rho0),
        stable_distance=float(s),
        particle_mass=float(m),
        smoothing_length=float(np.float32(h)),
        time_step=float(np.float32(1.0 / 120.0)),
        relaxation=600.0,
        xsph_parameter=0.01,
        vc_parameter=0.0005,
    )
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def dam_break(
    name: str,
    h: float,
    box_hi: Sequence[float],
    blocks: Sequence[Tuple[Sequence[float], Sequence[int]]],
    iters: int,
    seed: int = 1,
    jitter: float = 0.01,
    v_sigma: Optional[float] = None,
    **param_overrides,
) -> Scene:
    """Lattice blocks (origin, (nx, ny, nz)) of spacing 0.5h in the box
    [0, box_hi], each point jittered by U(-jitter*h, +jitter*h) per component."""
    params = dam_params(h, **param_overrides)
    s = params.stable_distance
    xs: List[np.ndarray] = [cartesian_volume(o, n, s) for o, n in blocks]
    x = np.concatenate(xs).astype(np.float32)
    n = x.shape[0]
    if jitter:
        u = uniform01_f32(seed, np.arange(3 * n, dtype=np.uint64)).reshape(n, 3)
        amp = np.float32(np.float32(jitter) * np.float32(h))
        x = (x + ((np.float32(2.0) * u - np.float32(1.0)) * amp).astype(np.float32)).astype(np.float32)
    if v_sigma is not None:
        v = normal_f32(seed ^ 0x5DEECE66D, 3 * n, v_sigma).reshape(n, 3)
    else:
        v = np.zeros_like(x)
    lo = np.zeros(3, np.float32)
    hi = np.asarray(box_hi, dtype=np.float32)
    grid = calc_grid_size(2.0 * float(np.float32(h)), hi)
    pts, nrm = box_planes(lo, hi)
    return Scene(
        name=name,
        x=np.ascontiguousarray(x),
        v=np.ascontiguousarray(v.astype(np.float32)),
        params=params,
        iters=iters,
        box_lo=lo,
        box_hi=hi,
        area_lo=np.full(3, -NO_DELETION, np.float32),
        area_hi=np.full(3, NO_DELETION, np.float32),
        grid=grid,
        plane_points=pts,
        plane_normals=nrm,
    )


def corner_block(h: float, n: Sequence[int], origin=None):
    s = 0.5 * h
    o = origin if origin is not None else (0.5 * s, 0.5 * s, 0.5 * s)
    return (tuple(np.float32(c) for c in o), tuple(n))


def config0(seed: int = 1) -> Scene:
    """32,768 particles (32^3, spacing 0.5h, h = 0.1) in the corner of
    [0,3.2]x[0,3.2]x[0,1.6]; 4 Jacobi iterations."""
    h = 0.1
    return dam_break("dam32k_h0.1", h, (3.2, 3.2, 1.6), [corner_block(h, (32, 32, 32))], iters=4, seed=seed)


def config1(seed: int = 1) -> Scene:
    """1,048,576 particles (a 128x128x64 lattice, h = 0.05) in a
    256x128x128-cell box [0,12.8]x[0,6.4]x[0,6.4]; 4 Jacobi iterations.
    (The workload is quoted as 1,048,576 particles on a "128^3 lattice";
    128^3 is 2,097,152, so the count is kept and the lattice is 128 wide,
    128 tall and 64 deep.)"""
    h = 0.05
    return dam_break("dam1M_h0.05", h, (12.8, 6.4, 6.4), [corner_block(h, (128, 128, 64))], iters=4, seed=seed)


def config2(seed: int = 1) -> Scene:
    """8,388,608 particles: two 128x256x128 lattice blocks at the x-ends of a
    1024x256x256-cell box [0,25.6]x[0,6.4]x[0,6.4], h = 0.025; 3 iterations."""
    h = 0.025
    s = 0.5 * h
    a = corner_block(h, (128, 256, 128))
    bx = 25.6 - 128 * s + 0.5 * s
    b = corner_block(h, (128, 256, 128), origin=(bx, 0.5 * s, 0.5 * s))
    return dam_break("double_dam8M_h0.025", h, (25.6, 6.4, 6.4), [a, b], iters=3, seed=seed)


def config4(seed: int = 1) -> Scene:
    """65,536 particles (64x32x32, h = 0.1) with N(0, 0.1^2) initial velocity."""
    h = 0.1
    return dam_break("dam64k_h0.1_v0", h, (3.2, 3.2, 1.6), [corner_block(h, (64, 32, 32))], iters=4, seed=seed,
                     v_sigma=0.1)


def batch_config3(n_scenes: int = 256, seed: int = 3, n_min: int = 4096, n_max: int = 200000,
                  iters: int = 4) -> List[Scene]:
    """The batched workload: n_scenes independent dam breaks with particle
    counts log-uniform in [n_min, n_max] (seeded), each the 32k dam's
    geometry scaled to its count (an a x a x c lattice of spacing 0.5h,
    h = 0.1, in the corner of a box 2a x 2a x c lattice spacings wide), and
    per-scene rest density scale U[0.9, 1.1], XSPH c U[0.005, 0.05] and
    vorticity eps U[0, 0.001]."""
    h = 0.1
    u = uniform01_f32(seed, np.arange(4 * n_scenes, dtype=np.uint64)).astype(np.float64).reshape(n_scenes, 4)
    out = []
    for k in range(n_scenes):
        n = int(round(math.exp(math.log(n_min) + u[k, 0] * (math.log(n_max) - math.log(n_min)))))
        a = max(2, int(round(n ** (1.0 / 3.0))))
        c = max(1, int(round(n / (a * a))))
        s = 0.5 * h
        box = (2 * a * s, 2 * a * s, c * s + s)
        sc = dam_break(f"batch{k}_{a}x{a}x{c}", h, box, [corner_block(h, (a, a, c))], iters=iters,
                       seed=seed * 1000003 + k)
        p = sc.params
        rho_scale = 0.9 + 0.2 * u[k, 1]
        p = dataclasses.replace(p, stable_density=float(np.float32(p.stable_density * rho_scale)),
                                xsph_parameter=float(np.float32(0.005 + 0.045 * u[k, 2])),
                                vc_parameter=float(np.float32(0.001 * u[k, 3])))
        out.append(dataclasses.replace(sc, params=p))
    return out


def small_dam(nx: int = 12, ny: int = 10, nz: int = 8, h: float = 0.1, box=(1.6, 1.6, 0.8), iters: int = 4,
              seed: int = 7, **overrides) -> Scene:
    """A small dam break for fast parity checks."""
    return dam_break(f"dam{nx}x{ny}x{nz}", h, box, [corner_block(h, (nx, ny, nz))], iters=iters, seed=seed,
                     **overrides)


# ---------------------------------------------------------------------------
# The reference's arrangement generators (pbf/util/pbf_arrangement.h) and the
# initial conditions of its two simulators, in FP32 as it computes them
# (dom_dim = glm::vec3, scalar_t = float; pbf/pbf_type.h:18-27).

_EPS10 = 10.0 * np.finfo(np.float64).eps


def _count(rng, space) -> int:
    """static_cast<int>(range / space + epsilon * 10.0) + 1 (float division,
    double sum)."""
    return int(float(np.float32(np.float32(rng) / np.float32(space))) + _EPS10) + 1


def cartesian_volume_range(origin, end, space) -> np.ndarray:
    """cartesian_volume(pos, origin, end, space) (pbf_arrangement.h:82-100):
    points origin + space*(x, y, z) up to end, x fastest."""
    o = np.asarray(origin, np.float32)
    e = np.asarray(end, np.float32)
    n = [_count(np.float32(e[a] - o[a]), space) for a in range(3)]
    return cartesian_volume(o, n, space)


def cartesian_line_inclusive(origin, end, space) -> np.ndarray:
    """cartesian_line_inclusive (pbf_arrangement.h:8-18): floor(|end -
    origin| / space) + 1 points from origin towards end, spaced `space`."""
    o = np.asarray(origin, np.float32)
    e = np.asarray(end, np.float32)
    d = (e - o).astype(np.float32)
    length = np.float32(np.sqrt(np.float32(np.sum(d.astype(np.float32) * d, dtype=np.float32))))
    step = (d / length).astype(np.float32) * np.float32(space)
    num = int(math.floor(np.float32(length / np.float32(space)))) + 1
    return np.stack([(o + step * np.float32(i)).astype(np.float32) for i in range(num)]).astype(np.float32)


def cartesian_rectangle(axis: str, axis_value, origin, end, space) -> np.ndarray:
    """cartesian_rectangle (pbf_arrangement.h:123-162): a lattice in the
    plane axis = axis_value; origin/end are 2-D in (y, z) for "x", (z, x) for
    "y" and (x, y) for "z", with the reference's loop orders."""
    s = np.float32(space)
    av = np.float32(axis_value)
    o = np.asarray(origin, np.float32)
    e = np.asarray(end, np.float32)
    n0, n1 = _count(np.float32(e[0] - o[0]), s), _count(np.float32(e[1] - o[1]), s)
    pts = []
    if axis == "x":
        for z in range(n1):
            for y in range(n0):
                pts.append((av, o[0] + s * np.float32(y), o[1] + s * np.float32(z)))
    elif axis == "y":
        for z in range(n0):
            for x in range(n1):
                pts.append((o[1] + s * np.float32(x), av, o[0] + s * np.float32(z)))
    elif axis == "z":
        for x in range(n0):
            for y in range(n1):
                pts.append((o[0] + s * np.float32(x), o[1] + s * np.float32(y), av))
    else:
        raise ValueError(axis)
    return np.asarray(pts, np.float32).reshape(-1, 3)


def cartesian_cylinder(axis: str, axis_value, origin, radius, height, space) -> np.ndarray:
    """cartesian_cylinder (pbf_arrangement.h:210-241): rings of n points,
    theta = 2 asin(space / 2r), n = floor(2 pi / theta), every `space` along
    the axis up to `height` (FP32 angles and accumulation; cos/sin in FP32)."""
    r = np.float32(radius)
    s = np.float32(space)
    theta = np.float32(2.0 * math.asin(float(s) / (2.0 * float(r))))
    pi_f = np.float32(math.pi)
    n = np.float32(math.floor(2.0 * float(pi_f) / float(theta)))
    theta_hat = np.float32(2.0 * float(pi_f) / float(n))
    av = np.float32(axis_value)
    o = np.asarray(origin, np.float32)
    pts = []
    h = np.float32(0.0)
    while h < np.float32(height):
        i = np.float32(0.0)
        while i < n:
            a = np.float32(i * theta_hat)
            c, sn = np.float32(r * np.cos(a)), np.float32(r * np.sin(a))
            if axis == "x":
                pts.append((av + h, o[0] + c, o[1] + sn))
            elif axis == "y":
                pts.append((o[1] + sn, av + h, o[0] + c))
            else:
                pts.append((o[0] + c, o[1] + sn, av + h))
            i = np.float32(i + np.float32(1.0))
        h = np.float32(h + s)
    return np.asarray(pts, np.float32).reshape(-1, 3)


def reference_params(space: float, time_step: float, relaxation_scale: float, **overrides) -> StepParams:
    """The parameter derivation of the reference's simulators
    (simulator/dam/pbf_dam_init_cond.cpp:11-27,
    simulator/sphere/pbf_sphere_init_cond.cpp:11-27): rho0 = 1000,
    m = space^3 * rho0 (double), h = 1.4 space, relaxation = scale / space^2,
    XSPH c = 0.1, vorticity eps = 0.005; the step as shipped (REFERENCE
    s_corr, no XSPH or vorticity)."""
    s = np.float32(space)
    m = np.float32(float(s) ** 3 * 1000.0)
    p = StepParams(
        stable_density=1000.0,
        stable_distance=float(s),
        particle_mass=float(m),
        smoothing_length=float(np.float32(np.float32(1.4) * s)),
        time_step=float(np.float32(time_step)),
        relaxation=float(np.float32(np.float32(relaxation_scale) / np.float32(s * s))),
        xsph_parameter=float(np.float32(0.1)),
        vc_parameter=float(np.float32(0.005)),
        scorr_mode=SCORR_REFERENCE,
        xsph_on=0,
        vort_on=0,
    )
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def reference_dam(space: float = 0.16) -> Scene:
    """The reference's dam break (simulator/dam/pbf_dam_sim.cpp:10-66 with
    pbf_dam_init_cond.cpp): the fluid block (2.25)..(5.65) as a lattice of
    `space`, domain (1)..(10), the no-boundary step (the fixed sphere at
    (3,3,3), radius 3, pbf_constraint.cu:262-315), 3 iterations, dt 0.015.
    The simulation area is the domain: the shipped sim passes a zero-size
    area (a shadowed member), which would delete every particle."""
    p = reference_params(space, 0.015, 1.0)
    x = cartesian_volume_range((2.25, 2.25, 2.25), (5.65, 5.65, 5.65), space)
    lo = np.full(3, 1.0, np.float32)
    hi = np.full(3, 10.0, np.float32)
    return Scene(name=f"reference_dam_s{space}", x=x, v=np.zeros_like(x), params=p, iters=3,
                 box_lo=lo, box_hi=hi, area_lo=lo, area_hi=hi,
                 grid=calc_grid_size(2.0 * float(np.float32(p.smoothing_length)), hi),
                 plane_points=np.zeros((0, 3), np.float32), plane_normals=np.zeros((0, 3), np.float32),
                 inner_spheres=np.array([[3.0, 3.0, 3.0, 3.0]], np.float32))


JET_CAPACITY = 60000


def reference_sphere_jet(space: float = 0.1) -> Tuple[Scene, np.ndarray, np.ndarray]:
    """The reference application's scene (main.cpp:81, pbf_sphere_sim(0.1);
    simulator/sphere/pbf_sphere_sim.cpp:281-327, pbf_sphere_init_cond.cpp):
    an inner sphere (3,3,3), radius 3, as the boundary, domain (0.05)..(7),
    dt 0.007, relaxation 2/space^2, 3 iterations, and a jet: the rectangle
    x = 1.25, (y, z) in (3.5, 2.5)..(4.5, 3.5) at `space`, velocity
    (2.8, 0, 0).  Returns the scene (holding one jet batch, the state after
    construction), and the jet's positions and velocities; see
    jet_emits()."""
    p = reference_params(space, 0.007, 2.0)
    jx = cartesian_rectangle("x", 1.25, (3.5, 2.5), (4.5, 3.5), space)
    jv = np.tile(np.array([2.8, 0.0, 0.0], np.float32), (jx.shape[0], 1))
    lo = np.full(3, 0.05, np.float32)
    hi = np.full(3, 7.0, np.float32)
    sc = Scene(name=f"reference_sphere_jet_s{space}", x=jx.copy(), v=jv.copy(), params=p, iters=3,
               box_lo=lo, box_hi=hi, area_lo=lo, area_hi=hi,
               grid=calc_grid_size(2.0 * float(np.float32(p.smoothing_length)), hi),
               plane_points=np.zeros((0, 3), np.float32), plane_normals=np.zeros((0, 3), np.float32),
               inner_spheres=np.array([[3.0, 3.0, 3.0, 3.0]], np.float32))
    return sc, jx, jv


def jet_emits(step: int, num: int, capacity: int = JET_CAPACITY) -> bool:
    """Whether the jet adds a batch before step `step` (1-based: the counter
    is raised before the test, pbf_sphere_sim.cpp:357-387): every 8th step
    while fewer than 80% of the capacity are live."""
    return step % 8 == 0 and num < capacity * 0.8


CONFIGS = {0: config0, 1: config1, 2: config2, 4: config4}
