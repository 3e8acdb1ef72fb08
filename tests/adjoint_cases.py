"""Scenes, losses and the torch.autograd wrapper of the differentiable form,
shared by the adjoint tests.  Test harness only: the library's C-ABI
(pbfx_adj_*) is the product; torch appears here to show the gradients
plugging into autograd."""
from __future__ import annotations

import dataclasses

import numpy as np

from ssf_pbf_cuda_amd import scenes


def _with(sc, **kw):
    return dataclasses.replace(sc, params=dataclasses.replace(sc.params, **kw))


def calm(**kw):
    """A 960-particle dam with N(0, 0.1^2) initial velocities and a weak
    tensile correction (k = 0.001): a flow whose step map stays in its linear
    regime over many steps, so central differences resolve its derivative."""
    return _with(scenes.small_dam(12, 10, 8, v_sigma=0.1), scorr_k=0.001, **kw)


def cases():
    """name -> (scene, steps): the variants the adjoint is checked on."""
    sph = dataclasses.replace(calm(), inner_spheres=np.array([[0.8, 0.8, 0.8, 0.75]], np.float32),
                              outer_spheres=np.array([[0.3, 0.3, 0.3, 0.2]], np.float32))
    return {
        "calm": (calm(), 5),
        "reference_scorr": (calm(scorr_mode=scenes.SCORR_REFERENCE, relaxation=6000.0), 3),
        "spheres": (sph, 2),
        "no_final_passes": (calm(xsph_on=0, vort_on=0), 3),
        "vorticity_only": (calm(xsph_on=0), 3),
        "exploding_1step": (scenes.small_dam(12, 10, 8, v_sigma=0.1), 1),
    }


def linear_loss_weights(n, seed=5):
    """Weights of L = sum wx . x_final + sum wv . v_final (final slot order)."""
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, 3)), rng.standard_normal((n, 3))


def rel_norm(got, ref):
    got = np.asarray(got, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    den = max(np.linalg.norm(ref), 1e-300)
    return float(np.linalg.norm(got - ref) / den)


def rel(got, ref):
    return abs(float(got) - float(ref)) / max(abs(float(ref)), 1e-300)


class Rollout:
    """A solver context with the differentiable form enabled."""

    def __init__(self, scene, steps, every):
        from ssf_pbf_cuda_amd import pbfx

        self.scene = scene
        self.steps = steps
        self.solver = pbfx.Solver.for_scene(scene)
        self.solver.adj_enable(steps, every)


def autograd_fn():
    """torch.autograd.Function over pbfx_adj_forward / pbfx_adj_backward:
    (x0, v0, c, eps) -> final positions indexed by initial particle."""
    import torch

    class PbfRollout(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x0, v0, c, eps, roll):
            sc = roll.scene
            p = dataclasses.replace(sc.params, xsph_parameter=float(c), vc_parameter=float(eps))
            roll.solver.set_params(p)
            x = x0.detach().clone().contiguous()
            v = v0.detach().clone().contiguous()
            n = roll.solver.adj_forward(x, v, x.shape[0], sc.area_lo, sc.area_hi, sc.iters, roll.steps)
            ident = torch.empty(n, dtype=torch.int32, device=x.device)
            roll.solver.adj_identity(ident)
            # float64, so that dL/d(output) reaches the FP64 adjoint unrounded
            out = torch.full(x0.shape, float("nan"), dtype=torch.float64, device=x0.device)
            out[ident.long()] = x[:n].double()
            ctx.roll, ctx.ident, ctx.n = roll, ident, n
            ctx.shape, ctx.dtype = x0.shape, x0.dtype
            return out

        @staticmethod
        def backward(ctx, g):
            roll, ident = ctx.roll, ctx.ident.long()
            gx = g[ident].contiguous().double()
            gv = torch.zeros_like(gx)
            gx0 = torch.empty(ctx.shape, dtype=torch.float64, device=g.device)
            gv0 = torch.empty_like(gx0)
            gc, ge, e = roll.solver.adj_backward(gx, gv, gx0, gv0)
            s = 2.0 ** e
            return (gx0 * s).to(ctx.dtype), (gv0 * s).to(ctx.dtype), \
                torch.tensor(gc * s, dtype=torch.float64), torch.tensor(ge * s, dtype=torch.float64), None

    return PbfRollout
