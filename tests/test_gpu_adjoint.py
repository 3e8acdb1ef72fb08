"""The differentiable form on the GPU (pbfx_adj_*): the forward it records is
the oracle's step bit for bit, and its FP64 gradients -- recomputed window by
window from checkpoints -- agree with the oracle's FP64 adjoint evaluated at
the same recorded values (the two differ in summation order only).
Gradients are compared as (values, power-of-two exponent) pairs, normwise."""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from oracle import pyoracle as orc
from ssf_pbf_cuda_amd import scenes
from tests.adjoint_cases import Rollout, autograd_fn, cases, linear_loss_weights, rel, rel_norm

CASES = cases()
# GPU FP64 adjoint vs oracle FP64 adjoint on identical values, normwise
# (they differ in summation order only)
TOL = 1e-9


def _gpu_grad(sc, steps, every, wx_fn):
    import torch

    roll = Rollout(sc, steps, every)
    x = torch.from_numpy(sc.x).cuda()
    v = torch.from_numpy(sc.v).cuda()
    n = roll.solver.adj_forward(x, v, sc.n, sc.area_lo, sc.area_hi, sc.iters, steps)
    ident = torch.empty(n, dtype=torch.int32, device="cuda")
    roll.solver.adj_identity(ident)
    xs, vs = x[:n].cpu().numpy(), v[:n].cpu().numpy()
    wx, wv = wx_fn(xs, ident.cpu().numpy().astype(np.int64))
    gx = torch.from_numpy(np.ascontiguousarray(wx, np.float64)).cuda()
    gv = torch.from_numpy(np.ascontiguousarray(wv, np.float64)).cuda()
    gx0 = torch.empty(x.shape, dtype=torch.float64, device="cuda")
    gv0 = torch.empty_like(gx0)
    gc, ge, e = roll.solver.adj_backward(gx, gv, gx0, gv0)
    out = dict(x=xs, v=vs, ident=ident.cpu().numpy(), gx0=gx0.cpu().numpy(), gv0=gv0.cpu().numpy(), gc=gc, ge=ge,
               e=e, wx=wx, wv=wv)
    roll.solver.close()
    return out


def _cpu_grad(sc, steps, wx, wv):
    T = orc.Tape(sc)
    x, v = T.run(steps)
    gx0, gv0, gc, ge, e = T.backward(wx, wv, with_exponent=True)
    return dict(x=x, v=v, ident=T.identity(), gx0=gx0, gv0=gv0, gc=gc, ge=ge, e=e)


def _compare(tag, g, c):
    """Normwise errors of the GPU gradients against the FP64 ones, with the
    exponents aligned."""
    d = g["e"] - c["e"]
    out = {
        "x0": rel_norm(np.ldexp(g["gx0"], d), c["gx0"]),
        "v0": rel_norm(np.ldexp(g["gv0"], d), c["gv0"]),
        "c": rel(np.ldexp(g["gc"], d), c["gc"]) if c["gc"] else abs(g["gc"]),
        "eps": rel(np.ldexp(g["ge"], d), c["ge"]) if c["ge"] else abs(g["ge"]),
    }
    print(tag, f"exp gpu {g['e']} cpu {c['e']}", {k: f"{x:.2e}" for k, x in out.items()}, flush=True)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_adjoint_matches_fp64_adjoint(name):
    sc, steps = CASES[name]
    every = 2 if steps > 1 else 1
    print(f"adjoint {name}: n={sc.n} steps={steps} every={every}", flush=True)
    g = _gpu_grad(sc, steps, every, lambda xs, ident: linear_loss_weights(xs.shape[0]))
    c = _cpu_grad(sc, steps, g["wx"], g["wv"])
    assert np.array_equal(g["x"].view(np.uint32), c["x"].view(np.uint32))
    assert np.array_equal(g["v"].view(np.uint32), c["v"].view(np.uint32))
    assert np.array_equal(g["ident"], c["ident"])
    err = _compare(name, g, c)
    for k, e in err.items():
        assert e < TOL, (name, k, e)


@pytest.mark.gpu
def test_checkpoint_interval_does_not_change_the_gradient():
    # recomputation from checkpoints replays the forward bit for bit, so the
    # gradient is the same whatever the window
    sc, steps = CASES["calm"]
    w = lambda xs, ident: linear_loss_weights(xs.shape[0])
    a = _gpu_grad(sc, steps, 1, w)
    b = _gpu_grad(sc, steps, steps, w)
    print("every=1 vs every=steps", a["gc"], b["gc"], flush=True)
    assert a["e"] == b["e"]
    for k in ("gx0", "gv0"):
        assert np.array_equal(a[k], b[k]), k
    assert a["gc"] == b["gc"] and a["ge"] == b["ge"]


@pytest.mark.gpu
def test_autograd_wrapper_against_fp64_adjoint():
    # L = mean |x_final - x_target|^2 by particle, the target run made with
    # XSPH c = 0.02 and vorticity epsilon = 0.0008; gradients through
    # torch.autograd with respect to x0, v0, c and epsilon
    import torch

    sc, steps = CASES["calm"]
    tgt = dataclasses.replace(sc, params=dataclasses.replace(sc.params, xsph_parameter=0.02, vc_parameter=0.0008))
    Tt = orc.Tape(tgt)
    xt, _ = Tt.run(steps)
    target = np.empty_like(xt)
    target[Tt.identity()] = xt
    print(f"autograd: n={sc.n} steps={steps}", flush=True)
    roll = Rollout(sc, steps, 2)
    x0 = torch.from_numpy(sc.x).cuda().requires_grad_(True)
    v0 = torch.from_numpy(sc.v).cuda().requires_grad_(True)
    c = torch.tensor(float(np.float32(sc.params.xsph_parameter)), dtype=torch.float64, requires_grad=True)
    eps = torch.tensor(float(np.float32(sc.params.vc_parameter)), dtype=torch.float64, requires_grad=True)
    xf = autograd_fn().apply(x0, v0, c, eps, roll)
    loss = ((xf - torch.from_numpy(target).double().cuda()) ** 2).sum() / sc.n
    loss.backward()
    roll.solver.close()
    # the same loss through the FP64 adjoint
    T = orc.Tape(sc)
    x, _ = T.run(steps)
    ident = T.identity()
    gx = 2.0 / sc.n * (x.astype(np.float64) - target[ident])
    gx0, gv0, gc, ge = T.backward(gx, np.zeros_like(gx))
    errs = {"x0": rel_norm(x0.grad.cpu().numpy(), gx0), "v0": rel_norm(v0.grad.cpu().numpy(), gv0),
            "c": rel(c.grad.item(), gc), "eps": rel(eps.grad.item(), ge)}
    print("autograd", f"loss {loss.item():.4e}", {k: f"{e:.2e}" for k, e in errs.items()}, flush=True)
    # x0.grad and v0.grad are stored as float32 (the dtype of x0, v0)
    tol = {"x0": 1e-6, "v0": 1e-6, "c": 1e-9, "eps": 1e-9}
    for k, e in errs.items():
        assert e < tol[k], (k, e)


@pytest.mark.gpu
def test_workload_65k_40_steps_against_fp64_adjoint():
    # The differentiable workload: 65,536 particles (64x32x32, h = 0.1,
    # v0 ~ N(0, 0.1^2)), 40 steps with checkpoints every 10; loss = mean
    # squared distance to the run with c = 0.02, epsilon = 0.0008.  The flow
    # explodes in its first steps and the adjoint grows by ~1e8 per step, so
    # both sides carry power-of-two exponents; what is compared is the
    # exponent and the normalised gradient.
    sc = scenes.config4()
    steps, every = 40, 10
    tgt = dataclasses.replace(sc, params=dataclasses.replace(sc.params, xsph_parameter=0.02, vc_parameter=0.0008))
    Tt = orc.Tape(tgt)
    xt, _ = Tt.run(steps)
    target = np.empty_like(xt)
    target[Tt.identity()] = xt
    print(f"workload adjoint: n={sc.n} steps={steps} every={every}", flush=True)

    def weights(xs, ident):
        gx = 2.0 / sc.n * (xs.astype(np.float64) - target[ident])
        return gx, np.zeros_like(gx)

    g = _gpu_grad(sc, steps, every, weights)
    c = _cpu_grad(sc, steps, g["wx"], g["wv"])
    assert np.array_equal(g["x"].view(np.uint32), c["x"].view(np.uint32))
    err = _compare("workload", g, c)
    assert np.all(np.isfinite(g["gx0"])) and np.all(np.isfinite(g["gv0"]))
    assert abs(g["e"] - c["e"]) <= 1
    # two FP64 evaluations in different summation orders agree to ~1e-9 on
    # this workload (its conditioning is ~1e7)
    for k, e in err.items():
        assert e < 1e-5, (k, e)
