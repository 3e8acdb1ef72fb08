"""The FP64 adjoint of the oracle (oracle/pbf_adjoint.cpp), pinned by central
finite differences of the step map with its discrete decisions frozen (the
map the adjoint differentiates), and the tape's FP32 run pinned bit for bit
to the oracle step.  Reference step: pbf::one_step
(pbf/solver/pbf_solve.cpp:182-216) in the full update order of :135-142."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import pyoracle as orc
from tests.adjoint_cases import calm, cases, linear_loss_weights, rel
from tests.helpers import oracle_run

CASES = cases()


@pytest.mark.parametrize("name", sorted(CASES))
def test_tape_run_is_the_oracle_step(name):
    sc, steps = CASES[name]
    T = orc.Tape(sc)
    x, v = T.run(steps)
    ref = oracle_run(sc, steps=steps)
    assert T.steps == steps and x.shape[0] == ref["num"]
    assert np.array_equal(x.view(np.uint32), ref["x"].view(np.uint32))
    assert np.array_equal(v.view(np.uint32), ref["v"].view(np.uint32))
    ident = T.identity()
    assert sorted(ident.tolist()) == list(range(sc.n))


def test_fp64_replay_tracks_the_fp32_run():
    sc, steps = CASES["calm"]
    T = orc.Tape(sc)
    x, v = T.run(steps)
    xo, vo = T.replay(sc.x.astype(np.float64), sc.v.astype(np.float64))
    assert np.abs(xo - x).max() < 1e-5 and np.abs(vo - v).max() < 1e-3


def _fd_setup(sc, steps):
    T = orc.Tape(sc)
    T.run(steps)
    x0 = sc.x.astype(np.float64)
    v0 = sc.v.astype(np.float64)
    c0 = [float(np.float32(sc.params.xsph_parameter)), float(np.float32(sc.params.vc_parameter))]
    xo, _ = T.replay(x0, v0, record=True)
    wx, wv = linear_loss_weights(xo.shape[0])

    def L(c=None, x=x0, v=v0):
        a, b = T.replay(x, v, coef=c if c is not None else c0)
        return float((wx * a).sum() + (wv * b).sum())

    return T, x0, v0, c0, wx, wv, L


@pytest.mark.parametrize("name", sorted(CASES))
def test_adjoint_matches_central_differences(name):
    # central differences of the FP64 frozen-decision map at two steps (the
    # better of truncation and rounding), against the FP64 adjoint at the
    # same values; the exploding case resolves only c and epsilon (its state
    # derivatives are far outside the linear regime at any FD step)
    sc, steps = CASES[name]
    T, x0, v0, c0, wx, wv, L = _fd_setup(sc, steps)
    gx0, gv0, gc, ge = T.backward(wx, wv)
    u = np.random.default_rng(9).standard_normal(x0.shape)
    errs = {"c": [], "eps": [], "v0": [], "x0": []}
    for h in (1e-5, 1e-7):
        fc = (L([c0[0] * (1 + h), c0[1]]) - L([c0[0] * (1 - h), c0[1]])) / (2 * h * c0[0])
        errs["c"].append(rel(gc, fc) if sc.params.xsph_on else abs(gc))
        if c0[1] and sc.params.vort_on:
            fe = (L([c0[0], c0[1] * (1 + h)]) - L([c0[0], c0[1] * (1 - h)])) / (2 * h * c0[1])
            errs["eps"].append(rel(ge, fe))
        else:
            errs["eps"].append(abs(ge))
        fv = (L(v=v0 + h * u) - L(v=v0 - h * u)) / (2 * h)
        errs["v0"].append(rel((gv0 * u).sum(), fv))
        fx = (L(x=x0 + h * u) - L(x=x0 - h * u)) / (2 * h)
        errs["x0"].append(rel((gx0 * u).sum(), fx))
    best = {k: min(e) for k, e in errs.items()}
    print(name, steps, {k: f"{e:.1e}" for k, e in best.items()}, flush=True)
    checked = ("c", "eps") if name.startswith("exploding") else ("c", "eps", "v0", "x0")
    for k in checked:
        assert best[k] < 1e-4, (name, k, best[k])


def test_parameter_gradients_at_the_workload_fd_step():
    # step 1e-3 (relative) on c and epsilon over 5 steps of the calm dam
    sc, steps = calm(), 5
    T, x0, v0, c0, wx, wv, L = _fd_setup(sc, steps)
    _, _, gc, ge = T.backward(wx, wv)
    h = 1e-3
    fc = (L([c0[0] * (1 + h), c0[1]]) - L([c0[0] * (1 - h), c0[1]])) / (2 * h * c0[0])
    fe = (L([c0[0], c0[1] * (1 + h)]) - L([c0[0], c0[1] * (1 - h)])) / (2 * h * c0[1])
    assert rel(gc, fc) < 1e-6 and rel(ge, fe) < 1e-5, (rel(gc, fc), rel(ge, fe))


def test_adjoint_is_linear_in_the_seed():
    # dL/dstate of (a L1 + b L2) = a dL1 + b dL2 (the recursion is linear;
    # catches state leaking between calls)
    sc, steps = CASES["calm"]
    T = orc.Tape(sc)
    T.run(steps)
    n = T.num(steps - 1)
    w1 = linear_loss_weights(n, seed=1)
    w2 = linear_loss_weights(n, seed=2)
    g1 = T.backward(*w1)
    g2 = T.backward(*w2)
    g = T.backward(2.0 * w1[0] - 3.0 * w2[0], 2.0 * w1[1] - 3.0 * w2[1])
    for a, b, c in zip(g1[:2], g2[:2], g[:2]):
        assert np.allclose(c, 2.0 * a - 3.0 * b, rtol=1e-9, atol=1e-9 * np.abs(c).max())
    assert g[2] == pytest.approx(2.0 * g1[2] - 3.0 * g2[2], rel=1e-9)
