"""The differentiable rollout (diff_2d_fem_amd.autograd.Rollout, a torch.autograd.Function
over the GPU adjoint) on the pneumatic fingers of the adjoint configuration
(diff_2d_fem_amd.workloads.pneumatic_fingers: chambered fingers under pressure, base held), with
the configuration's loss |tip - target|^2 summed over the fingers and gradients with respect
to the chamber pressures and every element's modulus scale.

* On the same trajectory, the GPU's dJ/dscale and dJ/dp equal the float64 oracle's
  (OracleAdjoint with the pressure tangent, scipy MINRES) to 1e-7 relative.
* Against central differences of GPU rollouts, along the gradient, within 1e-4 relative (the
  protocol of reference tests/test_adjoint.py: float64, dt 2e-4; pressure steps of 50 Pa).
* Checkpointing: recomputing segments of 3 steps gives the gradient of a rollout kept whole.
"""
import numpy as np
import pytest
import torch

from diff_2d_fem_amd import ImplicitIntegrator
from diff_2d_fem_amd import workloads as W
from diff_2d_fem_amd.autograd import Rollout
from oracle.adjoint import OracleAdjoint
from oracle.pncg import OracleIntegrator

pytestmark = pytest.mark.gpu

T, DT = 10, 2.0e-4
KW = dict(max_iter=4000, eps=1e-15, tol_x=1e-16, conv_check=8, self_collision=False, friction=False)


def _fingers(n=4):
    w = W.pneumatic_fingers(n)
    m = w["model"]
    tips = np.array([m.body_nodes(h)[np.argmax(m.rest_x(h)[:, 0])] for h in w["bodies"]])
    p0 = np.array([c.pressure for c in m.actuation.chambers])
    return w, m, tips, p0


def _loss(x, tips, targets):
    return ((x[tips] - targets) ** 2).sum()


def test_rollout_gradient_equals_oracle_on_the_same_trajectory():
    w, m, tips, p0 = _fingers()
    targets = torch.as_tensor(w["targets"], device="cuda")
    integ = ImplicitIntegrator(path="scene", precision="double", **KW)
    roll = Rollout(m, integ, DT, T, every=T)
    rng = np.random.default_rng(1)
    scale = torch.tensor(0.8 + 0.4 * rng.random(m.n_elements), requires_grad=True)
    p = torch.tensor(p0, requires_grad=True)
    x = roll(scale, p)
    loss = _loss(x, torch.as_tensor(tips, device="cuda"), targets)
    loss.backward()
    assert getattr(roll.adj, "minres_fail", 0) == 0
    # the same trajectory on the GPU, kept whole, for the oracle
    m.reset()
    roll2 = Rollout(m, integ, DT, T, every=1)
    roll2.forward(scale, p)
    xs = [roll2.checkpoints[k][0].cpu().numpy() for k in range(T)] + [roll2.final[0].cpu().numpy()]
    oi = OracleIntegrator(**KW)
    oi.attach(m, DT)
    oa = OracleAdjoint(oi.solver, m.me, Emin=0.0)
    oa.set_density(scale.detach().numpy())
    seed = np.zeros((m.n_nodes, 2))
    seed[tips] = 2.0 * (xs[-1][tips] - w["targets"])
    go = oa.transient_adjoint(xs, {T: seed}, oi.solver.drag, rtol=1e-14)
    gpo = oa.dJdp_steps.sum(0)
    e1 = np.abs(scale.grad.numpy() - go).max() / np.abs(go).max()
    e2 = np.abs(p.grad.numpy() - gpo).max() / np.abs(gpo).max()
    print(f"fingers: dJ/dscale vs oracle {e1:.2e}, dJ/dp vs oracle {e2:.2e}")
    assert e1 < 1e-7 and e2 < 1e-7


def test_rollout_gradient_matches_finite_differences():
    w, m, tips, p0 = _fingers()
    targets = torch.as_tensor(w["targets"], device="cuda")
    tip_t = torch.as_tensor(tips, device="cuda")
    integ = ImplicitIntegrator(path="scene", precision="double", **KW)
    roll = Rollout(m, integ, DT, T, every=4)
    rng = np.random.default_rng(2)
    s0 = 0.8 + 0.4 * rng.random(m.n_elements)
    scale = torch.tensor(s0, requires_grad=True)
    p = torch.tensor(p0, requires_grad=True)
    m.reset()
    _loss(roll(scale, p), tip_t, targets).backward()
    gs, gp = scale.grad.numpy().copy(), p.grad.numpy().copy()

    def J(sv, pv):
        m.reset()
        with torch.no_grad():
            return float(_loss(roll(torch.tensor(sv), torch.tensor(pv)), tip_t, targets))
    d = gs / np.abs(gs).max()
    fd = (J(s0 + 1e-3 * d, p0) - J(s0 - 1e-3 * d, p0)) / 2e-3
    rel_s = abs(gs @ d - fd) / abs(fd)
    dp = gp / np.abs(gp).max()
    fdp = (J(s0, p0 + 50.0 * dp) - J(s0, p0 - 50.0 * dp)) / 100.0
    rel_p = abs(gp @ dp - fdp) / abs(fdp)
    print(f"fingers: along the gradient, scale {rel_s:.2e}, pressure {rel_p:.2e}; "
          f"recomputed steps {roll.recomputed_steps}")
    assert rel_s < 1e-4 and rel_p < 1e-4
    assert roll.recomputed_steps > 0


def test_checkpointing_gives_the_whole_rollout_gradient():
    w, m, tips, p0 = _fingers(2)
    targets = torch.as_tensor(w["targets"], device="cuda")
    tip_t = torch.as_tensor(tips, device="cuda")
    integ = ImplicitIntegrator(path="scene", precision="double", **KW)
    grads = []
    for every in (T, 3):
        m.reset()
        roll = Rollout(m, integ, DT, T, every=every)
        scale = torch.ones(m.n_elements, dtype=torch.float64, requires_grad=True)
        p = torch.tensor(np.tile(p0, (T, 1)), requires_grad=True)       # a pressure per step
        _loss(roll(scale, p), tip_t, targets).backward()
        grads.append((scale.grad.numpy().copy(), p.grad.numpy().copy()))
        assert p.grad.shape == (T, len(p0))
    (a, pa), (b, pb) = grads
    print(f"checkpoints every 3 vs whole: {np.abs(a - b).max() / np.abs(a).max():.2e}")
    assert np.abs(a - b).max() <= 1e-9 * np.abs(a).max()
    assert np.abs(pa - pb).max() <= 1e-9 * np.abs(pa).max()
