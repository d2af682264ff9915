"""The GPU adjoint (diff_2d_fem_amd.adjoint, reference implicit/adjoint.py) and the contact
read-outs (reference implicit/solver.py:504-612).

* On the same trajectory, the GPU's dJ/drho equals the float64 oracle's (OracleAdjoint:
  the reference's operator with scipy's MINRES) to 1e-7 relative, for an unconstrained
  body, a disc in contact and a box in contact.
* Against central differences of GPU rollouts (the protocol of reference
  tests/test_adjoint.py, float64 with a small step): along the gradient within 1e-4
  relative, along a random direction within 1e-5 of |g| |d|.
* contact_force / collider_force / body_contact_force against sums the oracle computes.
"""
import numpy as np
import pytest

from diff_2d_fem_amd import ImplicitIntegrator
from oracle import terms as OT
from oracle.adjoint import OracleAdjoint
from oracle.pncg import OracleIntegrator
from tests.scenes import pressed_block, press_targets, grip, close_jaws

pytestmark = pytest.mark.gpu

T, DT = 6, 2.0e-4
KW = dict(max_iter=3000, eps=1e-14, tol_x=1e-10, conv_check=8, self_collision=False, friction=False)


def _setup(mode, path):
    m, h, x0, top = pressed_block(mode)
    integ = ImplicitIntegrator(path=path, precision="double", **KW)
    adj = integ.adjoint(m, DT)
    rng = np.random.default_rng(0)
    rho0 = 0.45 + 0.5 * rng.random(adj.n_tri)
    low = np.where(x0[:, 1] < np.median(x0[:, 1]))[0]
    c = np.zeros((m.n_nodes, 2))
    c[low, 0] = 1.0 / len(low)

    def rollout(rho):
        adj.set_density(rho)
        m.reset()
        xs = [m.get_x()]
        for t in range(1, T + 1):
            m.set_kinematic_target(h, press_targets(x0, top, t, T))
            integ.simulate(m, DT)
            xs.append(m.get_x())
        return xs
    return m, integ, adj, rho0, c, rollout, rng


@pytest.mark.parametrize("mode", ["elastic", "disc", "box"])
def test_adjoint_equals_oracle_on_the_same_trajectory(mode):
    m, integ, adj, rho0, c, rollout, _ = _setup(mode, "scene")
    xs = rollout(rho0)
    use = dict(use_disc=(mode == "disc"), use_sdf=(mode == "box"))
    if mode != "elastic":
        assert np.linalg.norm(integ.collider_force(m, friction=False)[0]) > 1e-2, "no contact"
    g = adj.transient_adjoint(xs, seed={T: c}, rtol=1e-13, maxiter=20000, **use)
    assert getattr(adj, "minres_fail", 0) == 0
    oi = OracleIntegrator(**KW)
    oi.attach(m, DT)
    oa = OracleAdjoint(oi.solver, m.me)
    oa.set_density(rho0)
    go = oa.transient_adjoint(xs, {T: c}, oi.solver.drag, rtol=1e-13, **use)
    err = np.abs(g - go).max() / np.abs(go).max()
    print(f"{mode}: |g_gpu - g_oracle| / |g| = {err:.2e}")
    assert err < 1e-7


@pytest.mark.parametrize("mode", ["elastic", "disc", "box"])
def test_adjoint_matches_finite_differences(mode):
    m, integ, adj, rho0, c, rollout, rng = _setup(mode, "scene")
    use = dict(use_disc=(mode == "disc"), use_sdf=(mode == "box"))
    xs = rollout(rho0)
    g = adj.transient_adjoint(xs, seed={T: c}, rtol=1e-13, maxiter=20000, **use)
    J = lambda rho: float((c * rollout(rho)[-1]).sum())
    hfd = 0.002
    d = g / np.abs(g).max()
    fd = (J(rho0 + hfd * d) - J(rho0 - hfd * d)) / (2 * hfd)
    rel = abs(g @ d - fd) / abs(fd)
    d2 = rng.standard_normal(adj.n_tri)
    fd2 = (J(rho0 + hfd * d2) - J(rho0 - hfd * d2)) / (2 * hfd)
    nerr = abs(g @ d2 - fd2) / (np.linalg.norm(g) * np.linalg.norm(d2))
    print(f"{mode}: along the gradient {rel:.2e}, random direction {nerr:.2e}")
    assert rel < 1e-4
    assert nerr < 1e-5


def test_adjoint_refuses_terms_without_a_tangent():
    m, pads = grip(n_env=1)
    integ = ImplicitIntegrator(path="scene", precision="double")
    adj = integ.adjoint(m, 1e-3)
    with pytest.raises(NotImplementedError):
        adj.transient_adjoint([m.get_x(), m.get_x()], seed={1: np.zeros((m.n_nodes, 2))})


@pytest.mark.parametrize("path", ["plain", "scene"])
def test_contact_read_outs(path):
    """collider_force and contact_force against the oracle's barrier/friction terms at the
    GPU's state, in the grip scene (colliders and IPC contact between bodies active)."""
    m, pads = grip(n_env=2)
    integ = ImplicitIntegrator(path=path, precision="double", max_iter=60, eps=0.0, conv_check=0)
    for k in range(1, 13):
        close_jaws(m, pads, k, n_close=8)
        integ.simulate(m, 1e-3)
    s = m._solver
    x, xn = s.get_x(), s.xn.cpu().numpy()
    oi = OracleIntegrator(max_iter=60, eps=0.0, conv_check=0)
    oi.attach(m, 1e-3)
    os_ = oi.solver
    os_.x, os_.xn = x.copy(), xn.copy()
    # the colliders' barrier + friction on every node (h^2 = 1, every node free)
    g = np.zeros_like(x)
    dg = np.zeros_like(x)
    ones = np.ones_like(x)
    for col in os_.colliders:
        OT.collider_grad_diag(col, x, os_.bnd, os_.dhat, os_.kappa, 1.0, 1, g, dg)
        OT.friction_grad_diag(col, x, xn, os_.bnd, os_.dhat, os_.kappa, col["mu"], os_.epsv,
                              os_.inv_dt, 1.0, ones, g, dg)
    f_col = integ.contact_force(m, bodies=False)
    assert np.abs(f_col + g).max() <= 1e-9 * np.abs(g).max()
    # between bodies: IPC gradient with h^2 = 1 on every node
    gi = np.zeros_like(x)
    os_.ipc.force(os_, gi, np.zeros_like(x), ones, friction=True)
    f_b = integ.contact_force(m, colliders=False)
    assert np.abs(gi).max() > 0.0
    assert np.abs(f_b + gi).max() <= 1e-9 * np.abs(gi).max()
    f_all = integ.contact_force(m)
    assert np.abs(f_all - f_col - f_b).max() <= 1e-9 * np.abs(f_all).max()
    # the reaction on the ground per environment equals minus the barrier force on the nodes
    F = integ.collider_force(m, friction=False)
    fn = integ.contact_force(m, bodies=False, friction=False)
    env = np.asarray(m.me.node2env)
    for e in range(m.n_envs):
        assert np.allclose(F[e], -fn[env == e].sum(0), rtol=1e-9, atol=1e-12)
    assert abs(F[:, 1]).min() > 0.0
    bcf = integ.body_contact_force(m, pads[0][0])
    assert np.allclose(bcf, f_all[m.body_nodes(pads[0][0])].sum(0))
