"""Pins of the oracle adjoint's tangents beyond the reference (oracle/adjoint.py): the IPC
barrier between bodies and the affine-body reduction.

* IPCContact.hvp equals central differences of the barrier gradient (friction off), which
  tests/test_oracle.py pins to the barrier energy, in a flat contact (interior feet only) and
  with a block corner over a slab node (end terms with their vertex-correction coefficients).
* The whole adjoint (IPC tangent, with and without a rigid region in the block) equals central
  differences of converged oracle rollouts of tests.scenes.ipc_press along the gradient
  (< 1e-4) and along a random direction (< 1e-5 of |g| |d|); without the IPC tangent it is
  off by more than 1e-3, so the check sees the term, and so is the adjoint that leaves the rigid
  region's nodes free instead of reducing them to the region's q.
"""
import numpy as np
import pytest

from oracle.adjoint import OracleAdjoint
from oracle.pncg import OracleIntegrator
from tests.scenes import ipc_press, ipc_press_targets
from tests.test_oracle import PX, _slab_and_block


def _barrier_grad(s, x):
    g, d = np.zeros_like(x), np.zeros_like(x)
    s.xn = x.copy()
    s.ipc.barrier_grad_diag(s, x, 1.0, g, d)
    return g


def _block_pose(m, s, block, corner):
    """Positions with the block tilted by 0.03 rad: its lowest corner 0.4 dhat above the slab,
    either between slab nodes (corner=False) or over one (corner=True, end terms)."""
    rng = np.random.default_rng(1)
    X = m.me.rest_pos.astype(np.float64)
    nodes = m.body_nodes(block)
    x = X.copy()
    th = 0.03
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    c = x[nodes].mean(0)
    x[nodes] = (x[nodes] - c) @ R.T + c
    i_low = nodes[np.argmin(x[nodes, 1])]
    if corner:
        slab = np.setdiff1d(np.arange(len(X)), nodes)
        top = slab[np.abs(X[slab, 1] - X[slab, 1].max()) < 1e-12]
        j = top[np.argmin(np.abs(X[top, 0] - x[i_low, 0]))]
        x[nodes] += X[j] - x[i_low] + np.array([0.05 * m.dhat, 0.4 * m.dhat])
    else:
        x[nodes] += np.array([0.37 * PX, -x[i_low, 1] + 0.4 * m.dhat])
    return x + 0.01 * m.dhat * rng.standard_normal(x.shape)


@pytest.mark.parametrize("corner", [False, True])
def test_ipc_hvp_is_the_gradient_derivative(corner):
    m, s, block = _slab_and_block(0.5)
    x = _block_pose(m, s, block, corner)
    kind = s.ipc.terms(x)[0]
    if corner:
        assert (kind == 1).sum() > 0 and (kind == 2).sum() > 0
    else:
        assert kind.size > 0
    rng = np.random.default_rng(0)
    act = np.where(np.abs(_barrier_grad(s, x)).sum(1) > 0)[0]
    for _ in range(3):
        v = np.zeros_like(x)
        v[act] = rng.standard_normal((len(act), 2))
        y = s.ipc.hvp(x, v, 1.0, np.zeros_like(x))
        h = 1e-6 * m.dhat
        fd = (_barrier_grad(s, x + h * v) - _barrier_grad(s, x - h * v)) / (2 * h)
        assert np.abs(y - fd).max() < 1e-7 * np.abs(fd).max()


@pytest.mark.parametrize("region", [False, True])
def test_oracle_adjoint_with_ipc_and_affine_matches_finite_differences(region):
    T, DT = 5, 2.0e-4
    m, h, x0, top = ipc_press(region=region)
    oi = OracleIntegrator(max_iter=6000, eps=1e-14, tol_x=1e-13, conv_check=8, friction=False)
    s = oi.attach(m, DT)
    assert (s.affine is not None) == region
    oa = OracleAdjoint(s, m.me)
    rng = np.random.default_rng(0)
    rho0 = 0.45 + 0.5 * rng.random(oa.n_tri)
    low = m.body_nodes(h)[np.where(x0[:, 1] < np.median(x0[:, 1]))[0]]
    c = np.zeros((m.n_nodes, 2))
    c[low, 0] = 1.0 / len(low)
    c[low, 1] = 0.5 / len(low)

    def rollout(rho):
        oa.set_density(rho)
        m.reset()
        s.x = np.asarray(m.me.rest_pos, np.float64).copy()
        s.v[:] = 0.0
        if s.affine is not None:
            s.affine.sync_from_x(s)
        xs = [s.x.copy()]
        for t in range(1, T + 1):
            m.set_kinematic_target(h, ipc_press_targets(x0, top, t, T))
            oi.simulate(m, DT)
            xs.append(s.x.copy())
        return xs
    xs = rollout(rho0)
    assert s.ipc.terms(xs[-1])[0].size > 0, "no contact"
    g = oa.transient_adjoint(xs, {T: c}, s.drag, rtol=1e-13)
    g0 = oa.transient_adjoint(xs, {T: c}, s.drag, rtol=1e-13, use_ipc=False)
    J = lambda rho: float((c * rollout(rho)[-1]).sum())
    hfd = 0.002
    dd = g / np.abs(g).max()
    fd = (J(rho0 + hfd * dd) - J(rho0 - hfd * dd)) / (2 * hfd)
    d2 = rng.standard_normal(oa.n_tri)
    fd2 = (J(rho0 + hfd * d2) - J(rho0 - hfd * d2)) / (2 * hfd)
    print(f"region={region}: along the gradient {abs(g @ dd - fd) / abs(fd):.2e}, "
          f"random {abs(g @ d2 - fd2) / (np.linalg.norm(g) * np.linalg.norm(d2)):.2e}, "
          f"without the IPC tangent {abs(g0 @ dd - fd) / abs(fd):.2e}")
    assert abs(g @ dd - fd) / abs(fd) < 1e-4
    assert abs(g @ d2 - fd2) / (np.linalg.norm(g) * np.linalg.norm(d2)) < 1e-5
    assert abs(g0 @ dd - fd) / abs(fd) > 1e-3
    if region:
        el = np.concatenate(s.affine.elems)          # the region's own elements get a gradient
        assert np.abs(g[el]).max() > 1e-6 * np.abs(g).max()
        aff, oa.affine = oa.affine, None             # planted mistake: the slaved nodes left free
        g_free = oa.transient_adjoint(xs, {T: c}, s.drag, rtol=1e-13)
        oa.affine = aff
        print(f"without the reduction {abs(g_free @ dd - fd) / abs(fd):.2e}")
        assert abs(g_free @ dd - fd) / abs(fd) > 1e-3
