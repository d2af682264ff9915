"""The oracle's "node" contact between bodies (oracle/node.py), pinned by what the mathematics
fixes rather than by retyping its formulas (the reference ships no vectors for this path):

* the gradient is the derivative of the energy (central differences);
* the unclamped diagonal is the diagonal of the energy's Hessian, the Hessian-vector product
  and p^T H p are directional derivatives of the gradient;
* the friction gradient is the derivative of its potential with the contact frame (f_n, n)
  frozen, and in the sticking regime (|u_t| < epsv dt) its p^T H p is that potential's exact
  second derivative;
* pairs are visited from both ends and act and react (the penalty forces sum to zero);
* a block resting on a held slab under gravity is held up by exactly its weight.
"""
import numpy as np
import pytest

from oracle.node import pair_mu
from oracle.pncg import OracleIntegrator
from tests.scenes import node_stack

PX = 0.5e-3


def _pressed(seed=0, mu=0.5, block_mu=None):
    """The node_stack scene with each block pushed 0.35 px into its slab and every node jittered,
    so that many pairs are active at various depths and angles."""
    m = node_stack(n_env=2, mu=mu, block_mu=block_mu)
    oi = OracleIntegrator(max_iter=40, eps=0.0, conv_check=0)
    s = oi.attach(m, 1e-3)
    rng = np.random.default_rng(seed)
    x = s.x.copy()
    for h in (1, 3):
        x[m.body_nodes(h), 1] -= 0.35 * PX
    x += rng.normal(scale=0.08 * PX, size=x.shape)
    s.x = x
    s.xn = x - rng.normal(scale=0.02 * PX, size=x.shape)     # slips over the step
    return m, s


def _energy(node, x):
    E = np.zeros(node.B)
    node.energy(x, 1.0, E)
    return E.sum()


def _grad(node, x, clamp=0):
    g, d = np.zeros_like(x), np.zeros_like(x)
    node.grad_diag(x, 1.0, clamp, g, d)
    return g, d


def test_pairs_are_visited_from_both_ends_and_forces_balance():
    m, s = _pressed()
    node, other, nh, d, env = s.node.visits(s.x)
    assert len(node) >= 20
    fwd = set(zip(node.tolist(), other.tolist()))
    assert fwd == set(zip(other.tolist(), node.tolist()))
    assert (d < s.node.r).all()
    g, _ = _grad(s.node, s.x)
    assert np.abs(g.sum(0)).max() < 1e-9 * np.abs(g).max()
    # friction with every node free: action and reaction too
    from oracle.node import BodyFriction
    bf = BodyFriction(s, s.node, 0.5)
    gf, df = np.zeros_like(s.x), np.zeros_like(s.x)
    bf.pair_force(s, gf, df, np.ones_like(s.x))
    assert np.abs(gf).max() > 0.0
    assert np.abs(gf.sum(0)).max() < 1e-9 * np.abs(gf).max()


def test_gradient_is_energy_derivative():
    m, s = _pressed(seed=1)
    x = s.x
    g, _ = _grad(s.node, x)
    nodes = np.unique(s.node.visits(x)[0])
    eps = 1e-9
    fd = np.zeros((len(nodes), 2))
    for a, n in enumerate(nodes):
        for c in range(2):
            xp, xm = x.copy(), x.copy()
            xp[n, c] += eps
            xm[n, c] -= eps
            fd[a, c] = (_energy(s.node, xp) - _energy(s.node, xm)) / (2 * eps)
    err = np.abs(fd - g[nodes]).max() / np.abs(g[nodes]).max()
    assert err < 1e-6, err


def test_diagonal_hvp_and_curvature_are_second_derivatives():
    m, s = _pressed(seed=2)
    x = s.x
    node = s.node
    nodes = np.unique(node.visits(x)[0])
    _, dg = _grad(node, x, clamp=0)
    eps = 1e-10
    for n in nodes[:12]:
        for c in range(2):
            xp, xm = x.copy(), x.copy()
            xp[n, c] += eps
            xm[n, c] -= eps
            h = (_grad(node, xp)[0][n, c] - _grad(node, xm)[0][n, c]) / (2 * eps)
            assert abs(h - dg[n, c]) < 1e-5 * max(abs(dg[n, c]), node.k * 1e-3), (n, c, h, dg[n, c])
    rng = np.random.default_rng(7)
    v = rng.normal(scale=PX, size=x.shape)
    y = np.zeros_like(x)
    node.hvp(x, v, 1.0, y)
    fd = (_grad(node, x + 1e-9 * v)[0] - _grad(node, x - 1e-9 * v)[0]) / 2e-9
    assert np.abs(y - fd).max() < 1e-6 * np.abs(fd).max()
    pHp = np.zeros(node.B)
    node.pHp(x, v, 1.0, 0, pHp)
    e_of = np.asarray(s.me.node2env)
    want = np.array([(v * fd)[e_of == e].sum() for e in range(node.B)])
    assert np.abs(pHp - want).max() < 1e-6 * np.abs(want).max()
    # clamping only ever drops negative parts
    pc = np.zeros(node.B)
    node.pHp(x, v, 1.0, 1, pc)
    assert (pc >= pHp - 1e-12 * np.abs(pHp).max()).all()
    _, dc = _grad(node, x, clamp=1)
    assert (dc >= 0.0).all() and (dc >= dg - 1e-9 * np.abs(dg).max()).all()


def _frozen_friction_potential(s, mu_of, x0):
    """Sum over visits of 1/2 mu f_n f0(|u_t|) with the frame (n, f_n) of each visit frozen at x0
    and only the slip free: the potential the lagged friction force derives from."""
    node, other, nh, d, env = s.node.visits(x0)
    f_n = s.node.k * (s.node.r - d)
    mu = mu_of(node, other)
    a = s.epsv * s.dt

    def phi(x):
        u = (x[node] - s.xn[node]) - (x[other] - s.xn[other])
        ut = np.abs(u[:, 1] * nh[:, 0] - u[:, 0] * nh[:, 1])
        f0 = np.where(ut < a, ut * ut / (2 * a), ut - 0.5 * a)
        return (0.5 * mu * f_n * f0).sum()
    return phi, node


@pytest.mark.parametrize("block_mu", [None, 0.9])
def test_friction_gradient_is_frozen_frame_derivative(block_mu, monkeypatch):
    """With the distance regulariser at zero (n exactly unit): the reference's friction force
    uses v - (v.n) n and its potential |u x n|, which part at O(1e-12 / d^2) when |n| < 1."""
    import oracle.node as ON
    from oracle.node import BodyFriction
    monkeypatch.setattr(ON, "REG", 0.0)
    m, s = _pressed(seed=3, block_mu=block_mu)
    mu_node = m._body_mu_per_node()
    bf = BodyFriction(s, s.node, 0.5, mu_node)
    x0 = s.x.copy()
    ones = np.ones_like(x0)
    g, dg = np.zeros_like(x0), np.zeros_like(x0)
    bf.pair_force(s, g, dg, ones)
    phi, visited = _frozen_friction_potential(s, lambda i, j: pair_mu(0.5, mu_node, i, j), x0)
    eps = 1e-11
    nodes = np.unique(visited)
    fd = np.zeros((len(nodes), 2))
    for a_, n in enumerate(nodes):
        for c in range(2):
            xp, xm = x0.copy(), x0.copy()
            xp[n, c] += eps
            xm[n, c] -= eps
            fd[a_, c] = (phi(xp) - phi(xm)) / (2 * eps)
    assert np.abs(fd - g[nodes]).max() < 1e-6 * np.abs(g[nodes]).max()
    if block_mu is not None:                 # the harmonic mean of 0.9 and the model's 0.5
        i, j = m.body_nodes(1)[:1], m.body_nodes(0)[:1]
        assert np.allclose(pair_mu(0.5, mu_node, i, j), 2 * 0.9 * 0.5 / 1.4)


def test_friction_curvature_in_the_sticking_regime(monkeypatch):
    import oracle.node as ON
    from oracle.node import BodyFriction
    monkeypatch.setattr(ON, "REG", 0.0)
    m, s = _pressed(seed=4)
    s.xn = s.x - 1e-3 * s.epsv * s.dt * np.random.default_rng(0).normal(size=s.x.shape)  # |u_t| << a
    bf = BodyFriction(s, s.node, 0.5)
    x0 = s.x.copy()
    phi, _ = _frozen_friction_potential(s, lambda i, j: np.full(len(i), 0.5), x0)
    p = np.random.default_rng(9).normal(scale=PX, size=x0.shape)
    s.p = p
    s.pHp = np.zeros(s.B)
    bf.pHp(s)
    want = (phi(x0 + 1e-6 * p) - 2 * phi(x0) + phi(x0 - 1e-6 * p)) / 1e-12 * s.h2
    assert abs(s.pHp.sum() - want) < 1e-5 * abs(want)


def test_block_resting_on_slab_is_held_by_its_weight():
    m = node_stack(n_env=1, tilt=0.0)
    oi = OracleIntegrator()
    for _ in range(40):
        oi.simulate(m, 1e-3)
    s = oi.solver
    g, _ = _grad(s.node, s.x)
    blk = m.body_nodes(1)
    force = -g[blk].sum(0)                         # contact force on the block
    weight = s.mass[blk].sum() * 9.81
    assert abs(force[1] - weight) < 1e-3 * weight, (force, weight)
    assert abs(force[0]) < 1e-3 * weight


def test_oracle_adjoint_with_node_contact_matches_finite_differences():
    """The oracle's adjoint with the node contact's tangent (use_self) against central
    differences of oracle rollouts of the pressed block (the same pairs act throughout; steps
    solved to tol_x = 1e-13 so that solver slack stays below the differences)."""
    from oracle.adjoint import OracleAdjoint
    from tests.scenes import node_press, node_press_targets
    T, DT = 6, 2.0e-4
    m, h, x0, top = node_press()
    oi = OracleIntegrator(max_iter=6000, eps=1e-14, tol_x=1e-13, conv_check=8, friction=False)
    s = oi.attach(m, DT)
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
        xs = [s.x.copy()]
        for t in range(1, T + 1):
            m.set_kinematic_target(h, node_press_targets(x0, top, t, T))
            oi.simulate(m, DT)
            xs.append(s.x.copy())
            assert len(s.node.visits(s.x)[0]) == 14
        return xs
    xs = rollout(rho0)
    g = oa.transient_adjoint(xs, {T: c}, s.drag, rtol=1e-13, use_self=True)
    g0 = oa.transient_adjoint(xs, {T: c}, s.drag, rtol=1e-13, use_self=False)
    J = lambda rho: float((c * rollout(rho)[-1]).sum())
    hfd = 0.002
    dd = g / np.abs(g).max()
    fd = (J(rho0 + hfd * dd) - J(rho0 - hfd * dd)) / (2 * hfd)
    d2 = rng.standard_normal(oa.n_tri)
    fd2 = (J(rho0 + hfd * d2) - J(rho0 - hfd * d2)) / (2 * hfd)
    assert abs(g @ dd - fd) / abs(fd) < 1e-4
    assert abs(g @ d2 - fd2) / (np.linalg.norm(g) * np.linalg.norm(d2)) < 2e-5
    assert abs(g0 @ dd - fd) / abs(fd) > 1e-2          # the contact tangent matters here
