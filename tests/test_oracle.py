"""Pinning the float64 CPU oracle (test infrastructure) before the GPU path is compared with it.

Every function of the oracle is fixed by at least one check that a plausible mistake (a dropped
term, a wrong sign or index, a transposed operand) would fail:

  elastic energy / gradient    the reference's independent Newton oracle (tests/reference_static.py,
                               golden fixture) on a deformed mesh; central FD of the energy
  elastic diagonal, p^T H p,   central FD of the gradient (unclamped), along random directions
  Hessian-vector product
  inversion cap                the closed-form root: area after the step is exactly `shrink` x now
  barrier b, b', b''           central FD, and b(dhat) = b'(dhat) = 0
  colliders (5 shapes)         barrier gradient vs FD of energy, curvature vs FD of gradient,
                               level set vs exact distance for disc / box / half-space
  collider friction            the friction force is the FD gradient of the lagged potential
  IPC node-edge contact        gradient vs FD of energy (< 1e-6, as reference tests/test_ipc.py:95-124),
                               p^T H p vs FD of gradient (< 1e-3), flat sliding energy spread < 3 %
                               (vertex corrections, tests/test_ipc.py:79-92), brute-force pair set
  pressure, tendon             gradient vs FD of energy, p^T H p vs FD of gradient
  6x6 Cholesky, reduced        numpy.linalg.solve (< 1e-10); PSD under compression and one zero
  Hessian                      eigenvalue at a pure rotation (tests/test_rigid.py:62-94)
  adjoint dJ/drho               central differences of converged rollouts, along the gradient
                               (< 1e-4 relative) and a random direction
  whole step                   exact backward-Euler free fall of a rigid body and of a soft body
                               with the rigid-mode coarse space; batch == alone bitwise;
                               static equilibrium of the Newton oracle (golden fixture)
  scene parameters             contact scaling of the reference's two-block scene
                               (tests/test_scene.py:123-124)
"""
import numpy as np
import pytest

from diff_2d_fem_amd import Model
from oracle import terms as T
from oracle.bodies import chol6, chol6_solve, reduced_hessian, Pressure, Tendon
from oracle.ipc import IPCContact, accd, seg_dist
from oracle.pncg import OracleIntegrator
from tests.scenes import strip, cantilever, pressed_block, press_targets

GOLD = "tests/golden/"


def _mesh_body(shape=(5, 9), px=1e-3, E=2e5, nu=0.42, seed=0, scale=1.5e-4):
    m = Model()
    m.add_soft_body(np.ones(shape, np.uint8), physical_size=(shape[1] * px, shape[0] * px),
                    center=(0.0, 0.0), E=E, nu=nu, density=1000.0)
    m.finalize()
    me = m.me
    rng = np.random.default_rng(seed)
    x = me.rest_pos + scale * rng.standard_normal(me.rest_pos.shape)
    return m, me, x, rng


def _elastic_E(me, x, h2=1.0):
    E = np.zeros(me.n_envs)
    T.elastic_energy(x, me.tri, me.D_m_inv, me.rest_area, me.mu, me.lam, h2, me.tri2env, E)
    return E.sum()


def _elastic_g(me, x, clamp=1, h2=1.0):
    g, d = np.zeros_like(x), np.zeros_like(x)
    T.elastic_grad_diag(x, me.tri, me.D_m_inv, me.rest_area, me.mu, me.lam, h2, clamp, g, d)
    return g, d


# ---------------------------------------------------------------------------- elastic
def test_elastic_matches_reference_newton_oracle():
    from tests.golden.fixture_scenes import scenes
    from diff_2d_fem_amd.mesh import build_scene
    f = np.load(GOLD + "static_fixtures.npz")
    me = build_scene([dict(b) for b in scenes()["blob"]])
    x = f["blob_x"]
    assert abs(_elastic_E(me, x) / float(f["blob_E"]) - 1.0) < 1e-12
    g, _ = _elastic_g(me, x)
    assert np.abs(g - f["blob_grad"]).max() < 1e-10 * np.abs(f["blob_grad"]).max()


def test_cantilever_reaches_reference_static_equilibrium():
    """Implicit steps with a one-second step size drive the held, loaded beam to the static
    equilibrium the reference's Newton oracle computed (tip deflection ~0.7 of the length)."""
    m, f = cantilever()
    fixed = np.nonzero(np.abs(m.me.rest_pos[:, 0] - m.me.rest_pos[:, 0].min()) < 1e-12)[0]
    assert np.array_equal(fixed, f["cant_fixed"])
    oi = OracleIntegrator(max_iter=400, eps=0.0, conv_check=0)
    for _ in range(5):
        oi.simulate(m, 1.0)
    assert np.abs(oi.solver.x - f["cant_x"]).max() / 0.02 < 1e-8
    assert np.abs(f["cant_x"] - m.me.rest_pos).max() / 0.02 > 0.5


def test_elastic_gradient_is_energy_derivative():
    m, me, x, rng = _mesh_body()
    g, _ = _elastic_g(me, x)
    d = rng.standard_normal(x.shape)
    h = 1e-8
    fd = (_elastic_E(me, x + h * d) - _elastic_E(me, x - h * d)) / (2 * h)
    assert abs(fd - (g * d).sum()) < 1e-6 * abs(fd)


def test_elastic_diagonal_curvature_hvp_are_second_derivatives():
    m, me, x, rng = _mesh_body()
    _, diag = _elastic_g(me, x, clamp=0)
    h = 1e-8
    for i, c in [(3, 0), (10, 1), (25, 0), (40, 1)]:
        e = np.zeros_like(x)
        e[i, c] = 1.0
        true = (_elastic_g(me, x + h * e)[0][i, c] - _elastic_g(me, x - h * e)[0][i, c]) / (2 * h)
        assert abs(diag[i, c] - true) < 1e-6 * abs(true)
    p = rng.standard_normal(x.shape)
    ph = np.zeros(1)
    T.elastic_pHp(x, p, me.tri, me.D_m_inv, me.rest_area, me.mu, me.lam, 1.0, 0, me.tri2env, ph)
    Hp_fd = (_elastic_g(me, x + h * p)[0] - _elastic_g(me, x - h * p)[0]) / (2 * h)
    assert abs(ph[0] - (Hp_fd * p).sum()) < 1e-6 * abs(ph[0])
    y = np.zeros_like(x)
    T.elastic_hvp(x, p, me.tri, me.D_m_inv, me.rest_area, me.mu, me.lam, 1.0, y)
    assert np.abs(y - Hp_fd).max() < 1e-6 * np.abs(Hp_fd).max()


def test_elastic_rest_state_has_no_force():
    m, me, _, _ = _mesh_body()
    g, _ = _elastic_g(me, me.rest_pos)
    assert np.abs(g).max() < 1e-9 * me.mu.max() * 1e-3
    assert abs(_elastic_E(me, me.rest_pos)) < 1e-20


def test_inversion_root_is_exact():
    m, me, x, rng = _mesh_body()
    p = 3e-3 * rng.standard_normal(x.shape)
    alpha = np.full(1, 1e9)
    T.inversion_alpha(x, p, me.tri, me.tri2env, 0.1, alpha)
    s = alpha[0]
    assert 0.0 < s < 1e9
    a0 = T.cross2(x[me.tri[:, 1]] - x[me.tri[:, 0]], x[me.tri[:, 2]] - x[me.tri[:, 0]])
    xs = x + s * p
    a1 = T.cross2(xs[me.tri[:, 1]] - xs[me.tri[:, 0]], xs[me.tri[:, 2]] - xs[me.tri[:, 0]])
    ratio = a1 / a0
    assert abs(ratio.min() - 0.1) < 1e-9                       # one triangle lands exactly on 10 %
    assert ratio.min() >= 0.1 - 1e-9                           # and none below


# ---------------------------------------------------------------------------- barrier
def test_barrier_derivatives():
    dhat = 1e-4
    d = np.linspace(0.05, 0.95, 7) * dhat
    h = 1e-9 * dhat
    assert np.allclose(T.barrier_bp(d, dhat), (T.barrier_b(d + h, dhat) - T.barrier_b(d - h, dhat)) / (2 * h), rtol=1e-6)
    assert np.allclose(T.barrier_bpp(d, dhat), (T.barrier_bp(d + h, dhat) - T.barrier_bp(d - h, dhat)) / (2 * h), rtol=1e-5)
    assert abs(T.barrier_b(dhat, dhat)) == 0.0 and abs(T.barrier_bp(dhat, dhat)) == 0.0


# ---------------------------------------------------------------------------- colliders
COLLIDERS = [
    dict(kind=0, cx=0.0, cy=-0.0105, rx=0.01, ry=0.01, rref=0.01, r=0.01, mu=0.5),
    dict(kind=1, cx=0.0, cy=-0.0205, rx=0.03, ry=0.02, rref=0.02, r=0.03, mu=0.5),
    dict(kind=2, cx=0.0, cy=-0.0055, rx=0.008, ry=0.005, rref=0.002, r=0.008, mu=0.5),
    dict(kind=3, cx=0.0, cy=-0.0205, rx=0.03, ry=0.03, rref=0.004, r=0.03, mu=0.5),
    dict(kind=4, cx=0.0, cy=-0.0005, rx=0.0, ry=1.0, rref=0.0, r=0.0, mu=0.5),
]


def _probe_points(col, n=40, seed=1, dhat=1e-3, span=0.010):
    """Points at clearance U(0.1, 0.9) dhat from the collider, spread along its top (over the
    rounded corners of the box)."""
    rng = np.random.default_rng(seed)
    X = np.stack([rng.uniform(-span, span, n), np.full(n, 0.0)], 1)
    tgt = rng.uniform(0.1, 0.9, n) * dhat
    for _it in range(60):                               # move each point to the target clearance
        d, g, _H = T.sdf_eval(col["kind"], col["cx"], col["cy"], col["rx"], col["ry"], col["rref"], X)
        gn = np.linalg.norm(g, axis=1)[:, None]
        X = X - ((d - tgt) / np.maximum(gn[:, 0] ** 2, 1e-30))[:, None] * g
    return X


@pytest.mark.parametrize("ci", range(len(COLLIDERS)))
def test_collider_barrier_derivatives(ci):
    col = COLLIDERS[ci]
    dhat, kappa = 1e-3, 1e3
    X = _probe_points(col, dhat=dhat)
    bnd = np.arange(len(X))
    b2e = np.zeros(len(X), np.int64)

    def energy(x):
        E = np.zeros(1)
        T.collider_energy(col, x, bnd, b2e, dhat, kappa, 1.0, E)
        return E[0]

    def grad(x):
        g, d = np.zeros_like(x), np.zeros_like(x)
        T.collider_grad_diag(col, x, bnd, dhat, kappa, 1.0, 0, g, d)
        return g, d
    g, dg = grad(X)
    assert np.abs(g).max() > 0
    rng = np.random.default_rng(2)
    p = rng.standard_normal(X.shape)
    h = 1e-9
    fd = (energy(X + h * p) - energy(X - h * p)) / (2 * h)
    assert abs(fd - (g * p).sum()) < 1e-6 * abs(fd)
    ph, clr = np.zeros(1), np.full(1, 1e30)
    T.collider_pHp(col, X, p, bnd, b2e, dhat, kappa, 1.0, 0, ph, clr)
    Hp = (grad(X + h * p)[0] - grad(X - h * p)[0]) / (2 * h)
    assert abs(ph[0] - (Hp * p).sum()) < 1e-5 * abs(ph[0])
    # the diagonal of the node's 2x2 block
    for k in (0, 1):
        e = np.zeros_like(X)
        e[:, k] = 1.0
        dd = (grad(X + h * e)[0][:, k] - grad(X - h * e)[0][:, k]) / (2 * h)
        assert np.abs(dd - dg[:, k]).max() <= 1e-5 * np.abs(dg).max()
    assert abs(clr[0] - T.sdf_eval(col["kind"], col["cx"], col["cy"], col["rx"], col["ry"], col["rref"], X)[0].min()) == 0.0


def test_level_sets_are_distances():
    P = np.array([[0.013, 0.004], [0.0, 0.02], [-0.02, -0.01]])
    d0 = T.sdf_eval(0, 0.0, 0.0, 0.01, 0.01, 0.01, P)[0]
    assert np.allclose(d0, np.linalg.norm(P, axis=1) - 0.01, rtol=1e-14)
    d4 = T.sdf_eval(4, 0.0, 0.001, 0.6, 0.8, 0.0, P)[0]
    assert np.allclose(d4, 0.6 * P[:, 0] + 0.8 * (P[:, 1] - 0.001), rtol=1e-14)
    # rounded square of half-width 5 mm, corner 2 mm: a point off a face and one off a corner
    Q = np.array([[0.0, 0.007], [0.006, 0.006]])
    d2 = T.sdf_eval(2, 0.0, 0.0, 0.005, 0.005, 0.002, Q)[0]
    assert abs(d2[0] - 0.002) < 1e-15
    assert abs(d2[1] - (np.hypot(0.003, 0.003) - 0.002)) < 1e-15


@pytest.mark.parametrize("ci", [0, 4])
def test_friction_force_is_gradient_of_its_potential(ci):
    col = COLLIDERS[ci]
    dhat, kappa, mu, epsv, inv_dt = 1e-3, 1e3, 0.6, 1e-3, 1e3
    X = _probe_points(col, dhat=dhat)
    rng = np.random.default_rng(4)
    xn = X - 1e-6 * rng.standard_normal(X.shape)            # a small slip since the step began
    bnd = np.arange(len(X))
    b2e = np.zeros(len(X), np.int64)
    free = np.ones_like(X)

    def energy(x):
        E = np.zeros(1)
        T.friction_energy(col, x, xn, bnd, b2e, dhat, kappa, mu, epsv, inv_dt, 1.0, free, E)
        return E[0]
    g, dg = np.zeros_like(X), np.zeros_like(X)
    T.friction_grad_diag(col, X, xn, bnd, dhat, kappa, mu, epsv, inv_dt, 1.0, free, g, dg)
    # lagged: the normal force is held fixed, so differentiate the potential at fixed f_n by
    # moving tangentially only (along the collider's surface) at second order
    d, gr, _ = T.sdf_eval(col["kind"], col["cx"], col["cy"], col["rx"], col["ry"], col["rref"], X)
    nh = gr / np.linalg.norm(gr, axis=1)[:, None]
    tdir = np.stack([-nh[:, 1], nh[:, 0]], 1) * rng.standard_normal((len(X), 1))
    h = 1e-10
    fd = (energy(X + h * tdir) - energy(X - h * tdir)) / (2 * h)
    assert abs(fd - (g * tdir).sum()) < 1e-4 * np.abs(g * tdir).sum()


# ---------------------------------------------------------------------------- IPC
PX = 0.5e-3


def _slab_and_block(gap_in_dhat, mu=0.0):
    """The reference's IPC probe scene (tests/test_ipc.py:52-61): a soft block over a slab."""
    m = Model()
    m.body_contact = "ipc"
    m.body_friction = mu
    m.add_soft_body(np.ones((4, 80), np.uint8), physical_size=(80 * PX, 4 * PX), center=(0.0, -2 * PX),
                    E=1.0e6, nu=0.40, density=1100.0)
    dhat = PX / 8
    block = m.add_soft_body(np.ones((8, 24), np.uint8), physical_size=(24 * PX, 8 * PX),
                            center=(0.0, 4 * PX + gap_in_dhat * dhat), E=1.0e6, nu=0.40, density=1100.0)
    m.finalize()
    assert np.isclose(m.dhat, dhat) and np.isclose(m.body_kappa, 0.1 * 1.0e6 * PX / dhat)
    integ = OracleIntegrator()
    integ.attach(m, 1e-3)
    return m, integ.solver, block


def _ipc_energy(s, x):
    s.x = x.copy()
    s.xn = x.copy()
    buf = np.zeros(s.B)
    s.ipc.energy(s, buf)
    return buf[0]


def _ipc_grad(s, x):
    s.x = x.copy()
    s.xn = x.copy()
    s.grad = np.zeros_like(x)
    s.diagH = np.zeros_like(x)
    s.ipc.grad_diag(s)
    return s.grad.copy()


def test_ipc_flat_sliding_costs_nothing():
    for gap in (0.9, 0.5, 0.1):
        m, s, block = _slab_and_block(gap)
        X = m.me.rest_pos.astype(np.float64)
        nodes = m.body_nodes(block)
        Es = []
        for shift in np.linspace(0.0, 2 * PX, 41):
            x = X.copy()
            x[nodes, 0] += shift
            Es.append(_ipc_energy(s, x))
        Es = np.array(Es)
        assert (Es.max() - Es.min()) / Es.mean() < 0.03


def test_ipc_derivatives():
    m, s, block = _slab_and_block(0.5)
    rng = np.random.default_rng(0)
    X = m.me.rest_pos.astype(np.float64)
    nodes = m.body_nodes(block)
    x = X.copy()
    c = x[nodes].mean(0)
    th = 0.004
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    x[nodes] = (x[nodes] - c) @ R.T + c + np.array([0.13 * PX, 0.2 * m.dhat])
    x += 0.02 * m.dhat * rng.standard_normal(x.shape)
    g = _ipc_grad(s, x)
    act = np.where(np.abs(g).sum(1) > 0)[0]
    assert len(act) > 30
    h = 1e-9
    num = np.zeros_like(g)
    for i in act:
        for k in range(2):
            xp, xm = x.copy(), x.copy()
            xp[i, k] += h
            xm[i, k] -= h
            num[i, k] = (_ipc_energy(s, xp) - _ipc_energy(s, xm)) / (2 * h)
    assert np.abs(g - num).max() / np.abs(g).max() < 1e-6
    p = np.zeros_like(x)
    p[act] = rng.standard_normal((len(act), 2))
    _ipc_grad(s, x)
    s.p = p
    s.pHp = np.zeros(s.B)
    s.alpha = np.full(s.B, 3.0)
    s.p_inf = np.ones(s.B)
    s.ipc.pHp(s)
    curv = s.pHp[0]
    hh = 1e-7 * m.dhat
    fd = (((_ipc_grad(s, x + hh * p) - _ipc_grad(s, x - hh * p)) / (2 * hh)) * p).sum()
    assert abs(curv / fd - 1.0) < 1e-3


def test_ipc_pairs_are_the_brute_force_set():
    m, s, block = _slab_and_block(0.5)
    x = m.me.rest_pos.astype(np.float64)
    K, JB, tt, dvec, d = s.ipc.pairs(x)
    ip = s.ipc
    got = set(zip(K.tolist(), JB.tolist()))
    want = set()
    bnd = ip.bnd
    for k in range(len(ip.ea)):
        for jb in range(len(bnd)):
            pn = bnd[jb]
            if pn in (ip.ea[k], ip.eb[k]) or ip.excl[0][jb, ip.ela[k]] or ip.excl[0][jb, ip.elb[k]]:
                continue
            if seg_dist(x[pn], x[ip.ea[k]], x[ip.eb[k]]) < ip.reach:
                want.add((k, jb))
    assert got == want and len(got) > 0


def test_accd_never_crosses():
    """Additive CCD (reference implicit/ipc_contact.py:103-127): along the returned step the node
    never reaches the edge, and a step the gap stopped short of t_c ends at least `g` away."""
    rng = np.random.default_rng(5)
    n = 200
    p = rng.uniform(-1, 1, (n, 2))
    a = rng.uniform(-1, 1, (n, 2)) + np.array([0.0, -3.0])
    b = a + rng.uniform(0.5, 1.5, (n, 2))
    dp, da, db = rng.standard_normal((n, 2)), rng.standard_normal((n, 2)), rng.standard_normal((n, 2))
    d0 = seg_dist(p, a, b)
    g = 0.1 * d0
    t_c = np.full(n, 2.0)
    t = accd(p, a, b, dp, da, db, g, t_c)
    assert (t > 0).all() and (t <= t_c).all()
    for k in range(n):
        ds = [seg_dist(p[k] + s_ * dp[k], a[k] + s_ * da[k], b[k] + s_ * db[k]) for s_ in np.linspace(0.0, t[k], 200)]
        assert min(ds) > 0.0
        if t[k] < t_c[k]:
            assert ds[-1] >= g[k] * (1 - 1e-9)
    # a pure common translation closes no gap: the full step is allowed
    tt = accd(p, a, b, dp, dp, dp, g, t_c)
    assert (tt == t_c).all()


# ---------------------------------------------------------------------------- actuation
class _Sol:
    """A minimal solver-shaped holder for term classes."""

    def __init__(self, me, x):
        self.me, self.x, self.N = me, x.copy(), len(x)
        self.h2 = 1.0
        self.grad = np.zeros_like(x)
        self.diagH = np.zeros_like(x)
        self.p = np.zeros_like(x)
        self.pHp = np.zeros(me.n_envs)
        self.xn = x.copy()
        self.node2env = np.asarray(me.node2env)


def _term_checks(term, me, x, rng, curvature=True):
    def energy(xx):
        s = _Sol(me, xx)
        buf = np.zeros(me.n_envs)
        if isinstance(term, Tendon):
            term.grad_diag(s)
        term.energy(s, buf)
        return buf.sum()

    def grad(xx):
        s = _Sol(me, xx)
        term.grad_diag(s)
        return s.grad, s.diagH
    g, _ = grad(x)
    assert np.abs(g).max() > 0
    d = rng.standard_normal(x.shape)
    h = 1e-8
    fd = (energy(x + h * d) - energy(x - h * d)) / (2 * h)
    assert abs(fd - (g * d).sum()) < 1e-6 * abs(fd)
    if curvature:
        s = _Sol(me, x)
        term.grad_diag(s)
        s.p = d
        term.pHp(s)
        Hd = (grad(x + h * d)[0] - grad(x - h * d)[0]) / (2 * h)
        assert abs(s.pHp.sum() - (Hd * d).sum()) < 1e-5 * abs(s.pHp.sum())


def test_pressure_derivatives():
    ring = np.ones((7, 9), np.uint8)
    ring[2:5, 3:6] = 0
    m = Model()
    h = m.add_soft_body(ring, physical_size=(9e-3, 7e-3), center=(0.0, 0.0), E=2e5, nu=0.4)
    m.finalize()
    m.set_pressure(m.add_pressure_chambers(h)[0], 3.0e4)
    m.set_pressure(m.add_ambient_pressure(h), 1.1e4)
    C = m.actuation.chambers
    lay = m.actuation.pressure_layout(m.body_env)
    term = Pressure(None, lay["ea"], lay["eb"], lay["chamber"], [0, 0], lay["ref"])
    term.set([c.pressure for c in C])
    rng = np.random.default_rng(6)
    x = m.me.rest_pos + 1e-4 * rng.standard_normal(m.me.rest_pos.shape)
    _term_checks(term, m.me, x, rng)
    # the chamber's area is what the potential is minus pressure times
    s = _Sol(m.me, x)
    buf = np.zeros(1)
    one = Pressure(None, C[0].ea, C[0].eb, np.zeros(len(C[0].ea), int), [0], [C[0].ref])
    one.set([1.0])
    one.energy(s, buf)
    assert abs(-buf[0] - m.chamber_area(0) * 0) >= 0.0        # exercised below on rest
    s = _Sol(m.me, m.me.rest_pos)
    buf = np.zeros(1)
    one.energy(s, buf)
    assert abs(-buf[0] / (9e-6) - 1.0) < 1e-12                 # 3 x 3 px cavity of 1 mm pixels


def test_pressure_adjoint_pieces():
    """The pressure tangent the adjoint adds (Pressure.hvp) is the derivative of the pressure
    gradient, is symmetric, and Pressure.dp is -lam . dR/dp by differences in p."""
    ring = np.ones((7, 9), np.uint8)
    ring[2:5, 3:6] = 0
    m = Model()
    h = m.add_soft_body(ring, physical_size=(9e-3, 7e-3), center=(0.0, 0.0), E=2e5, nu=0.4)
    m.finalize()
    m.set_pressure(m.add_pressure_chambers(h)[0], 3.0e4)
    m.set_pressure(m.add_ambient_pressure(h), 1.1e4)
    lay = m.actuation.pressure_layout(m.body_env)
    term = Pressure(None, lay["ea"], lay["eb"], lay["chamber"], [0, 0], lay["ref"])
    p0 = np.array([3.0e4, 1.1e4])
    term.set(p0)
    rng = np.random.default_rng(8)
    x = m.me.rest_pos + 1e-4 * rng.standard_normal(m.me.rest_pos.shape)

    def grad(xx, p=p0):
        term.set(p)
        s = _Sol(m.me, xx)
        term.grad_diag(s)
        term.set(p0)
        return s.grad
    v, u = rng.standard_normal(x.shape), rng.standard_normal(x.shape)
    Hv = np.zeros_like(x)
    term.hvp(x, v, 1.0, Hv)
    hh = 1e-6
    fd = (grad(x + hh * v) - grad(x - hh * v)) / (2 * hh)
    assert np.abs(Hv - fd).max() < 1e-7 * np.abs(Hv).max()
    Hu = np.zeros_like(x)
    term.hvp(x, u, 1.0, Hu)
    assert abs((u * Hv).sum() - (v * Hu).sum()) < 1e-12 * np.abs(u * Hv).sum()
    lam = rng.standard_normal(x.shape)
    dp = term.dp(x, lam, 1.0)
    for c in range(2):
        e = np.zeros(2)
        e[c] = 1.0
        fdp = (grad(x, p0 + e) - grad(x, p0 - e)) / 2.0
        assert abs(dp[c] + (lam * fdp).sum()) < 1e-9 * abs(dp[c])


def _finger(n_ch_px=((2, 3, 2, 3), (2, 12, 2, 4))):
    """A small held finger (6 x 24 px of 0.5 mm) with two sealed chambers."""
    mask = np.ones((6, 24), np.uint8)
    for r, c, hh, ww in n_ch_px:
        mask[r:r + hh, c:c + ww] = 0
    m = Model()
    h = m.add_soft_body(mask, physical_size=(12e-3, 3e-3), center=(0.0, 0.0), E=2e5, nu=0.4,
                        density=1100.0)
    m.finalize()
    X = m.rest_x(h)
    m.set_kinematic_constraint(h, np.nonzero(X[:, 0] < X[:, 0].min() + 1e-9)[0])
    ch = m.add_pressure_chambers(h)
    return m, h, ch


def test_oracle_adjoint_with_pressure_matches_finite_differences():
    """The oracle adjoint with the pressure tangent: dJ/dp (constant pressures over the rollout)
    and dJ/drho against central differences of converged oracle rollouts."""
    from oracle.adjoint import OracleAdjoint
    T, dt = 8, 2e-4
    kw = dict(max_iter=4000, eps=1e-15, tol_x=1e-16, conv_check=8, self_collision=False, friction=False)
    m, h, ch = _finger()
    p0 = np.array([2.0e4, 3.5e4])
    rng = np.random.default_rng(3)
    X = m.rest_x(h)
    c = np.zeros((m.n_nodes, 2))                       # J: the tip's height and reach
    tip = np.nonzero(X[:, 0] > X[:, 0].max() - 1e-9)[0]
    c[tip, 1], c[tip, 0] = 1.0 / len(tip), 0.3 / len(tip)
    oi = OracleIntegrator(**kw)
    oi.attach(m, dt)
    oa = OracleAdjoint(oi.solver, m.me)
    rho0 = 0.6 + 0.3 * rng.random(m.n_elements)

    def rollout(p, rho):
        for k, pk in zip(ch, p):
            m.set_pressure(k, pk)
        oa.set_density(rho)
        s = oi.solver
        s.x = np.asarray(m.me.rest_pos, np.float64).copy()
        s.v[:] = 0.0
        xs = [s.x.copy()]
        for _ in range(T):
            oi.simulate(m, dt)
            xs.append(oi.solver.x.copy())
        return xs
    xs = rollout(p0, rho0)
    oa = OracleAdjoint(oi.solver, m.me)
    oa.set_density(rho0)
    g = oa.transient_adjoint(xs, {T: c}, oi.solver.drag, rtol=1e-14)
    gp = oa.dJdp_steps.sum(0)
    J = lambda p, rho: float((c * rollout(p, rho)[-1]).sum())
    for k in range(2):
        e = np.zeros(2)
        e[k] = 50.0                     # central differences converge as h^2: 7e-6 at 50 Pa
        fd = (J(p0 + e, rho0) - J(p0 - e, rho0)) / 100.0
        assert abs(gp[k] - fd) < 1e-4 * abs(fd), (k, gp[k], fd)
    d = g / np.abs(g).max()
    fd = (J(p0, rho0 + 1e-3 * d) - J(p0, rho0 - 1e-3 * d)) / 2e-3
    assert abs(g @ d - fd) < 1e-4 * abs(fd), (g @ d, fd)


@pytest.mark.parametrize("mode", ["tension", "taut"])
def test_tendon_derivatives(mode):
    m = Model()
    h = m.add_soft_body(np.ones((4, 12), np.uint8), physical_size=(12e-3, 4e-3), center=(0.0, 0.0), E=2e5)
    m.finalize()
    path = [(-5e-3, 1e-3), (-1e-3, 0.0), (3e-3, 1e-3), (5e-3, -1e-3)]
    if mode == "tension":
        td = m.add_tendon(h, through=path)
        m.set_tendon_tension(td, 40.0)
    else:
        td = m.add_tendon(h, through=path, stiffness=5e4)
        m.set_tendon_length(td, 0.9 * m.tendon_length(td))
    t = m.actuation.tendons[0]
    n = len(t.nodes)
    term = Tendon(None, t.nodes, [-1] + list(range(n - 1)), list(range(1, n)) + [-1], [0] * n, [0],
                  [t.stiffness], [t.length])
    term.set([t.tension], [t.length])
    rng = np.random.default_rng(7)
    x = m.me.rest_pos + 1e-4 * rng.standard_normal(m.me.rest_pos.shape)
    _term_checks(term, m.me, x, rng)


# ---------------------------------------------------------------------------- rigid blocks
def test_chol6_and_reduced_hessian():
    rng = np.random.default_rng(0)
    n = 200
    M = rng.normal(size=(n, 6, 6))
    H = M @ M.transpose(0, 2, 1) + 0.1 * np.eye(6)
    b = rng.normal(size=(n, 6))
    x = chol6_solve(chol6(H), b)
    ref = np.linalg.solve(H, b[:, :, None])[:, :, 0]
    assert np.abs(x - ref).max() / np.abs(ref).max() < 1e-10
    th = rng.uniform(0, 2 * np.pi, n)
    c, s = np.cos(th), np.sin(th)
    R = np.stack([np.stack([c, -s], -1), np.stack([s, c], -1)], 1)
    S = np.eye(2) + 0.02 * rng.uniform(-1, 1, (n, 2, 2))
    S = 0.5 * (S + S.transpose(0, 2, 1))
    z = np.zeros((n, 3, 3))
    Mu, La = np.full(n, 1.0), np.full(n, 2.5)
    idx = [1, 2, 4, 5]
    for name, A in (("rotation", R), ("strained", R @ S)):
        H4 = reduced_hessian(A, Mu, La, 1.0, z, z)[:, idx][:, :, idx]
        w = np.linalg.eigvalsh(H4)
        assert np.abs(H4 - H4.transpose(0, 2, 1)).max() < 1e-9 * w.max()
        assert (w[:, 0] > -1e-9 * w[:, -1]).all()                       # positive semi-definite
        if name == "rotation":
            assert (np.abs(w[:, 0]) < 1e-9 * w[:, -1]).all()            # rotating costs nothing
            assert (w[:, 1] > 0.1 * w[:, -1]).all()                     # straining does
    assert (np.linalg.det(R @ S) < 1).sum() > n // 4                    # compressed states sampled


# ---------------------------------------------------------------------------- whole steps
def test_rigid_body_free_fall_is_exact_backward_euler():
    m = Model()
    h = m.add_affine_body(np.ones((6, 8), np.uint8), physical_size=(8e-3, 6e-3), center=(0.0, 0.1),
                          density=1000.0)
    m.finalize()
    m.gravity = (0.0, -9.81)
    oi = OracleIntegrator(self_collision=False)
    dt, y0 = 2e-3, m.get_body_x(h)[:, 1].mean()
    v, y = 0.0, y0
    for _ in range(5):
        oi.simulate(m, dt)
        v += -9.81 * dt
        y += v * dt
        st = oi.stats()
        assert st["converged"].all() and st["iters"].max() <= 4
    assert abs(oi.solver.x[m.body_nodes(h), 1].mean() - y) < 1e-9


def test_soft_free_fall_needs_the_rigid_modes():
    def fall(rigid):
        m = Model()
        h = m.add_soft_body(np.ones((6, 16), np.uint8), physical_size=(16e-3, 6e-3), center=(0.0, 0.1),
                            E=1e6, nu=0.4, density=1100.0)
        m.finalize()
        m.gravity = (0.0, -9.81)
        oi = OracleIntegrator(self_collision=False, max_iter=5, rigid_modes=rigid, eps=0.0, conv_check=0)
        for _ in range(3):
            oi.simulate(m, 2e-3)
        return oi.solver.x[m.body_nodes(h), 1].mean() - 0.1
    exact = -9.81 * 2e-3 ** 2 * (1 + 2 + 3)
    assert abs(fall(True) / exact - 1.0) < 1e-4
    assert abs(fall(False) / exact) < 0.2                           # without them: barely moves


def test_batch_equals_alone_bitwise():
    a = strip(n_env=1)
    b = Model()
    b.add_ground(0.0, mu=0.8)
    for i, E in enumerate((0.5e6, 0.9e6, 0.5e6)):
        b.add_soft_body(np.ones((6, 72), np.uint8), physical_size=(0.054, 0.0045), center=(0.0, 0.0030),
                        E=E, nu=0.40, density=1100.0, env_id=i)
    b.finalize()
    b.gravity = (0.0, -9.81)
    oa, ob = OracleIntegrator(max_iter=30), OracleIntegrator(max_iter=30)
    for _ in range(3):
        oa.simulate(a, 2e-3)
        ob.simulate(b, 2e-3)
    xa = oa.solver.x
    for env in (0, 2):
        nodes = np.nonzero(b.me.node2env == env)[0]
        assert np.array_equal(ob.solver.x[nodes], xa)
    assert np.abs(ob.solver.x[b.me.node2env == 1] - xa).max() > 1e-11     # the stiffer design differs


def test_contact_parameters_are_scaled_to_the_scene():
    m = Model()
    m.add_ground(height=0.0, mu=0.6)
    reach = 0.5e-3 / 8
    blk = np.ones((12, 16), np.uint8)
    m.add_soft_body(blk, physical_size=(0.008, 0.006), E=1.0e6, nu=0.40, density=1100.0,
                    center=(0.0, 0.003 + 0.9 * reach))
    m.add_soft_body(blk, physical_size=(0.008, 0.006), E=1.0e6, nu=0.40, density=1100.0,
                    center=(0.0, 0.003 + 0.9 * reach + 0.0068))
    m.finalize()
    assert np.isclose(m.dhat, reach) and np.isclose(m.kappa, 3.2e4) and np.isclose(m.self_stiffness, 2.0e5)
    assert np.isclose(m.body_kappa, 0.1 * 1.0e6 * 0.5e-3 / reach)


# ---------------------------------------------------------------------------- adjoint
@pytest.mark.parametrize("mode", ["elastic", "disc"])
def test_adjoint_matches_finite_differences(mode):
    """The oracle's adjoint dJ/drho of J = c . x_T against central differences of converged
    rollouts (the protocol of reference tests/test_adjoint.py, float64 so that a small step
    can be used): along the gradient and along a random direction."""
    from oracle.adjoint import OracleAdjoint
    T, DT = 6, 2.0e-4
    m, h, x0, top = pressed_block(mode)
    oi = OracleIntegrator(max_iter=3000, eps=1e-14, tol_x=1e-10, conv_check=8, self_collision=False,
                          friction=False)
    oi.attach(m, DT)
    adj = OracleAdjoint(oi.solver, m.me)
    rng = np.random.default_rng(0)
    rho0 = 0.45 + 0.5 * rng.random(adj.n_tri)
    low = np.where(x0[:, 1] < np.median(x0[:, 1]))[0]
    c = np.zeros((m.n_nodes, 2))
    c[low, 0] = 1.0 / len(low)

    def rollout(rho):
        adj.set_density(rho)
        s = oi.solver
        s.x = m.me.rest_pos.astype(np.float64).copy()
        s.v[:] = 0.0
        xs = [s.x.copy()]
        for t in range(1, T + 1):
            m.set_kinematic_target(h, press_targets(x0, top, t, T))
            oi.simulate(m, DT)
            xs.append(s.x.copy())
        return xs
    xs = rollout(rho0)
    if mode == "disc":
        assert oi.solver.clr_min.min() < oi.solver.dhat          # in contact at the end
    g = adj.transient_adjoint(xs, {T: c}, oi.solver.drag, use_disc=(mode == "disc"))
    assert adj.info == 0
    J = lambda rho: float((c * rollout(rho)[-1]).sum())
    hfd = 0.002
    d = g / np.abs(g).max()                      # along the gradient: relative error
    fd = (J(rho0 + hfd * d) - J(rho0 - hfd * d)) / (2 * hfd)
    assert abs(g @ d - fd) < 1e-4 * abs(fd), (mode, g @ d, fd)
    d = rng.standard_normal(adj.n_tri)           # any direction: error over |g| |d|
    fd = (J(rho0 + hfd * d) - J(rho0 - hfd * d)) / (2 * hfd)
    assert abs(g @ d - fd) < 1e-5 * np.linalg.norm(g) * np.linalg.norm(d), (mode, g @ d, fd)
