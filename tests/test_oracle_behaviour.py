"""What pins the oracle's iteration, friction frames and kinematic sweep beyond the reference's
fixtures: closed forms, the Dai-Kou method's own guarantees and the reference's behavioural
acceptance tests run on the oracle (the same scenes run on both GPU paths in
test_gpu_behaviour.py).

* Dai-Kou beta and the restart guard. On a convex quadratic with exact line search the method
  is preconditioned conjugate gradients: the oracle's iterates equal textbook Jacobi-PCG's (a
  wrong sign or preconditioner in beta's first term, or a guard that fires when the curvature
  condition holds, breaks the equality). On a nonlinear contact scene every direction
  descends by at least 3/4 of the preconditioned gradient norm, g.d <= -3/4 g.Pg, which the
  beta of Dai and Kou (2013) with tau = s.y / |s|^2 guarantees by Cauchy-Schwarz; dropping or
  flipping its second term voids the bound.
* Collider friction frames, every shape: the friction force is the derivative of its
  potential along the surface, and in the sliding branch the friction reaction is mu times
  the normal reaction (reference tests/test_objects.py:84-110; a missing |grad phi| factor
  fails the ellipse, whose level set is not a distance).
* The kinematic sweep (reference tests/test_ipc.py:256-295), no tunnelling and momentum
  (:216-253) and the Coulomb slide speed (:128-159, one of its four cases, fewer steps).
"""
import warnings

import numpy as np
import pytest

from oracle import terms as T
from oracle.pncg import OracleIntegrator
from tests import behaviour as BH
from tests.test_oracle import COLLIDERS, _probe_points


# ------------------------------------------------------------------------- Dai-Kou beta
class _Quadratic:
    """A test term E = h^2/2 (x - c)^T K (x - c) over the flattened coordinates; it reports
    dscale x diag(K) as its share of the preconditioner's diagonal."""

    def __init__(self, K, c, dscale=1.0):
        self.K, self.c, self.dscale = K, c, dscale

    def grad_diag(self, s):
        s.grad += s.h2 * (self.K @ (s.x - self.c).reshape(-1)).reshape(-1, 2)
        s.diagH += s.h2 * self.dscale * np.diag(self.K).reshape(-1, 2)

    def pHp(self, s):
        p = s.p.reshape(-1)
        s.pHp[0] += s.h2 * p @ self.K @ p

    def energy(self, s, buf):
        d = (s.x - self.c).reshape(-1)
        buf[0] += 0.5 * s.h2 * d @ self.K @ d


def _quadratic_solver(n_iter, seed=0, dscale=1.0):
    """The oracle's solver on one free body whose only potential (besides inertia) is a
    random sparse convex quadratic: elements removed, no contact, no caps that bind."""
    m, _ = _block()
    integ = OracleIntegrator(max_iter=n_iter, eps=0.0, conv_check=0, self_collision=False, friction=False,
                             rigid_modes=False, alpha_max=1.0e6, cg_restart=0.5)
    s = integ.attach(m, 1e-3)
    s.tri, s.tri2env = s.tri[:0], s.tri2env[:0]
    s.Bm, s.area, s.mu, s.lam = s.Bm[:0], s.area[:0], s.mu[:0], s.lam[:0]
    rng = np.random.default_rng(seed)
    n = 2 * s.N
    L = np.zeros((n, n))
    for _ in range(3 * n):                       # a random weighted graph Laplacian, then SPD
        i, j = rng.choice(n, 2, replace=False)
        w = rng.uniform(0.5, 2.0)
        L[i, i] += w
        L[j, j] += w
        L[i, j] -= w
        L[j, i] -= w
    K = (L + 0.1 * np.eye(n)) * s.mass.mean() / s.h2          # h^2 K of the order of M
    c = s.x + 1e-5 * rng.standard_normal(s.x.shape)
    s.terms.append(_Quadratic(K, c, dscale))
    s.v = 1e-3 * rng.standard_normal(s.x.shape)
    assert s.cg_guard == 0.5
    return s, K, c


def _block():
    from diff_2d_fem_amd import Model
    m = Model()
    h = m.add_soft_body(np.ones((3, 4), np.uint8), physical_size=(0.004, 0.003), center=(0.0, 0.0),
                        E=1.0e6, nu=0.4, density=1100.0)
    m.finalize()
    return m, h


@pytest.mark.parametrize("dscale", [1.0, 0.2])
def test_dai_kou_on_a_quadratic_is_preconditioned_cg(dscale):
    """dscale 1: Jacobi; 0.2: a larger diagonal preconditioner, whose steps are shorter than
    1/2, where a guard that left out the step length would fire."""
    n_iter = 12
    s, K, c = _quadratic_solver(n_iter, dscale=dscale)
    xs, alphas = [], []
    upd = s.update
    s.update = lambda: (upd(), xs.append(s.x.copy()), alphas.append(float(s.alpha[0])))
    s.step_targets()
    assert len(xs) == n_iter
    if dscale < 1.0:
        assert min(alphas) < 0.5
    # textbook diagonally preconditioned CG on (M + h^2 K) x = M xhat + h^2 K c from the step start
    M = np.repeat(s.mass, 2)
    A = np.diag(M) + s.h2 * K
    b = M * s.xhat.reshape(-1) + s.h2 * K @ c.reshape(-1)
    P = 1.0 / (M + s.h2 * dscale * np.diag(K))
    x = s.xn.reshape(-1).copy()
    r = b - A @ x
    z = P * r
    d = z.copy()
    ref = []
    for _ in range(n_iter):
        Ad = A @ d
        a = (r @ z) / (d @ Ad)
        x = x + a * d
        r_new = r - a * Ad
        z_new = P * r_new
        d = z_new + (r_new @ z_new) / (r @ z) * d
        r, z = r_new, z_new
        ref.append(x.copy())
    xstar = np.linalg.solve(A, b)
    scale = np.abs(s.xn.reshape(-1) - xstar).max()
    errs = [np.abs(a.reshape(-1) - b_).max() / scale for a, b_ in zip(xs, ref)]
    print("oracle vs PCG iterates:", " ".join(f"{e:.1e}" for e in errs))
    assert max(errs) < 1e-9
    # and the iterates do converge (not a trivially diagonal problem)
    prog = [np.abs(r_.reshape(-1) - xstar).max() / scale for r_ in ref]
    assert prog[0] > 1e-2 and prog[-1] < 1e-4


def test_dai_kou_directions_descend_by_three_quarters():
    """Every direction of a contact-heavy run (the reference's pusher against a wall) satisfies
    g.d <= -3/4 g.Pg, and beta's second term is active (g.p_prev far from zero) in many."""
    st = BH.OracleStepper()
    seen = []

    def run():
        s = st.integ.solver
        d0 = s.direction

        def direction(it):
            d0(it)
            seen.append((s.gTp.copy(), s.gPg.copy(), s.beta.copy(), s.g_p.copy(), s.y_p.copy()))
        s.direction = direction
    orig = st.step
    started = [False]

    def step(m, dt):
        if st.integ.solver is None:
            st.integ.attach(m, dt)
        if not started[0]:
            run()
            started[0] = True
        orig(m, dt)
    st.step = step
    BH.push(st, 1.5, wall=True, steps=6)
    gTp, gPg, beta, g_p, y_p = (np.array([v[k] for v in seen]) for k in range(5))
    worst = (gTp / (-gPg)).min()
    print(f"{len(seen)} directions; smallest -g.d / g.Pg = {worst:.4f}; "
          f"beta != 0 in {int((beta != 0).sum())}")
    assert (gTp <= -0.75 * gPg * (1.0 - 1e-12)).all()
    lagged = (beta != 0) & (np.abs(g_p) > 1e-3 * np.abs(gTp))
    assert lagged.sum() > 20, "the line search is exact everywhere: the bound was not exercised"


# ------------------------------------------------------------------------- friction frames
@pytest.mark.parametrize("ci", range(len(COLLIDERS)))
def test_friction_force_is_the_derivative_of_its_potential_along_the_surface(ci):
    """A slip along the surface since the step began; the friction gradient equals the central
    difference of the friction potential along the surface, for every collider kind."""
    col = COLLIDERS[ci]
    dhat, kappa, mu, epsv, inv_dt = 1e-3, 1e3, 0.6, 1e-3, 1e3
    X = _probe_points(col, dhat=dhat)
    rng = np.random.default_rng(5)
    d, gr, _ = T.sdf_eval(col["kind"], col["cx"], col["cy"], col["rx"], col["ry"], col["rref"], X)
    nh = gr / np.linalg.norm(gr, axis=1)[:, None]
    t = np.stack([-nh[:, 1], nh[:, 0]], 1)
    # half the points slide faster than epsv (the Coulomb branch), half slower (the smoothed one)
    speed = np.where(np.arange(len(X)) % 2 == 0, 3.0, 0.3) * epsv
    xn = X - (speed / inv_dt * rng.choice([-1.0, 1.0], len(X)))[:, None] * t
    bnd = np.arange(len(X))
    b2e = np.zeros(len(X), np.int64)
    free = np.ones_like(X)

    def energy(x):
        E = np.zeros(1)
        T.friction_energy(col, x, xn, bnd, b2e, dhat, kappa, mu, epsv, inv_dt, 1.0, free, E)
        return E[0]
    g, dg = np.zeros_like(X), np.zeros_like(X)
    T.friction_grad_diag(col, X, xn, bnd, dhat, kappa, mu, epsv, inv_dt, 1.0, free, g, dg)
    dirn = t * rng.standard_normal((len(X), 1))
    h = 1e-11
    fd = (energy(X + h * dirn) - energy(X - h * dirn)) / (2 * h)
    rel = abs(fd - (g * dirn).sum()) / np.abs(g * dirn).sum()
    print(f"kind {col['kind']}: friction gradient vs d(potential) along the surface: {rel:.1e}")
    assert rel < 1e-6
    # the force opposes the slip and has magnitude mu f_n in the Coulomb branch
    fast = speed > epsv
    slip = X - xn
    assert ((g * slip).sum(1) > 0.0).all()
    bp = np.abs(T.barrier_bp(d, dhat)) * kappa * np.linalg.norm(gr, axis=1)
    assert np.allclose(np.linalg.norm(g, axis=1)[fast], mu * bp[fast], rtol=1e-12)


def test_sliding_friction_is_mu_times_normal_for_every_shape():
    with warnings.catch_warnings(), np.errstate(all="ignore"):
        warnings.simplefilter("ignore")
        r = BH.friction_ratio(lambda: BH.OracleStepper())
    print(r)
    for shape, ratio in r.items():
        assert abs(ratio - BH.MU_OBJ) < 0.03 * BH.MU_OBJ, f"{shape}: {ratio:.4f}"


# ------------------------------------------------------------------------- IPC behaviour
def test_a_prescribed_body_is_swept():
    lag_px, squash = BH.check_swept(lambda: BH.OracleStepper())
    print(f"against the wall: {lag_px:.1f} px short, smallest area {squash:.3f}")


def test_nothing_passes_through():
    for other, diagonal in (("held", True), ("soft", False)):
        gap, inside, _ = BH.throw(BH.OracleStepper(), other, 5.0, diagonal)
        assert gap > 0.0 and not inside, f"thrown at a {other} block at 5 px/step: gap {gap:.2e} dhat"
    _, _, dp = BH.throw(BH.OracleStepper(), "soft", 0.5)
    print(f"momentum change through the impact: {dp:.1e}")
    assert dp < 2e-3


def test_slides_at_the_coulomb_speed():
    v, expect, gap = BH.slide(BH.OracleStepper(), "rigid", 0.3, 30.0, settle=20, run=50, average=20)
    print(f"rigid block, mu 0.3, 30 deg: {v:.5e} m/s, closed form {expect:.5e}; resting gap {gap:.2f} dhat")
    assert 0.5 < gap < 1.0
    assert abs(v / expect - 1.0) < 0.01
