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


# ---- IPC contact between bodies and affine bodies in the adjoint (beyond the reference, whose
# adjoint refuses both): the pressed block on a soft slab through the IPC barrier, frictionless,
# with and without a rigid region in the block (tests.scenes.ipc_press).
IPC_T, IPC_DT = 5, 2.0e-4
IPC_KW = dict(max_iter=6000, eps=1e-14, tol_x=1e-13, conv_check=8, friction=False)


def _ipc_setup(region, path="scene"):
    from tests.scenes import ipc_press, ipc_press_targets
    m, h, x0, top = ipc_press(region=region)
    integ = ImplicitIntegrator(path=path, precision="double", **IPC_KW)
    adj = integ.adjoint(m, IPC_DT)
    rng = np.random.default_rng(0)
    rho0 = 0.45 + 0.5 * rng.random(adj.n_tri)
    low = m.body_nodes(h)[np.where(x0[:, 1] < np.median(x0[:, 1]))[0]]
    c = np.zeros((m.n_nodes, 2))
    c[low, 0] = 1.0 / len(low)
    c[low, 1] = 0.5 / len(low)

    def command(model, t):
        model.set_kinematic_target(h, ipc_press_targets(x0, top, t, IPC_T))

    def rollout(rho):
        adj.set_density(rho)
        m.reset()
        xs = [m.get_x()]
        for t in range(1, IPC_T + 1):
            command(m, t)
            integ.simulate(m, IPC_DT)
            xs.append(m.get_x())
        return xs
    return m, integ, adj, rho0, c, rollout, rng, command


def _ipc_oracle(m, rho0):
    oi = OracleIntegrator(**IPC_KW)
    s = oi.attach(m, IPC_DT)
    oa = OracleAdjoint(s, m.me)
    oa.set_density(rho0)
    return s, oa


@pytest.mark.parametrize("region", [False, True])
def test_ipc_and_affine_operator_equals_oracle(region):
    """H v with the IPC tangent, and the IPC part of the preconditioner's diagonal, at the end
    state of a rollout, against the oracle's (node space, rigid-region elements with their
    moduli)."""
    m, integ, adj, rho0, c, rollout, rng, _ = _ipc_setup(region)
    xs = rollout(rho0)
    s, oa = _ipc_oracle(m, rho0)
    x = xs[-1]
    assert s.ipc.terms(x)[0].size > 0, "no contact"
    v = rng.standard_normal(x.shape)
    for use_ipc in (True, False):
        hv = adj.apply_H(x, v, use_ipc=use_ipc)
        ho = oa.apply_H(x, v, use_ipc=use_ipc)
        assert np.abs(hv - ho).max() <= 1e-12 * np.abs(ho).max(), use_ipc
    import torch
    xw = adj._t(x)
    adj._set_pairs(xw)
    assert adj._n_pairs == len(m._solver.ipc_pairs())
    d_on = adj._diag_precond(xw, use_ipc=True).cpu().numpy()
    d_off = adj._diag_precond(xw, use_ipc=False).cpu().numpy()
    go, do = np.zeros_like(x), np.zeros_like(x)
    s.ipc.barrier_grad_diag(s, x, adj.h2, go, do)
    assert np.abs((d_on - d_off) - do).max() <= 1e-9 * np.abs(do).max()


@pytest.mark.parametrize("region", [False, True])
def test_ipc_and_affine_adjoint_equals_oracle(region):
    m, integ, adj, rho0, c, rollout, rng, _ = _ipc_setup(region)
    xs = rollout(rho0)
    g = adj.transient_adjoint(xs, seed={IPC_T: c}, rtol=1e-13, maxiter=20000)
    assert getattr(adj, "minres_fail", 0) == 0
    s, oa = _ipc_oracle(m, rho0)
    go = oa.transient_adjoint(xs, {IPC_T: c}, s.drag, rtol=1e-13)
    err = np.abs(g - go).max() / np.abs(go).max()
    print(f"region={region}: |g_gpu - g_oracle| / |g| = {err:.2e}")
    assert err < 1e-7


@pytest.mark.parametrize("region", [False, True])
def test_ipc_and_affine_adjoint_matches_finite_differences(region):
    m, integ, adj, rho0, c, rollout, rng, _ = _ipc_setup(region)
    xs = rollout(rho0)
    g = adj.transient_adjoint(xs, seed={IPC_T: c}, rtol=1e-13, maxiter=20000)
    J = lambda rho: float((c * rollout(rho)[-1]).sum())
    hfd = 0.002
    d = g / np.abs(g).max()
    fd = (J(rho0 + hfd * d) - J(rho0 - hfd * d)) / (2 * hfd)
    rel = abs(g @ d - fd) / abs(fd)
    d2 = rng.standard_normal(adj.n_tri)
    fd2 = (J(rho0 + hfd * d2) - J(rho0 - hfd * d2)) / (2 * hfd)
    nerr = abs(g @ d2 - fd2) / (np.linalg.norm(g) * np.linalg.norm(d2))
    print(f"region={region}: along the gradient {rel:.2e}, random direction {nerr:.2e}")
    assert rel < 1e-4
    assert nerr < 1e-5


def test_ipc_and_affine_rollout_checkpointing():
    """Rollout (torch.autograd.Function) with checkpoints every 2 steps gives the gradient of
    transient_adjoint over the whole trajectory, with a rigid region and IPC contact: the
    recomputed segments restore the affine groups' state."""
    import torch
    from diff_2d_fem_amd.autograd import Rollout
    m, integ, adj, rho0, c, rollout, rng, command = _ipc_setup(True)
    xs = rollout(rho0)
    g = adj.transient_adjoint(xs, seed={IPC_T: c}, rtol=1e-13, maxiter=20000)
    m.reset()
    roll = Rollout(m, integ, IPC_DT, IPC_T, every=2, command=command, adjoint_kw=dict(rtol=1e-13, maxiter=20000))
    roll.adj.Emin = adj.Emin                     # density rho = Emin + scale (1 - Emin)
    scale = torch.tensor(rho0, requires_grad=True)
    x = roll(scale)
    assert np.abs(x.detach().cpu().numpy() - xs[-1]).max() <= 1e-9 * np.abs(xs[-1]).max()
    (x * torch.as_tensor(c, device=x.device)).sum().backward()
    assert roll.recomputed_steps == IPC_T
    err = np.abs(scale.grad.numpy() - g).max() / np.abs(g).max()
    print(f"checkpointed rollout vs whole trajectory: {err:.2e}")
    assert err < 1e-9


def test_minres_small_and_large_scenes():
    """The device MINRES keeps a scene's vectors in registers up to 4,096 degrees of freedom
    and passes them through memory above: one adjoint solve over a batch with a 25 x 97-node
    strip (4,850 dofs) beside a 7 x 73-node one, against scipy's MINRES on the oracle's
    operator, and the residual of H lam = rhs on the free degrees of freedom."""
    from diff_2d_fem_amd import Model
    m = Model()
    m.add_ground(0.0, mu=0.8)
    for env, shape, size in ((0, (24, 96), (0.054, 0.0135)), (1, (6, 72), (0.054, 0.0045))):
        m.add_soft_body(np.ones(shape, np.uint8), physical_size=size, center=(0.0, 0.009), E=0.5e6, nu=0.40,
                        density=1100.0, env_id=env)
    m.finalize()
    m.gravity = (0.0, -9.81)
    kw = dict(max_iter=200, eps=0.0, conv_check=0, friction=False)
    integ = ImplicitIntegrator(path="scene", precision="double", **kw)
    for _ in range(3):
        integ.simulate(m, 2e-3)
    x = m.get_x()
    adj = integ.adjoint(m, 2e-3)
    rhs = np.random.default_rng(4).standard_normal(x.shape)
    lam = adj.adjoint_solve(x, rhs, rtol=1e-12, maxiter=20000)
    assert getattr(adj, "minres_fail", 0) == 0
    oi = OracleIntegrator(**kw)
    oi.attach(m, 2e-3)
    oa = OracleAdjoint(oi.solver, m.me)
    lo = oa.adjoint_solve(x, rhs, rtol=1e-12, maxiter=20000)
    free = adj.free_np.astype(bool)
    assert np.abs(lam - lo).max() / np.abs(lo).max() < 1e-7
    r = (oa.apply_H(x, lam) - rhs)[free]
    assert np.abs(r).max() / np.abs(rhs[free]).max() < 1e-8


def test_elastic_and_pressure_product_in_one_tiled_pass():
    """d2f_elastic_pressure_hvp_f64 (elasticity and the chambers' pressure in one pass over the
    element tiles, the chamber edges dealt to the chunks) against the elastic product plus
    d2f_pressure_hvp_f64, on pneumatic fingers in a deformed state, and the adjoint's operator
    (which takes the fused pass) against the oracle's H v."""
    import ctypes as C
    import torch
    from diff_2d_fem_amd import _lib
    from diff_2d_fem_amd import workloads as W
    w = W.pneumatic_fingers(6)
    m = w["model"]
    integ = ImplicitIntegrator(path="scene", precision="double", self_collision=False)
    for _ in range(2):
        integ.simulate(m, w["dt"])
    x = m.get_x()
    adj = integ.adjoint(m, w["dt"])
    s = adj.s
    pt = adj._pressure_tiles()
    assert pt is not None
    rng = np.random.default_rng(9)
    v = rng.standard_normal(x.shape) * 1e-4
    d = torch.device("cuda")
    xt, vt = torch.tensor(x, device=d), torch.tensor(v, device=d)
    P = lambda t: C.c_void_p(t.data_ptr())
    L = _lib.lib()
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    y0 = torch.zeros_like(xt)
    _lib.check(L.d2f_elastic_hvp_f64(P(xt), P(vt), P(s.tri), P(s.Bm), P(s.area), P(adj._mu_el), P(adj._lam_el),
                                     adj.n_tri, adj.h2, C.byref(adj._tiles), P(y0), st), "hvp")
    _lib.check(L.d2f_pressure_hvp_f64(P(vt), P(s.pe_a), P(s.pe_b), P(s.pe_ch), P(s.ch_p), s.pressure["n_pe"],
                                      adj.h2, P(y0), st), "pressure hvp")
    y1 = torch.zeros_like(xt)
    _lib.check(L.d2f_elastic_pressure_hvp_f64(P(xt), P(vt), P(s.Bm), P(s.area), P(adj._mu_el), P(adj._lam_el),
                                              adj.n_tri, adj.h2, C.byref(adj._tiles), P(pt["pe_off"]),
                                              P(pt["pe_loc"]), P(pt["pe_ch"]), P(s.ch_p), P(y1), st), "fused")
    a, b = y0.cpu().numpy(), y1.cpu().numpy()
    assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()
    oi = OracleIntegrator(self_collision=False)
    oi.attach(m, w["dt"])
    oa = OracleAdjoint(oi.solver, m.me)
    Hv = adj.apply_H(x, v)
    Ho = oa.apply_H(x, v)
    assert np.abs(Hv - Ho).max() <= 1e-10 * np.abs(Ho).max()
