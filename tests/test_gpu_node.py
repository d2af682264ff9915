"""GPU parity of the "node" contact between bodies (model.body_contact = "node"; reference
implicit/kernels.py:755-869, implicit/body_contact.py, implicit/adjoint.py:115-143) against the
float64 oracle (oracle/node.py).

* The per-kernel entry points d2f_node_grad_diag_f64 and d2f_node_hvp_f64 agree with the oracle
  to 1e-12 relative (float64, summation order only).
* Whole steps on both device paths, with the model's or per-body friction and without, and the
  scene path's variants (exclusion bits read from global memory, rigid block, line search): positions
  within 1e-9 of the slab's length after the same fixed iteration count; mixed precision within
  1e-5. The friction coefficients and tilt are ones for which the oracle's own result moves by
  less than 1e-11 of the slab when its start is perturbed by one part in 1e15: where pairs stick
  near the smoothing kink of the friction law (|u_t| = epsv dt), a fixed-count iteration
  amplifies rounding to 1e-4 of the slab, on the oracle as much as on the GPU.
* contact_force between bodies equals the oracle's penalty + pair friction with h^2 = 1.
* The adjoint with use_self equals the oracle's on the same trajectory (1e-7) and matches
  central differences of GPU rollouts (1e-4 along the gradient, 2e-5 of |g| |d| along a random
  direction); without the contact tangent it does not.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from diff_2d_fem_amd import ImplicitIntegrator, _lib
from oracle.adjoint import OracleAdjoint
from oracle.node import BodyFriction
from oracle.pncg import OracleIntegrator
from tests.scenes import node_press, node_press_targets, node_stack
from tests.test_oracle_node import _pressed

pytestmark = pytest.mark.gpu

SIZE = 1.0e-2          # the slab's length


def _p(t):
    return C.c_void_p(t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_node_entry_points():
    m, s = _pressed(seed=5)
    me = m.me
    d = torch.device("cuda")
    L = _lib.lib()
    x = torch.tensor(s.x, device=d)
    bnd = torch.tensor(np.asarray(me.boundary_nodes), dtype=torch.int32, device=d)
    bst = torch.tensor(np.asarray(me.boundary_starts), dtype=torch.int32, device=d)
    sol = ImplicitIntegrator(path="plain", precision="double").attach(m, 1e-3)
    excl, eoff = sol.excl_bits, sol.excl_bits_off        # bit-packed from the mesh's exclusion lists
    nd = s.node
    rng = np.random.default_rng(2)
    for clamp in (0, 1):
        g0 = rng.normal(size=s.x.shape)
        g, dg = torch.tensor(g0, device=d), torch.tensor(g0, device=d)
        _lib.check(L.d2f_node_grad_diag_f64(_p(x), _p(bnd), _p(bst), me.n_envs, _p(excl), _p(eoff), nd.r, nd.k,
                                            4e-6, clamp, _p(g), _p(dg), _stream()), "node grad/diag")
        go, do = g0.copy(), g0.copy()
        nd.grad_diag(s.x, 4e-6, clamp, go, do)
        assert np.abs(go - g0).max() > 0.0
        assert _rel(g.cpu().numpy(), go) < 1e-12 and _rel(dg.cpu().numpy(), do) < 1e-12
    v = rng.normal(scale=1e-4, size=s.x.shape)
    y = torch.zeros_like(x)
    _lib.check(L.d2f_node_hvp_f64(_p(x), _p(torch.tensor(v, device=d)), _p(bnd), _p(bst), me.n_envs, _p(excl),
                                  _p(eoff), nd.r, nd.k, 4e-6, _p(y), _stream()), "node hvp")
    yo = np.zeros_like(s.x)
    nd.hvp(s.x, v, 4e-6, yo)
    assert np.abs(yo).max() > 0.0 and _rel(y.cpu().numpy(), yo) < 1e-12


def _run_pair(make, n_steps, dt, path, kw, precision="double", override=None):
    mo, mg = make(), make()
    oi = OracleIntegrator(**kw)
    gi = ImplicitIntegrator(path=path, precision=precision, **kw)
    for k in range(1, n_steps + 1):
        oi.simulate(mo, dt)
        if k == 1 and override:
            gi.attach(mg, dt)
            s = mg._solver
            s.lds_overrides = override
            s.build_layout()
        gi.simulate(mg, dt)
    return oi, gi, mo, mg


FRICTION = {"frictionless": dict(mu=0.0, tilt=3.0), "model": dict(mu=0.1, tilt=3.0),
            "per_body": dict(mu=0.05, block_mu=0.15, tilt=3.0)}


@pytest.mark.parametrize("fric", list(FRICTION))
@pytest.mark.parametrize("path", ["plain", "scene"])
def test_node_steps_fixed_iterations(path, fric):
    """Blocks falling onto held slabs and sliding under tilted gravity."""
    kw = dict(max_iter=60, eps=0.0, conv_check=0)
    oi, gi, mo, mg = _run_pair(lambda: node_stack(n_env=3, **FRICTION[fric]), 12, 1e-3, path, kw)
    s = oi.solver
    assert len(s.node.visits(s.x)[0]) >= 3 * 14            # every block rests on its slab
    assert (len([t for t in s.terms if isinstance(t, BodyFriction)]) == 1) == (fric != "frictionless")
    err = np.abs(mg.get_x() - s.x).max() / SIZE
    print(f"node {path} {fric}: max |dx| / size = {err:.3e}")
    assert err < 1e-9
    assert (gi.stats(mg)["iters"] == oi.stats()["iters"]).all()


@pytest.mark.parametrize("case", ["global_exclusion", "rigid_block", "line_search", "rigid_line_search",
                                  "converged"])
def test_node_scene_variants(case):
    kw = dict(max_iter=60, eps=0.0, conv_check=0)
    make = lambda: node_stack(n_env=2, mu=0.1, tilt=3.0)
    override = None
    if case == "global_exclusion":
        override = dict(excl_bits_lds=0)
    elif case == "rigid_block":
        make = lambda: node_stack(n_env=2, mu=0.05, tilt=3.0, rigid=True)
    elif case == "line_search":
        kw["line_search"] = True
    elif case == "rigid_line_search":       # backtracking restarts the affine block from q0
        make = lambda: node_stack(n_env=2, mu=0.05, tilt=3.0, rigid=True)
        kw.update(line_search=True, n_backtrack=4)
    else:
        kw = {}
    oi, gi, mo, mg = _run_pair(make, 10, 1e-3, "scene", kw, override=override)
    if override:
        assert mg._solver.excl_bits_lds == 0
    err = np.abs(mg.get_x() - oi.solver.x).max() / SIZE
    print(f"node scene {case}: max |dx| / size = {err:.3e}")
    assert err < (1e-7 if case == "converged" else 1e-9)


def test_node_scene_mixed_close_to_float64():
    kw = dict(max_iter=60, eps=0.0, conv_check=0)
    oi, gi, mo, mg = _run_pair(lambda: node_stack(n_env=2, mu=0.1, tilt=3.0), 10, 1e-3, "scene", kw,
                               precision="mixed")
    err = np.abs(mg.get_x() - oi.solver.x).max() / SIZE
    print(f"node scene mixed: max |dx| / size = {err:.3e}")
    assert err < 1e-5


@pytest.mark.parametrize("path", ["plain", "scene"])
def test_node_contact_force(path):
    m = node_stack(n_env=2, mu=0.5, block_mu=0.9)
    integ = ImplicitIntegrator(path=path, precision="double", max_iter=60, eps=0.0, conv_check=0)
    for _ in range(8):
        integ.simulate(m, 1e-3)
    s = m._solver
    x, xn = s.get_x(), s.xn.cpu().numpy()
    oi = OracleIntegrator(max_iter=60, eps=0.0, conv_check=0)
    os_ = oi.attach(m, 1e-3)
    os_.x, os_.xn = x.copy(), xn.copy()
    ones = np.ones_like(x)
    gp, dg = np.zeros_like(x), np.zeros_like(x)
    os_.node.grad_diag(x, 1.0, 1, gp, dg)
    gf = np.zeros_like(x)
    bf = [t for t in os_.terms if isinstance(t, BodyFriction)][0]
    bf.pair_force(os_, gf, np.zeros_like(x), ones)
    f_b = integ.contact_force(m, colliders=False)
    assert np.abs(gp).max() > 0.0 and np.abs(gf).max() > 0.0
    assert np.abs(f_b + gp + gf).max() <= 1e-9 * np.abs(gp + gf).max()
    fn = integ.contact_force(m, colliders=False, friction=False)
    assert np.abs(fn + gp).max() <= 1e-9 * np.abs(gp).max()
    # action and reaction: the blocks push their slabs exactly as hard as they are pushed
    assert np.abs(f_b.sum(0)).max() <= 1e-9 * np.abs(f_b).max()
    blk = integ.body_contact_force(m, 1, colliders=False)
    assert np.allclose(blk, f_b[m.body_nodes(1)].sum(0))


T, DT = 6, 2.0e-4
KW = dict(max_iter=6000, eps=1e-14, tol_x=1e-13, conv_check=8, friction=False)   # steps solved tight


def _press_setup(path="scene"):
    m, h, x0, top = node_press()
    integ = ImplicitIntegrator(path=path, precision="double", **KW)
    adj = integ.adjoint(m, DT)
    rng = np.random.default_rng(0)
    rho0 = 0.45 + 0.5 * rng.random(adj.n_tri)
    low = m.body_nodes(h)[np.where(x0[:, 1] < np.median(x0[:, 1]))[0]]
    c = np.zeros((m.n_nodes, 2))
    c[low, 0] = 1.0 / len(low)
    c[low, 1] = 0.5 / len(low)

    def rollout(rho):
        adj.set_density(rho)
        m.reset()
        xs = [m.get_x()]
        for t in range(1, T + 1):
            m.set_kinematic_target(h, node_press_targets(x0, top, t, T))
            integ.simulate(m, DT)
            xs.append(m.get_x())
        return xs
    return m, integ, adj, rho0, c, rollout, rng


def test_node_adjoint_equals_oracle_and_finite_differences():
    m, integ, adj, rho0, c, rollout, rng = _press_setup()
    xs = rollout(rho0)
    s = m._solver
    assert s.node is not None and not s.node["fric"]
    g = adj.transient_adjoint(xs, seed={T: c}, rtol=1e-13, maxiter=20000, use_self=True)
    assert getattr(adj, "minres_fail", 0) == 0
    oi = OracleIntegrator(**KW)
    oi.attach(m, DT)
    assert len(oi.solver.node.visits(xs[-1])[0]) == 14
    oa = OracleAdjoint(oi.solver, m.me)
    oa.set_density(rho0)
    go = oa.transient_adjoint(xs, {T: c}, oi.solver.drag, rtol=1e-13, use_self=True)
    err = np.abs(g - go).max() / np.abs(go).max()
    print(f"node adjoint: |g_gpu - g_oracle| / |g| = {err:.2e}")
    assert err < 1e-7
    J = lambda rho: float((c * rollout(rho)[-1]).sum())
    hfd = 0.002
    dd = g / np.abs(g).max()
    fd = (J(rho0 + hfd * dd) - J(rho0 - hfd * dd)) / (2 * hfd)
    rel = abs(g @ dd - fd) / abs(fd)
    d2 = rng.standard_normal(adj.n_tri)
    fd2 = (J(rho0 + hfd * d2) - J(rho0 - hfd * d2)) / (2 * hfd)
    nerr = abs(g @ d2 - fd2) / (np.linalg.norm(g) * np.linalg.norm(d2))
    # the same trajectory without the contact tangent gives a different (wrong) gradient
    xs = rollout(rho0)
    g0 = adj.transient_adjoint(xs, seed={T: c}, rtol=1e-13, maxiter=20000, use_self=False)
    rel0 = abs(g0 @ dd - fd) / abs(fd)
    print(f"node adjoint vs FD: along the gradient {rel:.2e} (without contact tangent {rel0:.2e}), "
          f"random direction {nerr:.2e}")
    assert rel < 1e-4 and nerr < 2e-5
    assert rel0 > 100 * rel


def test_node_adjoint_refuses_friction_between_bodies():
    m = node_stack(n_env=1, mu=0.5)
    integ = ImplicitIntegrator(path="scene", precision="double")
    adj = integ.adjoint(m, 1e-3)
    with pytest.raises(NotImplementedError, match="BodyFriction"):
        adj.transient_adjoint([m.get_x(), m.get_x()], seed={1: np.zeros((m.n_nodes, 2))}, use_self=True)
