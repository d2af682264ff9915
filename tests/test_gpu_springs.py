"""GPU parity and physics of springs, pins and active strain (reference
implicit/actuation.py:319-476, sim_model.py:988-1180):

* Whole steps with springs, pins and a moving anchor (the same fixed iteration count), and with
  an active layer whose activation changes every step (converged steps), on both device paths:
  positions within 1e-9 of the body size of the float64 oracle.
* The activation entry point against the oracle's rest-shape rewrite, in both element layouts.
* The reference's own protocols (tests/test_actuation.py:217-237, tests/test_rigid.py:225-248):
  a free body whose every element is activated takes the shape A, stress free; a bar pinned to
  the world swings with a physical pendulum's period and the pin does not let go.
* The adjoint refuses springs and active strain, which it has no tangent for.
"""
import numpy as np
import pytest
import torch

from diff_2d_fem_amd import ImplicitIntegrator, Model
from oracle.pncg import OracleIntegrator
from tests.scenes import bilayer, spring_rig

pytestmark = pytest.mark.gpu

PX = 0.5e-3


@pytest.mark.parametrize("path", ["plain", "scene"])
def test_springs_and_pins_steps(path):
    kw = dict(max_iter=60, eps=0.0, conv_check=0)
    mo, _, rig_o = spring_rig(n_env=2)
    mg, _, rig_g = spring_rig(n_env=2)
    oi = OracleIntegrator(**kw)
    gi = ImplicitIntegrator(path=path, precision="double", **kw)
    for k in range(1, 11):
        rig_o(mo, k)
        rig_g(mg, k)
        oi.simulate(mo, 1e-3)
        gi.simulate(mg, 1e-3)
    assert mg._solver.springs["n"] == 6
    err = np.abs(mg.get_x() - oi.solver.x).max() / (20 * PX)
    print(f"springs {path}: max |dx| / size = {err:.3e}")
    assert err < 1e-9
    assert (gi.stats(mg)["iters"] == oi.stats()["iters"]).all()
    xo = oi.solver.x
    for h, sp in enumerate(mo.actuation.springs):  # tension and length at the oracle's state
        xb = xo[sp.b] if sp.b >= 0 else sp.anchor
        ln = np.linalg.norm(xo[sp.a] - xb)
        assert np.allclose(mg.spring_force(h), (sp.stiffness * (ln - sp.length), ln), rtol=1e-7, atol=1e-12)


@pytest.mark.parametrize("fiber", [None, np.deg2rad(30.0)])
@pytest.mark.parametrize("path", ["plain", "scene"])
def test_active_layer_steps(path, fiber):
    """The bimetal strip, its activation commanded anew every step (after the scene layout
    exists, so both element layouts are rewritten on the device). Each step's rewritten rest
    shape matches the oracle's to a few ulps (the device fuses the multiply-adds of Bm0 A^-1).
    The steps are solved to convergence: a fixed 60-iteration iterate of this bending problem
    moves by 1e-7 of the strip when Bm changes by one ulp, the converged state by 1e-13."""
    kw = dict(max_iter=3000, eps=1e-14, tol_x=1e-12, conv_check=8)
    mo, ro = bilayer(n_env=2, fiber=fiber)
    mg, rg = bilayer(n_env=2, fiber=fiber)
    oi = OracleIntegrator(**kw)
    gi = ImplicitIntegrator(path=path, precision="double", **kw)
    ulp = 4 * np.finfo(np.float64).eps
    for k in range(1, 6):
        for m, regions in ((mo, ro), (mg, rg)):
            for i, r in enumerate(regions):
                m.set_activation(r, 1.0 - 0.006 * k * (i + 1), transverse=1.0 + 0.003 * k)
        s, so = gi.attach(mg, 2e-3), oi.attach(mo, 2e-3)
        assert np.abs(s.Bm.cpu().numpy().reshape(-1, 2, 2) - so.Bm).max() <= ulp * np.abs(so.Bm).max()
        assert np.array_equal(s.area.cpu().numpy(), so.area)
        oi.simulate(mo, 2e-3)
        gi.simulate(mg, 2e-3)
    err = np.abs(mg.get_x() - so.x).max() / (24 * PX)
    print(f"active layer {path} fiber={fiber}: max |dx| / size = {err:.3e}")
    assert err < 1e-9
    if path == "scene":                 # the scene layout holds the same rest shape
        perm = s.sc_perm
        assert np.abs(s.sc_Bm.cpu().numpy() - so.Bm.reshape(-1, 4)[perm]).max() <= ulp * np.abs(so.Bm).max()
        Amu = so.area[perm] * s.mu.cpu().numpy()[perm]
        assert np.array_equal(s.sc_Amu.cpu().numpy(), Amu)
        assert np.array_equal(s.sc_Amu32.cpu().numpy(), Amu.astype(np.float32))


def _free_body():
    """The reference's free-growth body (tests/test_actuation.py:34-40, 217-221)."""
    m = Model()
    m.add_disc(center=(0.0, -5.0), radius=0.01)
    m.damping = 2000.0
    mask = np.ones((16, 10), np.uint8)
    mask[5:9, 3:7] = 0
    h = m.add_soft_body(mask, physical_size=(0.010, 0.016), center=(0.0, 0.0), E=1.0e6, nu=0.40, density=1100.0)
    m.finalize()
    return m, h, mask


@pytest.mark.parametrize("path", ["plain", "scene"])
def test_active_material_free_growth(path):
    kw = dict(max_iter=3000, eps=1e-14, tol_x=2e-9, conv_check=8, self_collision=False, friction=False,
              alpha_max=1e9)
    m, h, mask = _free_body()
    m.set_activation(m.add_active_region(h, np.ones_like(mask)), 1.08)
    integ = ImplicitIntegrator(path=path, precision="double", **kw)
    for _ in range(10):
        integ.simulate(m, 2e-2)
    X = np.asarray(m.me.rest_pos, np.float64)
    x = m.get_x()
    assert np.abs((x - x.mean(0)) - 1.08 * (X - X.mean(0))).max() < 1e-4 * 0.016
    m, h, mask = _free_body()
    th = np.deg2rad(30.0)
    m.set_activation(m.add_active_region(h, np.ones_like(mask), fiber=th), 0.85, transverse=1.10)
    integ = ImplicitIntegrator(path=path, precision="double", **kw)
    for _ in range(10):
        integ.simulate(m, 2e-2)
    x = m.get_x()
    F = np.linalg.lstsq(X - X.mean(0), x - x.mean(0), rcond=None)[0].T
    assert np.allclose(np.linalg.svd(F)[1], [1.10, 0.85], atol=2e-4)
    f = np.array([np.cos(th), np.sin(th)])
    assert abs(np.linalg.norm(F @ f) - 0.85) < 2e-4


@pytest.mark.parametrize("path", ["plain", "scene"])
def test_pin_makes_a_pendulum(path):
    G0 = 9.81
    m = Model()
    m.damping = 0.0
    h = m.add_affine_body(np.ones((4, 24), np.uint8), physical_size=(0.012, 0.002), center=(0.006, 0.0),
                          density=1100.0)
    m.finalize()
    pin = m.add_pin(h, (0.0, 0.0))
    X = m.rest_x(h)
    mass = np.asarray(m.me.mass, np.float64)[m.body_nodes(h)]
    pivot = X[np.linalg.norm(X, axis=1).argmin()]
    d = np.linalg.norm((mass[:, None] * X).sum(0) / mass.sum() - pivot)
    Ip = (mass * ((X - pivot) ** 2).sum(1)).sum()
    th0 = np.deg2rad(10.0)
    period = 2 * np.pi * np.sqrt(Ip / (mass.sum() * G0 * d)) * (1 + th0 ** 2 / 16)
    m.gravity = (G0 * np.cos(th0), -G0 * np.sin(th0))
    dt = 5e-4
    integ = ImplicitIntegrator(path=path, precision="double", self_collision=False, friction=False)
    ang, gap = [], 0.0
    for _ in range(int(1.1 * period / dt)):
        integ.simulate(m, dt)
        ang.append(m.affine_state(0)["angle"] + th0)
        gap = max(gap, abs(m.spring_force(pin)[1]))
    ang = np.array(ang)
    t = dt * np.arange(1, len(ang) + 1)
    k = np.where(np.sign(ang[1:]) != np.sign(ang[:-1]))[0]
    assert len(k) >= 2, "the bar did not swing through its equilibrium twice"
    cross = [t[i] - ang[i] * dt / (ang[i + 1] - ang[i]) for i in k[:2]]
    got = 2.0 * (cross[1] - cross[0])
    print(f"pendulum {path}: period {got * 1e3:.2f} ms, physical {period * 1e3:.2f} ms, pin gap {gap:.1e} m")
    assert abs(got / period - 1.0) < 0.01
    assert gap < 1e-6


def test_activation_entry_point_matches_the_oracle():
    m, regions = bilayer(n_env=3, fiber=np.deg2rad(20.0))
    for i, r in enumerate(regions):
        m.set_activation(r, 0.9 + 0.05 * i, transverse=1.1)
    integ = ImplicitIntegrator(path="scene", precision="double", max_iter=5, eps=0.0, conv_check=0)
    integ.simulate(m, 1e-3)                      # builds the scene layout
    s = m._solver
    for i, r in enumerate(regions):              # a second command, rewritten on the device
        m.set_activation(r, 1.02 - 0.03 * i)
    integ.attach(m, 1e-3)
    torch.cuda.synchronize()
    oi = OracleIntegrator(max_iter=5, eps=0.0, conv_check=0)
    so = oi.attach(m, 1e-3)
    ulp = 4 * np.finfo(np.float64).eps               # the device fuses B0 Ainv's multiply-adds
    assert np.abs(s.Bm.cpu().numpy().reshape(-1, 2, 2) - so.Bm).max() <= ulp * np.abs(so.Bm).max()
    assert np.array_equal(s.area.cpu().numpy(), so.area)
    b32 = so.Bm.reshape(-1, 4)[s.sc_perm].astype(np.float32)
    assert np.abs(s.sc_Bm32.cpu().numpy() - b32).max() <= 2 * np.finfo(np.float32).eps * np.abs(b32).max()


def test_adjoint_refuses_springs_and_active_strain():
    m, _, rig = spring_rig(n_env=1)
    integ = ImplicitIntegrator(path="scene", precision="double", self_collision=False)
    adj = integ.adjoint(m, 1e-3)
    with pytest.raises(NotImplementedError, match="Springs"):
        adj.transient_adjoint([m.get_x(), m.get_x()], seed={1: np.zeros((m.n_nodes, 2))})
    m, regions = bilayer(n_env=1)
    integ = ImplicitIntegrator(path="scene", precision="double", self_collision=False)
    adj = integ.adjoint(m, 1e-3)
    with pytest.raises(NotImplementedError, match="ActiveStrain"):
        adj.transient_adjoint([m.get_x(), m.get_x()], seed={1: np.zeros((m.n_nodes, 2))})


def test_springs_and_active_layer_mixed_close_to_float64():
    """The scene path in mixed precision (float32 element arithmetic) with springs, pins and an
    active layer, against the float64 oracle, within 1e-5 of the body size. The active bilayer
    is a 24:6 cantilever that bends: the eigenvalues of its stiffness run from 1.1e2 (first
    bending mode) to 1.6e7, so any rounding of the element forces is amplified along the soft
    bending mode; the float32 stress is formed from a float64 displacement gradient, which
    keeps that rounding relative to the strain."""
    kw = dict(max_iter=60, eps=0.0, conv_check=0)
    mo, _, rig_o = spring_rig(n_env=2)
    mg, _, rig_g = spring_rig(n_env=2)
    oi = OracleIntegrator(**kw)
    gi = ImplicitIntegrator(path="scene", precision="mixed", **kw)
    for k in range(1, 11):
        rig_o(mo, k)
        rig_g(mg, k)
        oi.simulate(mo, 1e-3)
        gi.simulate(mg, 1e-3)
    err = np.abs(mg.get_x() - oi.solver.x).max() / (20 * PX)
    print(f"springs mixed: max |dx| / size = {err:.3e}")
    assert err < 1e-5
    kw = dict(max_iter=3000, eps=1e-14, tol_x=1e-12, conv_check=8)
    mo, ro = bilayer(n_env=2, fiber=np.deg2rad(30.0))
    mg, rg = bilayer(n_env=2, fiber=np.deg2rad(30.0))
    oi = OracleIntegrator(**kw)
    gi = ImplicitIntegrator(path="scene", precision="mixed", **kw)
    for k in range(1, 6):
        for m, regions in ((mo, ro), (mg, rg)):
            for i, r in enumerate(regions):
                m.set_activation(r, 1.0 - 0.006 * k * (i + 1), transverse=1.0 + 0.003 * k)
        oi.simulate(mo, 2e-3)
        gi.simulate(mg, 2e-3)
    err = np.abs(mg.get_x() - oi.solver.x).max() / (24 * PX)
    print(f"active layer mixed: max |dx| / size = {err:.3e}")
    assert err < 1e-5
